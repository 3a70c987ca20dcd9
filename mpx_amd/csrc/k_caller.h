// The caller-dynamics SQP iteration (model 2): the line search's trial merits
// from the caller's evaluations of its cost and dynamics at the trial points,
// and the export of the step the caller needs to form those points.
//
// The reference solver evaluates the caller's cost(W, reference, x, u, t) and
// dynamics(x, u, t, parameter) (mpx/utils/objectives.py:15, mpx/utils/models.py:6,
// bound at mpx/utils/mpc_wrapper.py:66-74) at every trial point of its merit
// line search.  Here the caller does that between mpxkkt_sqp_direction and
// mpxkkt_sqp_accept; the merit terms that are the library's own (V'c,
// |c|^2, the equality rows' augmented-Lagrangian terms, the box barrier) are
// added on the device, stage by stage, in the layout k_ls_decide reduces.
#pragma once
#include "common.h"
#include "k_trials.h"

namespace mpxk {

// Trials j < L of stage t = blockIdx.x * waves + wave of instance blockIdx.y:
//   t < N: trial_g[b][j][t] + V_{t+1}(a)'c(a) + AL(h(a)) - mu_b sum log s(a),
//          c(a) = trial_f[b][j][t] - (X_{t+1} + a dX_{t+1});
//   t = N: trial_g[b][j][N] + V_0(a)'c_init(a), c_init(a) = x0 - (X_0 + a dX_0);
// and |c(a)|^2, into tpart; the equality rows at the iterate and their slopes
// into hq (k_ls_update's multiplier update reads them).
template <typename T>
__global__ void __launch_bounds__(256) k_trials_caller(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, int L,
                                                       const T* __restrict__ trial_g,
                                                       const T* __restrict__ trial_f) {
  const int wpb = blockDim.x >> 6, wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = blockIdx.y, t = blockIdx.x * wpb + wv;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  if (t > N) return;
  const bool term = t == N;
  const T* X = it.X + (size_t)b * N1 * n;
  const T* V = it.V + (size_t)b * N1 * n;
  const T* dX = ws.dx + (size_t)b * N1 * n;
  const T* dV = ws.dv + (size_t)b * N1 * n;
  const int row = term ? 0 : t + 1;   // the constraint row this stage owns
  T* hq = (!term && p > 0) ? ws.hq + ((size_t)b * N + t) * 2 * p : nullptr;
  if (hq) {
    const T* D = P.D.at(b, t);
    const T* E = P.E.at(b, t);
    const T* x0 = X + (size_t)t * n;
    const T* x1 = dX + (size_t)t * n;
    const T* u0 = it.U + ((size_t)b * N + t) * m;
    const T* u1 = ws.du + ((size_t)b * N + t) * m;
    for (int k = lane; k < p; k += 64) {
      T h0 = P.e.at(b, t)[k], h1 = 0;
      for (int j = 0; j < n; ++j) { h0 += D[k * n + j] * x0[j]; h1 += D[k * n + j] * x1[j]; }
      for (int j = 0; j < m; ++j) { h0 += E[k * m + j] * u0[j]; h1 += E[k * m + j] * u1[j]; }
      hq[k] = h0;
      hq[p + k] = h1;
    }
  }
  const T* mu = hq ? it.Yeq + ((size_t)b * N + t) * p : nullptr;
  T a = ws.lsp[(size_t)b * NLS + LS_AMAX];
  T* tp = ws.tpart + ((size_t)b * N1 + t) * 2 * MAX_TRIALS;
  for (int j = 0; j < L; ++j, a *= prm.beta) {
    T s1 = 0, s2 = 0;
    const T* fj = term ? P.x0.at(b, 0) : trial_f + (((size_t)b * L + j) * N + t) * n;
    for (int i = lane; i < n; i += 64) {
      const size_t o = (size_t)row * n + i;
      const T c = fj[i] - (X[o] + a * dX[o]);
      s1 += (V[o] + a * dV[o]) * c;
      s2 += c * c;
    }
    if (!term) {
      if (P.box) {
        const T* u0 = it.U + ((size_t)b * N + t) * m;
        const T* u1 = ws.du + ((size_t)b * N + t) * m;
        for (int i = lane; i < m; i += 64) {
          const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
          const T u = u0[i] + a * u1[i];
          s1 -= prm.mu_b * (dev_log(hi - u) + dev_log(u - lo));
        }
      }
      for (int k = lane; k < p; k += 64) {
        const T hh = hq[k] + a * hq[p + k];
        s1 += mu[k] * hh + T(0.5) * prm.rho_c * hh * hh;
      }
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
      tp[2 * j] = s1 + trial_g[((size_t)b * L + j) * N1 + t];
      tp[2 * j + 1] = s2;
    }
  }
}

// dX, dU and a_max of the step to the caller (one block per instance).
template <typename T>
__global__ void k_export_dir(Prob<T> P, Ws<T> ws, T* dX, T* dU, T* amax) {
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const size_t ox = (size_t)b * N1 * n, ou = (size_t)b * N * m;
  for (int e = threadIdx.x; e < N1 * n; e += blockDim.x) dX[ox + e] = ws.dx[ox + e];
  for (int e = threadIdx.x; e < N * m; e += blockDim.x) dU[ou + e] = ws.du[ou + e];
  if (threadIdx.x == 0) amax[b] = ws.lsp[(size_t)b * NLS + LS_AMAX];
}

}  // namespace mpxk
