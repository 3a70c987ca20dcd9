// K4: merit penalty, fraction-to-boundary, merit line search and the
// primal-dual update, one block per instance.
//
// phi(a) = sum_t [stage cost + AL terms of the equality rows - mu_b sum log s]
//          + (V + a dV)^T c(a) + rho/2 |c(a)|^2
// at the trial point (X + a dX, U + a dU), with the dynamics re-evaluated on
// the device (tanh or LTV model).  rho follows the oracle's rule
// rho = clip(max(rho0, 2 (d_f + d_c + kappa/2) / |c|^2), rho_max); trials
// a_j = a_max beta^j are evaluated `group` at a time and the first that passes
// phi(a) <= phi(0) + c1 a phi'(0) + slack (1 + |phi(0)|) is taken.
#pragma once
#include "common.h"
#include "k_lqr.h"

namespace mpxk {

// contribution of (stage t, row i) to phi(a)
template <typename T>
__device__ T merit_item(const Prob<T>& P, const Prm<T>& prm, const Iter<T>& it, const Ws<T>& ws,
                        const LqrIn<T>& in, int b, int t, int i, T a, T rho) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const T* X = it.X + ((size_t)b * N1 + t) * n;
  const T* dX = ws.dx + ((size_t)b * N1 + t) * n;
  T val = 0;
  if (t == N) {
    if (i < n) {
      const T* QN = P.QN.at(b, 0);
      const T* xr = P.xref.at(b, N);
      T qx = 0;
      for (int j = 0; j < n; ++j) qx += QN[i * n + j] * (X[j] + a * dX[j] - xr[j]);
      val += T(0.5) * (X[i] + a * dX[i] - xr[i]) * qx;
      const T* X0 = it.X + (size_t)b * N1 * n;
      const T* dX0 = ws.dx + (size_t)b * N1 * n;
      const T c = P.x0.at(b, 0)[i] - (X0[i] + a * dX0[i]);
      const T v = it.V[(size_t)b * N1 * n + i] + a * ws.dv[(size_t)b * N1 * n + i];
      val += v * c + T(0.5) * rho * c * c;
    }
    return val;
  }
  const T* U = it.U + ((size_t)b * N + t) * m;
  const T* dU = ws.du + ((size_t)b * N + t) * m;
  if (i < n) {
    T f;
    if (P.model == 1) {
      const T* A = P.A.at(b, 0);
      const T* Bm = P.Bm.at(b, 0);
      const T* W = P.W.at(b, 0);
      T ax = 0, wx = 0, bu = 0;
      for (int j = 0; j < n; ++j) {
        const T xj = X[j] + a * dX[j];
        ax += A[i * n + j] * xj;
        wx += W[i * n + j] * xj;
      }
      for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * (U[j] + a * dU[j]);
      f = X[i] + a * dX[i] + P.h * (ax + bu + P.kappa * dev_tanh(wx));
    } else {
      const T* A = in.A.at(b, t);
      const T* Bm = in.Bm.at(b, t);
      T ax = 0, bu = 0;
      for (int j = 0; j < n; ++j) ax += A[i * n + j] * (X[j] + a * dX[j]);
      for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * (U[j] + a * dU[j]);
      f = ax + bu;
    }
    const size_t o = ((size_t)b * N1 + t + 1) * n + i;
    const T c = f - (it.X[o] + a * ws.dx[o]);
    const T v = it.V[o] + a * ws.dv[o];
    val += v * c + T(0.5) * rho * c * c;
    const T* Q = P.Q.at(b, t);
    const T* xr = P.xref.at(b, t);
    T qx = 0;
    for (int j = 0; j < n; ++j) qx += Q[i * n + j] * (X[j] + a * dX[j] - xr[j]);
    val += T(0.5) * (X[i] + a * dX[i] - xr[i]) * qx;
  }
  if (i < m) {
    const T* R = P.R.at(b, t);
    const T* ur = P.uref.at(b, t);
    T ru = 0;
    for (int j = 0; j < m; ++j) ru += R[i * m + j] * (U[j] + a * dU[j] - ur[j]);
    const T ui = U[i] + a * dU[i];
    val += T(0.5) * (ui - ur[i]) * ru;
    if (P.box) {
      const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
      val -= prm.mu_b * (dev_log(hi - ui) + dev_log(ui - lo));
    }
  }
  if (i < p) {
    const T* D = P.D.at(b, t);
    const T* E = P.E.at(b, t);
    T h = P.e.at(b, t)[i];
    for (int j = 0; j < n; ++j) h += D[i * n + j] * (X[j] + a * dX[j]);
    for (int j = 0; j < m; ++j) h += E[i * m + j] * (U[j] + a * dU[j]);
    const T mu = it.Yeq[((size_t)b * N + t) * p + i];
    val += mu * h + T(0.5) * prm.rho_c * h * h;
  }
  return val;
}

template <typename T>
__global__ void __launch_bounds__(512) k_linesearch(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, LqrIn<T> in, T* o_alpha,
                             T* o_rho, T* o_phi0, T* o_dphi, int* o_status, int* o_trials) {
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sred = reinterpret_cast<T*>(smem_raw);  // MAX_TRIALS * nt
  // 1) reduce the merit and slope partials
  T v6[6] = {0, 0, 0, 0, 0, 0};
  for (int t = tid; t < N1; t += nt) {
    const T* mp = ws.mpart + ((size_t)b * N1 + t) * NMP;
    v6[0] += mp[MP_COST];
    v6[1] += mp[MP_VC];
    v6[2] += mp[MP_CC];
  }
  for (int s = tid; s < N1; s += nt) {
    const T* fp = ws.fpart + ((size_t)b * N1 + s) * NFP;
    v6[3] += fp[FP_DF];
    v6[4] += fp[FP_DC];
    v6[5] += fp[FP_LC];
  }
  block_sum<T, 6>(v6, 6, sred);
  // curvature dz'H dz = sum lam'c - d_f (stationarity of the LQR step)
  const T cost = v6[0], vc = v6[1], cc = v6[2], df = v6[3], dc = v6[4], kap = v6[5] - v6[3];
  T rho = prm.rho0;
  if (cc > T(0)) {
    const T r2 = T(2) * (df + dc + T(0.5) * kap) / cc;
    rho = r2 > rho ? r2 : rho;
  }
  rho = rho < prm.rho_max ? rho : prm.rho_max;
  const T dphi = df + dc - rho * cc;
  const T phi0 = cost + vc + T(0.5) * rho * cc;
  // 2) fraction-to-boundary for slacks and box multipliers
  T amax = 1, adual = 1;
  if (P.box) {
    T ap = 1, ad = 1;
    for (int e = tid; e < N * m; e += nt) {
      const int t = e / m, i = e - t * m;
      const T u = it.U[((size_t)b * N + t) * m + i];
      const T du = ws.du[((size_t)b * N + t) * m + i];
      const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
      const T shi = hi - u, slo = u - lo;
      const T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
      const T nhi = nu[i], nlo = nu[m + i];
      if (du > T(0)) { const T r = prm.tau * shi / du; ap = r < ap ? r : ap; }
      if (du < T(0)) { const T r = prm.tau * slo / (-du); ap = r < ap ? r : ap; }
      const T dhi = prm.mu_b / shi - nhi + nhi * du / shi;
      const T dlo = prm.mu_b / slo - nlo - nlo * du / slo;
      if (dhi < T(0)) { const T r = -prm.tau * nhi / dhi; ad = r < ad ? r : ad; }
      if (dlo < T(0)) { const T r = -prm.tau * nlo / dlo; ad = r < ad ? r : ad; }
    }
    amax = block_min(ap, sred);
    adual = block_min(ad, sred);
  }
  // 3) trials
  const T slack = prm.slack * (T(1) + (phi0 < 0 ? -phi0 : phi0));
  const int Rr = n > m ? (n > p ? n : p) : (m > p ? m : p);
  const int per = N1 * Rr;
  T alpha = -1;
  int trials = 0;
  int j0 = 0;
  while (j0 < prm.max_trials && alpha < T(0)) {
    int g = prm.group;
    if (g > prm.max_trials - j0) g = prm.max_trials - j0;
    T acc[MAX_TRIALS];
    T av[MAX_TRIALS];
    for (int j = 0; j < g; ++j) {
      acc[j] = 0;
      T a = amax;
      for (int q = 0; q < j0 + j; ++q) a *= prm.beta;
      av[j] = a;
    }
    for (int e = tid; e < g * per; e += nt) {
      const int j = e / per, r = e - j * per;
      const int t = r / Rr, i = r - t * Rr;
      acc[j] += merit_item(P, prm, it, ws, in, b, t, i, av[j], rho);
    }
    block_sum<T, MAX_TRIALS>(acc, g, sred);
    trials += g;
    for (int j = 0; j < g; ++j) {
      if (acc[j] <= phi0 + prm.c1 * av[j] * dphi + slack) { alpha = av[j]; break; }
    }
    j0 += g;
  }
  int status = 0;
  if (alpha < T(0)) {
    alpha = amax;
    for (int q = 0; q < prm.max_trials - 1; ++q) alpha *= prm.beta;
    status = 2;
  }
  if (!isfinite(phi0) || !isfinite(dphi)) status = 1;
  // 4) multiplier updates (read the old iterate), then the primal-dual step
  if (p > 0) {
    for (int e = tid; e < N * p; e += nt) {
      const int t = e / p, k = e - t * p;
      const T* D = P.D.at(b, t);
      const T* E = P.E.at(b, t);
      const T* X = it.X + ((size_t)b * N1 + t) * n;
      const T* dX = ws.dx + ((size_t)b * N1 + t) * n;
      const T* U = it.U + ((size_t)b * N + t) * m;
      const T* dU = ws.du + ((size_t)b * N + t) * m;
      T h = P.e.at(b, t)[k];
      for (int j = 0; j < n; ++j) h += D[k * n + j] * (X[j] + alpha * dX[j]);
      for (int j = 0; j < m; ++j) h += E[k * m + j] * (U[j] + alpha * dU[j]);
      it.Yeq[((size_t)b * N + t) * p + k] += prm.rho_c * h;
    }
  }
  if (P.box) {
    for (int e = tid; e < N * m; e += nt) {
      const int t = e / m, i = e - t * m;
      const T u = it.U[((size_t)b * N + t) * m + i];
      const T du = ws.du[((size_t)b * N + t) * m + i];
      const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
      const T shi = hi - u, slo = u - lo;
      T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
      const T nhi = nu[i], nlo = nu[m + i];
      nu[i] = nhi + adual * (prm.mu_b / shi - nhi + nhi * du / shi);
      nu[m + i] = nlo + adual * (prm.mu_b / slo - nlo - nlo * du / slo);
    }
  }
  __syncthreads();
  for (int e = tid; e < N1 * n; e += nt) {
    const size_t o = (size_t)b * N1 * n + e;
    it.X[o] += alpha * ws.dx[o];
    it.V[o] += alpha * ws.dv[o];
  }
  for (int e = tid; e < N * m; e += nt) {
    const size_t o = (size_t)b * N * m + e;
    it.U[o] += alpha * ws.du[o];
  }
  if (tid == 0) {
    if (o_alpha) o_alpha[b] = alpha;
    if (o_rho) o_rho[b] = rho;
    if (o_phi0) o_phi0[b] = phi0;
    if (o_dphi) o_dphi[b] = dphi;
    if (o_status) o_status[b] = status;
    if (o_trials) o_trials[b] = trials;
  }
}

}  // namespace mpxk
