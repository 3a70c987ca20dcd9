// K4 as three stage-parallel kernels: the merit line search and the
// primal-dual update with one wave per (stage, instance).
//
//   k_ls_setup   one block per instance: rho, phi(0), phi'(0) from the stage
//                partials of K1 and of the forward pass, the fraction-to-
//                boundary steps of the box slacks and multipliers;
//   k_trials     one wave per stage: that stage's part of phi(a_j) for the
//                trials a_j = a_max beta^j, j0 <= j < j1;
//   k_ls_decide  one wave per instance: the first a_j that passes the Armijo
//                test (the stage parts summed in a fixed order);
//   k_ls_update  one thread per entry: the multiplier and iterate update at
//                the accepted step.
//
// At a trial point x = X + a dX, u = U + a dU every matrix product of the
// merit is affine in a: A x = A X + a A dX, W x, B u, Q (x - xr), R (u - ur),
// D x + E u.  A wave forms the products at a = 0 and their slopes once per
// stage (two matrix-vector products per matrix) and then evaluates each trial
// with O(n + m + p) work per stage: the tanh of the model, the box logarithms
// and the quadratic forms (d0 + a d1)'(Qd0 + a Qd1).  The terms are those of
// stage_terms / merit_item (mpx/utils/mpc_wrapper.py:244 calls the solver
// whose merit this is): stage t < N contributes its cost, the AL terms of its
// equality rows, the box barrier, V_{t+1}'c_t and |c_t|^2 with c_t = f(x_t,
// u_t) - x_{t+1}; t = N the terminal cost, V_0'c_init and |c_init|^2.  A
// trial's phi is the sum over stages of the first part plus rho/2 times the
// sum of |c|^2, so the trial kernel does not need rho.
//
// The trials are evaluated in two passes: j in [0, group) and, for the
// instances none of those passed, [group, max_trials); the accepted step and
// the count of examined trials are those of backtracking in groups of
// `group` (D4).
#pragma once
#include "common.h"
#include "k_lqr.h"

namespace mpxk {

enum { LS_RHO = 0, LS_PHI0 = 1, LS_DPHI = 2, LS_AMAX = 3, LS_ADUAL = 4, LS_SLACK = 5, LS_ALPHA = 6, NLS = 8 };
constexpr int kTrialWaves = 4;

// Elements of one wave's LDS vectors in k_trials: x0 = X_t, x1 = dX_t,
// d0 = X_t - xr_t (n each, rounded up to 4), u0, u1, e0 = U_t - ur_t (m each).
__host__ __device__ inline int trial_wave_elems(int n, int m) {
  const int n4 = (n + 3) & ~3, m4 = (m + 3) & ~3;
  return 3 * n4 + 3 * m4;
}

// With SH (tanh model, stage-invariant Q and R) the block's waves take
// consecutive stages of one instance and share its A, W, B, Q, R, staged once
// in LDS with rows padded to length + 1 (lane-per-row reads hit distinct
// banks); elements of that area, rounded up to 4.
__host__ __device__ inline int trial_shared_elems(int n, int m) {
  return ((3 * n * (n + 1) + n * (m + 1) + m * (m + 1)) + 3) & ~3;
}

// y0 = row . v0, y1 = row . v1 (row in global memory, v0 and v1 in LDS).  V4:
// the row is 16-byte aligned and L % 4 == 0; the sums run in the same order
// either way.
template <typename T, bool V4>
__device__ __forceinline__ void row_dot2(const T* __restrict__ row, int L, const T* v0, const T* v1, T& y0, T& y1) {
  T a0 = 0, a1 = 0;
  if constexpr (V4) {
    const float4* r4 = reinterpret_cast<const float4*>(row);
    const float4* p0 = reinterpret_cast<const float4*>(v0);
    const float4* p1 = reinterpret_cast<const float4*>(v1);
    for (int q = 0; q < (L >> 2); ++q) {
      const float4 w = r4[q], x = p0[q], y = p1[q];
      a0 += w.x * x.x; a1 += w.x * y.x;
      a0 += w.y * x.y; a1 += w.y * y.y;
      a0 += w.z * x.z; a1 += w.z * y.z;
      a0 += w.w * x.w; a1 += w.w * y.w;
    }
  } else {
    for (int j = 0; j < L; ++j) {
      const T w = row[j];
      a0 += w * v0[j];
      a1 += w * v1[j];
    }
  }
  y0 = a0;
  y1 = a1;
}

// row_dot2 for a padded LDS row and 16-byte aligned LDS vectors (FP32): the
// vectors are read four elements at a time, the sums run in the same order.
template <typename T>
__device__ __forceinline__ void row_dot2_lds(const T* row, int L, const T* v0, const T* v1, T& y0, T& y1) {
  if constexpr (IsF32<T>::value) {
    T a0 = 0, a1 = 0;
    const int L4 = L & ~3;
    const float4* p0 = reinterpret_cast<const float4*>(v0);
    const float4* p1 = reinterpret_cast<const float4*>(v1);
    for (int q = 0; q < (L4 >> 2); ++q) {
      const float4 x = p0[q], y = p1[q];
      const T* w = row + 4 * q;
      a0 += w[0] * x.x; a1 += w[0] * y.x;
      a0 += w[1] * x.y; a1 += w[1] * y.y;
      a0 += w[2] * x.z; a1 += w[2] * y.z;
      a0 += w[3] * x.w; a1 += w[3] * y.w;
    }
    for (int j = L4; j < L; ++j) {
      a0 += row[j] * v0[j];
      a1 += row[j] * v1[j];
    }
    y0 = a0;
    y1 = a1;
  } else {
    row_dot2<T, false>(row, L, v0, v1, y0, y1);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_ls_setup(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws) {
  const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  if (skip_instance(ws, b)) return;
  const int m = P.m, N = P.N, N1 = N + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sred = reinterpret_cast<T*>(smem_raw);   // 6 * blockDim.x
  T v6[6] = {0, 0, 0, 0, 0, 0};
  for (int t = tid; t < N1; t += nt) {
    const T* mp = ws.mpart + ((size_t)b * N1 + t) * NMP;
    const T* fp = ws.fpart + ((size_t)b * N1 + t) * NFP;
    v6[0] += mp[MP_COST];
    v6[1] += mp[MP_VC];
    v6[2] += mp[MP_CC];
    v6[3] += fp[FP_DF];
    v6[4] += fp[FP_DC];
    v6[5] += fp[FP_LC];
  }
  block_sum<T, 6>(v6, 6, sred);
  // rho = clip(max(rho0, 2 (d_f + d_c + kappa/2) / |c|^2), rho_max), kappa = sum lam'c - d_f
  const T cost = v6[0], vc = v6[1], cc = v6[2], df = v6[3], dc = v6[4], kap = v6[5] - v6[3];
  T rho = prm.rho0;
  if (cc > T(0)) {
    const T r2 = T(2) * (df + dc + T(0.5) * kap) / cc;
    rho = r2 > rho ? r2 : rho;
  }
  rho = rho < prm.rho_max ? rho : prm.rho_max;
  const T dphi = df + dc - rho * cc;
  const T phi0 = cost + vc + T(0.5) * rho * cc;
  T amax = 1, adual = 1;
  if (P.box) {
    T ap = 1, ad = 1;
    for (int e = tid; e < N * m; e += nt) {
      const int t = e / m, i = e - t * m;
      const T u = it.U[(size_t)b * N * m + e];
      const T du = ws.du[(size_t)b * N * m + e];
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
  if (tid == 0) {
    T* o = ws.lsp + (size_t)b * NLS;
    o[LS_RHO] = rho;
    o[LS_PHI0] = phi0;
    o[LS_DPHI] = dphi;
    o[LS_AMAX] = amax;
    o[LS_ADUAL] = adual;
    o[LS_SLACK] = prm.slack * (T(1) + (phi0 < 0 ? -phi0 : phi0));
  }
}

// First j < jend whose trial passes the Armijo test, or -1; every lane of the
// wave gets the same answer (sums over the stages in a fixed order).
template <typename T>
__device__ int ls_first_accept(const Prm<T>& prm, const Ws<T>& ws, int b, int N1, int jend) {
  const int lane = threadIdx.x & 63;
  const T* ls = ws.lsp + (size_t)b * NLS;
  const T rho = ls[LS_RHO], phi0 = ls[LS_PHI0], dphi = ls[LS_DPHI], slack = ls[LS_SLACK];
  const T* tp = ws.tpart + (size_t)b * N1 * 2 * MAX_TRIALS;
  T a = ls[LS_AMAX];
  for (int j = 0; j < jend; ++j, a *= prm.beta) {
    T s1 = 0, s2 = 0;
    for (int t = lane; t < N1; t += 64) {
      s1 += tp[(size_t)t * 2 * MAX_TRIALS + 2 * j];
      s2 += tp[(size_t)t * 2 * MAX_TRIALS + 2 * j + 1];
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    const T phi = s1 + T(0.5) * rho * s2;
    if (phi <= phi0 + prm.c1 * a * dphi + slack) return j;
  }
  return -1;
}

// Trials j0 <= j < j1 at stage t = t0 + blockIdx.x * waves + wave of instance
// blockIdx.y.  RPL: state (and equality) rows per lane, ceil(max(n, p) / 64).
template <typename T, int RPL, bool V4, bool SH>
__global__ void __launch_bounds__(512) k_trials(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, int j0, int j1,
                                                int t0) {
  static_assert(!V4 || IsF32<T>::value, "float4 rows are FP32 only");
  const int wpb = blockDim.x >> 6, wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = blockIdx.y, t = t0 + blockIdx.x * wpb + wv;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  // an earlier pass accepted: the same answer in every wave of the block (one instance)
  if (j0 > 0 && ls_first_accept(prm, ws, b, N1, j0) >= 0) return;
  const int n4 = (n + 3) & ~3, m4 = (m + 3) & ~3;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sh = reinterpret_cast<T*>(smem_raw);
  const int lA = 0, lW = n * (n + 1), lB = 2 * n * (n + 1), lQ = lB + n * (m + 1), lR = lQ + n * (n + 1);
  if constexpr (SH) {
    const T* A = P.A.at(b, 0);
    const T* W = P.W.at(b, 0);
    const T* Q = P.Q.at(b, 0);
    const T* Bm = P.Bm.at(b, 0);
    const T* R = P.R.at(b, 0);
    const int tid = threadIdx.x, nt = blockDim.x, nn = n * n, nm = n * m, mm = m * m;
    // element e = i ld + j of a row-major matrix goes to e + i in the padded rows
    if (nn <= 4 * nt && nm <= 2 * nt && mm <= nt) {
      // one pass with every load of the thread in flight before its first store
      // (a loop would wait on global memory once per iteration)
      T va[4], vw[4], vq[4], vb[2], vr;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = tid + k * nt;
        va[k] = e < nn ? A[e] : T(0);
        vw[k] = e < nn ? W[e] : T(0);
        vq[k] = e < nn ? Q[e] : T(0);
      }
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int e = tid + k * nt;
        vb[k] = e < nm ? Bm[e] : T(0);
      }
      vr = tid < mm ? R[tid] : T(0);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = tid + k * nt;
        if (e < nn) {
          const int o = e + e / n;
          sh[lA + o] = va[k];
          sh[lW + o] = vw[k];
          sh[lQ + o] = vq[k];
        }
      }
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int e = tid + k * nt;
        if (e < nm) sh[lB + e + e / m] = vb[k];
      }
      if (tid < mm) sh[lR + tid + tid / m] = vr;
    } else {
      for (int e = tid; e < nn; e += nt) {
        const int o = e + e / n;
        sh[lA + o] = A[e];
        sh[lW + o] = W[e];
        sh[lQ + o] = Q[e];
      }
      for (int e = tid; e < nm; e += nt) sh[lB + e + e / m] = Bm[e];
      for (int e = tid; e < mm; e += nt) sh[lR + e + e / m] = R[e];
    }
    __syncthreads();
  }
  if (t > N) return;                                          // whole wave: no block barrier follows
  T* x0 = sh + (SH ? trial_shared_elems(n, m) : 0) + (size_t)wv * trial_wave_elems(n, m);
  T* x1 = x0 + n4;
  T* d0 = x1 + n4;
  T* u0 = d0 + n4;
  T* u1 = u0 + m4;
  T* e0 = u1 + m4;
  const bool term = t == N;
  const size_t ox = ((size_t)b * N1 + t) * n;
  {
    const T* xr = P.xref.at(b, t);
    for (int i = lane; i < n; i += 64) {
      const T X = it.X[ox + i];
      x0[i] = X;
      x1[i] = ws.dx[ox + i];
      d0[i] = X - xr[i];
    }
    if (!term) {
      const T* ur = P.uref.at(b, t);
      const size_t ou = ((size_t)b * N + t) * m;
      for (int i = lane; i < m; i += 64) {
        const T U = it.U[ou + i];
        u0[i] = U;
        u1[i] = ws.du[ou + i];
        e0[i] = U - ur[i];
      }
    }
  }
  wave_lds_sync();
  // per state row: f(a) = F0 + a F1 (+ h kappa tanh(w0 + a w1)); c = f - (xn0 + a xn1)
  T F0[RPL], F1[RPL], w0[RPL], w1[RPL], q0[RPL], q1[RPL], dd0[RPL], dd1[RPL];
  T xn0[RPL], xn1[RPL], vn0[RPL], vn1[RPL];
  const bool tanh_model = P.model == 1;
  (void)lA;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int i = lane + 64 * r;
    F0[r] = F1[r] = w0[r] = w1[r] = q0[r] = q1[r] = dd0[r] = dd1[r] = 0;
    xn0[r] = xn1[r] = vn0[r] = vn1[r] = 0;
    if (i < n) {
      dd0[r] = d0[i];
      dd1[r] = x1[i];
      if (!term) {
        T za0, za1, zb0, zb1;
        if (SH) {
          row_dot2_lds<T>(sh + lA + i * (n + 1), n, x0, x1, za0, za1);
          row_dot2_lds<T>(sh + lW + i * (n + 1), n, x0, x1, w0[r], w1[r]);
          row_dot2_lds<T>(sh + lB + i * (m + 1), m, u0, u1, zb0, zb1);
          F0[r] = x0[i] + P.h * (za0 + zb0);
          F1[r] = x1[i] + P.h * (za1 + zb1);
        } else if (tanh_model) {
          row_dot2<T, V4>(P.A.at(b, 0) + (size_t)i * n, n, x0, x1, za0, za1);
          row_dot2<T, V4>(P.W.at(b, 0) + (size_t)i * n, n, x0, x1, w0[r], w1[r]);
          row_dot2<T, V4>(P.Bm.at(b, 0) + (size_t)i * m, m, u0, u1, zb0, zb1);
          F0[r] = x0[i] + P.h * (za0 + zb0);
          F1[r] = x1[i] + P.h * (za1 + zb1);
        } else {
          row_dot2<T, V4>(P.A.at(b, t) + (size_t)i * n, n, x0, x1, za0, za1);
          row_dot2<T, V4>(P.Bm.at(b, t) + (size_t)i * m, m, u0, u1, zb0, zb1);
          F0[r] = za0 + zb0;
          F1[r] = za1 + zb1;
        }
        if (SH) row_dot2_lds<T>(sh + lQ + i * (n + 1), n, d0, x1, q0[r], q1[r]);
        else row_dot2<T, V4>(P.Q.at(b, t) + (size_t)i * n, n, d0, x1, q0[r], q1[r]);
        const size_t on = ox + n + i;
        xn0[r] = it.X[on];
        xn1[r] = ws.dx[on];
        vn0[r] = it.V[on];
        vn1[r] = ws.dv[on];
      } else {
        // terminal cost and the initial-state row c_init = x0 - X_0
        row_dot2<T, V4>(P.QN.at(b, 0) + (size_t)i * n, n, d0, x1, q0[r], q1[r]);
        const size_t o0 = (size_t)b * N1 * n + i;
        F0[r] = P.x0.at(b, 0)[i];
        xn0[r] = it.X[o0];
        xn1[r] = ws.dx[o0];
        vn0[r] = it.V[o0];
        vn1[r] = ws.dv[o0];
      }
    }
  }
  // control row i = lane: R (u - ur), box bounds
  T ru0 = 0, ru1 = 0, ue0 = 0, ue1 = 0, uu0 = 0, lo = 0, hi = 0;
  if (!term && lane < m) {
    if (SH) row_dot2_lds<T>(sh + lR + lane * (m + 1), m, e0, u1, ru0, ru1);
    else row_dot2<T, V4>(P.R.at(b, t) + (size_t)lane * m, m, e0, u1, ru0, ru1);
    ue0 = e0[lane];
    ue1 = u1[lane];
    uu0 = u0[lane];
    if (P.box) {
      lo = P.ulo.at(b, t)[lane];
      hi = P.uhi.at(b, t)[lane];
    }
  }
  // equality rows k: h(a) = h0 + a h1
  T h0[RPL], h1[RPL], mu[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int k = lane + 64 * r;
    h0[r] = h1[r] = mu[r] = 0;
    if (!term && k < p) {
      T a0, a1, c0, c1;
      row_dot2<T, V4>(P.D.at(b, t) + (size_t)k * n, n, x0, x1, a0, a1);
      row_dot2<T, V4>(P.E.at(b, t) + (size_t)k * m, m, u0, u1, c0, c1);
      h0[r] = P.e.at(b, t)[k] + a0 + c0;
      h1[r] = a1 + c1;
      mu[r] = it.Yeq[((size_t)b * N + t) * p + k];
      if (j0 == 0) {
        T* hq = ws.hq + ((size_t)b * N + t) * 2 * p;
        hq[k] = h0[r];
        hq[p + k] = h1[r];
      }
    }
  }
  T a = ws.lsp[(size_t)b * NLS + LS_AMAX];
  for (int q = 0; q < j0; ++q) a *= prm.beta;
  T* tp = ws.tpart + ((size_t)b * N1 + t) * 2 * MAX_TRIALS;
  for (int j = j0; j < j1; ++j, a *= prm.beta) {
    T s1 = 0, s2 = 0;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      if (lane + 64 * r < n) {
        T f = F0[r] + a * F1[r];
        if (tanh_model && !term) f += P.h * P.kappa * dev_tanh(w0[r] + a * w1[r]);
        const T c = f - (xn0[r] + a * xn1[r]);
        const T v = vn0[r] + a * vn1[r];
        s1 += v * c + T(0.5) * (dd0[r] + a * dd1[r]) * (q0[r] + a * q1[r]);
        s2 += c * c;
      }
    }
    if (!term && lane < m) {
      s1 += T(0.5) * (ue0 + a * ue1) * (ru0 + a * ru1);
      if (P.box) {
        const T u = uu0 + a * ue1;
        s1 -= prm.mu_b * (dev_log(hi - u) + dev_log(u - lo));
      }
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      if (!term && lane + 64 * r < p) {
        const T hh = h0[r] + a * h1[r];
        s1 += mu[r] * hh + T(0.5) * prm.rho_c * hh * hh;
      }
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
      tp[2 * j] = s1;
      tp[2 * j + 1] = s2;
    }
  }
}

// The accepted step of instance blockIdx.x (one wave): the first trial that
// passes the Armijo test, else the last one (status 2); the trial count
// reported is that of backtracking in groups of prm.group.  Writes the step
// to lsp[LS_ALPHA] for k_ls_update and the instance's report.
template <typename T>
__global__ void __launch_bounds__(64) k_ls_decide(Prm<T> prm, Ws<T> ws, int N1, T* o_alpha, T* o_rho, T* o_phi0,
                                                  T* o_dphi, int* o_status, int* o_trials) {
  const int b = blockIdx.x, lane = threadIdx.x;
  if (skip_instance(ws, b)) return;
  T* ls = ws.lsp + (size_t)b * NLS;
  const int MT = prm.max_trials, g = prm.group;
  const int js = ls_first_accept(prm, ws, b, N1, MT);
  T alpha = ls[LS_AMAX];
  int status = 0, trials = MT;
  if (js >= 0) {
    for (int q = 0; q < js; ++q) alpha *= prm.beta;
    trials = (js / g + 1) * g;
    trials = trials < MT ? trials : MT;
  } else {
    for (int q = 0; q < MT - 1; ++q) alpha *= prm.beta;
    status = 2;
  }
  const T rho = ls[LS_RHO], phi0 = ls[LS_PHI0], dphi = ls[LS_DPHI];
  if (!isfinite(phi0) || !isfinite(dphi)) status = 1;
  if (lane == 0) {
    ls[LS_ALPHA] = alpha;
    if (o_alpha) o_alpha[b] = alpha;
    if (o_rho) o_rho[b] = rho;
    if (o_phi0) o_phi0[b] = phi0;
    if (o_dphi) o_dphi[b] = dphi;
    if (o_status) o_status[b] = status;
    if (o_trials) o_trials[b] = trials;
  }
}

// Iterate and multiplier update at the accepted step a = lsp[LS_ALPHA], one
// thread per entry: blocks [0, nbx) take X += a dX and V += a dV, the next
// nbu blocks U += a dU with the box multipliers' update (at u before the
// step), the rest Y_eq += rho_c (h + a dh) from the equality rows and slopes
// the trials stored.
template <typename T>
__global__ void __launch_bounds__(256) k_ls_update(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, int nbx, int nbu) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const size_t B = (size_t)P.B;
  int blk = blockIdx.x;
  if (blk < nbx) {
    const size_t e = (size_t)blk * 256 + threadIdx.x;
    if (e >= B * N1 * n) return;
    const int b = (int)(e / ((size_t)N1 * n));
    if (skip_instance(ws, b)) return;
    const T a = ws.lsp[(size_t)b * NLS + LS_ALPHA];
    it.X[e] += a * ws.dx[e];
    it.V[e] += a * ws.dv[e];
    return;
  }
  blk -= nbx;
  if (blk < nbu) {
    const size_t e = (size_t)blk * 256 + threadIdx.x;
    if (e >= B * N * m) return;
    const int b = (int)(e / ((size_t)N * m));
    if (skip_instance(ws, b)) return;
    const int r = (int)(e - (size_t)b * N * m), t = r / m, i = r - t * m;
    const T* ls = ws.lsp + (size_t)b * NLS;
    const T a = ls[LS_ALPHA];
    const T u = it.U[e], du = ws.du[e];
    if (P.box) {
      const T adual = ls[LS_ADUAL];
      const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
      const T shi = hi - u, slo = u - lo;
      T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
      const T nhi = nu[i], nlo = nu[m + i];
      nu[i] = nhi + adual * (prm.mu_b / shi - nhi + nhi * du / shi);
      nu[m + i] = nlo + adual * (prm.mu_b / slo - nlo - nlo * du / slo);
    }
    it.U[e] = u + a * du;
    return;
  }
  blk -= nbu;
  const size_t e = (size_t)blk * 256 + threadIdx.x;
  if (p == 0 || e >= B * N * p) return;
  const int b = (int)(e / ((size_t)N * p));
  if (skip_instance(ws, b)) return;
  const int r = (int)(e - (size_t)b * N * p), t = r / p, k = r - t * p;
  const T a = ws.lsp[(size_t)b * NLS + LS_ALPHA];
  const T* hq = ws.hq + ((size_t)b * N + t) * 2 * p;
  it.Yeq[e] += prm.rho_c * (hq[k] + a * hq[p + k]);
}

}  // namespace mpxk
