// K2/K3: the LQR subproblem of one SQP iteration.
//
// Stage data (per instance b, stage t < N): A_t, B_t (views), Q_t, R_t, D_t, E_t
// (problem), augmented gradients q_t, r_t, box diagonal rd_t and the defect
// c_t = cdef[t+1]; the QP Hessian is [[Q + rho_c D'D, rho_c D'E],
// [rho_c E'D, R + diag(rd) + rho_c E'E]] (the AL rows never materialise).
//
// Backward: value V_t(x) = x'S_t x/2 + s_t'x, by sequential Riccati steps
// (segments = 1) or by the chunked associative scan (segments > 1): every
// segment but the last reduces its stages to one element (A, b, C, eta, J)
// with leaf combines, the last runs Riccati from the terminal cost, the suffix
// values at segment starts follow from value combines, and every other segment
// then runs Riccati from its suffix value.  Forward: dx_0 = c_init,
// du = -K dx - k, dx' = A dx + B du + c, costates lam = S dx + s,
// dV = lam - V; again sequential or chunked (segment maps composed, states at
// segment starts propagated, then each segment rolled out).
#pragma once
#include "common.h"

namespace mpxk {

template <typename T>
struct LqrIn {
  View<T> A, Bm;
};

// LDS carve-up of the Riccati / combine working set.
template <typename T>
struct RicSmem {
  T *S, *A, *T1, *B, *SB, *G, *H, *D, *E, *s, *Sc, *c, *rd, *q, *r, *red;
};

template <typename T>
__host__ __device__ inline size_t ric_smem_elems(int n, int m, int p, int nt) {
  return (size_t)3 * n * n + 2 * (size_t)n * m + (size_t)m * (n + 1) + (size_t)m * m +
         (size_t)p * (n + m) + 4 * (size_t)n + 2 * (size_t)m + 2 * (size_t)n + 4 * (size_t)nt + 8;
}

template <typename T>
__device__ RicSmem<T> ric_carve(T* base, int n, int m, int p, int nt) {
  RicSmem<T> L;
  T* q = base;
  L.S = q; q += n * n;
  L.A = q; q += n * n;
  L.T1 = q; q += n * n;
  L.B = q; q += n * m;
  L.SB = q; q += n * m;
  L.G = q; q += m * (n + 1);
  L.H = q; q += m * m;
  L.D = q; q += p * n;
  L.E = q; q += p * m;
  L.s = q; q += n;
  L.Sc = q; q += n;
  L.c = q; q += n;
  L.q = q; q += n;
  L.rd = q; q += m;
  L.r = q; q += m;
  q += 2 * n;  // spare vectors (forward pass)
  L.red = q;
  return L;
}

// Load the stage's QP data into LDS (A, B, D, E, c, q, r, rd).
template <typename T>
__device__ void load_stage(const Prob<T>& P, const Ws<T>& ws, const LqrIn<T>& in, int b, int t,
                           RicSmem<T>& L) {
  const int n = P.n, m = P.m, p = P.p, N1 = P.N + 1;
  load_block(L.A, in.A.at(b, t), n, n, n);
  load_block(L.B, in.Bm.at(b, t), n, m, m);
  if (p > 0) {
    load_block(L.D, P.D.at(b, t), p, n, n);
    load_block(L.E, P.E.at(b, t), p, m, m);
  }
  const T* c = ws.cdef + ((size_t)b * N1 + t + 1) * n;
  const T* q = ws.qh + ((size_t)b * N1 + t) * n;
  const T* r = ws.rh + ((size_t)b * P.N + t) * m;
  const T* rd = ws.rd + ((size_t)b * P.N + t) * m;
  for (int i = threadIdx.x; i < n; i += blockDim.x) { L.c[i] = c[i]; L.q[i] = q[i]; }
  for (int i = threadIdx.x; i < m; i += blockDim.x) { L.r[i] = r[i]; L.rd[i] = rd[i]; }
}

// One backward Riccati step in LDS: (S, s) = V_{t+1} -> V_t, and K_t, k_t in
// G (m x (n+1); column n holds k).  The stage data must already be loaded.
template <typename T>
__device__ void riccati_step(const Prob<T>& P, const Prm<T>& prm, int b, int t, RicSmem<T>& L) {
  const int n = P.n, m = P.m, p = P.p, ld = n + 1;
  const int tid = threadIdx.x, nt = blockDim.x;
  const T rc = prm.rho_c;
  // Sc = S c + s ; T1 = S A ; SB = S B
  for (int i = tid; i < n; i += nt) {
    T v = L.s[i];
    for (int j = 0; j < n; ++j) v += L.S[i * n + j] * L.c[j];
    L.Sc[i] = v;
  }
  gemm_lds<T, false, false, 0>(n, n, n, L.S, n, L.A, n, L.T1, n);
  gemm_lds<T, false, false, 0>(n, m, n, L.S, n, L.B, m, L.SB, m);
  // H = R + diag(rd) + rc E'E ; G = rc E'D ; g = r
  const T* R = P.R.at(b, t);
  for (int e = tid; e < m * m; e += nt) {
    const int i = e / m, j = e - i * m;
    T v = R[e] + (i == j ? L.rd[i] : T(0));
    for (int k = 0; k < p; ++k) v += rc * L.E[k * m + i] * L.E[k * m + j];
    L.H[e] = v;
  }
  for (int e = tid; e < m * n; e += nt) {
    const int i = e / n, j = e - i * n;
    T v = 0;
    for (int k = 0; k < p; ++k) v += L.E[k * m + i] * L.D[k * n + j];
    L.G[i * ld + j] = rc * v;
  }
  for (int i = tid; i < m; i += nt) L.G[i * ld + n] = L.r[i];
  __syncthreads();
  // H += B'SB ; G += B'T1 ; g += B'Sc
  gemm_lds<T, true, false, 1>(m, m, n, L.B, m, L.SB, m, L.H, m);
  gemm_lds<T, true, false, 1>(m, n, n, L.B, m, L.T1, n, L.G, ld);
  for (int i = tid; i < m; i += nt) {
    T v = 0;
    for (int k = 0; k < n; ++k) v += L.B[k * m + i] * L.Sc[k];
    L.G[i * ld + n] += v;
  }
  __syncthreads();
  chol_lds(L.H, m);
  trsm_lower(L.H, m, L.G, n + 1, ld);  // Y = L^{-1} [G | g]
  __syncthreads();
  // S_t = Q + rc D'D + A'T1 - Y'Y (upper triangle mirrored) ; s_t = q + A'Sc - Y'y
  const T* Q = P.Q.at(b, t);
  for (int e = tid; e < n * n; e += nt) {
    const int i = e / n, j = e - i * n;
    if (i > j) continue;
    T v = Q[e];
    for (int k = 0; k < p; ++k) v += rc * L.D[k * n + i] * L.D[k * n + j];
    for (int k = 0; k < n; ++k) v += L.A[k * n + i] * L.T1[k * n + j];
    for (int r = 0; r < m; ++r) v -= L.G[r * ld + i] * L.G[r * ld + j];
    L.S[i * n + j] = v;
    L.S[j * n + i] = v;
  }
  for (int i = tid; i < n; i += nt) {
    T v = L.q[i];
    for (int k = 0; k < n; ++k) v += L.A[k * n + i] * L.Sc[k];
    for (int r = 0; r < m; ++r) v -= L.G[r * ld + i] * L.G[r * ld + n];
    L.s[i] = v;
  }
  __syncthreads();
  trsm_upper_t(L.H, m, L.G, n + 1, ld);  // [K | k] = L^{-T} Y
  __syncthreads();
}

template <typename T>
__device__ void store_value_and_gain(const Prob<T>& P, const Ws<T>& ws, int b, int t, const RicSmem<T>& L,
                                     bool with_gain) {
  const int n = P.n, m = P.m, N1 = P.N + 1, ld = n + 1;
  store_block(ws.S + ((size_t)b * N1 + t) * n * n, L.S, n * n);
  store_block(ws.s + ((size_t)b * N1 + t) * n, L.s, n);
  if (with_gain) {
    T* K = ws.K + ((size_t)b * P.N + t) * m * n;
    T* k = ws.k + ((size_t)b * P.N + t) * m;
    for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
      const int i = e / n, j = e - i * n;
      K[e] = L.G[i * ld + j];
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) k[i] = L.G[i * ld + n];
  }
}

// Riccati over stages [t_lo, t_hi) (t_hi <= N) starting from the value already in
// (L.S, L.s) at index t_hi; stores S, s for t in [t_lo, t_hi) (and t_hi when
// store_top) and K, k for t < t_hi.
template <typename T>
__device__ void riccati_range(const Prob<T>& P, const Prm<T>& prm, const Ws<T>& ws, const LqrIn<T>& in,
                              int b, int t_lo, int t_hi, RicSmem<T>& L, bool store_top) {
  if (store_top) store_value_and_gain(P, ws, b, t_hi, L, false);
  for (int t = t_hi - 1; t >= t_lo; --t) {
    load_stage(P, ws, in, b, t, L);
    __syncthreads();
    riccati_step(P, prm, b, t, L);
    store_value_and_gain(P, ws, b, t, L, true);
  }
  __syncthreads();
}

// Forward rollout over stages [t_lo, t_hi) from dx_{t_lo} in `x` (LDS, n);
// writes dx, du, dv for t in [t_lo, t_hi) (and t_hi when last) and accumulates
// d_f, d_c, kappa partials into acc.  On return `x` holds dx_{t_hi}.
template <typename T>
__device__ void forward_range(const Prob<T>& P, const Prm<T>& prm, const Iter<T>& it, const Ws<T>& ws,
                              const LqrIn<T>& in, int b, int t_lo, int t_hi, bool last, T* x, T* x2,
                              T* sdu, T* acc) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const int tid = threadIdx.x, nt = blockDim.x;
  const T rc = prm.rho_c;
  const T* Vb = it.V + (size_t)b * N1 * n;
  for (int t = t_lo; t <= t_hi; ++t) {
    const bool term = (t == t_hi);
    if (term && !last) break;
    const T* St = ws.S + ((size_t)b * N1 + t) * n * n;
    const T* st = ws.s + ((size_t)b * N1 + t) * n;
    const T* ct = ws.cdef + ((size_t)b * N1 + t) * n;  // row t pairs with V_t
    const T* qt = ws.qh + ((size_t)b * N1 + t) * n;
    T* dxo = ws.dx + ((size_t)b * N1 + t) * n;
    T* dvo = ws.dv + ((size_t)b * N1 + t) * n;
    // lam = S dx + s ; dV = lam - V ; d_c and the state part of d_f
    for (int i = tid; i < n; i += nt) {
      T lam = st[i];
      for (int j = 0; j < n; ++j) lam += St[i * n + j] * x[j];
      const T v = Vb[(size_t)t * n + i];
      const T dv = lam - v;
      dvo[i] = dv;
      dxo[i] = x[i];
      acc[FP_DC] += (dv - v) * ct[i];
      acc[FP_DF] += qt[i] * x[i];
    }
    if (term) {
      const T* QN = P.QN.at(b, 0);
      for (int i = tid; i < n; i += nt) {
        T v = 0;
        for (int j = 0; j < n; ++j) v += QN[i * n + j] * x[j];
        acc[FP_KAP] += x[i] * v;
      }
      break;
    }
    // du = -K dx - k
    const T* Kt = ws.K + ((size_t)b * N + t) * m * n;
    const T* kt = ws.k + ((size_t)b * N + t) * m;
    T* duo = ws.du + ((size_t)b * N + t) * m;
    for (int i = tid; i < m; i += nt) {
      T v = kt[i];
      for (int j = 0; j < n; ++j) v += Kt[i * n + j] * x[j];
      sdu[i] = -v;
      duo[i] = -v;
    }
    __syncthreads();
    // dx' = A dx + B du + c_t ; curvature and control part of d_f
    const T* A = in.A.at(b, t);
    const T* Bm = in.Bm.at(b, t);
    const T* cn = ws.cdef + ((size_t)b * N1 + t + 1) * n;
    const T* Q = P.Q.at(b, t);
    for (int i = tid; i < n; i += nt) {
      T v = cn[i], qx = 0;
      for (int j = 0; j < n; ++j) { v += A[i * n + j] * x[j]; qx += Q[i * n + j] * x[j]; }
      for (int j = 0; j < m; ++j) v += Bm[i * m + j] * sdu[j];
      x2[i] = v;
      acc[FP_KAP] += x[i] * qx;
    }
    {
      const T* R = P.R.at(b, t);
      const T* rt = ws.rh + ((size_t)b * N + t) * m;
      const T* rdt = ws.rd + ((size_t)b * N + t) * m;
      for (int i = tid; i < m; i += nt) {
        T ru = 0;
        for (int j = 0; j < m; ++j) ru += R[i * m + j] * sdu[j];
        acc[FP_KAP] += sdu[i] * ru + rdt[i] * sdu[i] * sdu[i];
        acc[FP_DF] += rt[i] * sdu[i];
      }
      if (p > 0) {
        const T* D = P.D.at(b, t);
        const T* E = P.E.at(b, t);
        for (int k = tid; k < p; k += nt) {
          T hd = 0;
          for (int j = 0; j < n; ++j) hd += D[k * n + j] * x[j];
          for (int j = 0; j < m; ++j) hd += E[k * m + j] * sdu[j];
          acc[FP_KAP] += rc * hd * hd;
        }
      }
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) x[i] = x2[i];
    __syncthreads();
  }
}

// segments == 1: sequential Riccati + forward pass, one block per instance.
template <typename T>
__global__ void k_lqr_seq(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, LqrIn<T> in) {
  const int b = blockIdx.x;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const int tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* base = reinterpret_cast<T*>(smem_raw);
  RicSmem<T> L = ric_carve(base, n, m, p, nt);
  load_block(L.S, P.QN.at(b, 0), n, n, n);
  for (int i = tid; i < n; i += nt) L.s[i] = ws.qh[((size_t)b * N1 + N) * n + i];
  __syncthreads();
  riccati_range(P, prm, ws, in, b, 0, N, L, true);
  // forward pass (reuse LDS: x in L.c, x2 in L.q, du in L.r)
  T acc[NFP] = {T(0), T(0), T(0), T(0)};
  T* x = L.Sc;
  T* x2 = L.c;
  T* sdu = L.r;
  for (int i = tid; i < n; i += nt) x[i] = ws.cdef[(size_t)b * N1 * n + i];
  __syncthreads();
  forward_range(P, prm, it, ws, in, b, 0, N, true, x, x2, sdu, acc);
  block_sum<T, NFP>(acc, 3, L.red);
  if (tid == 0) {
    T* fp = ws.fpart + (size_t)b * prm.P * NFP;
    fp[FP_DF] = acc[FP_DF];
    fp[FP_DC] = acc[FP_DC];
    fp[FP_KAP] = acc[FP_KAP];
    fp[3] = 0;
  }
}

}  // namespace mpxk
