// Chunked parallel-in-time LQR (segments P > 1), after Saerkkae &
// Garcia-Fernandez's associative scan of conditional value functions.
//
// Element e = (A, b, C, eta, J) is the conditional cost
//   V_e(x, y) = x'Jx/2 + eta'x + max_l l'(y - A x - b) - l'C l/2.
// Backward (elements 0..N, element N = terminal cost):
//   up    (grid P x B): segment s < P-1 reduces its leaves to one aggregate with
//         leaf combines e_t (x) E (rank-m Woodbury form, H = R + B'J B);
//         segment P-1 runs Riccati from the terminal cost (its value at a_{P-1}).
//   mid   (grid B): suffix values v_s = E_s (x) v_{s+1}, s = P-2..0, solving
//         (I + S C) Y = [S A | s + S b] by Gauss-Jordan with partial pivoting.
//   down  (grid (P-1) x B): segment s runs Riccati from v_{s+1}.
// Forward (stages 0..N-1 split the same way):
//   up    segment 0 rolls out from dx_0 = c_init; segments 1..P-2 compose their
//         affine maps dx_b = Phi dx_a + phi (F_t = A - B K, f_t = c - B k);
//   mid   states at segment starts; down: segments 1..P-1 roll out.
#pragma once
#include "common.h"
#include "k_lqr.h"

namespace mpxk {

__host__ __device__ inline void seg_bounds(int n_items, int P, int s, int* a, int* b) {
  const int base = n_items / P, rem = n_items % P;
  *a = s * base + (s < rem ? s : rem);
  *b = *a + base + (s < rem ? 1 : 0);
}

template <typename T>
__host__ __device__ inline size_t agg_elems(int n) { return (size_t)3 * n * n + 2 * (size_t)n; }

// LDS for the leaf-combine up-sweep
template <typename T>
__host__ __device__ inline size_t up_smem_elems(int n, int m, int p, int nt) {
  return (size_t)6 * n * n + 3 * (size_t)n * m + 2 * (size_t)m * (n + 1) + 2 * (size_t)m * m +
         (size_t)p * (n + m) + 8 * (size_t)n + 2 * (size_t)m + 4 * (size_t)nt + 8;
}

template <typename T>
__device__ void leaf_combine_up(const Prob<T>& P, const Prm<T>& prm, const Ws<T>& ws, const LqrIn<T>& in, int b,
                                int t_lo, int t_hi, T* base, T* out) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ld = n + 1;
  const T rc = prm.rho_c;
  T* q = base;
  T* gA = q; q += n * n;   // aggregate A
  T* gC = q; q += n * n;   // aggregate C
  T* gJ = q; q += n * n;   // aggregate J
  T* sA = q; q += n * n;   // stage A, becomes A~
  T* T1 = q; q += n * n;
  T* T2 = q; q += n * n;
  T* sB = q; q += n * m;
  T* JB = q; q += n * m;
  T* P2 = q; q += n * m;
  T* Z = q; q += m * (n + 1);  // R^{-1}[M' | r]  then reused
  T* G = q; q += m * (n + 1);
  T* LR = q; q += m * m;
  T* H = q; q += m * m;
  T* sD = q; q += p * n;
  T* sE = q; q += p * m;
  T* gb = q; q += n;
  T* ge = q; q += n;
  T* sc = q; q += n;
  T* sq = q; q += n;
  T* bl = q; q += n;   // b of the leaf
  T* el = q; q += n;   // eta of the leaf
  T* zv = q; q += n;
  T* v1 = q; q += n;
  T* sr = q; q += m;
  T* srd = q; q += m;
  for (int t = t_hi - 1; t >= t_lo; --t) {
    const bool first = (t == t_hi - 1);
    // stage data
    load_block(sA, in.A.at(b, t), n, n, n);
    load_block(sB, in.Bm.at(b, t), n, m, m);
    if (p > 0) {
      load_block(sD, P.D.at(b, t), p, n, n);
      load_block(sE, P.E.at(b, t), p, m, m);
    }
    {
      const T* c = ws.cdef + ((size_t)b * N1 + t + 1) * n;
      const T* qq = ws.qh + ((size_t)b * N1 + t) * n;
      const T* r = ws.rh + ((size_t)b * N + t) * m;
      const T* rd = ws.rd + ((size_t)b * N + t) * m;
      for (int i = tid; i < n; i += nt) { sc[i] = c[i]; sq[i] = qq[i]; }
      for (int i = tid; i < m; i += nt) { sr[i] = r[i]; srd[i] = rd[i]; }
    }
    __syncthreads();
    // R^ = R + diag(rd) + rc E'E  -> LR (factored), Z = [rc E'D | r]
    const T* R = P.R.at(b, t);
    for (int e = tid; e < m * m; e += nt) {
      const int i = e / m, j = e - i * m;
      T v = R[e] + (i == j ? srd[i] : T(0));
      for (int k = 0; k < p; ++k) v += rc * sE[k * m + i] * sE[k * m + j];
      LR[e] = v;
      H[e] = v;
    }
    for (int e = tid; e < m * n; e += nt) {
      const int i = e / n, j = e - i * n;
      T v = 0;
      for (int k = 0; k < p; ++k) v += sE[k * m + i] * sD[k * n + j];
      Z[i * ld + j] = rc * v;
    }
    for (int i = tid; i < m; i += nt) Z[i * ld + n] = sr[i];
    __syncthreads();
    chol_lds(LR, m, prm.hfloor);
    trsm_lower(LR, m, Z, n + 1, ld);
    __syncthreads();
    trsm_upper_t(LR, m, Z, n + 1, ld);
    __syncthreads();
    // leaf: A~ = A - B Z_M ; b_l = c - B z_r ; eta_l = q - M^ z_r (M^ = rc D'E)
    for (int e = tid; e < n * n; e += nt) {
      const int i = e / n, j = e - i * n;
      T v = 0;
      for (int k = 0; k < m; ++k) v += sB[i * m + k] * Z[k * ld + j];
      T1[e] = sA[e] - v;
    }
    for (int i = tid; i < n; i += nt) {
      T v = sc[i];
      for (int k = 0; k < m; ++k) v -= sB[i * m + k] * Z[k * ld + n];
      bl[i] = v;
      T w = sq[i];
      for (int r = 0; r < p; ++r) {
        T ez = 0;
        for (int k = 0; k < m; ++k) ez += sE[r * m + k] * Z[k * ld + n];
        w -= rc * sD[r * n + i] * ez;
      }
      el[i] = w;
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += nt) sA[e] = T1[e];
    __syncthreads();
    if (first) {
      // aggregate := leaf: A~, b_l, C = B R^{-1} B', eta_l, J_l = Q^ - M^ R^{-1} M^'
      const T* Qt = P.Q.at(b, t);
      // Y = L_R^{-1} B' (m x n) in G
      for (int e = tid; e < m * n; e += nt) {
        const int i = e / n, j = e - i * n;
        G[i * ld + j] = sB[j * m + i];
      }
      __syncthreads();
      trsm_lower(LR, m, G, n, ld);
      __syncthreads();
      for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e - i * n;
        if (i > j) continue;
        T c = 0;
        for (int k = 0; k < m; ++k) c += G[k * ld + i] * G[k * ld + j];
        gC[i * n + j] = c; gC[j * n + i] = c;
        T jv = Qt[e];
        for (int r = 0; r < p; ++r) {
          T ez = 0;  // (E Z_M)[r][j]
          for (int k = 0; k < m; ++k) ez += sE[r * m + k] * Z[k * ld + j];
          jv += rc * sD[r * n + i] * (sD[r * n + j] - ez);
        }
        gJ[i * n + j] = jv;
      }
      // symmetrise J (the r-sum above is D'(D - E Z_M), symmetric up to rounding)
      __syncthreads();
      for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e - i * n;
        if (i > j) gJ[e] = gJ[j * n + i];
      }
      for (int e = tid; e < n * n; e += nt) gA[e] = sA[e];
      for (int i = tid; i < n; i += nt) { gb[i] = bl[i]; ge[i] = el[i]; }
      __syncthreads();
      continue;
    }
    // combine e_t (x) (gA, gb, gC, ge, gJ)
    // JB = J2 B ; H = R^ + B'JB ; z = eta2 + J2 b_l
    gemm_lds<T, false, false, 0>(n, m, n, gJ, n, sB, m, JB, m);
    for (int i = tid; i < n; i += nt) {
      T v = ge[i];
      for (int k = 0; k < n; ++k) v += gJ[i * n + k] * bl[k];
      zv[i] = v;
    }
    __syncthreads();
    gemm_lds<T, true, false, 1>(m, m, n, sB, m, JB, m, H, m);
    // G = [JB' A~ | B'z]
    gemm_lds<T, true, false, 0>(m, n, n, JB, m, sA, n, G, ld);
    for (int i = tid; i < m; i += nt) {
      T v = 0;
      for (int k = 0; k < n; ++k) v += sB[k * m + i] * zv[k];
      G[i * ld + n] = v;
    }
    __syncthreads();
    chol_lds(H, m, prm.hfloor);
    // P2 = A2 B ; Y2 = L_H^{-1} P2' (into Z, m x n) ; Y3 = L_H^{-1} JB' (into LR area? use T2 temporarily)
    gemm_lds<T, false, false, 0>(n, m, n, gA, n, sB, m, P2, m);
    __syncthreads();
    trsm_lower(H, m, G, n + 1, ld);
    __syncthreads();
    trsm_upper_t(H, m, G, n + 1, ld);  // G = H^{-1}[JB'A~ | B'z]
    for (int e = tid; e < m * n; e += nt) {
      const int i = e / n, j = e - i * n;
      Z[i * ld + j] = P2[j * m + i];
    }
    __syncthreads();
    trsm_lower(H, m, Z, n, ld);  // Y2
    // XA~ = A~ - B G_A -> T1 ; v = b_l - B g
    for (int e = tid; e < n * n; e += nt) {
      const int i = e / n, j = e - i * n;
      T v = sA[e];
      for (int k = 0; k < m; ++k) v -= sB[i * m + k] * G[k * ld + j];
      T1[e] = v;
    }
    for (int i = tid; i < n; i += nt) {
      T v = bl[i];
      for (int k = 0; k < m; ++k) v -= sB[i * m + k] * G[k * ld + n];
      v1[i] = v;
    }
    __syncthreads();
    // A_new = A2 XA~ -> T2 ; b_new = A2 v + b2 (into sc as temp) ; C_new = C2 + Y2'Y2
    gemm_lds<T, false, false, 0>(n, n, n, gA, n, T1, n, T2, n);
    for (int i = tid; i < n; i += nt) {
      T v = gb[i];
      for (int k = 0; k < n; ++k) v += gA[i * n + k] * v1[k];
      sc[i] = v;
    }
    for (int e = tid; e < n * n; e += nt) {
      const int i = e / n, j = e - i * n;
      if (i > j) continue;
      T c = gC[e];
      for (int k = 0; k < m; ++k) c += Z[k * ld + i] * Z[k * ld + j];
      gC[i * n + j] = c; gC[j * n + i] = c;
    }
    // z <- z - JB g_z (each thread its own rows), for eta_new = eta_l + A~' z
    for (int i = tid; i < n; i += nt) {
      T v = zv[i];
      for (int k = 0; k < m; ++k) v -= JB[i * m + k] * G[k * ld + n];
      zv[i] = v;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
      T v = el[i];
      for (int k = 0; k < n; ++k) v += sA[k * n + i] * zv[k];
      ge[i] = v;
      gb[i] = sc[i];
    }
    // J~ = J2 - Y3'Y3 with Y3 = L_H^{-1} JB'  (Y3 into Z again)
    for (int e = tid; e < n * n; e += nt) gA[e] = T2[e];
    for (int e = tid; e < m * n; e += nt) {
      const int i = e / n, j = e - i * n;
      Z[i * ld + j] = JB[j * m + i];
    }
    __syncthreads();
    trsm_lower(H, m, Z, n, ld);
    __syncthreads();
    for (int e = tid; e < n * n; e += nt) {
      const int i = e / n, j = e - i * n;
      if (i > j) continue;
      T v = gJ[e];
      for (int k = 0; k < m; ++k) v -= Z[k * ld + i] * Z[k * ld + j];
      T1[i * n + j] = v; T1[j * n + i] = v;
    }
    __syncthreads();
    gemm_lds<T, false, false, 0>(n, n, n, T1, n, sA, n, T2, n);  // J~ A~
    __syncthreads();
    // J_new = J_l + A~' (J~ A~), J_l = Q + rc D'(D - E Z_M) with Z_M recomputed is
    // avoided: J_l = Q^ - M^ R^{-1} M^' = Q + rc D'D - rc D'E R^{-1} rc E'D.
    {
      const T* Qt = P.Q.at(b, t);
      // W = L_R^{-1} (rc E'D)  (m x n) into G
      for (int e = tid; e < m * n; e += nt) {
        const int i = e / n, j = e - i * n;
        T v = 0;
        for (int k = 0; k < p; ++k) v += sE[k * m + i] * sD[k * n + j];
        G[i * ld + j] = rc * v;
      }
      __syncthreads();
      trsm_lower(LR, m, G, n, ld);
      __syncthreads();
      for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e - i * n;
        if (i > j) continue;
        T v = Qt[e];
        for (int k = 0; k < p; ++k) v += rc * sD[k * n + i] * sD[k * n + j];
        for (int k = 0; k < m && p > 0; ++k) v -= G[k * ld + i] * G[k * ld + j];
        for (int k = 0; k < n; ++k) v += sA[k * n + i] * T2[k * n + j];
        gJ[i * n + j] = v; gJ[j * n + i] = v;
      }
    }
    __syncthreads();
  }
  // write the aggregate (A, C, J, b, eta)
  const size_t nn = (size_t)n * n;
  store_block(out, gA, n * n);
  store_block(out + nn, gC, n * n);
  store_block(out + 2 * nn, gJ, n * n);
  store_block(out + 3 * nn, gb, n);
  store_block(out + 3 * nn + n, ge, n);
}

template <typename T, int MMAX>
__global__ void __launch_bounds__(512) k_scan_up(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, LqrIn<T> in) {
  const int s = blockIdx.x, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* base = reinterpret_cast<T*>(smem_raw);
  int a, e;
  seg_bounds(N1, prm.P, s, &a, &e);
  if (s == prm.P - 1) {
    RicSmem<T> L = ric_carve(base, n, m, p, blockDim.x);
    load_block(L.S, P.QN.at(b, 0), n, n, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) L.s[i] = ws.qh[((size_t)b * N1 + N) * n + i];
    __syncthreads();
    riccati_range<T, MMAX>(P, prm, ws, in, b, a, N, L, true);
    return;
  }
  leaf_combine_up(P, prm, ws, in, b, a, e, base, ws.agg + ((size_t)b * prm.P + s) * agg_elems<T>(n));
}

template <typename T>
__host__ __device__ inline size_t mid_smem_elems(int n, int nt) {
  return (size_t)5 * n * n + (size_t)n * (n + 1) + 6 * (size_t)n + 2 * (size_t)nt + 8;
}

// Suffix values at segment starts: v_s = E_s (x) v_{s+1}.
template <typename T>
__global__ void __launch_bounds__(512) k_scan_mid(Prob<T> P, Prm<T> prm, Ws<T> ws) {
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int n = P.n, N = P.N, N1 = N + 1, Pn = prm.P;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ld = n + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* q = reinterpret_cast<T*>(smem_raw);
  T* S = q; q += n * n;
  T* A = q; q += n * n;
  T* C = q; q += n * n;
  T* J = q; q += n * n;
  T* M = q; q += n * n;
  T* Y = q; q += n * ld;
  T* sv = q; q += n;
  T* bb = q; q += n;
  T* eta = q; q += n;
  T* pv = q; q += 2 * n;
  int* piv = reinterpret_cast<int*>(q);
  int aL, eL;
  seg_bounds(N1, Pn, Pn - 1, &aL, &eL);
  load_block(S, ws.S + ((size_t)b * N1 + aL) * n * n, n, n, n);
  for (int i = tid; i < n; i += nt) sv[i] = ws.s[((size_t)b * N1 + aL) * n + i];
  __syncthreads();
  const size_t nn = (size_t)n * n;
  for (int s = Pn - 2; s >= 0; --s) {
    const T* E = ws.agg + ((size_t)b * Pn + s) * agg_elems<T>(n);
    load_block(A, E, n, n, n);
    load_block(C, E + nn, n, n, n);
    load_block(J, E + 2 * nn, n, n, n);
    for (int i = tid; i < n; i += nt) { bb[i] = E[3 * nn + i]; eta[i] = E[3 * nn + n + i]; }
    __syncthreads();
    // M = I + S C ; Y = [S A | s + S b]
    gemm_lds<T, false, false, 0>(n, n, n, S, n, C, n, M, n);
    gemm_lds<T, false, false, 0>(n, n, n, S, n, A, n, Y, ld);
    for (int i = tid; i < n; i += nt) {
      T v = sv[i];
      for (int k = 0; k < n; ++k) v += S[i * n + k] * bb[k];
      Y[i * ld + n] = v;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) M[i * n + i] += T(1);
    __syncthreads();
    // Gauss-Jordan with partial pivoting: Y <- M^{-1} Y
    for (int k = 0; k < n; ++k) {
      if (tid == 0) {
        int best = k;
        T bv = M[k * n + k] < 0 ? -M[k * n + k] : M[k * n + k];
        for (int i = k + 1; i < n; ++i) {
          const T v = M[i * n + k] < 0 ? -M[i * n + k] : M[i * n + k];
          if (v > bv) { bv = v; best = i; }
        }
        piv[0] = best;
      }
      __syncthreads();
      const int pr = piv[0];
      if (pr != k) {
        for (int j = tid; j < n; j += nt) { T t0 = M[k * n + j]; M[k * n + j] = M[pr * n + j]; M[pr * n + j] = t0; }
        for (int j = tid; j < ld; j += nt) { T t0 = Y[k * ld + j]; Y[k * ld + j] = Y[pr * ld + j]; Y[pr * ld + j] = t0; }
      }
      __syncthreads();
      const T d = M[k * n + k];
      for (int j = tid; j < n; j += nt) if (j > k) M[k * n + j] /= d;
      for (int j = tid; j < ld; j += nt) Y[k * ld + j] /= d;
      __syncthreads();
      for (int i = tid; i < n; i += nt) pv[i] = M[i * n + k];
      __syncthreads();
      for (int e = tid; e < n * (n + ld); e += nt) {
        const int i = e / (n + ld), j = e - i * (n + ld);
        if (i == k) continue;
        const T f = pv[i];
        if (j < n) { if (j > k) M[i * n + j] -= f * M[k * n + j]; }
        else Y[i * ld + (j - n)] -= f * Y[k * ld + (j - n)];
      }
      __syncthreads();
    }
    // S_new = J + A' Y_A (symmetric) ; s_new = eta + A' y
    for (int e = tid; e < n * n; e += nt) {
      const int i = e / n, j = e - i * n;
      if (i > j) continue;
      T v = J[e];
      for (int k = 0; k < n; ++k) v += A[k * n + i] * Y[k * ld + j];
      M[i * n + j] = v; M[j * n + i] = v;
    }
    for (int i = tid; i < n; i += nt) {
      T v = eta[i];
      for (int k = 0; k < n; ++k) v += A[k * n + i] * Y[k * ld + n];
      pv[n + i] = v;
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += nt) S[e] = M[e];
    for (int i = tid; i < n; i += nt) sv[i] = pv[n + i];
    __syncthreads();
    T* o = ws.sval + ((size_t)b * Pn + s) * (nn + n);
    store_block(o, S, n * n);
    store_block(o + nn, sv, n);
    __syncthreads();
  }
}

template <typename T, int MMAX>
__global__ void __launch_bounds__(512) k_scan_down(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, LqrIn<T> in) {
  const int s = blockIdx.x, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1, Pn = prm.P;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* base = reinterpret_cast<T*>(smem_raw);
  RicSmem<T> L = ric_carve(base, n, m, p, blockDim.x);
  int a, e;
  seg_bounds(N1, Pn, s, &a, &e);
  const size_t nn = (size_t)n * n;
  if (s + 1 <= Pn - 2) {
    const T* v = ws.sval + ((size_t)b * Pn + s + 1) * (nn + n);
    load_block(L.S, v, n, n, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) L.s[i] = v[nn + i];
  } else {
    load_block(L.S, ws.S + ((size_t)b * N1 + e) * nn, n, n, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) L.s[i] = ws.s[((size_t)b * N1 + e) * n + i];
  }
  __syncthreads();
  riccati_range<T, MMAX>(P, prm, ws, in, b, a, e, L, false);
}

// forward: segment 0 runs the dx chain; segments 1..P-2 compose their maps
// Phi = F_{e-1} ... F_a, phi (dx_e = Phi dx_a + phi) from the stored F_t, f_t.
template <typename T>
__host__ __device__ inline size_t fwd_smem_elems(int n, int m, int nt) {
  return (size_t)3 * n * n + 4 * (size_t)n + 8;
}

template <typename T>
__global__ void __launch_bounds__(512) k_fwd_up(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, LqrIn<T> in) {
  const int s = blockIdx.x, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int n = P.n, N = P.N, N1 = N + 1, Pn = prm.P;
  const int tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* q = reinterpret_cast<T*>(smem_raw);
  T* Phi = q; q += n * n;
  T* F = q; q += n * n;
  T* Tm = q; q += n * n;
  T* x = q; q += n;
  T* x2 = q; q += n;
  T* phi = q; q += n;
  T* w = q; q += n;
  int a, e;
  seg_bounds(N, Pn, s, &a, &e);
  if (s == 0) {
    for (int i = tid; i < n; i += nt) x[i] = ws.cdef[(size_t)b * N1 * n + i];
    __syncthreads();
    fwd_chain(P, ws, b, 0, e, Pn == 1, x, F, Tm, x2);
    for (int i = tid; i < n; i += nt) ws.xs[((size_t)b * Pn + 1) * n + i] = x[i];
    return;
  }
  for (int e2 = tid; e2 < n * n; e2 += nt) Phi[e2] = (e2 / n == e2 % n) ? T(1) : T(0);
  for (int i = tid; i < n; i += nt) phi[i] = 0;
  __syncthreads();
  for (int t = a; t < e; ++t) {
    load_block(F, ws.F + ((size_t)b * N + t) * n * n, n, n, n);
    const T* ft = ws.f + ((size_t)b * N + t) * n;
    __syncthreads();
    gemm_lds<T, false, false, 0>(n, n, n, F, n, Phi, n, Tm, n);
    for (int i = tid; i < n; i += nt) {
      T v = ft[i];
      for (int k = 0; k < n; ++k) v += F[i * n + k] * phi[k];
      w[i] = v;
    }
    __syncthreads();
    for (int e2 = tid; e2 < n * n; e2 += nt) Phi[e2] = Tm[e2];
    for (int i = tid; i < n; i += nt) phi[i] = w[i];
    __syncthreads();
  }
  T* o = ws.fmap + ((size_t)b * Pn + s) * ((size_t)n * n + n);
  store_block(o, Phi, n * n);
  store_block(o + (size_t)n * n, phi, n);
}

template <typename T>
__global__ void k_fwd_mid(Prob<T> P, Prm<T> prm, Ws<T> ws) {
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int n = P.n, Pn = prm.P;
  const int tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* x = reinterpret_cast<T*>(smem_raw);
  T* x2 = x + n;
  for (int i = tid; i < n; i += nt) x[i] = ws.xs[((size_t)b * Pn + 1) * n + i];
  __syncthreads();
  for (int s = 1; s <= Pn - 2; ++s) {
    const T* M = ws.fmap + ((size_t)b * Pn + s) * ((size_t)n * n + n);
    for (int i = tid; i < n; i += nt) {
      T v = M[(size_t)n * n + i];
      for (int k = 0; k < n; ++k) v += M[i * n + k] * x[k];
      x2[i] = v;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
      x[i] = x2[i];
      ws.xs[((size_t)b * Pn + s + 1) * n + i] = x2[i];
    }
    __syncthreads();
  }
}

// segments 1..P-1 run their dx chains from the propagated segment starts
template <typename T>
__global__ void __launch_bounds__(512) k_fwd_down(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, LqrIn<T> in) {
  const int s = blockIdx.x + 1, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int n = P.n, N = P.N, Pn = prm.P;
  const int tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* q = reinterpret_cast<T*>(smem_raw);
  T* F0 = q; q += n * n;
  T* F1 = q; q += n * n;
  T* x = q; q += n;
  T* x2 = q; q += n;
  int a, e;
  seg_bounds(N, Pn, s, &a, &e);
  for (int i = tid; i < n; i += nt) x[i] = ws.xs[((size_t)b * Pn + s) * n + i];
  __syncthreads();
  fwd_chain(P, ws, b, a, e, s == Pn - 1, x, F0, F1, x2);
}

}  // namespace mpxk
