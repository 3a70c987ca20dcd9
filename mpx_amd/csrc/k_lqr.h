// K2/K3: the LQR subproblem of one SQP iteration.
//
// Stage data (per instance b, stage t < N): A_t, B_t (views), Q_t, R_t, D_t, E_t
// (problem), augmented gradients q_t, r_t, box diagonal rd_t and the defect
// c_t = cdef[t+1]; the QP Hessian is [[Q + rho_c D'D, rho_c D'E],
// [rho_c E'D, R + diag(rd) + rho_c E'E]] (the AL rows never materialise).
//
// Backward: value V_t(x) = x'S_t x/2 + s_t'x, by sequential Riccati steps
// (segments = 1) or by the chunked associative scan (segments > 1, k_scan.h).
// Each step also stores the closed-loop map F_t = A_t - B_t K_t and
// f_t = c_t - B_t k_t, so the forward pass is the affine recursion
// dx_{t+1} = F_t dx_t + f_t (sequential, or chunked by segment maps), and the
// costates lam_t = S_t dx_t + s_t, du_t = -K_t dx_t - k_t, dV_t = lam_t - V_t
// and the merit slope partials follow in one pass parallel over all stages.
#pragma once
#include "common.h"

namespace mpxk {

template <typename T>
struct LqrIn {
  View<T> A, Bm;
};

// LDS carve-up of the Riccati working set.  A and B of the stage sit side by
// side in AB (n x (n+m)); the product S [A | B] lands in TB = [T1 | SB].
template <typename T>
struct RicSmem {
  T *S, *A, *T1, *B, *SB, *G, *Kb, *H, *D, *E, *s, *Sc, *c, *rd, *q, *r, *piv, *red;
  T* Mt;  // M_t' (m x n) of a caller-given cross Hessian, or nullptr
  int ldab;
};

// hasM: room for the transposed cross Hessian M_t' (m x n) after the rest.
template <typename T>
__host__ __device__ inline size_t ric_smem_elems(int n, int m, int p, int nt, bool hasM = false) {
  const size_t kb = (size_t)m * (n + 1 > m ? n + 1 : m);  // [K | k], or m x m scratch
  return (size_t)n * n + 2 * (size_t)n * (n + m) + (size_t)m * (n + 1) + kb + (size_t)m * m +
         (size_t)p * (n + m) + 6 * (size_t)n + 2 * (size_t)m + 128 + 4 * (size_t)nt + 8 +
         (hasM ? (size_t)m * n : 0);
}

template <typename T>
__device__ RicSmem<T> ric_carve(T* base, int n, int m, int p, int nt, bool hasM = false) {
  RicSmem<T> L;
  L.Mt = hasM ? base + ric_smem_elems<T>(n, m, p, nt, false) : nullptr;
  T* q = base;
  L.ldab = n + m;
  L.S = q; q += n * n;
  L.A = q; L.B = q + n; q += n * (n + m);
  L.T1 = q; L.SB = q + n; q += n * (n + m);
  L.G = q; q += m * (n + 1);
  L.Kb = q; q += m * (n + 1 > m ? n + 1 : m);
  L.H = q; q += m * m;
  L.D = q; q += p * n;
  L.E = q; q += p * m;
  L.s = q; q += n;
  L.Sc = q; q += n;
  L.c = q; q += n;
  L.q = q; q += n;
  L.rd = q; q += m;
  L.r = q; q += m;
  q += 2 * n;
  L.piv = q; q += 128;
  L.red = q;
  return L;
}

// e / ld for 0 <= e < 2^24: the float estimate is off by at most one
__device__ __forceinline__ int div_small(int e, int ld, float inv) {
  int i = (int)((float)e * inv);
  if (i * ld > e) --i;
  else if ((i + 1) * ld <= e) ++i;
  return i;
}

// Load the stage's QP data into LDS: [A | B], sqrt(rho_c) D, sqrt(rho_c) E, c, q, r, rd.
// The small parts (D .. rd, four entries per thread) are requested first and
// [A | B] twenty entries per thread at a time, every request of a round
// issued before its stores: at n = 76, m = 32 with 512 threads a stage costs
// one memory latency.
template <typename T>
__device__ __forceinline__ T stage_small(const T* D, const T* E, const T* c, const T* q, const T* r, const T* rd,
                                         T sr, int e, int o2, int o3, int o4, int o5, int o6, int tot) {
  if (e < o2) return sr * D[e];
  if (e < o3) return sr * E[e - o2];
  if (e < o4) return c[e - o3];
  if (e < o5) return q[e - o4];
  if (e < o6) return r[e - o5];
  if (e < tot) return rd[e - o6];
  return T(0);
}
template <typename T>
__device__ __forceinline__ void stage_small_store(RicSmem<T>& L, int e, T v, int o2, int o3, int o4, int o5, int o6,
                                                  int tot) {
  if (e < o2) L.D[e] = v;
  else if (e < o3) L.E[e - o2] = v;
  else if (e < o4) L.c[e - o3] = v;
  else if (e < o5) L.q[e - o4] = v;
  else if (e < o6) L.r[e - o5] = v;
  else if (e < tot) L.rd[e - o6] = v;
}

template <typename T>
__device__ void load_stage(const Prob<T>& P, const Prm<T>& prm, const Ws<T>& ws, const LqrIn<T>& in, int b,
                           int t, RicSmem<T>& L) {
  const int n = P.n, m = P.m, p = P.p, N1 = P.N + 1, ldab = L.ldab;
  const int tid = threadIdx.x, nt = blockDim.x;
  const T* A = in.A.at(b, t);
  const T* B = in.Bm.at(b, t);
  const T* D = p > 0 ? P.D.at(b, t) : nullptr;
  const T* E = p > 0 ? P.E.at(b, t) : nullptr;
  const T sr = dev_sqrt(prm.rho_c);
  const T* c = ws.cdef + ((size_t)b * N1 + t + 1) * n;
  const T* q = ws.qh + ((size_t)b * N1 + t) * n;
  const T* r = ws.rh + ((size_t)b * P.N + t) * m;
  const T* rd = ws.rd + ((size_t)b * P.N + t) * m;
  const int o2 = p * n, o3 = o2 + p * m, o4 = o3 + n, o5 = o4 + n, o6 = o5 + m, tot = o6 + m;
  T w[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) w[u] = stage_small(D, E, c, q, r, rd, sr, tid + u * nt, o2, o3, o4, o5, o6, tot);
  const int o1 = n * ldab;
  const float inv = 1.f / (float)ldab;
  for (int base = tid; base < o1; base += 20 * nt) {
    T v[20];
#pragma unroll
    for (int u = 0; u < 20; ++u) {
      const int e = base + u * nt;
      T x = 0;
      if (e < o1) {
        const int i = div_small(e, ldab, inv), j = e - i * ldab;
        x = j < n ? A[i * n + j] : B[i * m + (j - n)];
      }
      v[u] = x;
    }
#pragma unroll
    for (int u = 0; u < 20; ++u) {
      const int e = base + u * nt;
      if (e < o1) L.A[e] = v[u];
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) stage_small_store(L, tid + u * nt, w[u], o2, o3, o4, o5, o6, tot);
  for (int e = tid + 4 * nt; e < tot; e += nt)
    stage_small_store(L, e, stage_small(D, E, c, q, r, rd, sr, e, o2, o3, o4, o5, o6, tot), o2, o3, o4, o5, o6, tot);
  if (L.Mt) {
    // M_t (n x m, row-major) lands transposed: Mt[r][i] = M[i][r]
    const T* M = P.M.at(b, t);
    for (int e = tid; e < n * m; e += nt) {
      const int i = e / m, r = e - i * m;
      L.Mt[r * n + i] = M[e];
    }
  }
}

template <typename T>
struct IsF32 { static constexpr bool value = false; };
template <>
struct IsF32<float> { static constexpr bool value = true; };

// [G | g] <- W = Z [G | g] and Kb <- [K | k] = Z' W for H = Z^{-1} Z^{-T}
// (Z the inverse of H's lower Cholesky factor, pivots floored): H^{-1}[G | g]
// through the factor, never forming H^{-1} (whose FP32 rounding error is a
// factor cond(H)^(1/2) larger), and G'H^{-1}G = W'W.  16 < m <= 32: wave 0
// factors H = U'U and forms Y = U^{-1} = Z' in registers
// (chol_upper_inv_wave), then two products; other m: Cholesky and the two
// triangular solves block-wide in LDS (the register version spills at those
// template sizes inside the Riccati kernels).
// (during_chol: run by the waves other than wave 0 while wave 0 factors H,
// m <= 32; returns whether it did its work, else it is left to the caller)
template <typename T, int MMAX, typename F>
__device__ bool spd_apply_inverse(T* H, T* G, T* Kb, int m, int ncol, int ld, T floor, bool mfma, F&& during_chol) {
  bool done = false;
  if constexpr (MMAX == 32) {
    chol_upper_inv_wave<T, MMAX>(H, H, m, floor);   // Y = U^{-1} (upper) in H
    done = during_chol();
    lds_sync();
    if constexpr (IsF32<T>::value) {
      if (mfma) {
        const MSeg w0{H, G, m, ld, m, true, false, 1.f};
        mfma_gemm<false>(m, ncol, 1, w0, w0, w0, Kb, ld, nullptr, 0, nullptr);
      }
    }
    if (!mfma) gemm_lds<T, true, false, 0>(m, ncol, m, H, m, G, ld, Kb, ld);   // W = Y' G
    lds_sync();
    for (int e = threadIdx.x; e < m * ncol; e += blockDim.x) {
      const int i = e / ncol, j = e - i * ncol;
      G[i * ld + j] = Kb[i * ld + j];
    }
    lds_sync();
    if constexpr (IsF32<T>::value) {
      if (mfma) {
        const MSeg k0{H, G, m, ld, m, false, false, 1.f};
        mfma_gemm<false>(m, ncol, 1, k0, k0, k0, Kb, ld, nullptr, 0, nullptr);
      }
    }
    if (!mfma) gemm_lds<T, false, false, 0>(m, ncol, m, H, m, G, ld, Kb, ld);  // K = Y W
  } else {
    chol_lds(H, m, floor);                           // H = L L', L = Z^{-1}
    for (int e = threadIdx.x; e < m * ncol; e += blockDim.x) {
      const int i = e / ncol, j = e - i * ncol;
      Kb[i * ld + j] = G[i * ld + j];
    }
    lds_sync();
    trsm_lower(H, m, Kb, ncol, ld);                  // W = L^{-1} G
    lds_sync();
    for (int e = threadIdx.x; e < m * ncol; e += blockDim.x) {
      const int i = e / ncol, j = e - i * ncol;
      G[i * ld + j] = Kb[i * ld + j];
    }
    lds_sync();
    trsm_upper_t(H, m, Kb, ncol, ld);                // K = L^{-T} W
  }
  return done;
}

// Outputs i < cnt of a vector product, each summed over k by a group of
// four consecutive threads (k = s, s + 4, ... on lane s of the group) and
// reduced by two lane exchanges; fin(i, sum) on the group's first lane.
// Groups are aligned to four lanes, so a group is active or idle as a whole.
template <typename T, typename Part, typename Fin>
__device__ __forceinline__ void dot4(int cnt, Part&& part, Fin&& fin) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int base = tid; base < 4 * cnt; base += nt) {
    const int i = base >> 2, sl = base & 3;
    T v = part(i, sl);
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    if (sl == 0) fin(i, v);
  }
}

// One backward Riccati step in LDS: (S, s) = V_{t+1} -> V_t; [K | k] in Kb;
// F_t = A - B K and f_t = c - B k go straight to the workspace.
//   [T1 | SB] = S [A | B],  Sc = S c + s
//   H = R + diag(rd) + rho_c E'E + B'SB,  [G | g] = [M' + rho_c E'D + B'T1 | r + B'Sc]
//   (M: the caller's cross Hessian, 0 for the device model families)
//   [W | w] = Z [G | g], [K | k] = Z' [W | w]  (H^{-1} = Z' Z; W replaces G in LDS)
//   S_t = Q + rho_c D'D + A'T1 - W'W,  s_t = q + A'Sc - W'w
// FP32 forms the products on the matrix cores (unless flags bit 0), FP64 on VALU.
template <typename T, int MMAX>
__device__ void riccati_step(const Prob<T>& P, const Prm<T>& prm, const Ws<T>& ws, int b, int t,
                             RicSmem<T>& L) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, ld = n + 1, ldab = L.ldab;
  const int tid = threadIdx.x, nt = blockDim.x;
  const T* R = P.R.at(b, t);
  const T* Q = P.Q.at(b, t);
  T* Fo = ws.F + ((size_t)b * N + t) * n * n;
  T* fo = ws.f + ((size_t)b * N + t) * n;
  bool mfma = false;
  if constexpr (IsF32<T>::value) mfma = (prm.flags & 1) == 0;
  // A: Sc = S c + s ; [T1 | SB] = S [A | B]
  dot4<T>(n, [&](int i, int sl) {
    T v = 0;
    for (int j = sl; j < n; j += 4) v += L.S[i * n + j] * L.c[j];
    return v;
  }, [&](int i, T v) { L.Sc[i] = L.s[i] + v; });
  if constexpr (IsF32<T>::value) {
    if (mfma) {
      const MSeg sg{L.S, L.A, n, ldab, n, false, false, 1.f};
      mfma_gemm<false>(n, n + m, 1, sg, sg, sg, L.T1, ldab, nullptr, 0, nullptr);
    }
  }
  if (!mfma) gemm_lds<T, false, false, 0>(n, n + m, n, L.S, n, L.A, ldab, L.T1, ldab);
  lds_sync();
  // B: H, [G | g]
  if constexpr (IsF32<T>::value) {
    if (mfma) {
      const MSeg h0{L.B, L.SB, ldab, ldab, n, true, false, 1.f}, h1{L.E, L.E, m, m, p, true, false, 1.f};
      mfma_gemm<false>(m, m, p > 0 ? 2 : 1, h0, h1, h1, L.H, m, R, m, L.rd);
      const MSeg g0{L.B, L.T1, ldab, ldab, n, true, false, 1.f}, g1{L.E, L.D, m, n, p, true, false, 1.f};
      mfma_gemm<false>(m, n, p > 0 ? 2 : 1, g0, g1, g1, L.G, ld, L.Mt, n, nullptr);
    }
  }
  if (!mfma) {
    for (int e = tid; e < m * m; e += nt) {
      const int i = e / m, j = e - i * m;
      T v = R[e] + (i == j ? L.rd[i] : T(0));
      for (int k = 0; k < p; ++k) v += L.E[k * m + i] * L.E[k * m + j];
      for (int k = 0; k < n; ++k) v += L.B[k * ldab + i] * L.SB[k * ldab + j];
      L.H[e] = v;
    }
    for (int e = tid; e < m * n; e += nt) {
      const int i = e / n, j = e - i * n;
      T v = L.Mt ? L.Mt[e] : T(0);
      for (int k = 0; k < p; ++k) v += L.E[k * m + i] * L.D[k * n + j];
      for (int k = 0; k < n; ++k) v += L.B[k * ldab + i] * L.T1[k * ldab + j];
      L.G[i * ld + j] = v;
    }
  }
  dot4<T>(m, [&](int i, int sl) {
    T v = 0;
    for (int k = sl; k < n; k += 4) v += L.B[k * ldab + i] * L.Sc[k];
    return v;
  }, [&](int i, T v) { L.G[i * ld + n] = L.r[i] + v; });
  lds_sync();
  // C, D: [W | w] = Z [G | g] in G's place, [K | k] = Z' [W | w] in Kb.
  // While wave 0 factors H (m <= 32, FP32 products on the matrix cores), the
  // other waves form P = Q + A'T1 + rho_c D'D in S's place (S itself is no
  // longer read: T1 and Sc hold its products); S_t = P - W'W follows.
  const bool pre = spd_apply_inverse<T, MMAX>(L.H, L.G, L.Kb, m, n + 1, ld, prm.hfloor, mfma, [&]() -> bool {
    if constexpr (IsF32<T>::value) {
      if (mfma && nt > 64) {
        const MSeg s0{L.A, L.T1, ldab, ldab, n, true, false, 1.f};
        const MSeg s2{L.D, L.D, n, n, p, true, false, 1.f};
        mfma_gemm<true>(n, n, p > 0 ? 2 : 1, s0, s2, s2, L.S, n, Q, n, nullptr, 1);
        return true;
      }
    }
    return false;
  });
  lds_sync();
  // E: S_t, s_t ; F_t = A - B K and f_t = c - B k (to the workspace)
  if constexpr (IsF32<T>::value) {
    if (mfma) {
      const MSeg s0{L.A, L.T1, ldab, ldab, n, true, false, 1.f};
      const MSeg s1{L.G, L.G, ld, ld, m, true, false, -1.f};
      const MSeg s2{L.D, L.D, n, n, p, true, false, 1.f};
      if (pre) mfma_gemm<true>(n, n, 1, s1, s1, s1, L.S, n, L.S, n, nullptr);
      else mfma_gemm<true>(n, n, p > 0 ? 3 : 2, s0, s1, s2, L.S, n, Q, n, nullptr);
      const MSeg f0{L.B, L.Kb, ldab, ld, m, false, false, -1.f};
      mfma_gemm<false>(n, n, 1, f0, f0, f0, Fo, n, L.A, ldab, nullptr);
    }
  }
  if (!mfma) {
    for (int e = tid; e < n * n; e += nt) {
      const int i = e / n, j = e - i * n;
      T fv = L.A[i * ldab + j];
      for (int r = 0; r < m; ++r) fv -= L.B[i * ldab + r] * L.Kb[r * ld + j];
      Fo[e] = fv;
      if (i > j) continue;
      T v = Q[e];
      for (int k = 0; k < p; ++k) v += L.D[k * n + i] * L.D[k * n + j];
      for (int k = 0; k < n; ++k) v += L.A[k * ldab + i] * L.T1[k * ldab + j];
      for (int r = 0; r < m; ++r) v -= L.G[r * ld + i] * L.G[r * ld + j];
      L.S[i * n + j] = v;
      L.S[j * n + i] = v;
    }
  }
  dot4<T>(n, [&](int i, int sl) {
    T v = 0;
    for (int k = sl; k < n; k += 4) v += L.A[k * ldab + i] * L.Sc[k];
    for (int r = sl; r < m; r += 4) v -= L.G[r * ld + i] * L.G[r * ld + n];
    return v;
  }, [&](int i, T v) { L.s[i] = L.q[i] + v; });
  dot4<T>(n, [&](int i, int sl) {
    T w = 0;
    for (int r = sl; r < m; r += 4) w -= L.B[i * ldab + r] * L.Kb[r * ld + n];
    return w;
  }, [&](int i, T w) { fo[i] = L.c[i] + w; });
  lds_sync();
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
    // [K | k] as one flat pass over the m x (n + 1) entries of Kb
    const float inv = 1.f / (float)ld;
    for (int e = threadIdx.x; e < m * ld; e += blockDim.x) {
      const int i = div_small(e, ld, inv), j = e - i * ld;
      const T v = L.Kb[e];
      if (j < n) K[i * n + j] = v;
      else k[i] = v;
    }
  }
}

// Riccati over stages [t_lo, t_hi) (t_hi <= N) starting from the value already in
// (L.S, L.s) at index t_hi; stores S, s for t in [t_lo, t_hi) (and t_hi when
// store_top), K, k, F, f for t < t_hi.
template <typename T, int MMAX>
__device__ void riccati_range(const Prob<T>& P, const Prm<T>& prm, const Ws<T>& ws, const LqrIn<T>& in,
                              int b, int t_lo, int t_hi, RicSmem<T>& L, bool store_top) {
  if (store_top) store_value_and_gain(P, ws, b, t_hi, L, false);
  for (int t = t_hi - 1; t >= t_lo; --t) {
    load_stage(P, prm, ws, in, b, t, L);
    lds_sync();
    riccati_step<T, MMAX>(P, prm, ws, b, t, L);
    store_value_and_gain(P, ws, b, t, L, true);
  }
  lds_sync();
}

// dx over stages [t_lo, t_hi) from dx_{t_lo} in x (LDS, n): dx_{t+1} = F_t dx_t + f_t.
// F of the next stage is fetched into registers while the current one is applied.
// Writes dx_t for t in [t_lo, t_hi) (and dx_{t_hi} when last); x ends as dx_{t_hi}.
// F0, F1: LDS buffers of n*n; x2: LDS n.
template <typename T>
__device__ void fwd_chain(const Prob<T>& P, const Ws<T>& ws, int b, int t_lo, int t_hi, bool last, T* x, T* F0,
                          T* F1, T* x2) {
  const int n = P.n, N = P.N, N1 = N + 1;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nn = n * n;
  constexpr int PF = 16;  // prefetch registers per thread
  const bool fast = nn <= PF * nt;
  T* cur = F0;
  T* nxt = F1;
  if (t_lo < t_hi) load_block(cur, ws.F + ((size_t)b * N + t_lo) * nn, n, n, n);
  lds_sync();
  for (int t = t_lo; t < t_hi; ++t) {
    T pre[PF];
    const bool more = t + 1 < t_hi;
    if (fast && more) {
      const T* Fn = ws.F + ((size_t)b * N + t + 1) * nn;
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int e = tid + u * nt;
        pre[u] = e < nn ? Fn[e] : T(0);
      }
    }
    const T* f = ws.f + ((size_t)b * N + t) * n;
    T* dxo = ws.dx + ((size_t)b * N1 + t) * n;
    for (int i = tid; i < n; i += nt) {
      T v = f[i];
#pragma unroll 8
      for (int j = 0; j < n; ++j) v += cur[i * n + j] * x[j];
      x2[i] = v;
      dxo[i] = x[i];
    }
    if (more) {
      if (fast) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int e = tid + u * nt;
          if (e < nn) nxt[e] = pre[u];
        }
      } else {
        load_block(nxt, ws.F + ((size_t)b * N + t + 1) * nn, n, n, n);
      }
    }
    lds_sync();
    for (int i = tid; i < n; i += nt) x[i] = x2[i];
    T* tmp = cur; cur = nxt; nxt = tmp;
    lds_sync();
  }
  if (last)
    for (int i = tid; i < n; i += nt) ws.dx[((size_t)b * N1 + t_hi) * n + i] = x[i];
  lds_sync();
}

template <typename T>
__host__ __device__ inline size_t stage_smem_elems(int n, int nt) {
  return (size_t)n * (n + 1) + (size_t)n + (size_t)NFP * nt + 8;
}

// Parallel over (stage t, instance b): lam = S dx + s, dV = lam - V,
// du = -K dx - k, and the partials q'dx + r'du, (dV - V)'c, lam'c of stage t
// (the line search forms the curvature dz'H dz = sum lam'c - d_f from them).
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_stage(Prob<T> P, Iter<T> it, Ws<T> ws) {
  const int t = blockIdx.x, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const int tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sS = reinterpret_cast<T*>(smem_raw);   // n x (n+1) (padded rows)
  T* sx = sS + n * (n + 1);
  T* red = sx + n;
  const int lds = n + 1;
  const T* S = ws.S + ((size_t)b * N1 + t) * n * n;
  for (int i = 0; i < n; ++i)
    for (int j = tid; j < n; j += nt) sS[i * lds + j] = S[i * n + j];
  const T* dx = ws.dx + ((size_t)b * N1 + t) * n;
  for (int i = tid; i < n; i += nt) sx[i] = dx[i];
  __syncthreads();
  T acc[NFP] = {0, 0, 0, 0};
  const T* s = ws.s + ((size_t)b * N1 + t) * n;
  const T* V = it.V + ((size_t)b * N1 + t) * n;
  const T* c = ws.cdef + ((size_t)b * N1 + t) * n;
  const T* q = ws.qh + ((size_t)b * N1 + t) * n;
  T* dv = ws.dv + ((size_t)b * N1 + t) * n;
  for (int i = tid; i < n; i += nt) {
    T lam = s[i];
#pragma unroll 8
    for (int j = 0; j < n; ++j) lam += sS[i * lds + j] * sx[j];
    const T v = V[i];
    const T d = lam - v;
    dv[i] = d;
    acc[FP_DC] += (d - v) * c[i];
    acc[FP_LC] += lam * c[i];
    acc[FP_DF] += q[i] * sx[i];
  }
  if (t < N) {
    const T* K = ws.K + ((size_t)b * N + t) * m * n;
    const T* k = ws.k + ((size_t)b * N + t) * m;
    const T* r = ws.rh + ((size_t)b * N + t) * m;
    T* du = ws.du + ((size_t)b * N + t) * m;
    for (int i = tid; i < m; i += nt) {
      T v = k[i];
      for (int j = 0; j < n; ++j) v += K[i * n + j] * sx[j];
      du[i] = -v;
      acc[FP_DF] -= r[i] * v;
    }
  }
  block_sum<T, NFP>(acc, 3, red);
  if (tid == 0) {
    T* fp = ws.fpart + ((size_t)b * N1 + t) * NFP;
    fp[FP_DF] = acc[FP_DF];
    fp[FP_DC] = acc[FP_DC];
    fp[FP_LC] = acc[FP_LC];
    fp[3] = 0;
  }
}

// segments == 1: sequential Riccati, then the dx chain, one block per instance.
template <typename T, int MMAX>
__global__ void __launch_bounds__(512) k_lqr_seq(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, LqrIn<T> in) {
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const int tid = threadIdx.x, nt = blockDim.x;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* base = reinterpret_cast<T*>(smem_raw);
  RicSmem<T> L = ric_carve(base, n, m, p, nt, P.M.p != nullptr);
  load_block(L.S, P.QN.at(b, 0), n, n, n);
  for (int i = tid; i < n; i += nt) L.s[i] = ws.qh[((size_t)b * N1 + N) * n + i];
  __syncthreads();
  riccati_range<T, MMAX>(P, prm, ws, in, b, 0, N, L, true);
  __syncthreads();  // F_t, f_t stored by every wave are read below
  // dx chain: F buffers in A/T1 (each >= n*n), x in Sc, x2 in c
  for (int i = tid; i < n; i += nt) L.Sc[i] = ws.cdef[(size_t)b * N1 * n + i];
  __syncthreads();
  fwd_chain(P, ws, b, 0, N, true, L.Sc, L.A, L.T1, L.c);
}

}  // namespace mpxk
