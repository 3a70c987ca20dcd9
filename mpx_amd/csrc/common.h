// Device-side types and helpers shared by the KKT kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mpxk {

// strided view: element (b, t) starts at p + b*sb + t*st
template <typename T>
struct View {
  const T* p;
  long long sb, st;
  __device__ __forceinline__ const T* at(int b, int t) const {
    return p + (long long)b * sb + (long long)t * st;
  }
};

template <typename T>
struct Prob {
  int B, n, m, N, p, box, model;
  T h, kappa;
  View<T> A, Bm, W, Q, R, QN, xref, uref, D, E, e, ulo, uhi, x0;
  View<T> fbar, M, q, r, g;  // model 2 (caller linearisation): f(x_t,u_t), cross Hessian, gradients, costs
};

enum { MODEL_LTV = 0, MODEL_TANH = 1, MODEL_CALLER = 2 };

template <typename T>
struct Prm {
  T rho0, rho_max, rho_c, mu_b, tau, beta, c1, slack;
  T hfloor;  // Cholesky pivot floor, relative to the pivot's diagonal entry
  int max_trials, group, P;
  int flags;  // bit 0: form FP32 products on VALU instead of the matrix cores
};

template <typename T>
struct Iter {
  T *X, *U, *V, *Yeq, *Ybox;
};

// per-stage merit partials written by the linearisation kernel
enum { MP_COST = 0, MP_VC = 1, MP_CC = 2, NMP = 3 };
// per-segment partials of the forward pass: q'dx + r'du, (dV - V)'c, lam'c
enum { FP_DF = 0, FP_DC = 1, FP_LC = 2, NFP = 4 };

constexpr int MAX_TRIALS = 16;

template <typename T>
struct Ws {
  T* Alin;   // [B][N][n][n]   Jacobians f_x of the tanh model
  T* Blin;   // [B][n][m]      h*B of the tanh model
  T* sdh;    // [B][N][n]      sech^2(W x_t) of the tanh model (the matrix-core linearisation)
  T* qh;     // [B][N1][n]     augmented state gradients (row N: terminal)
  T* rh;     // [B][N][m]      augmented control gradients
  T* rd;     // [B][N][m]      box Hessian diagonal nu/s
  T* cdef;   // [B][N1][n]     constraint rows [x0 - X_0, f - x_{t+1}]
  T* S;      // [B][N1][n][n]  value Hessians
  T* s;      // [B][N1][n]     value gradients
  T* K;      // [B][N][m][n]   feedback gains (du = -K dx - k)
  T* k;      // [B][N][m]
  T* F;      // [B][N][n][n]   closed-loop maps A - B K
  T* f;      // [B][N][n]      offsets c - B k
  T* dx;     // [B][N1][n]
  T* du;     // [B][N][m]
  T* dv;     // [B][N1][n]
  T* mpart;  // [B][N1][NMP]
  T* fpart;  // [B][N1][NFP]   per-stage slope partials
  T* agg;    // [B][P][3n^2+2n] segment aggregates (A, C, J, b, eta)
  T* fmap;   // [B][P][n^2+n]   segment forward maps (Phi, phi)
  T* sval;   // [B][P][n^2+n]   values (S, s) at segment starts
  T* xs;     // [B][P][n]       forward states at segment starts
  T* kpart;  // [B][N1]         squared KKT residual per stage
  T* lsp;    // [B][8]          line-search scalars (rho, phi(0), phi'(0), step caps, slack)
  T* tpart;  // [B][N1][2*MAX_TRIALS] per-stage parts of phi at each trial step
  T* hq;     // [B][N][2p]      equality rows at the iterate and their slopes along the step
  T* gpart;  // [B][N1][2]      per-stage cost g_t and squared defect rows (stop test)
  T* last;   // [B]             previous iteration's sum g_t^2 (stop test)
  int* run;  // [B]             1 while the instance iterates; nullptr: every instance runs
  int* nrun; // [1]             instances still running (host check of the solve loop)
  T* trial_g; // [B][L][N1]     trial stage costs of a device-evaluated caller model (SRBD)
  T* trial_f; // [B][L][N][n]   trial next states of that model
};

// Instances whose run flag is 0 (converged in a multi-iteration solve) leave
// every kernel at once; b is uniform over the block, so no barrier is skipped
// by part of a block.
template <typename T>
__device__ __forceinline__ bool skip_instance(const Ws<T>& ws, int b) {
  return ws.run != nullptr && ws.run[b] == 0;
}

template <typename T>
__device__ __forceinline__ T dev_tanh(T x) { return tanh(x); }
template <>
__device__ __forceinline__ float dev_tanh<float>(float x) { return tanhf(x); }

template <typename T>
__device__ __forceinline__ T dev_log(T x) { return log(x); }
template <>
__device__ __forceinline__ float dev_log<float>(float x) { return logf(x); }

template <typename T>
__device__ __forceinline__ T dev_sqrt(T x) { return sqrt(x); }
template <>
__device__ __forceinline__ float dev_sqrt<float>(float x) { return sqrtf(x); }

// Block barrier that waits only for the block's LDS traffic, not for its global
// stores: phases whose global writes are read by later kernels (or after a full
// __syncthreads) need not stall on them.
__device__ __forceinline__ void lds_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// LDS written and read by lanes of one wave: wait for the wave's LDS traffic
// and keep the compiler from moving memory operations across.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Block copies from global memory into LDS with K loads of each thread in
// flight before its first store (a plain strided loop waits on global memory
// once per iteration).  lds_copy: dst[e] = src[e]; lds_copy_padded: the
// row-major rows of length ld land in rows of length ld + 1, dst[e + e / ld].
template <typename T, int K>
__device__ __forceinline__ void lds_copy(T* dst, const T* src, int cnt) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int base = 0; base < cnt; base += K * nt) {
    T v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = base + tid + k * nt;
      v[k] = e < cnt ? src[e] : T(0);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = base + tid + k * nt;
      if (e < cnt) dst[e] = v[k];
    }
  }
}
template <typename T, int K>
__device__ __forceinline__ void lds_copy_padded(T* dst, const T* src, int rows, int ld) {
  const int tid = threadIdx.x, nt = blockDim.x, cnt = rows * ld;
  for (int base = 0; base < cnt; base += K * nt) {
    T v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = base + tid + k * nt;
      v[k] = e < cnt ? src[e] : T(0);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int e = base + tid + k * nt;
      if (e < cnt) dst[e + e / ld] = v[k];
    }
  }
}

// value of v in lane `src` (uniform), for 32- and 64-bit types
__device__ __forceinline__ float rlane(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double rlane(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// v from the lane CTRL rotates onto this one within its row of 16 (DPP
// row_ror, a modifier of the add that consumes it: no LDS round trip).
template <int CTRL>
__device__ __forceinline__ float dpp_ror(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_ror(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wave (every lane gets the result; call with every
// lane active): rotations by 8, 4, 2, 1 within each row of 16 lanes leave the
// row's sum in every lane of the row, then the four row sums are read out.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp_ror<0x128>(v);
  v += dpp_ror<0x124>(v);
  v += dpp_ror<0x122>(v);
  v += dpp_ror<0x121>(v);
  return (rlane(v, 0) + rlane(v, 16)) + (rlane(v, 32) + rlane(v, 48));
}

// Block-wide sum of `cnt` values per thread (cnt <= NV); blockDim.x must be a
// power of two; sred holds NV*blockDim.x elements.  Every thread gets the sums.
template <typename T, int NV>
__device__ void block_sum(T* v, int cnt, T* sred) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = 0; j < cnt; ++j) sred[j * nt + tid] = v[j];
  __syncthreads();
  for (int off = nt >> 1; off > 0; off >>= 1) {
    if (tid < off)
      for (int j = 0; j < cnt; ++j) sred[j * nt + tid] += sred[j * nt + tid + off];
    __syncthreads();
  }
  for (int j = 0; j < cnt; ++j) v[j] = sred[j * nt];
  __syncthreads();
}

template <typename T>
__device__ T block_min(T v, T* sred) {
  const int tid = threadIdx.x, nt = blockDim.x;
  sred[tid] = v;
  __syncthreads();
  for (int off = nt >> 1; off > 0; off >>= 1) {
    if (tid < off) sred[tid] = sred[tid] < sred[tid + off] ? sred[tid] : sred[tid + off];
    __syncthreads();
  }
  T r = sred[0];
  __syncthreads();
  return r;
}

// Copy a rows x cols row-major block (source leading dimension lds) into LDS (dense).
template <typename T>
__device__ __forceinline__ void load_block(T* dst, const T* src, int rows, int cols, int lds) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tot = rows * cols;
  for (int e = tid; e < tot; e += nt) {
    int i = e / cols, j = e - i * cols;
    dst[e] = src[(long long)i * lds + j];
  }
}

template <typename T>
__device__ __forceinline__ void store_block(T* dst, const T* src, int tot) {
  for (int e = threadIdx.x; e < tot; e += blockDim.x) dst[e] = src[e];
}

// C (M x N, ldc) MODE= op(A) op(B), A: M x K (or K x M if TA), B: K x N (or N x K if TB).
// MODE 0: set, 1: add, 2: subtract.  2x2 register tile per thread.
template <typename T, bool TA, bool TB, int MODE>
__device__ void gemm_lds(int M, int N, int K, const T* A, int lda, const T* B, int ldb,
                         T* C, int ldc) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tm = (M + 1) >> 1, tn = (N + 1) >> 1;
  for (int tile = tid; tile < tm * tn; tile += nt) {
    const int i0 = (tile / tn) * 2, j0 = (tile % tn) * 2;
    const bool i1 = i0 + 1 < M, j1 = j0 + 1 < N;
    const int ia = i1 ? i0 + 1 : i0, jb = j1 ? j0 + 1 : j0;
    T c00 = 0, c01 = 0, c10 = 0, c11 = 0;
    for (int k = 0; k < K; ++k) {
      const T a0 = TA ? A[k * lda + i0] : A[i0 * lda + k];
      const T a1 = TA ? A[k * lda + ia] : A[ia * lda + k];
      const T b0 = TB ? B[j0 * ldb + k] : B[k * ldb + j0];
      const T b1 = TB ? B[jb * ldb + k] : B[k * ldb + jb];
      c00 += a0 * b0; c01 += a0 * b1; c10 += a1 * b0; c11 += a1 * b1;
    }
#define MPXK_ST(ii, jj, v)                                  \
  do {                                                      \
    T& d = C[(ii) * ldc + (jj)];                            \
    if (MODE == 0) d = (v); else if (MODE == 1) d += (v); else d -= (v); \
  } while (0)
    MPXK_ST(i0, j0, c00);
    if (j1) MPXK_ST(i0, j0 + 1, c01);
    if (i1) MPXK_ST(i0 + 1, j0, c10);
    if (i1 && j1) MPXK_ST(i0 + 1, j0 + 1, c11);
#undef MPXK_ST
  }
}

// Pivot j of a Cholesky factorisation: a (the pivot before its square root)
// becomes max(|a|, floor |h0|) (h0 = the matrix's original diagonal entry), so
// the factor is that of the matrix plus a non-negative diagonal; NaN stays NaN.
template <typename T>
__device__ __forceinline__ T floored_pivot(T a, T h0, T floor) {
  const T fl = floor * (h0 < T(0) ? -h0 : h0);
  const T aa = a < T(0) ? -a : a;
  return (aa >= fl || aa != aa) ? aa : fl;
}

// In-place lower Cholesky of the m x m symmetric matrix H (row-major, LDS),
// pivots modified by floored_pivot with the original diagonal entry, which
// thread 0 recovers as H_jj + sum_k L_jk^2.  A NaN pivot
// propagates (reported per instance).
template <typename T>
__device__ void chol_lds(T* H, int m, T floor) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = 0; j < m; ++j) {
    if (tid == 0) {
      const T a = H[j * m + j];
      T h0 = a;
      for (int k = 0; k < j; ++k) h0 += H[j * m + k] * H[j * m + k];
      H[j * m + j] = dev_sqrt(floored_pivot(a, h0, floor));
    }
    lds_sync();
    const T d = H[j * m + j];
    for (int i = j + 1 + tid; i < m; i += nt) H[i * m + j] /= d;
    lds_sync();
    const int rem = m - j - 1;
    for (int e = tid; e < rem * rem; e += nt) {
      const int i = j + 1 + e / rem, kk = j + 1 + e % rem;
      if (kk <= i) H[i * m + kk] -= H[i * m + j] * H[kk * m + j];
    }
    lds_sync();
  }
}

// Forward substitution L Y = G for the columns of G (m x ncol, ld), in place.
template <typename T>
__device__ void trsm_lower(const T* L, int m, T* G, int ncol, int ld) {
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    for (int i = 0; i < m; ++i) {
      T v = G[i * ld + c];
      for (int kk = 0; kk < i; ++kk) v -= L[i * m + kk] * G[kk * ld + c];
      G[i * ld + c] = v / L[i * m + i];
    }
  }
}

// Backward substitution L^T X = Y for the columns of Y, in place.
template <typename T>
__device__ void trsm_upper_t(const T* L, int m, T* G, int ncol, int ld) {
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    for (int i = m - 1; i >= 0; --i) {
      T v = G[i * ld + c];
      for (int kk = i + 1; kk < m; ++kk) v -= L[kk * m + i] * G[kk * ld + c];
      G[i * ld + c] = v / L[i * m + i];
    }
  }
}

}  // namespace mpxk

namespace mpxk {

// ---------------------------------------------------------------------------
// FP32 matrix-core GEMM on LDS operands (v_mfma_f32_16x16x4f32, wave64).
// Operand layouts of the instruction: A (16x4) lane l holds A[l%16][l/16],
// B (4x16) lane l holds B[l/16][l%16], result lane l holds D[4(l/16)+v][l%16].
// The product is a sum of up to 3 K-segments (each op(A_s) op(B_s) scaled by
// sign_s), so rank updates fold into one pass; waves of the block take the
// 16x16 output tiles round-robin.  With SYM only tiles ti <= tj are formed
// and every value goes to (i, j) and (j, i) (the result must be symmetric).
// At store: C = [C +] acc + addend[i*lda_add + j] + (i == j ? diag[i] : 0).
// ---------------------------------------------------------------------------
typedef float mpxk_f32x4 __attribute__((ext_vector_type(4)));

struct MSeg {
  const float* A;
  const float* B;
  int lda, ldb, K;
  bool ta, tb;
  float sign;
};

// acc += sign * op(A)[i, :] op(B)[:, j] over one K-segment.  Operands of the
// next four k-steps are loaded while the current four MFMAs issue.  No operand
// load runs under an execution mask: a row i >= M (column j >= N) reads row
// M - 1 (column N - 1) instead, which only feeds output entries that are never
// stored, and k is clamped to K - 1 with the entries past K replaced by 0
// after the load (kmask).
#define MPXK_MF_LOAD(K0, A0, A1, A2, A3, B0, B1, B2, B3)                       \
  do {                                                                         \
    const int k_ = (K0) + q;                                                   \
    const int k0c_ = k_ < g.K ? k_ : g.K - 1;                                  \
    const int k1c_ = k_ + 4 < g.K ? k_ + 4 : g.K - 1;                          \
    const int k2c_ = k_ + 8 < g.K ? k_ + 8 : g.K - 1;                          \
    const int k3c_ = k_ + 12 < g.K ? k_ + 12 : g.K - 1;                        \
    const float a0_ = pa[k0c_ * sak], b0_ = pb[k0c_ * sbk];                    \
    const float a1_ = pa[k1c_ * sak], b1_ = pb[k1c_ * sbk];                    \
    const float a2_ = pa[k2c_ * sak], b2_ = pb[k2c_ * sbk];                    \
    const float a3_ = pa[k3c_ * sak], b3_ = pb[k3c_ * sbk];                    \
    A0 = kmask(a0_ * g.sign, k_ - g.K);                                        \
    B0 = kmask(b0_, k_ - g.K);                                                 \
    A1 = kmask(a1_ * g.sign, k_ + 4 - g.K);                                    \
    B1 = kmask(b1_, k_ + 4 - g.K);                                             \
    A2 = kmask(a2_ * g.sign, k_ + 8 - g.K);                                    \
    B2 = kmask(b2_, k_ + 8 - g.K);                                             \
    A3 = kmask(a3_ * g.sign, k_ + 12 - g.K);                                   \
    B3 = kmask(b3_, k_ + 12 - g.K);                                            \
  } while (0)

// x where d < 0, else +0 (whatever x holds): the sign of d spread over a word
// and ANDed into x's bits, on the vector ALU (no lane-mask registers)
__device__ __forceinline__ float kmask(float x, int d) { return __int_as_float(__float_as_int(x) & (d >> 31)); }

__device__ __forceinline__ void mfma_seg(mpxk_f32x4& acc, const MSeg& g, int i, int j, int M, int N, int q) {
  if (g.K <= 0) return;
  const int sak = g.ta ? g.lda : 1, sai = g.ta ? 1 : g.lda;
  const int sbk = g.tb ? 1 : g.ldb, sbj = g.tb ? g.ldb : 1;
  const float* pa = g.A + (i < M ? i : M - 1) * sai;
  const float* pb = g.B + (j < N ? j : N - 1) * sbj;
  float a0, a1, a2, a3, b0, b1, b2, b3;
  MPXK_MF_LOAD(0, a0, a1, a2, a3, b0, b1, b2, b3);
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f, d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f;
    if (k0 + 16 < g.K) MPXK_MF_LOAD(k0 + 16, c0, c1, c2, c3, d0, d1, d2, d3);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc, 0, 0, 0);
    if (k0 + 4 < g.K) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc, 0, 0, 0);
    if (k0 + 8 < g.K) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b2, acc, 0, 0, 0);
    if (k0 + 12 < g.K) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a3, b3, acc, 0, 0, 0);
    a0 = c0; a1 = c1; a2 = c2; a3 = c3; b0 = d0; b1 = d1; b2 = d2; b3 = d3;
  }
}
#undef MPXK_MF_LOAD

// tile index -> (ti, tj); with SYM only the upper triangle of tiles is enumerated
template <bool SYM>
__device__ __forceinline__ void tile_of(int tile, int tn, int* ti, int* tj) {
  if (!SYM) { *ti = tile / tn; *tj = tile - *ti * tn; return; }
  int r = 0, rem = tile;
  while (rem >= tn - r) { rem -= tn - r; ++r; }
  *ti = r; *tj = r + rem;
}

template <bool SYM>
__device__ __forceinline__ void mfma_store(const mpxk_f32x4& acc, const float* add, int ti, int tj, int M, int N,
                                           int q, int r, float* C, int ldc) {
  const int cj = tj * 16 + r;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int ci = ti * 16 + 4 * q + v;
    if (ci >= M || cj >= N) continue;
    if (SYM && ti == tj && ci > cj) continue;
    const float val = acc[v] + add[v];
    C[ci * ldc + cj] = val;
    if (SYM && ci != cj) C[cj * ldc + ci] = val;
  }
}

// The addend of the lane's entries; an entry outside M x N reads the clamped
// one instead (no execution mask) and is never stored.
template <bool SYM>
__device__ __forceinline__ void mfma_addend(float* add, int ti, int tj, int M, int N, int q, int r,
                                            const float* addend, int ld_add, const float* diag) {
  const int cj = tj * 16 + r, cjc = cj < N ? cj : N - 1;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int ci = ti * 16 + 4 * q + v, cic = ci < M ? ci : M - 1;
    add[v] = addend ? addend[cic * ld_add + cjc] : 0.f;
    if (diag) {
      const float dv = diag[cic];
      add[v] += ci == cj ? dv : 0.f;
    }
  }
}

// C = sum_s sign_s op(A_s) op(B_s) + addend + diag, 16x16 tiles taken by the
// waves w0.. of the block round-robin (C may be the addend itself: each entry
// is read before its tile is stored, and a mirrored store lands in a tile
// below the diagonal, which no tile reads).
template <bool SYM>
__device__ void mfma_gemm(int M, int N, int nseg, MSeg s0, MSeg s1, MSeg s2, float* C, int ldc,
                          const float* addend, int ld_add, const float* diag, int w0 = 0) {
  // the waves w0.. of the block take part
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) - w0, nw = (blockDim.x >> 6) - w0;
  if (wave < 0) return;
  const int tm = (M + 15) >> 4, tn = (N + 15) >> 4;
  const int ntile = SYM ? tn * (tn + 1) / 2 : tm * tn;
  const int r = lane & 15, q = lane >> 4;
  // the next tile's addend is requested before this tile's products (it may
  // sit in global memory)
  float add[4] = {0.f, 0.f, 0.f, 0.f};
  int ti = 0, tj = 0;
  if (wave < ntile) {
    tile_of<SYM>(wave, tn, &ti, &tj);
    mfma_addend<SYM>(add, ti, tj, M, N, q, r, addend, ld_add, diag);
  }
  for (int t0 = wave; t0 < ntile; t0 += nw) {
    float addn[4] = {0.f, 0.f, 0.f, 0.f};
    int tin = 0, tjn = 0;
    if (t0 + nw < ntile) {
      tile_of<SYM>(t0 + nw, tn, &tin, &tjn);
      mfma_addend<SYM>(addn, tin, tjn, M, N, q, r, addend, ld_add, diag);
    }
    mpxk_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int i = ti * 16 + r, j = tj * 16 + r;
    mfma_seg(acc, s0, i, j, M, N, q);
    if (nseg > 1) mfma_seg(acc, s1, i, j, M, N, q);
    if (nseg > 2) mfma_seg(acc, s2, i, j, M, N, q);
    mfma_store<SYM>(acc, add, ti, tj, M, N, q, r, C, ldc);
#pragma unroll
    for (int v = 0; v < 4; ++v) add[v] = addn[v];
    ti = tin;
    tj = tjn;
  }
}

__device__ inline void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}



// Wave 0 alone, no block barrier inside (other waves return at once): for the
// symmetric m x m matrix H (LDS, row-major, m <= MM <= 64) the upper Cholesky
// factor U (H = U'U, pivots floored as in chol_lds) and Y = U^{-1} (upper
// triangular, zero below), written to Y.
// Lane k keeps column k of H in registers, padded to MM with the identity so
// that no step needs a bound, and column k of Z = L^{-1} (L = U'), starting
// as e_k.  Elimination step j forms row j of U (lane k's h[j] becomes
// U[j][k]), takes its entries L[i][j] = U[j][i] from the lanes by readlane and
// with each one updates the trailing column (H[i][k] -= L[i][j] U[j][k]) and
// the forward substitution of Z (Z[i][k] -= L[i][j] Z[j][k], after Z[j][k] /=
// L[j][j]); Y = Z' is written row k from lane k.  LDS is touched only to load
// H and store Y.
// (not inlined: it gets registers of its own; inlined into the block
// kernels, its readlane results spill with their scalar registers)
template <typename T, int MM>
__device__ __attribute__((noinline)) void chol_upper_inv_wave(const T* H, T* Y, int m, T floor) {
  if (threadIdx.x >= 64) return;
  const int k = threadIdx.x;
  const bool own = k < m;
  const int kc = own ? k : 0;
  T h[MM], z[MM];
  T hkk = 0;   // the lane's original diagonal entry (pivot floors)
#pragma unroll
  for (int i = 0; i < MM; ++i) {
    const T v = H[(i < m ? i : 0) * m + kc];
    h[i] = (own && i < m) ? v : T(i == k ? 1 : 0);
    z[i] = T(i == k ? 1 : 0);
    if (i == k) hkk = h[i];
  }
#pragma unroll
  for (int j = 0; j < MM; ++j) {
    const T piv = floored_pivot(rlane(h[j], j), rlane(hkk, j), floor);
    T d, lk, zj;
    if constexpr (sizeof(T) == 4) {   // FP32: one reciprocal square root, no divisions
      const T id = __builtin_amdgcn_rsqf(piv);
      d = piv * id;
      lk = k > j ? h[j] * id : T(0);
      zj = z[j] * id;
    } else {
      d = dev_sqrt(piv);
      lk = k > j ? h[j] / d : T(0);
      zj = z[j] / d;
    }
    h[j] = k == j ? d : lk;
    z[j] = zj;
    // the readlanes of lk start once zj is known (an opaque copy that takes
    // zj in), so that each L[i][j] serves both updates at once instead of
    // waiting in a scalar register, which spills
    asm volatile("" : "+v"(lk) : "v"(zj));
#pragma unroll
    for (int i = j + 1; i < MM; ++i) {
      const T lij = rlane(lk, i);
      h[i] -= lij * lk;
      z[i] -= lij * zj;
    }
  }
  wave_lds_sync();   // every lane has read H before Y (H itself in the block kernels) is written
  if (own) {
#pragma unroll
    for (int i = 0; i < MM; ++i)
      if (i < m) Y[k * m + i] = z[i];
  }
  wave_lds_sync();
}

}  // namespace mpxk
