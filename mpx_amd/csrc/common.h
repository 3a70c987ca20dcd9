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
};

template <typename T>
struct Prm {
  T rho0, rho_max, rho_c, mu_b, tau, beta, c1, slack;
  int max_trials, group, P;
};

template <typename T>
struct Iter {
  T *X, *U, *V, *Yeq, *Ybox;
};

// per-stage merit partials written by the linearisation kernel
enum { MP_COST = 0, MP_VC = 1, MP_CC = 2, NMP = 3 };
// per-segment partials of the forward pass
enum { FP_DF = 0, FP_DC = 1, FP_KAP = 2, NFP = 4 };

constexpr int MAX_TRIALS = 16;

template <typename T>
struct Ws {
  T* Alin;   // [B][N][n][n]   Jacobians f_x of the tanh model
  T* Blin;   // [B][n][m]      h*B of the tanh model
  T* qh;     // [B][N1][n]     augmented state gradients (row N: terminal)
  T* rh;     // [B][N][m]      augmented control gradients
  T* rd;     // [B][N][m]      box Hessian diagonal nu/s
  T* cdef;   // [B][N1][n]     constraint rows [x0 - X_0, f - x_{t+1}]
  T* S;      // [B][N1][n][n]  value Hessians
  T* s;      // [B][N1][n]     value gradients
  T* K;      // [B][N][m][n]   feedback gains (du = -K dx - k)
  T* k;      // [B][N][m]
  T* dx;     // [B][N1][n]
  T* du;     // [B][N][m]
  T* dv;     // [B][N1][n]
  T* mpart;  // [B][N1][NMP]
  T* fpart;  // [B][P][NFP]
  T* agg;    // [B][P][3n^2+2n] segment aggregates (A, C, J, b, eta)
  T* fmap;   // [B][P][n^2+n]   segment forward maps (Phi, phi)
  T* sval;   // [B][P][n^2+n]   values (S, s) at segment starts
  T* xs;     // [B][P][n]       forward states at segment starts
  T* kpart;  // [B][N1]         squared KKT residual per stage
};

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

// In-place lower Cholesky of the m x m SPD matrix H (row-major, LDS).
// A non-positive pivot leaves NaN, which then propagates (reported per instance).
template <typename T>
__device__ void chol_lds(T* H, int m) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = 0; j < m; ++j) {
    if (tid == 0) H[j * m + j] = dev_sqrt(H[j * m + j]);
    __syncthreads();
    const T d = H[j * m + j];
    for (int i = j + 1 + tid; i < m; i += nt) H[i * m + j] /= d;
    __syncthreads();
    const int rem = m - j - 1;
    for (int e = tid; e < rem * rem; e += nt) {
      const int i = j + 1 + e / rem, kk = j + 1 + e % rem;
      if (kk <= i) H[i * m + kk] -= H[i * m + j] * H[kk * m + j];
    }
    __syncthreads();
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
