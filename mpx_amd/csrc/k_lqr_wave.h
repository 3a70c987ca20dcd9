// Register-resident FP32 LQR solve: one wave64 per instance, every matrix held
// in VGPRs as 16x16 tiles of the v_mfma_f32_16x16x4f32 result layout.
//
// Tile layout ("C layout"): lane l = 16 g + c (g = l / 16, c = l % 16) holds
// X[4 g + v][c] in register v = 0..3 of the tile.  The instruction's operand
// layouts are A: lane l holds A[l % 16][l / 16], B: lane l holds B[l / 16][l % 16].
// The sum over k is order-free, so step s of a 16-wide k tile lets lane group
// g carry k = 4 g + s: the B operand is then register s of a C-layout tile of
// X, and the A operand is register s of a C-layout tile of A^T.  A product
// M X therefore runs straight from registers when M^T and X are held in C
// layout -- for the symmetric value matrix, S^T = S.
//
// Homogeneous coordinates fold the affine parts into the matrices: with
// x~ = [x; 1] (n + 1 <= 16 NT),
//   A~ = [[A, c], [0, 1]],  B~ = [B; 0],  S~ = [[S, s], [s', 0]],
//   Q~ = [[Q + rho_c D'D, q], [q', 0]],   G~ = [G | g],  K~ = [K | k],
// and one Riccati step
//   H  = R + diag(rd) + rho_c E'E + B~' S~ B~
//   G~ = B~' S~ A~ + rho_c E' [D | 0] + [0 | r]
//   K~ = H^{-1} G~
//   S~ <- Q~ + A~' S~ A~ - G~' K~ ,   F~ = A~ - B~ K~ = [[A - B K, c - B k], [0, 1]]
// is the affine step of k_lqr.h (S_t = Q + rho_c D'D + A'SA - G'K,
// s_t = q + A'(S c + s) - G'k, F_t = A - BK, f_t = c - Bk) with the vectors
// riding in column n.  The corner of S~ is reset to 0 every step.
//
// The m x m inverse is an in-place Gauss-Jordan sweep in one tile (H is SPD:
// no pivoting), with ds_bpermute broadcasts of the pivot row and column; the
// symmetric tiles are mirrored through a 1 KB LDS transpose.
// The forward chain x~_{t+1} = F~_t x~_t runs on the matrix cores too, with x~
// as column 0 of a B-operand tile: the result tile is again in that layout.
#pragma once

#include "common.h"
#include "k_lqr.h"

namespace mpxk {

typedef mpxk_f32x4 f4;

__device__ __forceinline__ f4 mma(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float f4_at(const f4& x, int v) {
  return v == 0 ? x[0] : (v == 1 ? x[1] : (v == 2 ? x[2] : x[3]));
}

__device__ __forceinline__ f4 zero4() { return f4{0.f, 0.f, 0.f, 0.f}; }

// Tile transpose through LDS (one wave per block: the barrier is the wave's own).
__device__ __forceinline__ f4 tile_transpose(const f4& x, float* buf, int g, int c) {
#pragma unroll
  for (int v = 0; v < 4; ++v) buf[(4 * g + v) * 17 + c] = x[v];
  __syncthreads();
  f4 y;
#pragma unroll
  for (int v = 0; v < 4; ++v) y[v] = buf[c * 17 + 4 * g + v];
  __syncthreads();
  return y;
}

// In-place inverse of the SPD tile H (m x m in the top-left corner, identity
// elsewhere): Gauss-Jordan without pivoting, pivot row j broadcast down the
// columns, pivot column j broadcast along the rows.
__device__ __forceinline__ void gj_inverse(f4& h, int m, int g, int c) {
  for (int j = 0; j < m; ++j) {
    const int jg = j >> 2, jv = j & 3;
    const float rowj = __shfl(f4_at(h, jv), jg * 16 + c);   // H[j][c]
    const float a = __shfl(rowj, j);                         // H[j][j] (lane j: g = 0, c = j)
    f4 colj;
#pragma unroll
    for (int v = 0; v < 4; ++v) colj[v] = __shfl(h[v], g * 16 + j);   // H[4g+v][j]
    const float rnew = (c == j ? 1.f : rowj) / a;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float hv = (c == j) ? 0.f : h[v];
      h[v] = (4 * g + v == j) ? rnew : hv - colj[v] * rnew;
    }
  }
}

// C-layout tile (it, jt) of the homogeneous A~ = [[A, c], [0, 1]].
__device__ __forceinline__ f4 load_abar(const float* A, const float* cd, int n, int it, int jt, int g, int c) {
  f4 x;
  const int col = 16 * jt + c;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = 16 * it + 4 * g + v;
    float e = 0.f;
    if (row < n) {
      if (col < n) e = A[row * n + col];
      else if (col == n) e = cd[row];
    } else if (row == n && col == n) {
      e = 1.f;
    }
    x[v] = e;
  }
  return x;
}

// C-layout tile (it, jt) of Q~ = [[Q, q], [q', 0]].
__device__ __forceinline__ f4 load_qbar(const float* Q, const float* q, int n, int it, int jt, int g, int c) {
  f4 x;
  const int col = 16 * jt + c;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = 16 * it + 4 * g + v;
    float e = 0.f;
    if (row < n) {
      if (col < n) e = Q[row * n + col];
      else if (col == n) e = q[row];
    } else if (row == n && col < n) {
      e = q[col];
    }
    x[v] = e;
  }
  return x;
}

// Store a C-layout tile of an (n+1)-wide homogeneous matrix: the n x n block to
// M (row-major, ld n) and column n to vec; rows >= rows_max are dropped.
__device__ __forceinline__ void store_hom(const f4& x, float* M, float* vec, int n, int rows_max, int it, int jt,
                                          int g, int c) {
  const int col = 16 * jt + c;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = 16 * it + 4 * g + v;
    if (row < rows_max) {
      if (col < n) M[row * n + col] = x[v];
      else if (col == n) vec[row] = x[v];
    }
  }
}

template <int NT>
__device__ __forceinline__ void store_value_wave(const f4 (&S)[NT][NT], const Ws<float>& ws, int b, int t, int N1,
                                                 int n, int g, int c) {
  float* So = ws.S + ((size_t)b * N1 + t) * n * n;
  float* so = ws.s + ((size_t)b * N1 + t) * n;
#pragma unroll
  for (int it = 0; it < NT; ++it)
#pragma unroll
    for (int jt = 0; jt < NT; ++jt) store_hom(S[it][jt], So, so, n, n, it, jt, g, c);
}

// segments == 1, FP32, n + 1 <= 16 NT, m <= 16, p <= 16: backward Riccati and
// the dx chain of one instance per wave.
template <int NT>
__global__ void __launch_bounds__(64, 2) k_lqr_wave(Prob<float> P, Prm<float> prm, Iter<float> it_unused,
                                                 Ws<float> ws, LqrIn<float> in) {
  (void)it_unused;
  const int b = blockIdx.x;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  __shared__ float tbuf[16 * 17];
  const float sr = sqrtf(prm.rho_c);

  f4 S[NT][NT];
  {
    const float* QN = P.QN.at(b, 0);
    const float* qN = ws.qh + ((size_t)b * N1 + N) * n;
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) S[i][j] = load_qbar(QN, qN, n, i, j, g, c);
    store_value_wave<NT>(S, ws, b, N, N1, n, g, c);
  }

  for (int t = N - 1; t >= 0; --t) {
    const float* A = in.A.at(b, t);
    const float* Bm = in.Bm.at(b, t);
    const float* cd = ws.cdef + ((size_t)b * N1 + t + 1) * n;
    f4 Ab[NT][NT], Bb[NT], Db[NT], Eb;
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) Ab[i][j] = load_abar(A, cd, n, i, j, g, c);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 x;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * i + 4 * g + v;          // B~ tile (i, 0): rows of B
        x[v] = (row < n && c < m) ? Bm[row * m + c] : 0.f;
      }
      Bb[i] = x;
    }
    if (p > 0) {
      const float* D = P.D.at(b, t);
      const float* E = P.E.at(b, t);
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        f4 x;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 4 * g + v, col = 16 * j + c;
          x[v] = (row < p && col < n) ? sr * D[row * n + col] : 0.f;
        }
        Db[j] = x;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 4 * g + v;
        Eb[v] = (row < p && c < m) ? sr * E[row * m + c] : 0.f;
      }
    }
    // H init: R + diag(rd), identity in the padding
    f4 H;
    {
      const float* R = P.R.at(b, t);
      const float* rd = ws.rd + ((size_t)b * N + t) * m;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 4 * g + v;
        float e = (row == c) ? 1.f : 0.f;
        if (row < m && c < m) e = R[row * m + c] + (row == c ? rd[row] : 0.f);
        H[v] = e;
      }
    }
    // G~ init: [0 | r]
    f4 G[NT];
    {
      const float* r = ws.rh + ((size_t)b * N + t) * m;
#pragma unroll
      for (int j = 0; j < NT; ++j) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 4 * g + v;
          G[j][v] = (16 * j + c == n && row < m) ? r[row] : 0.f;
        }
      }
    }

    // SA = S~ A~ (consumed below, tile by tile), SB = S~ B~
    f4 SB[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 acc = zero4();
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(S[k][i][s], Bb[k][s], acc);
      SB[i] = acc;
    }
    f4 SA[NT][NT];
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        f4 acc = zero4();
#pragma unroll
        for (int k = 0; k < NT; ++k)
#pragma unroll
          for (int s = 0; s < 4; ++s) acc = mma(S[k][i][s], Ab[k][j][s], acc);
        SA[i][j] = acc;
      }
    // H += B~' S~ B~ + rho_c E'E ;  G~ += B~' S~ A~ + rho_c E'[D | 0] ;
    // S~ <- A~' S~ A~ (tiles i <= j; S~ itself is no longer needed)
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) H = mma(Bb[k][s], SB[k][s], H);
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) G[j] = mma(Bb[k][s], SA[k][j][s], G[j]);
    if (p > 0) {
#pragma unroll
      for (int s = 0; s < 4; ++s) H = mma(Eb[s], Eb[s], H);
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int s = 0; s < 4; ++s) G[j] = mma(Eb[s], Db[j][s], G[j]);
    }
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = i; j < NT; ++j) {
        f4 acc = zero4();
#pragma unroll
        for (int k = 0; k < NT; ++k)
#pragma unroll
          for (int s = 0; s < 4; ++s) acc = mma(Ab[k][i][s], SA[k][j][s], acc);
        S[i][j] = acc;
      }

    // H <- H^{-1}, symmetrised so that it serves as its own transpose
    gj_inverse(H, m, g, c);
    {
      const f4 Ht = tile_transpose(H, tbuf, g, c);
#pragma unroll
      for (int v = 0; v < 4; ++v) H[v] = 0.5f * (H[v] + Ht[v]);
    }
    // K~ = H^{-1} G~
    f4 K[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      f4 acc = zero4();
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mma(H[s], G[j][s], acc);
      K[j] = acc;
    }

    // S~ <- Q~ + rho_c [D|0]'[D|0] + A~' S~ A~ - G~' K~ (tiles i <= j, then mirrored)
    {
      const float* Q = P.Q.at(b, t);
      const float* q = ws.qh + ((size_t)b * N1 + t) * n;
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = i; j < NT; ++j) {
          f4 acc = S[i][j] + load_qbar(Q, q, n, i, j, g, c);
#pragma unroll
          for (int s = 0; s < 4; ++s) acc = mma(-G[i][s], K[j][s], acc);
          if (p > 0) {
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = mma(Db[i][s], Db[j][s], acc);
          }
          S[i][j] = acc;
        }
#pragma unroll
      for (int i = 0; i < NT; ++i) {
        const f4 d = tile_transpose(S[i][i], tbuf, g, c);
#pragma unroll
        for (int v = 0; v < 4; ++v) S[i][i][v] = 0.5f * (S[i][i][v] + d[v]);
#pragma unroll
        for (int j = i + 1; j < NT; ++j) S[j][i] = tile_transpose(S[i][j], tbuf, g, c);
      }
      // corner of S~ (the constant of the value function) is not carried
      const int it_n = n >> 4, cn = n & 15;
#pragma unroll
      for (int i = 0; i < NT; ++i)
        if (i == it_n) {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (4 * g + v == cn && c == cn) S[i][i][v] = 0.f;
        }
    }

    // B~' tiles for F~ = A~ - B~ K~: lane holds B[16 i + c][4 g + v]
    f4 BT[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rb = 16 * i + c, cb = 4 * g + v;
        BT[i][v] = (rb < n && cb < m) ? Bm[rb * m + cb] : 0.f;
      }
    // F~ = A~ - B~ K~
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        f4 acc = Ab[i][j];
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(-BT[i][s], K[j][s], acc);
        Ab[i][j] = acc;
      }

    // stores: S_t, s_t ; K_t, k_t ; F_t, f_t
    store_value_wave<NT>(S, ws, b, t, N1, n, g, c);
    {
      float* Ko = ws.K + ((size_t)b * N + t) * m * n;
      float* ko = ws.k + ((size_t)b * N + t) * m;
#pragma unroll
      for (int j = 0; j < NT; ++j) store_hom(K[j], Ko, ko, n, m, 0, j, g, c);
      float* Fo = ws.F + ((size_t)b * N + t) * n * n;
      float* fo = ws.f + ((size_t)b * N + t) * n;
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) store_hom(Ab[i][j], Fo, fo, n, n, i, j, g, c);
    }
  }
  __syncthreads();  // F_t, f_t stored above are read back below

  // dx chain: x~ in column 0 of B-operand tiles; A operand = F~ read transposed
  f4 x[NT];
  {
    const float* x0 = ws.cdef + (size_t)b * N1 * n;
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * k + 4 * g + v;
        x[k][v] = (c != 0) ? 0.f : (row < n ? x0[row] : (row == n ? 1.f : 0.f));
      }
  }
  for (int t = 0; t < N; ++t) {
    const float* Fi = ws.F + ((size_t)b * N + t) * n * n;
    const float* fi = ws.f + ((size_t)b * N + t) * n;
    float* dxo = ws.dx + ((size_t)b * N1 + t) * n;
    f4 y[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 acc = zero4();
      const int row = 16 * i + c;                  // A operand: lane holds F~[row][4g + s]
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        f4 a;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int col = 16 * k + 4 * g + s;
          float e = 0.f;
          if (row < n) {
            if (col < n) e = Fi[row * n + col];
            else if (col == n) e = fi[row];
          } else if (row == n && col == n) {
            e = 1.f;
          }
          a[s] = e;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(a[s], x[k][s], acc);
      }
      y[i] = acc;
    }
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      if (c == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * k + 4 * g + v;
          if (row < n) dxo[row] = x[k][v];
        }
      }
      x[k] = y[k];
    }
  }
  {
    float* dxo = ws.dx + ((size_t)b * N1 + N) * n;
#pragma unroll
    for (int k = 0; k < NT; ++k)
      if (c == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * k + 4 * g + v;
          if (row < n) dxo[row] = x[k][v];
        }
      }
  }
}

}  // namespace mpxk
