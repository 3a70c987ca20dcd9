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
//
// Memory goes through raw buffer descriptors (one per array, in SGPRs) with a
// 32-bit offset per lane: an offset past the descriptor's record count reads 0
// and drops a store, which masks the ragged tile edges without branches or
// 64-bit addresses.
#pragma once

#include "common.h"
#include "k_lqr.h"

namespace mpxk {

typedef mpxk_f32x4 f4;
typedef __amdgpu_buffer_rsrc_t rsrc_t;

// byte offset past every record count used here: loads return 0, stores drop
constexpr int kOOB = 0x40000000;

__device__ __forceinline__ rsrc_t make_rsrc(const float* p, int nelem) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, nelem * 4, 0x00020000);
}
// (the b32 buffer builtins move raw 32-bit words: bit-cast, never convert)
__device__ __forceinline__ float bload(rsrc_t r, int idx, bool ok) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, ok ? idx * 4 : kOOB, 0, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, int idx, bool ok, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, ok ? idx * 4 : kOOB, 0, 0);
}

// An opaque copy: offsets computed from it inside a loop stay inside the loop
// (loop-invariant per-lane offsets hoisted out of the stage loop would hold
// hundreds of VGPRs for its whole length).
__device__ __forceinline__ int fresh(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

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

// Vector in tile-row layout: lane (g, c) holds v[16 it + 4 g + k] (0 past len).
template <int NT>
__device__ __forceinline__ void load_rows(f4 (&out)[NT], rsrc_t r, int g) {
#pragma unroll
  for (int it = 0; it < NT; ++it)
#pragma unroll
    for (int k = 0; k < 4; ++k) out[it][k] = bload(r, 16 * it + 4 * g + k, true);
}

// C-layout tile (it, jt) of an n x n row-major matrix (descriptor of n*n records;
// rows past n fall outside it), with `col_n` added in column n and `corner` at
// (n, n): the homogeneous A~ = [[A, c], [0, 1]] and Q~ = [[Q, q], [q', 0]].
__device__ __forceinline__ f4 load_hom(rsrc_t M, const f4& col_n, float row_n, float corner, int n, int it, int jt,
                                       int g, int c) {
  f4 x;
  const int col = 16 * jt + c;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = 16 * it + 4 * g + v;
    const float a = bload(M, row * n + col, col < n);
    const float e = (col == n && row < n) ? col_n[v] : 0.f;
    const float f = (row == n) ? (col < n ? row_n : (col == n ? corner : 0.f)) : 0.f;
    x[v] = a + e + f;
  }
  return x;
}

// Store a C-layout tile of an (n+1)-wide homogeneous matrix: the leading n
// columns to M (row-major, ld n) and column n to vec; rows past the
// descriptors' record counts are dropped.
__device__ __forceinline__ void store_hom(const f4& x, rsrc_t M, rsrc_t vec, int n, int it, int jt, int g, int c) {
  const int col = 16 * jt + c;
#pragma unroll
  for (int v = 0; v < 4; ++v) bstore(M, (16 * it + 4 * g + v) * n + col, col < n, x[v]);
  if (jt == (n >> 4)) {   // the tile holding column n (uniform)
#pragma unroll
    for (int v = 0; v < 4; ++v) bstore(vec, 16 * it + 4 * g + v, col == n, x[v]);
  }
}

template <int NT>
__device__ __forceinline__ void load_qbar(f4 (&S)[NT][NT], const float* Q, const float* q, int n, int g, int c,
                                          bool upper_only) {
  const rsrc_t rQ = make_rsrc(Q, n * n), rq = make_rsrc(q, n);
  f4 qrow[NT];
  load_rows<NT>(qrow, rq, g);
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
      if (!upper_only || j >= i) {
        const float qc = bload(rq, 16 * j + c, true);
        S[i][j] = load_hom(rQ, qrow[i], qc, 0.f, n, i, j, g, c);
      }
}

template <int NT>
__device__ __forceinline__ void store_value_wave(const f4 (&S)[NT][NT], const Ws<float>& ws, int b, int t, int N1,
                                                 int n, int g, int c) {
  const rsrc_t rS = make_rsrc(ws.S + ((size_t)b * N1 + t) * n * n, n * n);
  const rsrc_t rs = make_rsrc(ws.s + ((size_t)b * N1 + t) * n, n);
#pragma unroll
  for (int it = 0; it < NT; ++it)
#pragma unroll
    for (int jt = 0; jt < NT; ++jt) store_hom(S[it][jt], rS, rs, n, it, jt, g, c);
}

// A-operand tiles of F~ for one forward step: lane (g, c) of tile (i, k) holds
// F~[16 i + c][16 k + 4 g + s] in register s.
template <int NT>
__device__ __forceinline__ void load_fbar_t(f4 (&Fa)[NT][NT], const Ws<float>& ws, int b, int t, int N, int n,
                                            int g, int c) {
  const rsrc_t rF = make_rsrc(ws.F + ((size_t)b * N + t) * n * n, n * n);
  const rsrc_t rf = make_rsrc(ws.f + ((size_t)b * N + t) * n, n);
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    const int row = 16 * i + c;
    const float fr = bload(rf, row, true);
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int col = 16 * k + 4 * g + s;
        const float e = bload(rF, row * n + col, col < n);
        Fa[i][k][s] = e + ((row < n && col == n) ? fr : 0.f) + ((row == n && col == n) ? 1.f : 0.f);
      }
  }
}

// One stage's data in registers (C-layout tiles).
template <int NT>
struct WaveStage {
  f4 A[NT][NT];   // A~ = [[A, c], [0, 1]]
  f4 B[NT];       // B~ tiles (i, 0)
  f4 BT[NT];      // B~' tiles (0, i): lane (g, c) holds B[16 i + c][4 g + v]
  f4 D[NT];       // sqrt(rho_c) [D | 0]
  f4 E;           // sqrt(rho_c) E
  f4 H;           // R + diag(rd), identity in the padding
  f4 r;           // r: lane (g, c) holds r[4 g + v]
};

template <int NT>
__device__ __forceinline__ void load_wave_stage(WaveStage<NT>& d, const Prob<float>& P, const Ws<float>& ws,
                                                const LqrIn<float>& in, int b, int t, float sr, int g, int c) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const rsrc_t rA = make_rsrc(in.A.at(b, t), n * n);
  const rsrc_t rB = make_rsrc(in.Bm.at(b, t), n * m);
  {
    f4 cvec[NT];
    load_rows<NT>(cvec, make_rsrc(ws.cdef + ((size_t)b * N1 + t + 1) * n, n), g);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) d.A[i][j] = load_hom(rA, cvec[i], 0.f, 1.f, n, i, j, g, c);
  }
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      d.B[i][v] = bload(rB, (16 * i + 4 * g + v) * m + c, c < m);
      d.BT[i][v] = bload(rB, (16 * i + c) * m + 4 * g + v, 4 * g + v < m);
    }
  if (p > 0) {
    const rsrc_t rD = make_rsrc(P.D.at(b, t), p * n), rE = make_rsrc(P.E.at(b, t), p * m);
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) d.D[j][v] = sr * bload(rD, (4 * g + v) * n + 16 * j + c, 16 * j + c < n);
#pragma unroll
    for (int v = 0; v < 4; ++v) d.E[v] = sr * bload(rE, (4 * g + v) * m + c, c < m);
  } else {
#pragma unroll
    for (int j = 0; j < NT; ++j) d.D[j] = zero4();
    d.E = zero4();
  }
  const rsrc_t rR = make_rsrc(P.R.at(b, t), m * m);
  const rsrc_t rd = make_rsrc(ws.rd + ((size_t)b * N + t) * m, m);
  const rsrc_t rr = make_rsrc(ws.rh + ((size_t)b * N + t) * m, m);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = 4 * g + v;
    const bool in_m = row < m && c < m;
    d.H[v] = bload(rR, row * m + c, c < m) + bload(rd, row, row == c) + ((!in_m && row == c) ? 1.f : 0.f);
    d.r[v] = bload(rr, row, true);
  }
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
  load_qbar<NT>(S, P.QN.at(b, 0), ws.qh + ((size_t)b * N1 + N) * n, n, g, c, false);
  store_value_wave<NT>(S, ws, b, N, N1, n, g, c);

  WaveStage<NT> cur;
  load_wave_stage<NT>(cur, P, ws, in, b, N - 1, sr, g, c);
  for (int t = N - 1; t >= 0; --t) {
    const int g = fresh(lane >> 4), c = fresh(lane & 15);
    // SB = S~ B~ ; H = R + diag(rd) + rho_c E'E + B~' SB
    f4 SB[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 acc = zero4();
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(S[k][i][s], cur.B[k][s], acc);
      SB[i] = acc;
    }
    f4 H = cur.H;
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) H = mma(cur.B[k][s], SB[k][s], H);
    // G~ = [0 | r] + rho_c E'[D | 0] + B~' S~ A~ ; Pm = A~' S~ A~ (tiles i <= j),
    // one column tile of S~ A~ at a time
    f4 G[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) G[j][v] = (16 * j + c == n) ? cur.r[v] : 0.f;
    if (p > 0) {
#pragma unroll
      for (int s = 0; s < 4; ++s) H = mma(cur.E[s], cur.E[s], H);
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int s = 0; s < 4; ++s) G[j] = mma(cur.E[s], cur.D[j][s], G[j]);
    }
    f4 Pm[NT][NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      f4 SAc[NT];
#pragma unroll
      for (int i = 0; i < NT; ++i) {
        f4 acc = zero4();
#pragma unroll
        for (int k = 0; k < NT; ++k)
#pragma unroll
          for (int s = 0; s < 4; ++s) acc = mma(S[k][i][s], cur.A[k][j][s], acc);
        SAc[i] = acc;
      }
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) G[j] = mma(cur.B[k][s], SAc[k][s], G[j]);
#pragma unroll
      for (int i = 0; i <= j; ++i) {
        f4 acc = zero4();
#pragma unroll
        for (int k = 0; k < NT; ++k)
#pragma unroll
          for (int s = 0; s < 4; ++s) acc = mma(cur.A[k][i][s], SAc[k][s], acc);
        Pm[i][j] = acc;
      }
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
      f4 Qb[NT][NT];
      load_qbar<NT>(Qb, P.Q.at(b, t), ws.qh + ((size_t)b * N1 + t) * n, n, g, c, true);
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = i; j < NT; ++j) {
          f4 acc = Pm[i][j] + Qb[i][j];
#pragma unroll
          for (int s = 0; s < 4; ++s) acc = mma(-G[i][s], K[j][s], acc);
          if (p > 0) {
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = mma(cur.D[i][s], cur.D[j][s], acc);
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

    // F~ = A~ - B~ K~ (in place of A~)
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        f4 acc = cur.A[i][j];
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(-cur.BT[i][s], K[j][s], acc);
        cur.A[i][j] = acc;
      }

    // the next stage's data is requested before this stage's stores, so that
    // waiting for it never waits for the stores (vmcnt counts both, in order)
    WaveStage<NT> nxt;
    load_wave_stage<NT>(nxt, P, ws, in, b, t > 0 ? t - 1 : 0, sr, g, c);

    // stores: S_t, s_t ; K_t, k_t ; F_t, f_t
    store_value_wave<NT>(S, ws, b, t, N1, n, g, c);
    {
      const rsrc_t rK = make_rsrc(ws.K + ((size_t)b * N + t) * m * n, m * n);
      const rsrc_t rk = make_rsrc(ws.k + ((size_t)b * N + t) * m, m);
#pragma unroll
      for (int j = 0; j < NT; ++j) store_hom(K[j], rK, rk, n, 0, j, g, c);
      const rsrc_t rF = make_rsrc(ws.F + ((size_t)b * N + t) * n * n, n * n);
      const rsrc_t rf = make_rsrc(ws.f + ((size_t)b * N + t) * n, n);
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) store_hom(cur.A[i][j], rF, rf, n, i, j, g, c);
    }
    cur = nxt;
  }
  __syncthreads();  // F_t, f_t stored above are read back below

  // dx chain: x~ in column 0 of B-operand tiles; A operand = F~ read transposed,
  // the next stage's tiles fetched while the current one is applied
  f4 x[NT];
  {
    const rsrc_t r0 = make_rsrc(ws.cdef + (size_t)b * N1 * n, n);
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * k + 4 * g + v;
        x[k][v] = bload(r0, row, c == 0) + ((c == 0 && row == n) ? 1.f : 0.f);
      }
  }
  f4 Fa[NT][NT];
  load_fbar_t<NT>(Fa, ws, b, 0, N, n, g, c);
  for (int t = 0; t < N; ++t) {
    const int g = fresh(lane >> 4), c = fresh(lane & 15);
    f4 Fn[NT][NT];
    if (t + 1 < N) load_fbar_t<NT>(Fn, ws, b, t + 1, N, n, g, c);
    f4 y[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 acc = zero4();
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(Fa[i][k][s], x[k][s], acc);
      y[i] = acc;
    }
    const rsrc_t rdx = make_rsrc(ws.dx + ((size_t)b * N1 + t) * n, n);
#pragma unroll
    for (int k = 0; k < NT; ++k) {
#pragma unroll
      for (int v = 0; v < 4; ++v) bstore(rdx, 16 * k + 4 * g + v, c == 0, x[k][v]);
      x[k] = y[k];
    }
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int k = 0; k < NT; ++k) Fa[i][k] = Fn[i][k];
  }
  {
    const rsrc_t rdx = make_rsrc(ws.dx + ((size_t)b * N1 + N) * n, n);
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) bstore(rdx, 16 * k + 4 * g + v, c == 0, x[k][v]);
  }
}

}  // namespace mpxk
