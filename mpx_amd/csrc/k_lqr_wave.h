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
//   Q~ = [[Q, q], [q', 0]],  G~ = [G | g],  K~ = [K | k],
// and one Riccati step
//   H  = R + diag(rd) + rho_c E'E + B~' S~ B~
//   G~ = B~' S~ A~ + rho_c E' [D | 0] + [0 | r]
//   W~ = Z G~,  K~ = Z' W~   (H = L L', Z = L^{-1}, H^{-1} = Z' Z)
//   S~ <- Q~ + rho_c [D|0]'[D|0] + A~' S~ A~ - W~' W~ ,
//   F~ = A~ - B~ K~ = [[A - B K, c - B k], [0, 1]]
// is the affine step of k_lqr.h (S_t = Q + rho_c D'D + A'SA - G'K,
// s_t = q + A'(S c + s) - G'k, F_t = A - BK, f_t = c - Bk) with the vectors
// riding in column n.  The corner of S~ is reset to 0 every step.
//
// The m x m factor is a Cholesky factorisation with the forward substitution
// of Z = L^{-1} in the same sweep over one tile (broadcasts of the pivot column
// by DPP and of the pivot row by ds_bpermute); H^{-1} is applied as Z' (Z G~)
// on the matrix cores and never formed (in FP32 forming it first costs a
// factor of about cond(H)^(1/2) in accuracy); the symmetric value tiles are
// mirrored through a 1 KB LDS transpose.  The forward pass
// (k_fwd_wave) runs on the matrix cores too, with x~ as column 0 of a
// B-operand tile: every product's result tile is again in that layout, so
// x~_{t+1} = A~ x~ + B~ du and du = -K~ x~ chain from registers and
// F = A - B K is never formed; the costates lam = S~ x~ follow in parallel over
// the stages (k_costate_wave).
//
// Memory goes through raw buffer descriptors, one per matrix and stage, whose
// record count ends at the matrix's last row.  The lane's byte offset of
// element (row, col) is (row ld + col) * 4: a row past the matrix lands past
// the record count, and a column past it gets a lane part set past every
// record count.  Out-of-range loads return 0 and out-of-range stores are
// dropped, so the ragged tile edges cost no branches.  The lane parts are
// recomputed every stage from an opaque copy of the lane id: hoisted out of
// the stage loop, the hundreds of loop-invariant offsets would occupy VGPRs
// for its whole length.  Each stage's data (including Q~) is requested before
// the previous stage's results are stored: vmcnt counts loads and stores in
// issue order, so waiting for the data never waits for the stores.
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
// An opaque copy: offsets computed from it inside a loop stay inside the loop.
__device__ __forceinline__ int fresh(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
// (the b32 buffer builtins move raw 32-bit words: bit-cast, never convert)
__device__ __forceinline__ float rload(rsrc_t r, int voff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
}
// four consecutive floats at a 16-byte aligned offset (all in range or all past it)
__device__ __forceinline__ f4 rload4(rsrc_t r, int voff) {
  return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
}
__device__ __forceinline__ void rstore(rsrc_t r, int voff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, 0, 0);
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

// Lane j of each 16-lane row of the wave, broadcast to the row (DPP row_newbcast).
template <int J>
__device__ __forceinline__ float row_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + J, 0xf, 0xf, false));
}

// x from lane 16 (J / 4) + c of the wave (row J of a C-layout tile, for every
// row of 16 lanes); c4 = 4 c.
template <int J>
__device__ __forceinline__ float tile_row(float x, int c4) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(c4 + 64 * (J >> 2), __float_as_int(x)));
}

// y + bcast_J(x) s on the rows of 16 lanes set in RM (lane J of each row
// broadcast to the row); the other rows keep y.
template <int J, int RM>
__device__ __forceinline__ float fma_bcast(float x, float s, float y) {
  if constexpr (RM == 0xf) {
    return fmaf(row_bcast<J>(x), s, y);
  } else {
    const float xb = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x150 + J, RM, 0xf, false));
    return fmaf(xb, s, y);
  }
}

// First row of 16 lanes g whose row 4 g + v of a tile lies below row J.
__host__ __device__ constexpr int chol_gmin(int J, int v) { return J >= v ? (J - v) / 4 + 1 : 0; }

// Z[r][:] += H[r][J] nzj for the rows r = 4 g + v > J of register v.
template <int J, int V>
__device__ __forceinline__ float chol_zrow(float h, float z, float nzj) {
  constexpr int gm = chol_gmin(J, V);
  if constexpr (gm < 4) return fma_bcast<J, (0xf << gm) & 0xf>(h, nzj, z);
  return z;
}

// Step J of chol_lower_inv.  Lane (g, c) holds H[4 g + v][c] and Z[4 g + v][c]
// in register v.  The pivot a = |H[J][J]| comes from one lane (a scalar read)
// and is raised to fl = floor |H0[J][J]| (H0: the tile before the
// factorisation; a NaN pivot stays NaN); id = 1 / L[J][J] = a^{-1/2}.
// H[r][J] reaches the lanes of row r by a DPP broadcast from lane J of the
// lane's row of 16, H[J][c] and Z[J][c] from row J (tile_row).  The trailing
// matrix takes H[r][c] -= H[r][J] H[J][c] / a, i.e. H - L[:, J] L[:, J]'; this
// update runs over every entry: the entries in a row or a column <= J are
// never read again.  The forward substitution of L Z = I advances by the same
// column: Z[J][:] *= id, then Z[r][:] -= L[r][J] Z[J][:] for r > J only (DPP
// row masks select the rows of 16 lanes whose row 4 g + v lies below J).
template <int J>
__device__ __forceinline__ void chol_step(f4& h, f4& z, const float (&fl)[16], int c4) {
  constexpr int jg = J >> 2, jv = J & 3;
  const float a = fabsf(rlane(h[jv], jg * 16 + J));
  const float piv = fl[J] > a ? fl[J] : a;
  const float id = __builtin_amdgcn_rsqf(piv);
  const float nid2 = -(id * id);
  const float hJ = tile_row<J>(h[jv], c4);   // H[J][c]
  const float zJ = tile_row<J>(z[jv], c4);   // Z[J][c] before the step
  const float nlc = hJ * nid2;               // -H[J][c] / a
  const float nzj = zJ * nid2;               // -L[J][J]^{-1} Z[J][c] (after the step)
  // (Z first: it reads column J of H before the trailing update)
  z[0] = chol_zrow<J, 0>(h[0], z[0], nzj);
  z[1] = chol_zrow<J, 1>(h[1], z[1], nzj);
  z[2] = chol_zrow<J, 2>(h[2], z[2], nzj);
  z[3] = chol_zrow<J, 3>(h[3], z[3], nzj);
  // row J of Z: Z[J][c] id into the row of 16 lanes jg of register jv
  z[jv] = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(z[jv]), __float_as_int(zJ * id), 0xE4,
                                                     1 << jg, 0xf, false));
#pragma unroll
  for (int v = 0; v < 4; ++v)
    if (chol_gmin(J, v) < 4) h[v] = fma_bcast<J, 0xf>(h[v], nlc, h[v]);
}

template <int J, int NS>
__device__ __forceinline__ void chol_steps(f4& h, f4& z, const float (&fl)[16], int c4) {
  if constexpr (J < NS) {
    chol_step<J>(h, z, fl, c4);
    chol_steps<J + 1, NS>(h, z, fl, c4);
  }
}

// Z = L^{-1} for the symmetric tile H (m x m in the top-left corner, the
// identity elsewhere; H = L L' with floored pivots), in one sweep of NS >= m
// steps (straight-line code that the scheduler can place among independent
// matrix products; past the m x m block every step is the identity, so the
// sweep may stop at any NS >= m).  H is consumed.
template <int NS>
__device__ __forceinline__ f4 chol_lower_inv(f4& h, float floor, int g, int c) {
  float fl[16];
#pragma unroll
  for (int j = 0; j < NS; ++j) fl[j] = floor * fabsf(rlane(h[j & 3], (j >> 2) * 16 + j));
  f4 z;
#pragma unroll
  for (int v = 0; v < 4; ++v) z[v] = (4 * g + v == c) ? 1.f : 0.f;
  chol_steps<0, NS>(h, z, fl, 4 * c);
  return z;
}

// Loop-invariant lane offsets (bytes) of one instance's shapes.
template <int NT>
struct LaneOff {
  int nn[NT];   // ld n, column 16 jt + c < n   : (4 g n + 16 jt + c) * 4
  int mm;       // ld m, column c < m           : (4 g m + c) * 4
  int bt[4];    // B' tiles, column 4 g + v < m : (c m + 4 g + v) * 4
  int g16;      // vector in tile-row layout    : 16 g
  int c4;       // vector along the columns     : 4 c
  int hom;      // column n to its vector       : 16 g on lane c == n % 16
  int col0;     // column 0 of a B operand      : 16 g on lane c == 0
  int cn;       // n % 16
};

template <int NT>
__device__ __forceinline__ LaneOff<NT> lane_offsets(int n, int m, int g, int c) {
  LaneOff<NT> L;
#pragma unroll
  for (int jt = 0; jt < NT; ++jt) L.nn[jt] = (16 * jt + c < n) ? (4 * g * n + 16 * jt + c) * 4 : kOOB;
  L.mm = (c < m) ? (4 * g * m + c) * 4 : kOOB;
#pragma unroll
  for (int v = 0; v < 4; ++v) L.bt[v] = (4 * g + v < m) ? (c * m + 4 * g + v) * 4 : kOOB;
  L.g16 = 16 * g;
  L.c4 = 4 * c;
  L.cn = n & 15;
  L.hom = (c == L.cn) ? 16 * g : kOOB;
  L.col0 = (c == 0) ? 16 * g : kOOB;
  return L;
}

// C-layout tile (it, jt) of a row-major matrix with ld n (descriptor rM).
template <int NT>
__device__ __forceinline__ f4 tile_nn(rsrc_t rM, int n, int it, int jt, const LaneOff<NT>& L) {
  f4 x;
#pragma unroll
  for (int v = 0; v < 4; ++v) x[v] = rload(rM, L.nn[jt] + (16 * it + v) * n * 4);
  return x;
}

// vector entries v[16 it + 4 g + k] (0 past the descriptor) in tile-row layout
template <int NT>
__device__ __forceinline__ f4 rows_vec(rsrc_t r, int it, const LaneOff<NT>& L) {
  f4 x;
#pragma unroll
  for (int k = 0; k < 4; ++k) x[k] = rload(r, L.g16 + (16 * it + k) * 4);
  return x;
}

// Store tile (it, jt) of an (n+1)-wide homogeneous matrix: the first n columns
// to rM (ld n), column n to rvec; rows past the descriptors are dropped.
template <int NT>
__device__ __forceinline__ void store_hom(const f4& x, rsrc_t rM, rsrc_t rvec, int n, int it, int jt,
                                          const LaneOff<NT>& L) {
#pragma unroll
  for (int v = 0; v < 4; ++v) rstore(rM, L.nn[jt] + (16 * it + v) * n * 4, x[v]);
  if (jt == (n >> 4)) {   // the tile holding column n (uniform)
#pragma unroll
    for (int v = 0; v < 4; ++v) rstore(rvec, L.hom + (16 * it + v) * 4, x[v]);
  }
}

// One stage's data in registers (C-layout tiles).
template <int NT>
struct WaveStage {
  f4 A[NT][NT];   // A~ = [[A, c], [0, 1]]
  f4 Q[NT][NT];   // Q~ = [[Q, q], [q', 0]], tiles i <= j
  f4 B[NT];       // B~ tiles (i, 0)
  f4 D[NT];       // sqrt(rho_c) [D | 0]
  f4 E;           // sqrt(rho_c) E
  f4 H;           // R + diag(rd), identity in the padding
  f4 r;           // r: lane (g, c) holds r[4 g + v]
  // with stage-invariant Q, B, R (held by the kernel) only the stage's vectors:
  f4 qrow[NT];    // q[16 i + 4 g + v]: column n of tile row i
  float qcol[NT]; // q[16 j + c]: row n of tile column j
  f4 rdv;         // rd[4 g + v]
};

// Q~ tiles (i <= j, or all) of Q (ld n) and q
template <int NT>
__device__ __forceinline__ void load_qbar(f4 (&S)[NT][NT], const float* Q, const float* q, int n, int g, int c,
                                          bool upper_only, const LaneOff<NT>& L) {
  const int tn = n >> 4;
  const rsrc_t rQ = make_rsrc(Q, n * n), rq = make_rsrc(q, n);
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      if (upper_only && j < i) continue;
      f4 x = tile_nn<NT>(rQ, n, i, j, L);
      if (j == tn) {   // column n: q[16 i + 4 g + v]
        const f4 qr = rows_vec<NT>(rq, i, L);
#pragma unroll
        for (int v = 0; v < 4; ++v) x[v] += (c == L.cn) ? qr[v] : 0.f;
      }
      if (i == tn) {   // row n: q[16 j + c]
        const float qc = rload(rq, L.c4 + 64 * j);
#pragma unroll
        for (int v = 0; v < 4; ++v) x[v] += (4 * g + v == L.cn) ? qc : 0.f;
      }
      S[i][j] = x;
    }
}

// CQ: Q, B and R do not vary over the stages and the kernel holds them; the
// stage brings only q and rd for them.
template <int NT, bool CQ>
__device__ __forceinline__ void load_wave_stage(WaveStage<NT>& d, const Prob<float>& P, const Ws<float>& ws,
                                                const LqrIn<float>& in, int b, int t, float sr, int g, int c,
                                                const LaneOff<NT>& L) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1, tn = n >> 4;
  const rsrc_t rA = make_rsrc(in.A.at(b, t), n * n);
  const rsrc_t rB = make_rsrc(in.Bm.at(b, t), n * m);
  const rsrc_t rc = make_rsrc(ws.cdef + ((size_t)b * N1 + t + 1) * n, n);
#pragma unroll
  for (int i = 0; i < NT; ++i) {
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      f4 x = tile_nn<NT>(rA, n, i, j, L);
      if (j == tn) {   // column n: c ; corner 1
        const f4 cv = rows_vec<NT>(rc, i, L);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          x[v] += (c == L.cn) ? cv[v] : 0.f;
          if (i == tn) x[v] += (c == L.cn && 4 * g + v == L.cn) ? 1.f : 0.f;
        }
      }
      d.A[i][j] = x;
    }
  }
  if constexpr (CQ) {
    const rsrc_t rq = make_rsrc(ws.qh + ((size_t)b * N1 + t) * n, n);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      d.qrow[i] = rows_vec<NT>(rq, i, L);
      d.qcol[i] = rload(rq, L.c4 + 64 * i);
    }
  } else {
    load_qbar<NT>(d.Q, P.Q.at(b, t), ws.qh + ((size_t)b * N1 + t) * n, n, g, c, true, L);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) d.B[i][v] = rload(rB, L.mm + (16 * i + v) * m * 4);
  }
  if (p > 0) {
    const rsrc_t rD = make_rsrc(P.D.at(b, t), p * n), rE = make_rsrc(P.E.at(b, t), p * m);
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) d.D[j][v] = sr * rload(rD, L.nn[j] + v * n * 4);
#pragma unroll
    for (int v = 0; v < 4; ++v) d.E[v] = sr * rload(rE, L.mm + v * m * 4);
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
    const float dv = rload(rd, L.g16 + v * 4);
    if constexpr (CQ) d.rdv[v] = dv;
    else d.H[v] = rload(rR, L.mm + v * m * 4) + (row == c ? (row < m ? dv : 1.f) : 0.f);
    d.r[v] = rload(rr, L.g16 + v * 4);
  }
}

template <int NT>
__device__ __forceinline__ void store_value_wave(const f4 (&S)[NT][NT], const Ws<float>& ws, int b, int t, int N1,
                                                 int n, const LaneOff<NT>& L) {
  const rsrc_t rS = make_rsrc(ws.S + ((size_t)b * N1 + t) * n * n, n * n);
  const rsrc_t rs = make_rsrc(ws.s + ((size_t)b * N1 + t) * n, n);
#pragma unroll
  for (int it = 0; it < NT; ++it)
#pragma unroll
    for (int jt = 0; jt < NT; ++jt) store_hom<NT>(S[it][jt], rS, rs, n, it, jt, L);
}

// segments == 1, FP32, n + 1 <= 16 NT, m <= 16, p <= 16: backward Riccati of
// one instance per wave; stores S~_t and K~_t for k_fwd_wave.
// CQ: Q, B, R stage-invariant (views of stage stride 0): held in registers
// for the whole recursion instead of being re-read every stage.
template <int NT, bool CQ>
__global__ void __launch_bounds__(64, 2) k_lqr_wave(Prob<float> P, Prm<float> prm, Iter<float> it_unused,
                                                    Ws<float> ws, LqrIn<float> in) {
  (void)it_unused;
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  __shared__ float tbuf[16 * 17];
  const float sr = sqrtf(prm.rho_c);
  const LaneOff<NT> L = lane_offsets<NT>(n, m, g, c);

  f4 S[NT][NT];
  load_qbar<NT>(S, P.QN.at(b, 0), ws.qh + ((size_t)b * N1 + N) * n, n, g, c, false, L);
  store_value_wave<NT>(S, ws, b, N, N1, n, L);

  // stage-invariant Q (tiles i <= j, no q), B~ and R (CQ)
  f4 Qc[NT][NT], Bc[NT], Rc;
  if constexpr (CQ) {
    const rsrc_t rQ = make_rsrc(P.Q.at(b, 0), n * n), rB = make_rsrc(in.Bm.at(b, 0), n * m);
    const rsrc_t rR = make_rsrc(P.R.at(b, 0), m * m);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = i; j < NT; ++j) Qc[i][j] = tile_nn<NT>(rQ, n, i, j, L);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) Bc[i][v] = rload(rB, L.mm + (16 * i + v) * m * 4);
#pragma unroll
    for (int v = 0; v < 4; ++v) Rc[v] = rload(rR, L.mm + v * m * 4);
  }
  WaveStage<NT> cur;
  load_wave_stage<NT, CQ>(cur, P, ws, in, b, N - 1, sr, g, c, L);
  for (int t = N - 1; t >= 0; --t) {
    // lane conditions and offsets from opaque copies of the lane coordinates:
    // recomputed every stage instead of held (as lane masks in scalar
    // registers) across the loop
    const int gf = fresh(g), cf = fresh(c);
    const LaneOff<NT> Ls = lane_offsets<NT>(n, m, gf, cf);
    // SB = S~ B~ ; H = R + diag(rd) + rho_c E'E + B~' SB
    f4 SB[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 acc = zero4();
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(S[k][i][s], CQ ? Bc[k][s] : cur.B[k][s], acc);
      SB[i] = acc;
    }
    f4 H;
    if constexpr (CQ) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 4 * gf + v;
        H[v] = Rc[v] + (row == cf ? (row < m ? cur.rdv[v] : 1.f) : 0.f);
      }
    } else {
      H = cur.H;
    }
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) H = mma(CQ ? Bc[k][s] : cur.B[k][s], SB[k][s], H);
    // G~ = [0 | r] + rho_c E'[D | 0] + B~' S~ A~ ; Pm = Q~ + A~' S~ A~ (tiles
    // i <= j), one column tile of S~ A~ at a time
    f4 G[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) G[j][v] = (16 * j + cf == n) ? cur.r[v] : 0.f;
    if (p > 0) {
#pragma unroll
      for (int s = 0; s < 4; ++s) H = mma(cur.E[s], cur.E[s], H);
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int s = 0; s < 4; ++s) G[j] = mma(cur.E[s], cur.D[j][s], G[j]);
    }
    // Z = L^{-1} (H = L L'); Z' in C layout is the A operand of Z G~
    const f4 Z = chol_lower_inv<16>(H, prm.hfloor, gf, cf);
    const f4 Zt = tile_transpose(Z, tbuf, gf, cf);
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
        for (int s = 0; s < 4; ++s) G[j] = mma(CQ ? Bc[k][s] : cur.B[k][s], SAc[k][s], G[j]);
#pragma unroll
      for (int i = 0; i <= j; ++i) {
        f4 acc;
        if constexpr (CQ) {   // Q~ = [[Q, q], [q', 0]] from the held Q and this stage's q
          const int tn = n >> 4;
          acc = Qc[i][j];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            if (j == tn) acc[v] += (cf == Ls.cn) ? cur.qrow[i][v] : 0.f;
            if (i == tn) acc[v] += (4 * gf + v == Ls.cn) ? cur.qcol[j] : 0.f;
          }
        } else {
          acc = cur.Q[i][j];
        }
#pragma unroll
        for (int k = 0; k < NT; ++k)
#pragma unroll
          for (int s = 0; s < 4; ++s) acc = mma(cur.A[k][i][s], SAc[k][s], acc);
        Pm[i][j] = acc;
      }
    }

    // W~ = Z G~ (in G's registers), K~ = Z' W~
    f4 K[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      f4 w = zero4();
#pragma unroll
      for (int s = 0; s < 4; ++s) w = mma(Zt[s], G[j][s], w);
      G[j] = w;
      f4 acc = zero4();
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mma(Z[s], w[s], acc);
      K[j] = acc;
    }

    // S~ <- Pm + rho_c [D|0]'[D|0] - W~' W~ (tiles i <= j, then mirrored)
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = i; j < NT; ++j) {
        f4 acc = Pm[i][j];
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(-G[i][s], G[j][s], acc);
        if (p > 0) {
#pragma unroll
          for (int s = 0; s < 4; ++s) acc = mma(cur.D[i][s], cur.D[j][s], acc);
        }
        S[i][j] = acc;
      }
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const f4 d = tile_transpose(S[i][i], tbuf, gf, cf);
#pragma unroll
      for (int v = 0; v < 4; ++v) S[i][i][v] = 0.5f * (S[i][i][v] + d[v]);
#pragma unroll
      for (int j = i + 1; j < NT; ++j) S[j][i] = tile_transpose(S[i][j], tbuf, gf, cf);
    }
    {  // the corner of S~ (the constant of the value function) is not carried
      const int tn = n >> 4;
#pragma unroll
      for (int i = 0; i < NT; ++i)
        if (i == tn) {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (4 * gf + v == Ls.cn && cf == Ls.cn) S[i][i][v] = 0.f;
        }
    }

    // the next stage's data is requested before this stage's stores
    WaveStage<NT> nxt;
    // (the stage index is pinned to a scalar register: selected into a vector
    // register it makes every descriptor built from it look divergent)
    load_wave_stage<NT, CQ>(nxt, P, ws, in, b, __builtin_amdgcn_readfirstlane(t > 0 ? t - 1 : 0), sr, gf, cf, Ls);

    // stores: S_t, s_t ; K_t, k_t
    store_value_wave<NT>(S, ws, b, t, N1, n, Ls);
    {
      const rsrc_t rK = make_rsrc(ws.K + ((size_t)b * N + t) * m * n, m * n);
      const rsrc_t rk = make_rsrc(ws.k + ((size_t)b * N + t) * m, m);
#pragma unroll
      for (int j = 0; j < NT; ++j) store_hom<NT>(K[j], rK, rk, n, 0, j, Ls);
    }
    cur = nxt;
  }
}

// A-operand tiles of an (n+1)-column homogeneous matrix with row-major block M
// (ld n, nrows rows) and column n = colv: lane (g, c) of tile (i, k) holds
// M~[16 i + c][16 k + 4 g + s] in register s, with 1 at (n, n) when corner.
// off[k][s] is the lane's byte offset of column 16 k + 4 g + s within a row
// (kOOB past column n - 1).
template <int NT, int RT>
__device__ __forceinline__ void load_aop(f4 (&out)[RT][NT], const float* M, const float* colv, int n, int nrows,
                                         bool corner, int g, int c, const int (&off)[NT][4], const LaneOff<NT>& L) {
  const rsrc_t rM = make_rsrc(M, nrows * n), rv = make_rsrc(colv, nrows);
  // n % 4 == 0: a lane's four columns 16 k + 4 g .. + 3 of a row are contiguous,
  // 16-byte aligned, and either all below n or all past it: one 128-bit load
  const bool wide = (n & 3) == 0;
#pragma unroll
  for (int i = 0; i < RT; ++i) {
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      if (wide) {
        out[i][k] = rload4(rM, off[k][0] + 16 * i * n * 4);
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) out[i][k][s] = rload(rM, off[k][s] + 16 * i * n * 4);
      }
    }
    const float cv = rload(rv, L.c4 + 64 * i);   // colv[16 i + c]
    const float cc = (corner && 16 * i + c == n) ? 1.f : 0.f;
    // column n by a per-lane test (a test on the tile index alone lets the
    // compiler turn the update into a run-time index into the tile array)
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) out[i][k][s] += (16 * k + 4 * g + s == n) ? cv + cc : 0.f;
  }
}

// One forward step's A-operand tiles: A~_t, K~_t and (unless the kernel holds
// a stage-invariant B, CB) B~_t.
template <int NT, bool CB>
__device__ __forceinline__ void load_chain_stage(f4 (&A)[NT][NT], f4 (&K)[1][NT], f4 (&BT)[NT], const Ws<float>& ws,
                                                 const LqrIn<float>& in, int b, int t, int N, int n, int m, int g,
                                                 int c, const int (&off)[NT][4], const LaneOff<NT>& L) {
  const int N1 = N + 1;
  load_aop<NT, NT>(A, in.A.at(b, t), ws.cdef + ((size_t)b * N1 + t + 1) * n, n, n, true, g, c, off, L);
  load_aop<NT, 1>(K, ws.K + ((size_t)b * N + t) * m * n, ws.k + ((size_t)b * N + t) * m, n, m, false, g, c, off,
                  L);
  if constexpr (CB) return;
  const rsrc_t rB = make_rsrc(in.Bm.at(b, t), n * m);
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int s = 0; s < 4; ++s) BT[i][s] = rload(rB, L.bt[s] + 16 * i * m * 4);   // B[16 i + c][4 g + s]
}

// lane's byte offsets of columns 16 k + 4 g + s within a row of ld n (kOOB
// past column n - 1), for A-operand tiles
template <int NT>
__device__ __forceinline__ void aop_offsets(int (&off)[NT][4], int n, int g, int c) {
#pragma unroll
  for (int k = 0; k < NT; ++k)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int col = 16 * k + 4 * g + s;
      off[k][s] = col < n ? (c * n + col) * 4 : kOOB;
    }
}

// Forward pass of the LQR solve, one wave per instance (the gains K~_t of
// k_lqr_wave): with x~_t = [dx_t; 1] as column 0 of B-operand tiles,
//   du_t = -K~_t x~_t,  x~_{t+1} = A~_t x~_t + B~_t du_t
// on the matrix cores; writes dx and du.  The next stage's tiles are requested
// before this stage's stores.
// CB: B~ stage-invariant (stride 0), loaded once.
template <int NT, bool CB>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 3))) k_fwd_wave(Prob<float> P, Ws<float> ws, LqrIn<float> in) {
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const LaneOff<NT> L = lane_offsets<NT>(n, m, g, c);
  int off[NT][4];
  aop_offsets<NT>(off, n, g, c);
  f4 x[NT];
  {
    const rsrc_t r0 = make_rsrc(ws.cdef + (size_t)b * N1 * n, n);
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        x[k][v] = rload(r0, L.col0 + (16 * k + v) * 4) + ((c == 0 && 16 * k + 4 * g + v == n) ? 1.f : 0.f);
  }
  f4 cA[NT][NT], cK[1][NT], cB[NT];
  if constexpr (CB) {
    const rsrc_t rB = make_rsrc(in.Bm.at(b, 0), n * m);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int s = 0; s < 4; ++s) cB[i][s] = rload(rB, L.bt[s] + 16 * i * m * 4);
  }
  load_chain_stage<NT, CB>(cA, cK, cB, ws, in, b, 0, N, n, m, g, c, off, L);
  for (int t = 0; t < N; ++t) {
    // K~ x~ (tile 0, column 0), then x~_{t+1} = A~ x~ - B~ (K~ x~)
    f4 kx = zero4();
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) kx = mma(cK[0][k][s], x[k][s], kx);
    f4 y[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 acc = zero4();
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(cA[i][k][s], x[k][s], acc);
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mma(-cB[i][s], kx[s], acc);
      y[i] = acc;
    }
    f4 nA[NT][NT], nK[1][NT], nB[NT];
    {
      const int gf = fresh(g), cf = fresh(c);
      const LaneOff<NT> Ls = lane_offsets<NT>(n, m, gf, cf);
      int offs[NT][4];
      aop_offsets<NT>(offs, n, gf, cf);
      load_chain_stage<NT, CB>(nA, nK, nB, ws, in, b, __builtin_amdgcn_readfirstlane(t + 1 < N ? t + 1 : t), N, n, m,
                               gf, cf, offs, Ls);
    }
    {
      const rsrc_t rdx = make_rsrc(ws.dx + ((size_t)b * N1 + t) * n, n);
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int v = 0; v < 4; ++v) rstore(rdx, L.col0 + (16 * k + v) * 4, x[k][v]);
      const rsrc_t rdu = make_rsrc(ws.du + ((size_t)b * N + t) * m, m);
#pragma unroll
      for (int v = 0; v < 4; ++v) rstore(rdu, L.col0 + v * 4, -kx[v]);
    }
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      x[k] = y[k];
      cK[0][k] = nK[0][k];
      if constexpr (!CB) cB[k] = nB[k];
#pragma unroll
      for (int i = 0; i < NT; ++i) cA[i][k] = nA[i][k];
    }
  }
  {
    const rsrc_t rdx = make_rsrc(ws.dx + ((size_t)b * N1 + N) * n, n);
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) rstore(rdx, L.col0 + (16 * k + v) * 4, x[k][v]);
  }
}

// Costates and slope partials, one wave per (stage t, instance b):
//   lam_t = (S~_t x~_t)[:n] on the matrix cores (x~_t = [dx_t; 1] as column 0 of
//   B-operand tiles), dv_t = lam_t - V_t, and the partials q_t'dx_t + r_t'du_t,
//   (dv_t - V_t)'c_t, lam_t'c_t of k_fwd_stage (c_t is the constraint row paired
//   with V_t; c_0 = x0 - X_0).
template <int NT>
__global__ void __launch_bounds__(64) k_costate_wave(Prob<float> P, Iter<float> it, Ws<float> ws) {
  const int t = blockIdx.x, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const LaneOff<NT> L = lane_offsets<NT>(n, m, g, c);
  int off[NT][4];
  aop_offsets<NT>(off, n, g, c);
  f4 Sa[NT][NT];
  load_aop<NT, NT>(Sa, ws.S + ((size_t)b * N1 + t) * n * n, ws.s + ((size_t)b * N1 + t) * n, n, n, false, g, c,
                   off, L);
  f4 x[NT], qv[NT], vv[NT], cv[NT];
  const rsrc_t rx = make_rsrc(ws.dx + ((size_t)b * N1 + t) * n, n);
  const rsrc_t rq = make_rsrc(ws.qh + ((size_t)b * N1 + t) * n, n);
  const rsrc_t rV = make_rsrc(it.V + ((size_t)b * N1 + t) * n, n);
  const rsrc_t rc = make_rsrc(ws.cdef + ((size_t)b * N1 + t) * n, n);
#pragma unroll
  for (int k = 0; k < NT; ++k) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
      x[k][v] = rload(rx, L.col0 + (16 * k + v) * 4) + ((c == 0 && 16 * k + 4 * g + v == n) ? 1.f : 0.f);
    qv[k] = rows_vec<NT>(rq, k, L);
    vv[k] = rows_vec<NT>(rV, k, L);
    cv[k] = rows_vec<NT>(rc, k, L);
  }
  float df = 0.f, dc = 0.f, lc = 0.f;
  const bool col0 = c == 0;
  if (t < N) {   // r_t'du_t
    const rsrc_t rr = make_rsrc(ws.rh + ((size_t)b * N + t) * m, m);
    const rsrc_t ru = make_rsrc(ws.du + ((size_t)b * N + t) * m, m);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float rv = rload(rr, (4 * g + v) * 4), uv = rload(ru, (4 * g + v) * 4);
      df += col0 ? rv * uv : 0.f;
    }
  }
  f4 lam[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    f4 acc = zero4();
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mma(Sa[i][k][s], x[k][s], acc);
    lam[i] = acc;
  }
  const rsrc_t rdv = make_rsrc(ws.dv + ((size_t)b * N1 + t) * n, n);
#pragma unroll
  for (int k = 0; k < NT; ++k)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const bool row_ok = col0 && 16 * k + 4 * g + v < n;
      const float dv = lam[k][v] - vv[k][v];
      df += row_ok ? qv[k][v] * x[k][v] : 0.f;
      dc += row_ok ? (dv - vv[k][v]) * cv[k][v] : 0.f;
      lc += row_ok ? lam[k][v] * cv[k][v] : 0.f;
      rstore(rdv, L.col0 + (16 * k + v) * 4, dv);
    }
  df = wave_sum(df);
  dc = wave_sum(dc);
  lc = wave_sum(lc);
  if (lane == 0) {
    float* fp = ws.fpart + ((size_t)b * N1 + t) * NFP;
    fp[FP_DF] = df;
    fp[FP_DC] = dc;
    fp[FP_LC] = lc;
    fp[3] = 0.f;
  }
}

}  // namespace mpxk
