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
// F = A - B K is never formed (for the tanh model k_fwd_tanh, k_fwd_tanh.h,
// runs the chain on the vector ALU); the costates lam = S x + s follow in
// parallel over the stages from the stored upper tiles of S (k_costate_rows,
// k_costate.h).
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

// vector entries v[16 it + 4 g + k] (0 past the descriptor) in tile-row layout
template <int NT>
__device__ __forceinline__ f4 rows_vec(rsrc_t r, int it, const LaneOff<NT>& L) {
  f4 x;
#pragma unroll
  for (int k = 0; k < 4; ++k) x[k] = rload(r, L.g16 + (16 * it + k) * 4);
  return x;
}

// ---------------------------------------------------------------------------
// The backward recursion (k_lqr_wave) holds its tiles in the interleaved
// order: lane (g, c), register v of tile (it, jt) holds
//   X[16 it + 4 v + g][16 jt + pc],   pc = 4 (c % 4) + c / 4,
// i.e. tile entry (p, q) = (4 g + v, c) holds the entry at the offsets
// (pi(p), pi(q)) within the tile, pi(4 a + b) = 4 b + a (its own inverse).
// Products and transposes of the permuted matrices are the permuted products
// and transposes, so the matrix-core code is that of the natural order.  What
// the order buys: step s of a k tile carries the offsets k = 4 s + g
// (g = 0..3), so the steps past the end of a dimension (n + 1 within the last
// state tile, m, p) meet only zero rows and are skipped (276 instead of 352
// matrix instructions per stage at n = 36, m = 12, p = 8); and register v
// holds the rows 4 v .. 4 v + 3, so the factorisation stops updating a
// register once its rows are done.  (On gfx950 the FP32 matrix instructions
// and the vector ALU do not overlap, within a wave or between the waves of a
// SIMD: every instruction saved is time saved.)

// lane offsets (bytes) of the interleaved order
template <int NT>
struct IlOff {
  int nn[NT];   // ld n, column 16 jt + pc < n : (g n + 16 jt + pc) * 4
  int mm;       // ld m, column pc < m          : (g m + pc) * 4
  int g4;       // a vector along the rows      : 4 g
  int pc4;      // a vector along the columns   : 4 pc
  int hom;      // column n's vector            : 4 g on the lanes pc == n % 16
};

template <int NT>
__device__ __forceinline__ IlOff<NT> il_offsets(int n, int m, int g, int c) {
  IlOff<NT> L;
  const int pc = 4 * (c & 3) + (c >> 2);
#pragma unroll
  for (int jt = 0; jt < NT; ++jt) L.nn[jt] = (16 * jt + pc < n) ? (g * n + 16 * jt + pc) * 4 : kOOB;
  L.mm = (pc < m) ? (g * m + pc) * 4 : kOOB;
  L.g4 = 4 * g;
  L.pc4 = 4 * pc;
  L.hom = (pc == (n & 15)) ? 4 * g : kOOB;
  return L;
}

// tile (it, jt) of a row-major matrix with ld n
template <int NT>
__device__ __forceinline__ f4 il_tile_n(rsrc_t r, int n, int it, int jt, const IlOff<NT>& L) {
  f4 x;
#pragma unroll
  for (int v = 0; v < 4; ++v) x[v] = rload(r, L.nn[jt] + (16 * it + 4 * v) * n * 4);
  return x;
}
// tile (it, 0) of a row-major matrix with ld m (columns < m)
template <int NT>
__device__ __forceinline__ f4 il_tile_m(rsrc_t r, int m, int it, const IlOff<NT>& L) {
  f4 x;
#pragma unroll
  for (int v = 0; v < 4; ++v) x[v] = rload(r, L.mm + (16 * it + 4 * v) * m * 4);
  return x;
}
// x[16 it + 4 v + g] in register v, on the lanes whose part off is in range
__device__ __forceinline__ f4 il_rows(rsrc_t r, int off, int it) {
  f4 x;
#pragma unroll
  for (int v = 0; v < 4; ++v) x[v] = rload(r, off + (16 * it + 4 * v) * 4);
  return x;
}

// acc + a b over the first K steps of one k tile (K = 4: the whole tile)
template <int K>
__device__ __forceinline__ f4 mma_k(const f4& a, const f4& b, f4 acc) {
#pragma unroll
  for (int s = 0; s < K; ++s) acc = mma(a[s], b[s], acc);
  return acc;
}
// the same with a uniform run-time K (1..4)
__device__ __forceinline__ f4 mma_ks(const f4& a, const f4& b, f4 acc, int ks) {
  acc = mma(a[0], b[0], acc);
  if (ks > 1) acc = mma(a[1], b[1], acc);
  if (ks > 2) acc = mma(a[2], b[2], acc);
  if (ks > 3) acc = mma(a[3], b[3], acc);
  return acc;
}
// f(integral_constant<int, k>) for the uniform k = ks in KMIN..KMAX: one
// branch per section of code instead of one per step
template <int KMIN, int KMAX, typename F>
__device__ __forceinline__ void with_steps(int ks, F&& f) {
  if constexpr (KMIN == KMAX) {
    f(std::integral_constant<int, KMAX>{});
  } else {
    if (ks <= KMIN) f(std::integral_constant<int, KMIN>{});
    else with_steps<KMIN + 1, KMAX>(ks, f);
  }
}

// f(integral_constant<int, i>) for i = B .. E - 1
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Lane J of each 16-lane row of the wave, broadcast to the row (DPP row_newbcast).
template <int J>
__device__ __forceinline__ float row_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x150 + J, 0xf, 0xf, true));
}
// y + x[lane J of the row of 16] s on the rows of 16 lanes set in RM (a DPP
// broadcast fused into the multiply-add; the other rows keep y).  The DPP
// source must not have been written in the last two instructions: the nop
// covers that, whatever the compiler places before the block.
template <int J, int RM>
__device__ __forceinline__ float fmac_bcast(float y, float x, float s) {
  asm("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:%4 bank_mask:0xf"
      : "+v"(y)
      : "v"(x), "v"(s), "n"(J), "n"(RM));
  return y;
}

// Lane l gets x from lane 16 JG + l % 16: the row of 16 lanes JG broadcast to
// all four.  v_permlane16_swap exchanges the odd rows of its first operand with
// the even rows of its second, v_permlane32_swap the upper half of the first
// with the lower half of the second; with x as both operands the first gives
// the rows (r0, r0, r2, r2) and (r1, r1, r3, r3), the second of those
// (y0, y1, y0, y1) and (y2, y3, y2, y3).
template <int JG>
__device__ __forceinline__ float row16_bcast(float x) {
  const unsigned u = __float_as_uint(x);
  const auto p = __builtin_amdgcn_permlane16_swap(u, u, false, false);
  const unsigned y = (JG & 1) ? p[1] : p[0];
  const auto q = __builtin_amdgcn_permlane32_swap(y, y, false, false);
  return __uint_as_float((JG & 2) ? q[1] : q[0]);
}

// Step J (logical pivot J) of the factorisation H = L L' with the forward
// substitution of Z = L^{-1}, on one tile in the interleaved order, for
// m <= 4 KM: the rows >= 4 KM (registers >= KM) are the identity padding and
// are never touched.  Row (and column) J sits at tile position
// pJ = 4 (J % 4) + J / 4, i.e. in register J / 4 of the lanes of row-of-16
// J % 4, column lane pJ.  Row J of H and of Z reaches every lane (row16_bcast:
// lane (g, c) gets H[J][pc]); the pivot a = |H[J][J]| is lane pJ of it and is
// raised to fl = floor |H0[J][J]| (H0: the tile before the factorisation; a NaN
// pivot stays NaN); id = 1 / L[J][J] = a^{-1/2}.  H[r][J] reaches the lanes of
// row r by a DPP broadcast from lane pJ of the lane's row of 16.  The trailing
// matrix takes H[r][c] -= H[r][J] H[J][c] / a (= H - L[:, J] L[:, J]') on every
// register that still holds a row > J; entries in a row or a column <= J are
// never read again.  The forward substitution advances by the same column:
// Z[J][:] *= id, Z[r][:] -= L[r][J] Z[J][:] for the rows r > J only (DPP row
// masks).  Steps past m (pivot 1, zero column) change nothing that is read.
template <int J, int KM>
__device__ __forceinline__ void chol_step(f4& h, f4& z, const float (&fl)[16]) {
  constexpr int jg = J & 3, jv = J >> 2, pJ = 4 * (J & 3) + (J >> 2);
  const float hJ = row16_bcast<jg>(h[jv]);
  const float zJ = row16_bcast<jg>(z[jv]);
  const float a = fabsf(row_bcast<pJ>(hJ));
  const float piv = fl[J] > a ? fl[J] : a;
  const float id = __builtin_amdgcn_rsqf(piv);
  const float nid2 = -(id * id);
  const float nlc = hJ * nid2;   // -H[J][c] / a
  const float nzj = zJ * nid2;   // -id Z[J][c] (Z's row J after the step)
  // Z first: it reads column J of H before the trailing update.  Register v
  // holds the rows 4 v + g: those > J are the rows of 16 g >= J - 4 v + 1.
  static_for<0, KM>([&](auto V) {
    constexpr int v = decltype(V)::value;
    constexpr int gm = J - 4 * v + 1 < 0 ? 0 : J - 4 * v + 1;
    if constexpr (gm < 4) z[v] = fmac_bcast<pJ, (0xf << gm) & 0xf>(z[v], h[v], nzj);
  });
  z[jv] = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(z[jv]), __float_as_int(zJ * id), 0xE4,
                                                     1 << jg, 0xf, false));
  // the registers holding a row > J: 4 v + 3 > J
  static_for<0, KM>([&](auto V) {
    constexpr int v = decltype(V)::value;
    if constexpr (4 * v + 3 > J) h[v] = fmac_bcast<pJ, 0xf>(h[v], h[v], nlc);
  });
}

// Z = L^{-1} for the tile H (m x m in the top-left corner, m <= 4 KM, the
// identity elsewhere; H = L L' with floored pivots); z enters as the identity.
// H is consumed.
template <int KM>
__device__ __forceinline__ void chol_lower_inv(f4& h, f4& z, float floor) {
  float fl[16];
  static_for<0, 4 * KM>([&](auto JJ) {
    constexpr int J = decltype(JJ)::value;
    fl[J] = floor * fabsf(rlane(h[J >> 2], 16 * (J & 3) + 4 * (J & 3) + (J >> 2)));
  });
  static_for<0, 4 * KM>([&](auto JJ) { chol_step<decltype(JJ)::value, KM>(h, z, fl); });
}

// One stage's data in the interleaved order.
template <int NT>
struct IlStage {
  f4 A[NT][NT];   // A~ = [[A, c], [0, 1]] without the corner 1 (added where used)
  f4 D[NT];       // [D | 0]
  f4 E;           // E
  f4 rdv;         // rd by rows
  f4 rh;          // r by rows on the lanes of column n (G~'s column n)
  f4 R;           // without CQ: R
  f4 Q[NT][NT];   // without CQ: Q~ tiles i <= j
  f4 B[NT];       // without CQ: B~ tiles (i, 0)
  f4 qrow[NT];    // with CQ: q by rows on the lanes of column n (Q~'s column n)
  float qcol[NT]; // with CQ: q[16 j + pc] (Q~'s row n)
  f4 sdr[NT];     // with FA: sech^2(W x_t) by rows (A~ is formed where it is used)
};

// The loads below do no arithmetic on what they load: an instruction that
// uses a loaded register waits for the load, so the values are combined
// where the stage uses them.  Entries that come from another array (column n
// of A~ and Q~, row n of Q~) are loaded into the same registers under an
// execution mask of their lanes (the tile's own load returns 0 there).

// Q~ tiles (i <= j, or all) of Q (ld n) and q: column n holds q by rows, row n
// q along the columns (the corner stays 0: q has n entries).
template <int NT>
__device__ __forceinline__ void il_load_qbar(f4 (&S)[NT][NT], const float* Q, const float* q, int n,
                                             bool upper_only, int g, bool homlane, const IlOff<NT>& L) {
  const rsrc_t rQ = make_rsrc(Q, n * n), rq = make_rsrc(q, n);
  const int cn = n & 15;
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      if (upper_only && j < i) continue;
      f4 x = il_tile_n<NT>(rQ, n, i, j, L);
      if (j == NT - 1 && homlane) x = il_rows(rq, L.g4, i);
      if (i == NT - 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (4 * v + g == cn) x[v] = rload(rq, L.pc4 + 64 * j);
      }
      S[i][j] = x;
    }
}

// The stage's vectors and constraint rows E (needed when the stage starts).
template <int NT, bool CQ>
__device__ __forceinline__ void il_load_vecs(IlStage<NT>& d, const Prob<float>& P, const Ws<float>& ws, int b, int t,
                                             const IlOff<NT>& L) {
  const int m = P.m, p = P.p, N = P.N;
  const rsrc_t rE = make_rsrc(P.E.at(b, t), p * m);   // p == 0: no records, every load returns 0
#pragma unroll
  for (int v = 0; v < 4; ++v) d.E[v] = rload(rE, L.mm + 4 * v * m * 4);
  d.rdv = il_rows(make_rsrc(ws.rd + ((size_t)b * N + t) * m, m), L.g4, 0);
  d.rh = il_rows(make_rsrc(ws.rh + ((size_t)b * N + t) * m, m), L.hom, 0);
  if constexpr (!CQ) d.R = il_tile_m<NT>(make_rsrc(P.R.at(b, t), m * m), m, 0, L);
}

// The stage's matrices: A~ without its corner (column n = c of the next
// stage's defect), [D | 0], and q (with CQ) or Q~ and B~.
// FA (the tanh model): sech^2(W x_t) by rows instead of A_t, and column n
// (c) alone into the last tile column of A~.
template <int NT, bool CQ, bool FA>
__device__ __forceinline__ void il_load_mats(IlStage<NT>& d, const Prob<float>& P, const Ws<float>& ws,
                                             const LqrIn<float>& in, int b, int t, int g, bool homlane,
                                             const IlOff<NT>& L) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const rsrc_t rc = make_rsrc(ws.cdef + ((size_t)b * N1 + t + 1) * n, n);
  if constexpr (FA) {
    const rsrc_t rs = make_rsrc(ws.sdh + ((size_t)b * N + t) * n, n);
#pragma unroll
    for (int i = 0; i < NT; ++i) d.sdr[i] = il_rows(rs, L.g4, i);
    if (homlane) {
#pragma unroll
      for (int i = 0; i < NT; ++i) d.A[i][NT - 1] = il_rows(rc, L.g4, i);
    }
  } else {
    const rsrc_t rA = make_rsrc(in.A.at(b, t), n * n);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        f4 x = il_tile_n<NT>(rA, n, i, j, L);
        if (j == NT - 1 && homlane) x = il_rows(rc, L.g4, i);
        d.A[i][j] = x;
      }
  }
  const rsrc_t rD = make_rsrc(P.D.at(b, t), p * n);
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) d.D[j][v] = rload(rD, L.nn[j] + 4 * v * n * 4);
  if constexpr (CQ) {
    const rsrc_t rq = make_rsrc(ws.qh + ((size_t)b * N1 + t) * n, n);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      d.qrow[i] = il_rows(rq, L.hom, i);
      d.qcol[i] = rload(rq, L.pc4 + 64 * i);
    }
  } else {
    il_load_qbar<NT>(d.Q, P.Q.at(b, t), ws.qh + ((size_t)b * N1 + t) * n, n, true, g, homlane, L);
    const rsrc_t rB = make_rsrc(in.Bm.at(b, t), n * m);
#pragma unroll
    for (int i = 0; i < NT; ++i) d.B[i] = il_tile_m<NT>(rB, m, i, L);
  }
}

// Store tile (it, jt) of an (n+1)-wide homogeneous matrix: the first n columns
// to rM (ld n), column n (in the last tile column) to rvec; rows past the
// descriptors are dropped.
template <int NT>
__device__ __forceinline__ void il_store_hom(const f4& x, rsrc_t rM, rsrc_t rvec, int n, int it, int jt,
                                             const IlOff<NT>& L) {
#pragma unroll
  for (int v = 0; v < 4; ++v) rstore(rM, L.nn[jt] + (16 * it + 4 * v) * n * 4, x[v]);
  if (jt == NT - 1) {
#pragma unroll
    for (int v = 0; v < 4; ++v) rstore(rvec, L.hom + (16 * it + 4 * v) * 4, x[v]);
  }
}

// S~_t: the tiles (it, jt), it <= jt, of the symmetric S_t (the lower ones
// are read as their transposes) and s_t (column n, in the last tile column)
template <int NT>
__device__ __forceinline__ void il_store_value(const f4 (&S)[NT][NT], const Ws<float>& ws, int b, int t, int N1,
                                               int n, const IlOff<NT>& L) {
  const rsrc_t rS = make_rsrc(ws.S + ((size_t)b * N1 + t) * n * n, n * n);
  const rsrc_t rs = make_rsrc(ws.s + ((size_t)b * N1 + t) * n, n);
#pragma unroll
  for (int it = 0; it < NT; ++it)
#pragma unroll
    for (int jt = it; jt < NT; ++jt) il_store_hom<NT>(S[it][jt], rS, rs, n, it, jt, L);
}

// segments == 1, FP32, n + 1 <= 16 NT (NT minimal: column n lies in the last
// tile), m <= 16, p <= 16: backward Riccati of one instance per wave; stores
// S~_t and K~_t for k_fwd_wave.  Each register of a stage's data is reloaded
// with the next stage's as soon as its last use in the stage has passed.
// CQ: Q, B, R stage-invariant (views of stage stride 0): held in registers
// for the whole recursion instead of being re-read every stage.
// KSN: the k steps of the last state tile, (n + 1 - 16 (NT - 1) + 3) / 4 or
// more (steps past n + 1 add zeros).
// AS (n % 4 == 0, KSN exact): the offset n, the homogeneous index of A~, S~ and B~, sits
// alone in step KSN - 1 of the last state tile (lanes g == 0, the rest of the
// step is padding).  On that step B~ has its zero row n and A~ its row
// n = e_n', so the contractions over the state stop a step short (KSS steps)
// and row n's part is added on the vector ALU: column n of S~ A~ gains column
// n of S~, and row n of A~' (S~ A~) gains row n of S~ A~ (254 instead of 276
// matrix instructions per stage at n = 36, m = 12, p = 8).
template <int NT, bool CQ, int KSN, bool FA, bool AS>
__global__ void __launch_bounds__(64, 2) k_lqr_wave(Prob<float> P, Prm<float> prm, Iter<float> it_unused,
                                                    Ws<float> ws, LqrIn<float> in) {
  static_assert(!FA || CQ, "A~ is formed from stage-invariant A, W only with stage-invariant Q, B, R");
  static_assert(!AS || CQ, "the homogeneous step is split off only with the step count of n");
  constexpr int KSS = AS ? KSN - 1 : KSN;   // steps of the last state tile in a contraction over the state
  (void)it_unused;
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  __shared__ float tbuf[16 * 17];
  // FA: the lane's entries of the tiles of I + h A and h kappa W (rows and
  // columns past n - 1 are 0), each lane reading back only its own
  __shared__ f4 fab[FA ? 2 * NT * NT * 64 : 1];
  const float sr = sqrtf(prm.rho_c);
  // k steps over m and over p
  const int ksm = (m + 3) >> 2, ksp = (p + 3) >> 2;
  const int cn = n & 15, pc = 4 * (c & 3) + (c >> 2);
  const IlOff<NT> L = il_offsets<NT>(n, m, g, c);
  if constexpr (FA) {
    const rsrc_t rA = make_rsrc(P.A.at(b, 0), n * n), rW = make_rsrc(P.W.at(b, 0), n * n);
    const float hs = P.h, hk = P.h * P.kappa;
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        f4 a = il_tile_n<NT>(rA, n, i, j, L), w = il_tile_n<NT>(rW, n, i, j, L);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * i + 4 * v + g, col = 16 * j + pc;
          a[v] = (row == col && row < n ? 1.f : 0.f) + hs * a[v];
          w[v] = hk * w[v];
        }
        fab[(i * NT + j) * 64 + lane] = a;
        fab[(NT * NT + i * NT + j) * 64 + lane] = w;
      }
  }

  f4 S[NT][NT];
  il_load_qbar<NT>(S, P.QN.at(b, 0), ws.qh + ((size_t)b * N1 + N) * n, n, false, g, pc == cn, L);

  // stage-invariant Q (tiles i <= j, no q), B~ and R + the padding (CQ)
  f4 Qc[NT][NT], Bc[NT], Rc;
  if constexpr (CQ) {
    const rsrc_t rQ = make_rsrc(P.Q.at(b, 0), n * n), rB = make_rsrc(in.Bm.at(b, 0), n * m);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = i; j < NT; ++j) Qc[i][j] = il_tile_n<NT>(rQ, n, i, j, L);
#pragma unroll
    for (int i = 0; i < NT; ++i) Bc[i] = il_tile_m<NT>(rB, m, i, L);
    Rc = il_tile_m<NT>(make_rsrc(P.R.at(b, 0), m * m), m, 0, L);
#pragma unroll
    for (int v = 0; v < 4; ++v) Rc[v] += (4 * g + v == c && 4 * v + g >= m) ? 1.f : 0.f;
  }
  IlStage<NT> cur;
  il_load_vecs<NT, CQ>(cur, P, ws, b, N - 1, L);
  il_load_mats<NT, CQ, FA>(cur, P, ws, in, b, N - 1, g, pc == cn, L);
  // S~_N is stored after the first stage's loads are requested, as every
  // stage's stores follow the next stage's loads: the waits at the top of the
  // loop then count the same requests on entry as on the way round
  if constexpr (!FA) il_store_value<NT>(S, ws, b, N, N1, n, L);   // (FA: the costates come from k_adjoint_tanh)
  for (int t = N - 1; t >= 0; --t) {
    // lane offsets from opaque copies of the lane coordinates: recomputed every
    // stage instead of held across the loop
    const int gf = fresh(g), cf = fresh(c);
    const IlOff<NT> Ls = il_offsets<NT>(n, m, gf, cf);
    const bool homlane = 4 * (cf & 3) + (cf >> 2) == cn;
    // masks (0 / 1), computed where used: the tile diagonal and row n
    f4 dmask, rown;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      dmask[v] = 4 * gf + v == cf ? 1.f : 0.f;
      rown[v] = 4 * v + gf == cn ? 1.f : 0.f;
    }
    // the next stage (the stage index pinned to a scalar register: selected
    // into a vector register it makes every descriptor built from it look
    // divergent)
    const int tp = __builtin_amdgcn_readfirstlane(t > 0 ? t - 1 : 0);
    f4 Bt[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) Bt[i] = CQ ? Bc[i] : cur.B[i];
    // SB = S~ B~ ; H = R + diag(rd) + rho_c E'E + B~' SB
    f4 SB[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 acc = zero4();
#pragma unroll
      for (int k = 0; k < NT - 1; ++k) acc = mma_k<4>(S[k][i], Bt[k], acc);
      SB[i] = mma_k<KSS>(S[NT - 1][i], Bt[NT - 1], acc);
    }
    f4 H;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      H[v] = fmaf(dmask[v], cur.rdv[v], CQ ? Rc[v] : cur.R[v] + (dmask[v] != 0.f && 4 * v + gf >= m ? 1.f : 0.f));
#pragma unroll
    for (int k = 0; k < NT - 1; ++k) H = mma_k<4>(Bt[k], SB[k], H);
    H = mma_k<KSS>(Bt[NT - 1], SB[NT - 1], H);
    // G~ = [0 | r] + rho_c E'[D | 0] (+ B~' S~ A~ below)
    f4 G[NT];
#pragma unroll
    for (int j = 0; j < NT - 1; ++j) G[j] = zero4();
    G[NT - 1] = cur.rh;
    if (p > 0) {   // E and D scaled by sqrt(rho_c) in place
#pragma unroll
      for (int v = 0; v < 4; ++v) cur.E[v] *= sr;
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) cur.D[j][v] *= sr;
      with_steps<1, 4>(ksp, [&](auto KP) {
        constexpr int kp = decltype(KP)::value;
        H = mma_k<kp>(cur.E, cur.E, H);
#pragma unroll
        for (int j = 0; j < NT; ++j) G[j] = mma_k<kp>(cur.E, cur.D[j], G[j]);
      });
    }
    // E, rd, r (and R) of this stage are dead: the next stage's are requested
    il_load_vecs<NT, CQ>(cur, P, ws, b, tp, Ls);

    // Z = L^{-1} (H = L L'); Z' is the A operand of Z G~
    f4 Z;
#pragma unroll
    for (int v = 0; v < 4; ++v) Z[v] = dmask[v];
    with_steps<1, 4>(ksm, [&](auto KM) { chol_lower_inv<decltype(KM)::value>(H, Z, prm.hfloor); });
    const f4 Zt = tile_transpose(Z, tbuf, g, c);

    // G~ += B~' S~ A~ ; Pm = Q~ + rho_c [D|0]'[D|0] + A~' S~ A~ (tiles i <= j),
    // a column tile of S~ A~ at a time
    if constexpr (FA) {   // A~ = I + h A + diag(sd) h kappa W, column n = c on its lanes
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const f4 a = fab[(i * NT + j) * 64 + lane], w = fab[(NT * NT + i * NT + j) * 64 + lane];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float x = fmaf(cur.sdr[i][v], w[v], a[v]);
            cur.A[i][j][v] = (j == NT - 1 && homlane) ? cur.A[i][j][v] : x;
          }
        }
    }
    if constexpr (!AS) {   // the corner 1 (with AS no matrix instruction reads row n of A~)
#pragma unroll
      for (int v = 0; v < 4; ++v) cur.A[NT - 1][NT - 1][v] += homlane ? rown[v] : 0.f;
    }
    f4 Pm[NT][NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      f4 SAc[NT];
#pragma unroll
      for (int i = 0; i < NT; ++i) {
        f4 acc = zero4();
#pragma unroll
        for (int k = 0; k < NT - 1; ++k) acc = mma_k<4>(S[k][i], cur.A[k][j], acc);
        SAc[i] = mma_k<KSS>(S[NT - 1][i], cur.A[NT - 1][j], acc);
        if constexpr (AS) {
          if (j == NT - 1) {   // + S~[:, n] A~[n][n], on the lanes of column n
            const float hm = homlane ? 1.f : 0.f;
#pragma unroll
            for (int v = 0; v < 4; ++v) SAc[i][v] = fmaf(hm, S[i][NT - 1][v], SAc[i][v]);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < NT - 1; ++k) G[j] = mma_k<4>(Bt[k], SAc[k], G[j]);
      G[j] = mma_k<KSS>(Bt[NT - 1], SAc[NT - 1], G[j]);
#pragma unroll
      for (int i = 0; i <= j; ++i) {
        f4 acc;
        if constexpr (CQ) {   // Q~ = [[Q, q], [q', 0]] from the held Q and this stage's q
          acc = Qc[i][j];
          if (j == NT - 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[v] += cur.qrow[i][v];
          }
          if (i == NT - 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[v] = fmaf(rown[v], cur.qcol[j], acc[v]);
          }
        } else {
          acc = cur.Q[i][j];
        }
        if (p > 0) acc = mma_ks(cur.D[i], cur.D[j], acc, ksp);
#pragma unroll
        for (int k = 0; k < NT - 1; ++k) acc = mma_k<4>(cur.A[k][i], SAc[k], acc);
        acc = mma_k<KSS>(cur.A[NT - 1][i], SAc[NT - 1], acc);
        if constexpr (AS) {
          if (i == NT - 1) {   // + A~[n][n] (S~ A~)[n, :]: row n = 4 (KSN - 1) + g, g = 0
            constexpr int v = KSN - 1;
            acc[v] = fmaf(rown[v], SAc[NT - 1][v], acc[v]);
          }
        }
        Pm[i][j] = acc;
      }
    }
    // A~, D and q (or Q~ and B~) of this stage are dead
    il_load_mats<NT, CQ, FA>(cur, P, ws, in, b, tp, gf, homlane, Ls);

    // W~ = Z G~, K~ = Z' W~ (sums over the m rows of G~ and W~);
    // S~ <- Pm - W~' W~ (tiles i <= j, then mirrored)
    f4 K[NT];
    with_steps<1, 4>(ksm, [&](auto KM) {
      constexpr int km = decltype(KM)::value;
      f4 Wn[NT];
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const f4 w = mma_k<km>(Zt, G[j], zero4());
        G[j] = w;
        K[j] = mma_k<km>(Z, w, zero4());
#pragma unroll
        for (int v = 0; v < 4; ++v) Wn[j][v] = -w[v];
      }
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = i; j < NT; ++j) S[i][j] = mma_k<km>(Wn[i], G[j], Pm[i][j]);
    });
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const f4 d = tile_transpose(S[i][i], tbuf, g, c);
#pragma unroll
      for (int v = 0; v < 4; ++v) S[i][i][v] = 0.5f * (S[i][i][v] + d[v]);
#pragma unroll
      for (int j = i + 1; j < NT; ++j) S[j][i] = tile_transpose(S[i][j], tbuf, g, c);
    }
    // the corner of S~ (the constant of the value function) is not carried
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (homlane && 4 * v + gf == cn) S[NT - 1][NT - 1][v] = 0.f;

    // stores: S_t, s_t (not with FA) ; K_t, k_t
    if constexpr (!FA) il_store_value<NT>(S, ws, b, t, N1, n, Ls);
    {
      const rsrc_t rK = make_rsrc(ws.K + ((size_t)b * N + t) * m * n, m * n);
      const rsrc_t rk = make_rsrc(ws.k + ((size_t)b * N + t) * m, m);
#pragma unroll
      for (int j = 0; j < NT; ++j) il_store_hom<NT>(K[j], rK, rk, n, 0, j, Ls);
    }
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

// A-operand tiles of a row-major block M (ld n, nrows rows) without any
// column n: lane (g, c) of tile (i, k) holds M[16 i + c][16 k + 4 g + s] in
// register s, 0 past column n - 1 or row nrows - 1 (no arithmetic on the
// loaded values: a use of a loaded register waits for its load).  WIDE
// (n % 4 == 0): a lane's four columns 16 k + 4 g .. + 3 of a row are
// contiguous, 16-byte aligned, and either all below n or all past it: one
// 128-bit load.  (A compile-time choice: after a run-time branch between the
// two forms the wait counts of the loads in flight are not known.)
template <int NT, int RT, bool WIDE>
__device__ __forceinline__ void load_aop_raw(f4 (&out)[RT][NT], const float* M, int n, int nrows,
                                             const int (&off)[NT][4]) {
  const rsrc_t rM = make_rsrc(M, nrows * n);
  if constexpr (WIDE) {
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
      for (int k = 0; k < NT; ++k) out[i][k] = rload4(rM, off[k][0] + 16 * i * n * 4);
  } else {
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) out[i][k][s] = rload(rM, off[k][s] + 16 * i * n * 4);
  }
}

// One forward step's data: the A-operand tiles of A_t and K_t (column n of
// A~ and K~ left out: their loads return 0 there), and the columns n as
// vectors: lane column i < NT of colc holds c_{t+1}[16 i + 4 g + v] in
// register v (the other lanes 0), every lane of colk holds k_t[4 g + v].  A
// product's accumulator starts from its lane column broadcast along the row of
// 16 (DPP), or from colk: every column of the result then carries the term,
// column 0 (dx's) included; the other columns are never read.  (No branches
// and no arithmetic on loaded values: nothing waits for these loads before the
// next stage uses them.)
template <int NT>
struct FwdA {
  f4 A[NT][NT];
  f4 colc;
};
template <int NT>
struct FwdK {
  f4 K[1][NT];
  f4 colk;
};

template <int NT, bool WIDE>
__device__ __forceinline__ void fwd_load_a(FwdA<NT>& d, const Ws<float>& ws, const LqrIn<float>& in, int b, int t,
                                           int N, int n, int g, int c, const int (&off)[NT][4]) {
  load_aop_raw<NT, NT, WIDE>(d.A, in.A.at(b, t), n, n, off);
  const rsrc_t rc = make_rsrc(ws.cdef + ((size_t)b * (N + 1) + t + 1) * n, n);
  const int oc = c < NT ? (16 * c + 4 * g) * 4 : kOOB;
#pragma unroll
  for (int v = 0; v < 4; ++v) d.colc[v] = rload(rc, oc + v * 4);
}

template <int NT, bool WIDE>
__device__ __forceinline__ void fwd_load_k(FwdK<NT>& d, const Ws<float>& ws, int b, int t, int N, int n, int m,
                                           int g, const int (&off)[NT][4]) {
  load_aop_raw<NT, 1, WIDE>(d.K, ws.K + ((size_t)b * N + t) * m * n, n, m, off);
  const rsrc_t rk = make_rsrc(ws.k + ((size_t)b * N + t) * m, m);
#pragma unroll
  for (int v = 0; v < 4; ++v) d.colk[v] = rload(rk, (4 * g + v) * 4);
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
// k_lqr_wave): with dx_t as column 0 of B-operand tiles,
//   du_t = -(K_t dx_t + k_t),  dx_{t+1} = A_t dx_t + c_{t+1} + B_t du_t
// on the matrix cores (the affine terms start the accumulators, so the
// homogeneous coordinate of the backward kernel is not needed); writes dx and
// du.  Stage t + 1's data is requested before stage t's products, into the
// other of two buffers (the loop is unrolled by two, so no buffer is copied).
// CB: B~ stage-invariant (stride 0), loaded once; else loaded every stage.
// WIDE: n % 4 == 0 (128-bit loads of the A-operand tiles).
template <int NT, bool CB, bool WIDE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3))) k_fwd_wave(Prob<float> P, Ws<float> ws, LqrIn<float> in) {
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const LaneOff<NT> L = lane_offsets<NT>(n, m, g, c);
  int off[NT][4];
  aop_offsets<NT>(off, n, g, c);
  f4 x[NT];   // dx_0 = c_init
  {
    const rsrc_t r0 = make_rsrc(ws.cdef + (size_t)b * N1 * n, n);
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) x[k][v] = rload(r0, L.col0 + (16 * k + v) * 4);
  }
  // -B~ (stage-invariant)
  f4 nB[NT];
  auto load_nb = [&](int t) {
    const rsrc_t rB = make_rsrc(in.Bm.at(b, t), n * m);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int s = 0; s < 4; ++s) nB[i][s] = -rload(rB, L.bt[s] + 16 * i * m * 4);   // B[16 i + c][4 g + s]
  };
  if constexpr (CB) load_nb(0);
  FwdA<NT> s0, s1;
  FwdK<NT> kk;
  fwd_load_a<NT, WIDE>(s0, ws, in, b, 0, N, n, g, c, off);
  fwd_load_k<NT, WIDE>(kk, ws, b, 0, N, n, m, g, off);
  auto step = [&](FwdA<NT>& cur, FwdA<NT>& nxt, int t) {
    const int tn = __builtin_amdgcn_readfirstlane(t + 1 < N ? t + 1 : t);   // the next stage
    const int gf = fresh(g), cf = fresh(c);
    int offs[NT][4];
    aop_offsets<NT>(offs, n, gf, cf);
    fwd_load_a<NT, WIDE>(nxt, ws, in, b, tn, N, n, gf, cf, offs);
    if constexpr (!CB) load_nb(t);
    // the loads stay here, a stage ahead of their use (under register pressure
    // the scheduler would sink them towards their uses in the next step)
    __builtin_amdgcn_sched_barrier(0);
    // K dx + k (tile 0, column 0), then dx_{t+1} = A dx + c - B (K dx + k)
    f4 kx = kk.colk;
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) kx = mma(kk.K[0][k][s], x[k][s], kx);
    // K and k of this stage are dead: the next stage's, in place
    fwd_load_k<NT, WIDE>(kk, ws, b, tn, N, n, m, gf, offs);
    __builtin_amdgcn_sched_barrier(0);
    f4 y[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 acc;
      auto bc = [&](auto I) {
        constexpr int ii = decltype(I)::value;
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] = row_bcast<ii>(cur.colc[v]);
      };
      if (i == 0) bc(std::integral_constant<int, 0>{});
      else if (i == 1) bc(std::integral_constant<int, 1>{});
      else bc(std::integral_constant<int, 2>{});
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = mma(cur.A[i][k][s], x[k][s], acc);
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = mma(nB[i][s], kx[s], acc);
      y[i] = acc;
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
    for (int k = 0; k < NT; ++k) x[k] = y[k];
  };
  int t = 0;
  for (; t + 1 < N; t += 2) {
    step(s0, s1, t);
    step(s1, s0, t + 1);
  }
  if (t < N) step(s0, s1, t);
  {
    const rsrc_t rdx = make_rsrc(ws.dx + ((size_t)b * N1 + N) * n, n);
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) rstore(rdx, L.col0 + (16 * k + v) * 4, x[k][v]);
  }
}

}  // namespace mpxk
