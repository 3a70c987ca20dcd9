// Costates and slope partials after the wave forward pass, one wave per
// (stage t, instance b), on the vector ALU:
//   lam_t = S_t dx_t + s_t,  dv_t = lam_t - V_t,
// and the partials q_t'dx_t + r_t'du_t, (dv_t - V_t)'c_t, lam_t'c_t of
// k_fwd_stage (c_t is the constraint row paired with V_t; c_0 = x0 - X_0).
// k_lqr_wave stores the tiles (i, k), i <= k, of the symmetric S_t (and s_t
// apart); tile (i, k), i > k, is read as the transpose of tile (k, i).  Lane
// (g, c) holds rows 16 i + c, columns 16 k + 4 g + s of S_t and dx_t's
// entries 16 k + 4 g + s: 4 NT^2 products per lane and a sum over the four
// lanes of a column.  (The matrix cores would use one of the 16 columns of
// each product.)  WIDE: n % 4 == 0 and 16-byte aligned rows (128-bit loads of
// the stored tiles and of dx).
#pragma once
#include "k_trials_mma.h"

namespace mpxk {

// stages per wave of k_costate_rows
constexpr int kCostateStages = 4;

// One stage's data, lane (g, c): S_t rows 16 i + c, columns 16 k + 4 g + s;
// dx_t entries 16 k + 4 g + s; dx_t, q_t, s_t, V_t, c_t rows 16 i + c;
// r_t[c], du_t[c] (0 for t = N).
template <int NT>
struct CsStage {
  f4 S[NT][NT];
  f4 x[NT];
  float xv[NT], qv[NT], sv[NT], vv[NT], cv[NT];
  float rr, uu;
};

template <int NT, bool WIDE>
__device__ __forceinline__ void cs_load(CsStage<NT>& d, const Iter<float>& it, const Ws<float>& ws, int b, int t,
                                        int n, int m, int N, int g, int c) {
  const size_t st = (size_t)b * (N + 1) + t;
  const rsrc_t rS = make_rsrc(ws.S + st * n * n, n * n);
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      if (k >= i) {   // stored: row 16 i + c, columns 16 k + 4 g .. + 3 (rows past n - 1 past the record)
        d.S[i][k] = vload4<WIDE>(rS, (16 * i + c) * n + 16 * k + 4 * g, 16 * k + 4 * g, n);
      } else {        // S[16 i + c][16 k + 4 g + s] = S[16 k + 4 g + s][16 i + c], in tile (k, i)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int r = 16 * k + 4 * g + s, col = 16 * i + c;
          d.S[i][k][s] = rload(rS, r < n && col < n ? (r * n + col) * 4 : kOOB);
        }
      }
    }
  const rsrc_t rx = make_rsrc(ws.dx + st * n, n), rq = make_rsrc(ws.qh + st * n, n);
  const rsrc_t rs = make_rsrc(ws.s + st * n, n), rV = make_rsrc(it.V + st * n, n);
  const rsrc_t rc = make_rsrc(ws.cdef + st * n, n);
#pragma unroll
  for (int k = 0; k < NT; ++k) d.x[k] = vload4<WIDE>(rx, 16 * k + 4 * g, 16 * k + 4 * g, n);
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    d.xv[i] = rload(rx, (16 * i + c) * 4);
    d.qv[i] = rload(rq, (16 * i + c) * 4);
    d.sv[i] = rload(rs, (16 * i + c) * 4);
    d.vv[i] = rload(rV, (16 * i + c) * 4);
    d.cv[i] = rload(rc, (16 * i + c) * 4);
  }
  // stage N has no control: its r and du offsets fall past the records
  const int tu = t < N ? t : N - 1;
  const int ou = t < N ? c * 4 : kOOB;
  d.rr = rload(make_rsrc(ws.rh + ((size_t)b * N + tu) * m, m), ou);
  d.uu = rload(make_rsrc(ws.du + ((size_t)b * N + tu) * m, m), ou);
}

template <int NT>
__device__ __forceinline__ void cs_stage(const CsStage<NT>& d, const Ws<float>& ws, int b, int t, int n, int N,
                                         int g, int c, int lane) {
  const size_t st = (size_t)b * (N + 1) + t;
  // q'dx and r'du on row 0 of the lanes
  float df = g == 0 ? d.rr * d.uu : 0.f, dc = 0.f, lc = 0.f;
#pragma unroll
  for (int i = 0; i < NT; ++i) df += g == 0 ? d.qv[i] * d.xv[i] : 0.f;
  const rsrc_t rdv = make_rsrc(ws.dv + st * n, n);
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    float pa = 0.f;
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) pa += d.S[i][k][s] * d.x[k][s];
    const float lam = col_sum4(pa) + d.sv[i];
    const float dv = lam - d.vv[i];
    const bool row_ok = g == 0 && 16 * i + c < n;
    dc += row_ok ? (dv - d.vv[i]) * d.cv[i] : 0.f;
    lc += row_ok ? lam * d.cv[i] : 0.f;
    rstore(rdv, row_ok ? (16 * i + c) * 4 : kOOB, dv);
  }
  df = wave_sum(df);
  dc = wave_sum(dc);
  lc = wave_sum(lc);
  if (lane == 0) {
    float* fp = ws.fpart + st * NFP;
    fp[FP_DF] = df;
    fp[FP_DC] = dc;
    fp[FP_LC] = lc;
    fp[3] = 0.f;
  }
}

// SPW consecutive stages per wave (stage SPW blockIdx.x + j, j < SPW, while
// <= N), the next stage's data requested before the current stage's
// arithmetic (two buffers; the loop is unrolled by two).
template <int NT, bool WIDE, int SPW>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WIDE ? 3 : 2))) k_costate_rows(Prob<float> P, Iter<float> it, Ws<float> ws) {
  const int b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, N = P.N;
  const int t0 = blockIdx.x * SPW, t1 = min(t0 + SPW, N + 1);
  CsStage<NT> d0, d1;
  cs_load<NT, WIDE>(d0, it, ws, b, t0, n, m, N, g, c);
  for (int t = t0; t < t1; t += 2) {
    cs_load<NT, WIDE>(d1, it, ws, b, t + 1 < t1 ? t + 1 : t, n, m, N, g, c);
    cs_stage<NT>(d0, ws, b, t, n, N, g, c, lane);
    if (t + 1 >= t1) break;
    cs_load<NT, WIDE>(d0, it, ws, b, t + 2 < t1 ? t + 2 : t + 1, n, m, N, g, c);
    cs_stage<NT>(d1, ws, b, t + 1, n, N, g, c, lane);
  }
}

}  // namespace mpxk
