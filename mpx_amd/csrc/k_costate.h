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
// the stored tiles and of the column vectors).
#pragma once
#include "k_trials_mma.h"

namespace mpxk {

template <int NT, bool WIDE>
__global__ void __launch_bounds__(64) k_costate_rows(Prob<float> P, Iter<float> it, Ws<float> ws) {
  const int t = blockIdx.x, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const size_t st = (size_t)b * N1 + t;
  const rsrc_t rS = make_rsrc(ws.S + st * n * n, n * n);
  f4 Sa[NT][NT];
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      if (k >= i) {   // stored: row 16 i + c, columns 16 k + 4 g .. + 3 (rows past n - 1 past the record)
        Sa[i][k] = vload4<WIDE>(rS, (16 * i + c) * n + 16 * k + 4 * g, 16 * k + 4 * g, n);
      } else {        // S[16 i + c][16 k + 4 g + s] = S[16 k + 4 g + s][16 i + c], in tile (k, i)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int r = 16 * k + 4 * g + s, col = 16 * i + c;
          Sa[i][k][s] = rload(rS, r < n && col < n ? (r * n + col) * 4 : kOOB);
        }
      }
    }
  const rsrc_t rx = make_rsrc(ws.dx + st * n, n), rq = make_rsrc(ws.qh + st * n, n);
  const rsrc_t rs = make_rsrc(ws.s + st * n, n), rV = make_rsrc(it.V + st * n, n);
  const rsrc_t rc = make_rsrc(ws.cdef + st * n, n);
  f4 x[NT], q[NT];
  float sv[NT], vv[NT], cv[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    x[k] = vload4<WIDE>(rx, 16 * k + 4 * g, 16 * k + 4 * g, n);
    q[k] = vload4<WIDE>(rq, 16 * k + 4 * g, 16 * k + 4 * g, n);
  }
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    sv[i] = rload(rs, (16 * i + c) * 4);
    vv[i] = rload(rV, (16 * i + c) * 4);
    cv[i] = rload(rc, (16 * i + c) * 4);
  }
  float rr = 0.f, uu = 0.f;
  if (t < N) {
    const rsrc_t rrh = make_rsrc(ws.rh + ((size_t)b * N + t) * m, m);
    const rsrc_t rdu = make_rsrc(ws.du + ((size_t)b * N + t) * m, m);
    rr = rload(rrh, c * 4);
    uu = rload(rdu, c * 4);
  }
  // q'dx on the lanes of column 0 (their entries are dx's, each once), r'du
  // on row 0
  float df = g == 0 ? rr * uu : 0.f, dc = 0.f, lc = 0.f;
#pragma unroll
  for (int k = 0; k < NT; ++k)
#pragma unroll
    for (int s = 0; s < 4; ++s) df += c == 0 ? q[k][s] * x[k][s] : 0.f;
  const rsrc_t rdv = make_rsrc(ws.dv + st * n, n);
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    float pa = 0.f;
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) pa += Sa[i][k][s] * x[k][s];
    const float lam = col_sum4(pa) + sv[i];
    const float dv = lam - vv[i];
    const bool row_ok = g == 0 && 16 * i + c < n;
    dc += row_ok ? (dv - vv[i]) * cv[i] : 0.f;
    lc += row_ok ? lam * cv[i] : 0.f;
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

}  // namespace mpxk
