// Costates of the tanh model by the adjoint recursion of the LQR subproblem,
// one wave per instance, after k_fwd_tanh:
//   lam_N = Q_N dx_N + q_N,
//   lam_t = Q dx_t + rho_c D_t'(D_t dx_t + E_t du_t) + q_t + A_t' lam_{t+1},
//   A_t' lam = (I + h A)' lam + (h kappa W)'(sd_t * lam),
// i.e. stationarity in dx_t of the subproblem (Q + rho_c D'D the state
// Hessian, rho_c D'E the cross term, q_t with the equality terms; the
// multipliers of the dense KKT system, the same as S_t dx_t + s_t of the
// Riccati solution, which k_lqr_wave then never stores); then dv_t = lam_t -
// V_t and the slope partials q_t'dx_t + r_t'du_t, (dv_t - V_t)'c_t, lam_t'c_t
// (c_t the constraint row paired with V_t; c_0 = x0 - X_0); the equality
// rows' slope D_t dx_t + E_t du_t goes to hq for the trials.
//
// Lane (g, c) holds rows 16 i + c, columns 16 k + 4 g + s of Q (registers),
// (I + h A)' and (h kappa W)' (LDS) for the whole sweep, lam_{t+1},
// sd_t * lam_{t+1} and dx_t by the entries 16 k + 4 g + s (gathered from the
// rows by ds_bpermute), and D_t' (rows 16 i + c, columns 4 g + s < p).
// D dx + E du is a row sum over the 16 lanes (every lane of row g then holds
// the entries 4 g + s); each lam_t row is 4 NT^2 products per matrix per lane
// and a sum over the four lanes of a column; lam_t's rows are gathered back
// into the column layout by ds_bpermute.  Each stage's data is requested two
// stages ahead (three register buffers).
#pragma once
#include "k_fwd_tanh.h"

namespace mpxk {

// sum over the 16 lanes of the row, in every lane of the row
__device__ __forceinline__ float row_sum16(float v) {
  v += dpp_ror<0x128>(v);
  v += dpp_ror<0x124>(v);
  v += dpp_ror<0x122>(v);
  v += dpp_ror<0x121>(v);
  return v;
}

// One stage's data, lane (g, c): dx_t, sd_t, q_t, V_t, c_t rows 16 i + c;
// du_t[c], r_t[c]; D_t[4 g + s][16 i + c] (0 past p - 1 and n - 1) and
// E_t[4 g + s][c] (0 past p - 1 and m - 1).
template <int NT>
struct AdjT {
  float xr[NT], sr[NT], qr[NT], vr[NT], cr[NT];
  float du, rh;
  f4 Dt[NT], Et;
};

// The instance's arrays (one descriptor each for all of its stages).
struct AdjR {
  rsrc_t x, sd, q, V, c, du, rh, D, E, dv, hq;
  long long dst, est;   // stage strides of D and E (elements)
};

__device__ __forceinline__ AdjR adj_rsrc(const Prob<float>& P, const Iter<float>& it, const Ws<float>& ws, int b) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  AdjR r;
  r.x = make_rsrc(ws.dx + (size_t)b * N1 * n, N1 * n);
  r.sd = make_rsrc(ws.sdh + (size_t)b * N * n, N * n);
  r.q = make_rsrc(ws.qh + (size_t)b * N1 * n, N1 * n);
  r.V = make_rsrc(it.V + (size_t)b * N1 * n, N1 * n);
  r.c = make_rsrc(ws.cdef + (size_t)b * N1 * n, N1 * n);
  r.du = make_rsrc(ws.du + (size_t)b * N * m, N * m);
  r.rh = make_rsrc(ws.rh + (size_t)b * N * m, N * m);
  r.D = view_rsrc(P.D, b, N, p * n);   // p = 0: no records
  r.E = view_rsrc(P.E, b, N, p * m);
  r.dv = make_rsrc(ws.dv + (size_t)b * N1 * n, N1 * n);
  r.hq = make_rsrc(ws.hq + (size_t)b * N * 2 * p, N * 2 * p);
  r.dst = P.D.st;
  r.est = P.E.st;
  return r;
}

template <int NT>
__device__ __forceinline__ void adj_load(AdjT<NT>& d, const AdjR& R, int n, int m, int p, int t, int g, int c) {
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    const int row = 16 * i + c, o = (t * n + row) * 4;
    const bool in = row < n;
    d.xr[i] = rload(R.x, in ? o : kOOB);
    d.sr[i] = rload(R.sd, in ? o : kOOB);   // t < N
    d.qr[i] = rload(R.q, in ? o : kOOB);
    d.vr[i] = rload(R.V, in ? o : kOOB);
    d.cr[i] = rload(R.c, in ? o : kOOB);
  }
  d.du = rload(R.du, c < m ? (t * m + c) * 4 : kOOB);
  d.rh = rload(R.rh, c < m ? (t * m + c) * 4 : kOOB);
  const int od = (int)(t * R.dst), oe = (int)(t * R.est);
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int r = 4 * g + s, col = 16 * i + c;
      d.Dt[i][s] = rload(R.D, r < p && col < n ? (od + r * n + col) * 4 : kOOB);
    }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int r = 4 * g + s;
    d.Et[s] = rload(R.E, r < p && c < m ? (oe + r * m + c) * 4 : kOOB);
  }
}

// a vector's rows 16 i + c (every g) into the column layout (entries 16 k + 4 g + s)
template <int NT>
__device__ __forceinline__ void rows_to_cols(const float (&yr)[NT], f4 (&yc)[NT], int g) {
#pragma unroll
  for (int k = 0; k < NT; ++k)
#pragma unroll
    for (int s = 0; s < 4; ++s) yc[k][s] = lane_read(yr[k], 20 * g + s);
}

// dv_t = lam_t - V_t (stored) and the lane's parts of the slope partials of
// stage t, added to acc (q'dx + r'du, (dv - V)'c, lam'c)
template <int NT>
__device__ __forceinline__ void adj_out(const float (&lam)[NT], const float (&xr)[NT], const float (&qr)[NT],
                                        const float (&vr)[NT], const float (&cr)[NT], float du, float rh,
                                        rsrc_t rdv, int t, int n, int g, int c, float (&acc)[3]) {
  acc[0] += g == 0 ? rh * du : 0.f;
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    const float dv = lam[i] - vr[i];
    const bool row_ok = g == 0 && 16 * i + c < n;
    acc[0] += row_ok ? qr[i] * xr[i] : 0.f;
    acc[1] += row_ok ? (dv - vr[i]) * cr[i] : 0.f;
    acc[2] += row_ok ? lam[i] * cr[i] : 0.f;
    rstore(rdv, row_ok ? (t * n + 16 * i + c) * 4 : kOOB, dv);
  }
}

template <int NT, bool WIDE>
__global__ void __launch_bounds__(64) k_adjoint_tanh(Prob<float> P, Prm<float> prm, Iter<float> it, Ws<float> ws) {
  const int b = blockIdx.x;
  if (skip_instance(ws, b)) return;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const float hs = P.h, hk = P.h * P.kappa, rc = prm.rho_c;
  const AdjR R = adj_rsrc(P, it, ws, b);
  int off[NT][4];
  aop_offsets<NT>(off, n, g, c);
  // the slope partials are summed over the stages by the line search: the
  // instance's totals go to stage N's entry, the other stages' are 0
  float acc[3] = {0.f, 0.f, 0.f};

  // the terminal stage: lam_N = Q_N dx_N + q_N
  float lamr[NT];
  f4 lamc[NT];
  {
    f4 QN[NT][NT], xc[NT];
    load_aop_raw<NT, NT, WIDE>(QN, P.QN.at(b, 0), n, n, off);
    const size_t st = (size_t)b * N1 + N;
    const rsrc_t rx = make_rsrc(ws.dx + st * n, n), rq = make_rsrc(ws.qh + st * n, n);
    const rsrc_t rV = make_rsrc(it.V + st * n, n), rcd = make_rsrc(ws.cdef + st * n, n);
    float xr[NT], qr[NT], vr[NT], cr[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) xc[k] = vload4<WIDE>(rx, 16 * k + 4 * g, 16 * k + 4 * g, n);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const int o = (16 * i + c) * 4;
      xr[i] = rload(rx, o);
      qr[i] = rload(rq, o);
      vr[i] = rload(rV, o);
      cr[i] = rload(rcd, o);
    }
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      float pa = 0.f;
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) pa += QN[i][k][s] * xc[k][s];
      lamr[i] = col_sum4(pa) + qr[i];
    }
    adj_out<NT>(lamr, xr, qr, vr, cr, 0.f, 0.f, R.dv, N, n, g, c, acc);
    rows_to_cols<NT>(lamr, lamc, g);
  }

  // Q (rows 16 i + c, columns 16 k + 4 g + s) in registers; (I + h A)' and
  // (h kappa W)' in the same layout in LDS, each lane reading back only its
  // own entries
  __shared__ f4 atw[2 * NT * NT * 64];
  f4 Qt[NT][NT];
  {
    f4 At[NT][NT], Wt[NT][NT];
    const rsrc_t rA = make_rsrc(P.A.at(b, 0), n * n), rW = make_rsrc(P.W.at(b, 0), n * n);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int row = 16 * k + 4 * g + s, col = 16 * i + c;   // the transposed entry
          const int o = row < n && col < n ? (row * n + col) * 4 : kOOB;
          At[i][k][s] = rload(rA, o);
          Wt[i][k][s] = rload(rW, o);
        }
    load_aop_raw<NT, NT, WIDE>(Qt, P.Q.at(b, 0), n, n, off);
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int k = 0; k < NT; ++k) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int row = 16 * k + 4 * g + s, col = 16 * i + c;
          At[i][k][s] = (row == col && row < n ? 1.f : 0.f) + hs * At[i][k][s];
          Wt[i][k][s] = hk * Wt[i][k][s];
        }
        atw[(i * NT + k) * 64 + lane] = At[i][k];
        atw[(NT * NT + i * NT + k) * 64 + lane] = Wt[i][k];
      }
  }
  // stage t from lam_{t+1} (in lamc)
  auto step = [&](const AdjT<NT>& d, int t) {
    // dx_t and sd_t * lam_{t+1} by the entries 16 k + 4 g + s
    f4 xc[NT], zc[NT];
    rows_to_cols<NT>(d.xr, xc, g);
    rows_to_cols<NT>(d.sr, zc, g);
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) zc[k][s] *= lamc[k][s];
    // h = D dx + E du: the lane's rows 4 g + s, summed over the row of lanes
    // (also the equality rows' slope along the step, for the trials)
    f4 w;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      float hp = d.Et[s] * d.du;
#pragma unroll
      for (int i = 0; i < NT; ++i) hp += d.Dt[i][s] * d.xr[i];
      const float h = row_sum16(hp);
      rstore(R.hq, c == 0 && 4 * g + s < p ? (t * 2 * p + p + 4 * g + s) * 4 : kOOB, h);
      w[s] = rc * h;
    }
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      float pa = 0.f;
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const f4 a = atw[(i * NT + k) * 64 + lane], wt = atw[(NT * NT + i * NT + k) * 64 + lane];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          pa += a[s] * lamc[k][s];
          pa += wt[s] * zc[k][s];
          pa += Qt[i][k][s] * xc[k][s];
        }
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) pa += d.Dt[i][s] * w[s];
      lamr[i] = col_sum4(pa) + d.qr[i];
    }
    adj_out<NT>(lamr, d.xr, d.qr, d.vr, d.cr, d.du, d.rh, R.dv, t, n, g, c, acc);
    rows_to_cols<NT>(lamr, lamc, g);
  };

  // the stages N - 1 .. 0, each one's data requested two stages ahead
  AdjT<NT> d0, d1, d2;
  adj_load<NT>(d0, R, n, m, p, N - 1, g, c);
  adj_load<NT>(d1, R, n, m, p, N > 1 ? N - 2 : 0, g, c);
  for (int t = N - 1; t >= 0; t -= 3) {
    adj_load<NT>(d2, R, n, m, p, t > 1 ? t - 2 : 0, g, c);
    __builtin_amdgcn_sched_barrier(0);
    step(d0, t);
    if (t == 0) break;
    adj_load<NT>(d0, R, n, m, p, t > 2 ? t - 3 : 0, g, c);
    __builtin_amdgcn_sched_barrier(0);
    step(d1, t - 1);
    if (t == 1) break;
    adj_load<NT>(d1, R, n, m, p, t > 3 ? t - 4 : 0, g, c);
    __builtin_amdgcn_sched_barrier(0);
    step(d2, t - 2);
  }
  const float df = wave_sum(acc[0]), dc = wave_sum(acc[1]), lc = wave_sum(acc[2]);
  // stage N's entry carries the totals, the others 0
  const rsrc_t rfp = make_rsrc(ws.fpart + (size_t)b * N1 * NFP, N1 * NFP);
  for (int t = lane; t < N; t += 64) {
#pragma unroll
    for (int q = 0; q < NFP; ++q) rstore(rfp, (t * NFP + q) * 4, 0.f);
  }
  if (lane == 0) {
    rstore(rfp, (N * NFP + FP_DF) * 4, df);
    rstore(rfp, (N * NFP + FP_DC) * 4, dc);
    rstore(rfp, (N * NFP + FP_LC) * 4, lc);
    rstore(rfp, (N * NFP + 3) * 4, 0.f);
  }
}

}  // namespace mpxk
