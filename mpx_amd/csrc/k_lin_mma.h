// The linearisation (K1) on the matrix cores, for the tanh model with
// stage-invariant A, W, B, Q, R (FP32, n <= 48, m <= 16, p <= 16): one wave
// takes 16 consecutive stages t < N of one instance, stage 16 blockIdx.x + c
// in lane column c, in the layout of k_trials_mma (k_trials_mma.h): the stage
// vectors are the columns of B-operand tiles, the matrices A-operand tiles,
// and every per-row quantity of stage c sits in lane (g, c), register v of
// tile i for row 16 i + 4 g + v.  Per stage it writes what stage_terms
// (k_inst.h) writes in linearisation mode -- q (with D'(mu + rho_c h)), r
// (with E'(mu + rho_c h) and the box barrier gradient), the box Hessian
// diagonal, the defect c_t = f(x_t, u_t) - x_{t+1}, the merit parts and the
// box multipliers that are not positive replaced -- and the projection of u_t
// into the box interior, from the products A x, W x, B u, Q (x - x_ref) and
// R (u - u_ref) on the matrix cores; sech^2(W x_t) (k_fwd_tanh's form of
// the Jacobian) and the equality rows h_t at the iterate (the trials').  The Jacobian
//   f_x = I + h (A + kappa diag(sech^2(W x)) W)
// is written last (wfx != 0) by the whole wave, stage by stage: each lane keeps the
// entries e = 256 k + 4 lane .. + 3 of I + h A and h kappa W (the same for
// every stage) and reads the stage's sech^2 of row e / n from LDS, so every
// store is 1 KiB of consecutive floats.  The terminal stage, the initial-state
// row and f_u = h B are k_linearize_term's.
#pragma once
#include "k_inst.h"
#include "k_trials_mma.h"

namespace mpxk {

typedef unsigned int mpxk_u32x4 __attribute__((ext_vector_type(4)));

// Store entries e0 .. e0 + 3 of a vector at element offset o, those with
// e0 + v < lim when ok (the others get kOOB offsets and are dropped).  WIDE as
// for vload4.
template <bool WIDE>
__device__ __forceinline__ void vstore4(rsrc_t r, int o, int e0, int lim, bool ok, const f4& x) {
  if constexpr (WIDE) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(mpxk_u32x4, x), r, ok && e0 < lim ? o * 4 : kOOB, 0,
                                           0);
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) rstore(r, ok && e0 + v < lim ? (o + v) * 4 : kOOB, x[v]);
  }
}

template <int NT, bool WIDE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NT == 1 ? 3 : 2)))
k_linearize_mma(Prob<float> P, Prm<float> prm, Iter<float> it, Ws<float> ws, int wfx) {
  constexpr int SD = 16 * NT + 1;   // one stage's sech^2(W x) in LDS
  __shared__ float sdl[16 * SD];
  const int b = blockIdx.y, lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const int t = blockIdx.x * 16 + c;
  const bool live = t < N;
  const int tc = live ? t : N - 1;   // a stage with data for the columns past N - 1 (nothing of theirs is kept)
  constexpr int BIG = 1 << 30;       // no guard: entries past the vector's end meet zero matrix columns
  constexpr int RQ = NT == 3 ? 2 : 4;
  const float hs = P.h, kap = P.kappa;
  const bool box = P.box != 0;
  const int mb = box ? m : 0;        // the box loads and stores of a problem without boxes go nowhere
  int off[NT][4];
  aop_offsets<NT>(off, n, g, c);

  // (1) the stage vectors x_t, u_t (B-operand tiles), the box, the A-operand
  // tiles of A, B, W, and x_{t+1}; issued together (the scheduler would sink
  // each load to its use)
  const rsrc_t rX = make_rsrc(it.X + (size_t)b * N1 * n, N1 * n);
  const rsrc_t rU = make_rsrc(it.U + (size_t)b * N * m, N * m);
  const rsrc_t rlo = view_rsrc(P.ulo, b, N, m), rhi = view_rsrc(P.uhi, b, N, m);
  f4 X0[NT], U0, lo, hi, Mt[NT][NT], Wt[NT][NT], Bt[NT], Xn[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) X0[k] = vload4<WIDE>(rX, tc * n + 16 * k + 4 * g, 0, BIG);
  U0 = vload4<WIDE>(rU, tc * m + 4 * g, 0, BIG);
  lo = vload4<WIDE>(rlo, (int)(tc * P.ulo.st) + 4 * g, 4 * g, mb);
  hi = vload4<WIDE>(rhi, (int)(tc * P.uhi.st) + 4 * g, 4 * g, mb);
  load_aop_raw<NT, NT, WIDE>(Mt, P.A.at(b, 0), n, n, off);
  {
    const rsrc_t rB = make_rsrc(P.Bm.at(b, 0), n * m);
#pragma unroll
    for (int i = 0; i < NT; ++i) Bt[i] = vload4<WIDE>(rB, (16 * i + c) * m + 4 * g, 4 * g, m);
  }
  load_aop_raw<NT, NT, WIDE>(Wt, P.W.at(b, 0), n, n, off);
#pragma unroll
  for (int i = 0; i < NT; ++i) Xn[i] = vload4<WIDE>(rX, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
  __builtin_amdgcn_sched_barrier(0);

  // u_t projected into the box interior (written back in place)
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float w = 1e-3f * (hi[v] - lo[v]);
    float u = U0[v];
    u = u >= hi[v] ? hi[v] - w : u;
    u = u <= lo[v] ? lo[v] + w : u;
    U0[v] = 4 * g + v < mb ? u : U0[v];
  }
  vstore4<WIDE>(rU, tc * m + 4 * g, 4 * g, mb, live, U0);

  // f = x + h (A x + B u + kappa tanh(W x)), the defect c = f - x_{t+1} (0 past
  // row n) and sech^2(W x) = 1 - tanh^2
  f4 C[NT], SDv[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    f4 a = zero4(), w = zero4();
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a = mma(Mt[i][k][s], X0[k][s], a);
        w = mma(Wt[i][k][s], X0[k][s], w);
      }
#pragma unroll
    for (int s = 0; s < 4; ++s) a = mma(Bt[i][s], U0[s], a);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float th = dev_tanh(w[v]);
      SDv[i][v] = 1.f - th * th;
      C[i][v] = 16 * i + 4 * g + v < n ? X0[i][v] + hs * (a[v] + kap * th) - Xn[i][v] : 0.f;
    }
  }

  // (2) Q, R, x_ref, u_ref, V_{t+1} and the box multipliers
  load_aop_raw<NT, NT, WIDE>(Mt, P.Q.at(b, 0), n, n, off);
  f4 xr[NT], ur, Rt, Vn[NT], nhi, nlo;
  const rsrc_t rYb = make_rsrc(it.Ybox ? it.Ybox + (size_t)b * N * 2 * m : nullptr, N * 2 * m);
  {
    const rsrc_t rxr = view_rsrc(P.xref, b, N1, n), rur = view_rsrc(P.uref, b, N, m);
#pragma unroll
    for (int k = 0; k < NT; ++k)
      xr[k] = vload4<WIDE>(rxr, (int)(tc * P.xref.st) + 16 * k + 4 * g, 16 * k + 4 * g, n);
    ur = vload4<WIDE>(rur, (int)(tc * P.uref.st) + 4 * g, 4 * g, m);
    const rsrc_t rR = make_rsrc(P.R.at(b, 0), m * m);
    Rt = vload4<WIDE>(rR, c * m + 4 * g, 4 * g, m);
    const rsrc_t rV = make_rsrc(it.V + (size_t)b * N1 * n, N1 * n);
#pragma unroll
    for (int i = 0; i < NT; ++i) Vn[i] = vload4<WIDE>(rV, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
    nhi = vload4<WIDE>(rYb, tc * 2 * m + 4 * g, 4 * g, mb);
    nlo = vload4<WIDE>(rYb, tc * 2 * m + m + 4 * g, 4 * g, mb);
  }
  __builtin_amdgcn_sched_barrier(0);

  // the stage's sech^2 rows for the Jacobian
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) sdl[c * SD + 16 * i + 4 * g + v] = SDv[i][v];

  // merit parts: V'^T c, |c|^2, the tracking cost 1/2 d' Q d (d = x - x_ref)
  // and the control cost 1/2 e' R e (e = u - u_ref); q = Q d, r = R e
  float cost = 0.f, vc = 0.f, cc = 0.f;
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      vc += Vn[i][v] * C[i][v];
      cc += C[i][v] * C[i][v];
    }
  f4 qa[NT], ra;
  {
    f4 D0[NT], E0;
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) D0[k][v] = X0[k][v] - xr[k][v];
#pragma unroll
    for (int v = 0; v < 4; ++v) E0[v] = U0[v] - ur[v];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      f4 q = zero4();
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) q = mma(Mt[i][k][s], D0[k][s], q);
#pragma unroll
      for (int v = 0; v < 4; ++v) cost += 0.5f * D0[i][v] * q[v];   // Q's rows past n are 0
      qa[i] = q;
    }
    ra = zero4();
#pragma unroll
    for (int s = 0; s < 4; ++s) ra = mma(Rt[s], E0[s], ra);
#pragma unroll
    for (int v = 0; v < 4; ++v) cost += 0.5f * E0[v] * ra[v];   // R's rows past m are 0
  }

  // equality rows h = e + D x + E u, w = mu + rho_c h: q += D' w, r += E' w,
  // cost += mu h + rho_c / 2 h^2 (once per column); RQ rows per pass, the rows
  // past p of the last pass read other rows' data (or 0) and are dropped
  if (p > 0) {
    const rsrc_t rD = view_rsrc(P.D, b, N, p * n), rE = view_rsrc(P.E, b, N, p * m);
    const rsrc_t re = view_rsrc(P.e, b, N, p);
    const rsrc_t rmu = make_rsrc(it.Yeq + (size_t)b * N * p, N * p);
    const int od = (int)(tc * P.D.st) + 4 * g, oe = (int)(tc * P.E.st) + 4 * g;
    // h at the iterate, for the trials (one descriptor for the instance: a
    // per-lane base would make the descriptor divergent)
    const rsrc_t rhq = make_rsrc(ws.hq + (size_t)b * N * 2 * p, N * 2 * p);
    for (int r0 = 0; r0 < p; r0 += RQ) {
      f4 Dr[RQ][NT], Er[RQ];
      float ev[RQ], mv[RQ];
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        const int r = r0 + q;
#pragma unroll
        for (int k = 0; k < NT; ++k) Dr[q][k] = vload4<WIDE>(rD, od + r * n + 16 * k, 16 * k + 4 * g, n);
        Er[q] = vload4<WIDE>(rE, oe + r * m, 4 * g, m);
        ev[q] = rload(re, (int)(tc * P.e.st + r) * 4);
        mv[q] = rload(rmu, (tc * p + r) * 4);
      }
#pragma unroll
      for (int q = 0; q < RQ; ++q) {
        float h = 0.f;
#pragma unroll
        for (int k = 0; k < NT; ++k)
#pragma unroll
          for (int v = 0; v < 4; ++v) h += Dr[q][k][v] * X0[k][v];
#pragma unroll
        for (int v = 0; v < 4; ++v) h += Er[q][v] * U0[v];
        h = col_sum4(h) + ev[q];
        const bool in = r0 + q < p;
        rstore(rhq, live && in && ((r0 + q) & 3) == g ? (tc * 2 * p + r0 + q) * 4 : kOOB, h);
        const float w = in ? mv[q] + prm.rho_c * h : 0.f;
        cost += in && g == 0 ? mv[q] * h + 0.5f * prm.rho_c * h * h : 0.f;
#pragma unroll
        for (int k = 0; k < NT; ++k)
#pragma unroll
          for (int v = 0; v < 4; ++v) qa[k][v] += Dr[q][k][v] * w;
#pragma unroll
        for (int v = 0; v < 4; ++v) ra[v] += Er[q][v] * w;
      }
    }
  }

  // the box: barrier cost and gradient, the Hessian diagonal, and the
  // multipliers that are not positive replaced by mu_b / slack
  f4 rd = zero4();
  if (box) {
    f4 nh = nhi, nl = nlo;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const bool ok = 4 * g + v < m;
      const float shi = hi[v] - U0[v], slo = U0[v] - lo[v];
      cost -= ok ? prm.mu_b * (dev_log(shi) + dev_log(slo)) : 0.f;
      const bool rh_ = ok && !(nh[v] > 0.f), rl_ = ok && !(nl[v] > 0.f);
      nh[v] = rh_ ? prm.mu_b / shi : nh[v];
      nl[v] = rl_ ? prm.mu_b / slo : nl[v];
      rstore(rYb, live && rh_ ? (tc * 2 * m + 4 * g + v) * 4 : kOOB, nh[v]);
      rstore(rYb, live && rl_ ? (tc * 2 * m + m + 4 * g + v) * 4 : kOOB, nl[v]);
      ra[v] += ok ? prm.mu_b / shi - prm.mu_b / slo : 0.f;
      rd[v] = ok ? nh[v] / shi + nl[v] / slo : 0.f;
    }
  }

  // the stage's outputs
  {
    const rsrc_t rqh = make_rsrc(ws.qh + (size_t)b * N1 * n, N1 * n);
    const rsrc_t rcd = make_rsrc(ws.cdef + (size_t)b * N1 * n, N1 * n);
    const rsrc_t rrh = make_rsrc(ws.rh + (size_t)b * N * m, N * m);
    const rsrc_t rrd = make_rsrc(ws.rd + (size_t)b * N * m, N * m);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      vstore4<WIDE>(rqh, t * n + 16 * i + 4 * g, 16 * i + 4 * g, n, live, qa[i]);
      vstore4<WIDE>(rcd, (t + 1) * n + 16 * i + 4 * g, 16 * i + 4 * g, n, live, C[i]);
    }
    const rsrc_t rsd = make_rsrc(ws.sdh + (size_t)b * N * n, N * n);
#pragma unroll
    for (int i = 0; i < NT; ++i) vstore4<WIDE>(rsd, t * n + 16 * i + 4 * g, 16 * i + 4 * g, n, live, SDv[i]);
    vstore4<WIDE>(rrh, t * m + 4 * g, 4 * g, m, live, ra);
    vstore4<WIDE>(rrd, t * m + 4 * g, 4 * g, m, live, rd);
    cost = col_sum4(cost);
    vc = col_sum4(vc);
    cc = col_sum4(cc);
    if (live && g == 0) {
      float* mp = ws.mpart + ((size_t)b * N1 + t) * NMP;
      mp[MP_COST] = cost;
      mp[MP_VC] = vc;
      mp[MP_CC] = cc;
    }
  }

  // the Jacobians of the wave's stages (wfx = 0: the LQR kernels form them)
  if (!wfx) return;
  __syncthreads();   // (one wave per block) the sech^2 rows in LDS
  const int nn = n * n;
  constexpr int KC = NT * NT;   // ceil((16 NT)^2 / 256) chunks of 256 entries
  constexpr int VR = WIDE ? 1 : 4;
  const rsrc_t rA = make_rsrc(P.A.at(b, 0), nn), rW = make_rsrc(P.W.at(b, 0), nn);
  f4 hA[KC], hW[KC];
  int row[KC][VR];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int e = 256 * k + 4 * lane;
    hA[k] = vload4<WIDE>(rA, e, e, nn);
    hW[k] = vload4<WIDE>(rW, e, e, nn);
#pragma unroll
    for (int v = 0; v < VR; ++v) row[k][v] = min((e + v) / n, n - 1);   // (entries past n^2 are not stored)
  }
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int e = 256 * k + 4 * lane;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = row[k][WIDE ? 0 : v];
      hA[k][v] = (e + v - i * n == i ? 1.f : 0.f) + hs * hA[k][v];
      hW[k][v] = hs * kap * hW[k][v];
    }
  }
  const int t0 = blockIdx.x * 16, ns = min(16, N - t0);
  const rsrc_t rAl = make_rsrc(ws.Alin + (size_t)b * N * nn, N * nn);
  for (int tt = 0; tt < ns; ++tt) {
    const float* sd = sdl + tt * SD;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int e = 256 * k + 4 * lane;
      f4 y;
#pragma unroll
      for (int v = 0; v < 4; ++v) y[v] = hA[k][v] + sd[row[k][WIDE ? 0 : v]] * hW[k][v];
      vstore4<WIDE>(rAl, (t0 + tt) * nn + e, e, nn, true, y);
    }
  }
}

// The terminal stage (the cost of x_N, q_N), the initial-state row
// c_init = x0 - X_0 and the merit parts of both, and f_u = h B of the tanh
// model: one wave per instance, the arithmetic of terminal_terms (k_inst.h)
// with Q_N (x_N - x_ref,N) as A-operand tiles of Q_N (rows 16 i + c, columns
// 16 k + 4 g + s) times the entries 16 k + 4 g + s of the difference: 4 NT^2
// products per lane and a sum over the four lanes of a column.
template <int NT, bool WIDE>
__global__ void __launch_bounds__(64) k_linearize_term(Prob<float> P, Iter<float> it, Ws<float> ws) {
  const int b = blockIdx.x, lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  int off[NT][4];
  aop_offsets<NT>(off, n, g, c);
  f4 QN[NT][NT], xc[NT], rc[NT];
  load_aop_raw<NT, NT, WIDE>(QN, P.QN.at(b, 0), n, n, off);
  const rsrc_t rX = make_rsrc(it.X + (size_t)b * N1 * n, N1 * n);
  const rsrc_t rxr = make_rsrc(P.xref.at(b, N), n);
  const rsrc_t rx0 = make_rsrc(P.x0.at(b, 0), n), rV = make_rsrc(it.V + (size_t)b * N1 * n, n);
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    xc[k] = vload4<WIDE>(rX, N * n + 16 * k + 4 * g, 16 * k + 4 * g, n);
    rc[k] = vload4<WIDE>(rxr, 16 * k + 4 * g, 16 * k + 4 * g, n);
  }
  float dr[NT], x0r[NT], X0r[NT], V0r[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    const int row = 16 * i + c, o = row < n ? row * 4 : kOOB;
    dr[i] = rload(rX, row < n ? (N * n + row) * 4 : kOOB) - rload(rxr, o);
    x0r[i] = rload(rx0, o);
    X0r[i] = rload(rX, o);
    V0r[i] = rload(rV, o);
  }
  const rsrc_t rqh = make_rsrc(ws.qh + (size_t)b * N1 * n, N1 * n);
  const rsrc_t rcd = make_rsrc(ws.cdef + (size_t)b * N1 * n, n);
  float cost = 0.f, vc = 0.f, cc = 0.f;
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    float pa = 0.f;
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) pa += QN[i][k][s] * (xc[k][s] - rc[k][s]);
    const float q = col_sum4(pa);
    const int row = 16 * i + c;
    const bool ok = g == 0 && row < n;
    const float c0 = x0r[i] - X0r[i];
    cost += ok ? 0.5f * dr[i] * q : 0.f;
    vc += ok ? V0r[i] * c0 : 0.f;
    cc += ok ? c0 * c0 : 0.f;
    rstore(rqh, ok ? (N * n + row) * 4 : kOOB, q);
    rstore(rcd, ok ? row * 4 : kOOB, c0);
  }
  cost = wave_sum(cost);
  vc = wave_sum(vc);
  cc = wave_sum(cc);
  if (lane == 0) {
    float* mp = ws.mpart + ((size_t)b * N1 + N) * NMP;
    mp[MP_COST] = cost;
    mp[MP_VC] = vc;
    mp[MP_CC] = cc;
  }
  const rsrc_t rB = make_rsrc(P.Bm.at(b, 0), n * m), rBl = make_rsrc(ws.Blin + (size_t)b * n * m, n * m);
  for (int e = lane; e < n * m; e += 64) rstore(rBl, e * 4, P.h * rload(rB, e * 4));
}

}  // namespace mpxk
