// The trial kernel of the line search on the matrix cores, for the tanh model
// with stage-invariant A, W, B, Q, R (FP32, n <= 48, m <= 16, p <= 16): one
// wave takes 16 consecutive stages t < N of one instance, stage 16 blockIdx.x +
// c in lane column c.  The merit parts of k_trials (k_trials.h; the terminal
// stage stays there) need, per stage, the products of A, W, Q, B and R with the
// stage's vectors at the iterate and along the step; here they are products of
// the stage-invariant matrices with 16 stages at once: the vectors of the
// wave's stages are the columns of B-operand tiles (lane (g, c), register v
// of tile k: entry 16 k + 4 g + v of stage c's vector), the matrices the
// A-operand tiles, and each result tile holds rows 16 i + 4 g + v of stage c
// in lane (g, c), register v -- the layout of the stage's other vectors, so
// the per-trial arithmetic runs on every lane with no rearrangement.  The
// equality rows (D_t, E_t vary over the stages) are dot products: each lane
// takes the entries of its stage's vectors it holds, the four lanes of a
// column add their parts (col_sum4).  The quadratic terms of the merit
// (tracking cost, control cost, equality rows) are polynomials in the step a
// whose coefficients are summed once; only the model's tanh, the defect and
// the box logarithms are evaluated per trial.
//
// The kernel is bound by the latency of its loads (each one gathers 16
// stages): WIDE (n, m multiples of 4, every row 16-byte aligned) moves a
// lane's four consecutive entries with one 128-bit load, and the equality rows
// go four at a time, so that one wait covers the loads of four rows.
#pragma once
#include "k_lqr_wave.h"
#include "k_trials.h"

namespace mpxk {

// Sum over the four rows of 16 lanes of lane column c, in every lane of the
// column (the upper and lower halves of the wave, then the pairs of rows).
__device__ __forceinline__ float col_sum4(float x) {
  const unsigned u = __float_as_uint(x);
  const auto q = __builtin_amdgcn_permlane32_swap(u, u, false, false);
  const float y = __uint_as_float(q[0]) + __uint_as_float(q[1]);   // r0 + r2, r1 + r3 in every row
  const unsigned w = __float_as_uint(y);
  const auto r = __builtin_amdgcn_permlane16_swap(w, w, false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Descriptor over the stages 0 .. nst - 1 of view v for instance b (stage
// stride st, len elements per stage); element e of stage t at byte (t st + e) 4.
__device__ __forceinline__ rsrc_t view_rsrc(const View<float>& v, int b, int nst, int len) {
  return make_rsrc(v.at(b, 0), (int)((nst - 1) * v.st + len));
}

// Entries e0 .. e0 + 3 of a vector at element offset o (e0 = 16 k + 4 g), 0
// where e0 + v >= lim when guarded (kOOB offsets: no arithmetic on the loads).
// WIDE: lim % 4 == 0 and o % 4 == 0, so the four are all below lim or all
// past it and one 128-bit load takes them.
template <bool WIDE>
__device__ __forceinline__ f4 vload4(rsrc_t r, int o, int e0, int lim) {
  if constexpr (WIDE) {
    return rload4(r, e0 < lim ? o * 4 : kOOB);
  } else {
    f4 x;
#pragma unroll
    for (int v = 0; v < 4; ++v) x[v] = rload(r, e0 + v < lim ? (o + v) * 4 : kOOB);
    return x;
  }
}

// Trials j0 <= j < j1 at the stages 16 blockIdx.x + c < N of instance blockIdx.y.
template <int NT, bool WIDE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NT == 1 ? 3 : 2)))
k_trials_mma(Prob<float> P, Prm<float> prm, Iter<float> it, Ws<float> ws, int j0, int j1) {
  const int b = blockIdx.y, lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  // an earlier pass accepted: the same answer in every lane (one instance)
  if (j0 > 0 && ls_first_accept(prm, ws, b, N1, j0) >= 0) return;
  const int t = blockIdx.x * 16 + c;
  const bool live = t < N;
  const int tc = live ? t : N - 1;   // a stage with data for the columns past N - 1 (nothing of theirs is kept)
  constexpr int BIG = 1 << 30;       // no guard: entries past the vector's end meet zero matrix columns

  // The loads go in three groups, each issued before the arithmetic of the
  // one before it (pinned by scheduling barriers: the scheduler would sink
  // each load to its use, one wait per tile): (1) the stage vectors and the
  // data of the defect and of W x, (2) the data of the tracking and control
  // costs, (3) the costate at t + 1 and the box; the equality rows follow in
  // passes of RQ rows.
  constexpr int RQ = NT == 3 ? 2 : 4;
  const float hs = P.h;
  int off[NT][4];
  aop_offsets<NT>(off, n, g, c);

  // (1) B-operand tiles of the stage vectors (entries past the vector's end
  // read the next stage's or 0: they meet the zero columns of the matrix
  // tiles); the A-operand tiles of A, B, W (rows 16 i + c, columns
  // 16 k + 4 g + s; 0 past their ends); x_{t+1} and dx_{t+1}
  const rsrc_t rX = make_rsrc(it.X + (size_t)b * N1 * n, N1 * n);
  const rsrc_t rdX = make_rsrc(ws.dx + (size_t)b * N1 * n, N1 * n);
  const rsrc_t rU = make_rsrc(it.U + (size_t)b * N * m, N * m);
  const rsrc_t rdU = make_rsrc(ws.du + (size_t)b * N * m, N * m);
  f4 X0[NT], X1[NT], U0, U1;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    X0[k] = vload4<WIDE>(rX, tc * n + 16 * k + 4 * g, 0, BIG);
    X1[k] = vload4<WIDE>(rdX, tc * n + 16 * k + 4 * g, 0, BIG);
  }
  U0 = vload4<WIDE>(rU, tc * m + 4 * g, 0, BIG);
  U1 = vload4<WIDE>(rdU, tc * m + 4 * g, 0, BIG);
  f4 Mt[NT][NT], Wt[NT][NT], Bt[NT], Xn[NT], dXn[NT];
  load_aop_raw<NT, NT, WIDE>(Mt, P.A.at(b, 0), n, n, off);
  {
    const rsrc_t rB = make_rsrc(P.Bm.at(b, 0), n * m);
#pragma unroll
    for (int i = 0; i < NT; ++i) Bt[i] = vload4<WIDE>(rB, (16 * i + c) * m + 4 * g, 4 * g, m);
  }
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    Xn[i] = vload4<WIDE>(rX, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
    dXn[i] = vload4<WIDE>(rdX, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
  }
  load_aop_raw<NT, NT, WIDE>(Wt, P.W.at(b, 0), n, n, off);
  __builtin_amdgcn_sched_barrier(0);

  // the defect c = x + h (A x + B u) - x_{t+1} at the iterate (0) and along the
  // step (1); rows past n contribute nothing
  f4 C0[NT], C1[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    f4 a0 = zero4(), a1 = zero4();
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a0 = mma(Mt[i][k][s], X0[k][s], a0);
        a1 = mma(Mt[i][k][s], X1[k][s], a1);
      }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      a0 = mma(Bt[i][s], U0[s], a0);
      a1 = mma(Bt[i][s], U1[s], a1);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const bool ok = 16 * i + 4 * g + v < n;
      C0[i][v] = ok ? X0[i][v] + hs * a0[v] - Xn[i][v] : 0.f;
      C1[i][v] = ok ? X1[i][v] + hs * a1[v] - dXn[i][v] : 0.f;
    }
  }

  // (2) Q, R, x_ref, u_ref
  load_aop_raw<NT, NT, WIDE>(Mt, P.Q.at(b, 0), n, n, off);
  f4 xr[NT], ur, Rt;
  {
    const rsrc_t rxr = view_rsrc(P.xref, b, N1, n), rur = view_rsrc(P.uref, b, N, m);
#pragma unroll
    for (int k = 0; k < NT; ++k)
      xr[k] = vload4<WIDE>(rxr, (int)(tc * P.xref.st) + 16 * k + 4 * g, 16 * k + 4 * g, n);
    ur = vload4<WIDE>(rur, (int)(tc * P.uref.st) + 4 * g, 4 * g, m);
    const rsrc_t rR = make_rsrc(P.R.at(b, 0), m * m);
    Rt = vload4<WIDE>(rR, c * m + 4 * g, 4 * g, m);
  }
  __builtin_amdgcn_sched_barrier(0);

  // W x: the argument of the model's tanh (W's rows past n are 0)
  f4 w0[NT], w1[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    f4 a0 = zero4(), a1 = zero4();
#pragma unroll
    for (int k = 0; k < NT; ++k)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a0 = mma(Wt[i][k][s], X0[k][s], a0);
        a1 = mma(Wt[i][k][s], X1[k][s], a1);
      }
    w0[i] = a0;
    w1[i] = a1;
  }

  // (3) the multiplier V_{t+1} and its step (rows past n: their defect is 0);
  // the box rows u = U + a dU within (lo, hi) (rows past m: a term of 0)
  f4 vn0[NT], vn1[NT], lo, hi;
  {
    const rsrc_t rV = make_rsrc(it.V + (size_t)b * N1 * n, N1 * n);
    const rsrc_t rdV = make_rsrc(ws.dv + (size_t)b * N1 * n, N1 * n);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      vn0[i] = vload4<WIDE>(rV, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
      vn1[i] = vload4<WIDE>(rdV, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
    }
  }
  const bool box = P.box != 0;
  if (box) {
    const rsrc_t rlo = view_rsrc(P.ulo, b, N, m), rhi = view_rsrc(P.uhi, b, N, m);
    lo = vload4<WIDE>(rlo, (int)(tc * P.ulo.st) + 4 * g, 4 * g, m);
    hi = vload4<WIDE>(rhi, (int)(tc * P.uhi.st) + 4 * g, 4 * g, m);
  }
  __builtin_amdgcn_sched_barrier(0);

  // tracking cost 1/2 (d0 + a x1)' Q (d0 + a x1) and control cost
  // 1/2 (e0 + a u1)' R (e0 + a u1), d0 = x - x_ref, e0 = u - u_ref: the lane's
  // rows' parts of the polynomial (Q's rows past n and R's past m are 0)
  float pq0 = 0.f, pq1 = 0.f, pq2 = 0.f;
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
      f4 q0 = zero4(), q1 = zero4();
#pragma unroll
      for (int k = 0; k < NT; ++k)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          q0 = mma(Mt[i][k][s], D0[k][s], q0);
          q1 = mma(Mt[i][k][s], X1[k][s], q1);
        }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        pq0 += D0[i][v] * q0[v];
        pq1 += D0[i][v] * q1[v] + X1[i][v] * q0[v];
        pq2 += X1[i][v] * q1[v];
      }
    }
    f4 r0 = zero4(), r1 = zero4();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      r0 = mma(Rt[s], E0[s], r0);
      r1 = mma(Rt[s], U1[s], r1);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      pq0 += E0[v] * r0[v];
      pq1 += E0[v] * r1[v] + U1[v] * r0[v];
      pq2 += U1[v] * r1[v];
    }
  }

  // equality rows h = e + D x + E u at the iterate and along the step, and
  // their part of phi, mu h + rho_c / 2 h^2, as a polynomial in a (every lane
  // of a column holds the stage's coefficients); RQ rows per pass, the rows
  // past p of the last pass read other rows' data (or 0) and are dropped
  float pe0 = 0.f, pe1 = 0.f, pe2 = 0.f;
  if (p > 0) {
    const rsrc_t rD = view_rsrc(P.D, b, N, p * n), rE = view_rsrc(P.E, b, N, p * m);
    const rsrc_t re = view_rsrc(P.e, b, N, p);
    const rsrc_t rmu = make_rsrc(it.Yeq + (size_t)b * N * p, N * p);
    const int od = (int)(tc * P.D.st) + 4 * g, oe = (int)(tc * P.E.st) + 4 * g;
    float* hq = ws.hq + ((size_t)b * N + tc) * 2 * p;
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
        const int r = r0 + q;
        float h0 = 0.f, h1 = 0.f;
#pragma unroll
        for (int k = 0; k < NT; ++k)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            h0 += Dr[q][k][v] * X0[k][v];
            h1 += Dr[q][k][v] * X1[k][v];
          }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          h0 += Er[q][v] * U0[v];
          h1 += Er[q][v] * U1[v];
        }
        h0 = col_sum4(h0) + ev[q];
        h1 = col_sum4(h1);
        const float mk = r < p ? 1.f : 0.f;
        pe0 += mk * (mv[q] * h0 + 0.5f * prm.rho_c * h0 * h0);
        pe1 += mk * (mv[q] * h1 + prm.rho_c * h0 * h1);
        pe2 += mk * (0.5f * prm.rho_c * h1 * h1);
        if (j0 == 0 && live && r < p && (r & 3) == g) {
          hq[r] = h0;
          hq[p + r] = h1;
        }
      }
    }
  }
  // the quadratic parts of the lane (the equality rows' once per column)
  const float c0 = 0.5f * pq0 + (g == 0 ? pe0 : 0.f);
  const float c1 = 0.5f * pq1 + (g == 0 ? pe1 : 0.f);
  const float c2 = 0.5f * pq2 + (g == 0 ? pe2 : 0.f);

  f4 bu0, bu1, blo, bhi;
  if (box) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const bool ok = 4 * g + v < m;
      bu0[v] = ok ? U0[v] : 0.f;
      bu1[v] = ok ? U1[v] : 0.f;
      blo[v] = ok ? lo[v] : -1.f;
      bhi[v] = ok ? hi[v] : 1.f;
    }
  }

  const float hk = P.h * P.kappa;
  float a = ws.lsp[(size_t)b * NLS + LS_AMAX];
  for (int q = 0; q < j0; ++q) a *= prm.beta;
  float* tp = ws.tpart + ((size_t)b * N1 + tc) * 2 * MAX_TRIALS;
  for (int j = j0; j < j1; ++j, a *= prm.beta) {
    float s1 = c0 + a * (c1 + a * c2), s2 = 0.f;
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float cc = C0[i][v] + a * C1[i][v] + hk * dev_tanh(w0[i][v] + a * w1[i][v]);
        s1 += (vn0[i][v] + a * vn1[i][v]) * cc;
        s2 += cc * cc;
      }
    if (box) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float u = bu0[v] + a * bu1[v];
        s1 -= prm.mu_b * (dev_log(bhi[v] - u) + dev_log(u - blo[v]));
      }
    }
    s1 = col_sum4(s1);
    s2 = col_sum4(s2);
    if (live && g == 0) {
      tp[2 * j] = s1;
      tp[2 * j + 1] = s2;
    }
  }
}

}  // namespace mpxk
