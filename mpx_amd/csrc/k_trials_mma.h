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
// stages): the block's four waves (64 stages) stage the instance's matrices
// in LDS once, WIDE (n, m multiples of 4, every row 16-byte aligned) moves a
// lane's four consecutive entries with one 128-bit load, and the equality rows
// go several at a time, so that one wait covers the loads of several rows.
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

// The stage-invariant matrices of one instance in LDS for a block of waves:
// rows of LD = 16 NT + 4 floats (LD / 4 odd: the rows 16 i + c of a 128-bit
// A-operand read fall on distinct banks), 16 NT rows, 0 past the matrix.
// (each array is allocated in whole passes of the block's 256 threads, so
// that the staging stores need no guard)
template <int NT>
struct MmaLds {
  static constexpr int LD = 16 * NT + 4, LDM = 20, ROWS = 16 * NT;
  static constexpr int pad(int e) { return (e + 255) / 256 * 256; }
  static constexpr int NN = pad(ROWS * LD), NM = pad(ROWS * LDM), MM = pad(16 * LDM);
};

// rows x cols row-major M into sm (ROWS x LD, allocated in whole 256-entry
// passes), every entry of sm written; the block's 256 threads request all
// their loads before the first store
template <int ROWS, int LD>
__device__ __forceinline__ void lds_stage(float* sm, const float* M, int rows, int cols, int tid) {
  constexpr int PER = (ROWS * LD + 255) / 256;
  const rsrc_t r = make_rsrc(M, rows * cols);
  float v[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = tid + 256 * q, row = e / LD, col = e - row * LD;
    v[q] = rload(r, row < rows && col < cols ? (row * cols + col) * 4 : kOOB);
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) sm[tid + 256 * q] = v[q];
}

// A-operand tile (i, k) of an LDS matrix: row 16 i + c, columns 16 k + 4 g .. + 3
template <int LD>
__device__ __forceinline__ f4 lds_aop(const float* sm, int i, int k, int g, int c) {
  return *reinterpret_cast<const f4*>(sm + (16 * i + c) * LD + 16 * k + 4 * g);
}

// Trials j0 <= j < j1 at the stages 64 blockIdx.x + 16 w + c < N of instance
// blockIdx.y (wave w of the block).  The block stages A, W, Q, B, R in LDS
// once (MmaLds); each wave reads its A-operand tiles from there where a
// product uses them.
// HQ: the equality rows at the iterate and their slope along the step are
// in hq already (k_linearize_mma, k_adjoint_tanh): the stage reads 2 p values
// instead of D_t and E_t.
template <int NT, bool WIDE, bool HQ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NT == 1 || (HQ && !(WIDE && NT == 3)) ? 3 : 2)))
k_trials_mma(Prob<float> P, Prm<float> prm, Iter<float> it, Ws<float> ws, int j0, int j1) {
  using L = MmaLds<NT>;
  __shared__ __align__(16) float smA[L::NN], smW[L::NN], smQ[L::NN], smB[L::NM], smR[L::MM];
  const int b = blockIdx.y, tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, g = lane >> 4, c = lane & 15;
  if (skip_instance(ws, b)) return;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  // an earlier pass accepted: the same answer in every wave (one instance)
  if (j0 > 0 && ls_first_accept(prm, ws, b, N1, j0) >= 0) return;
  lds_stage<L::ROWS, L::LD>(smA, P.A.at(b, 0), n, n, tid);
  lds_stage<L::ROWS, L::LD>(smW, P.W.at(b, 0), n, n, tid);
  lds_stage<L::ROWS, L::LD>(smQ, P.Q.at(b, 0), n, n, tid);
  lds_stage<L::ROWS, L::LDM>(smB, P.Bm.at(b, 0), n, m, tid);
  lds_stage<16, L::LDM>(smR, P.R.at(b, 0), m, m, tid);
  __syncthreads();
  const int t0 = blockIdx.x * 64 + wv * 16;
  if (t0 >= N) return;   // no live stage (after the barrier)
  const int t = t0 + c;
  const bool live = t < N;
  const int tc = live ? t : N - 1;   // a stage with data for the columns past N - 1 (nothing of theirs is kept)
  constexpr int BIG = 1 << 30;       // no guard: entries past the vector's end meet zero matrix columns
  constexpr int RQ = NT == 3 ? 2 : 4;
  const float hs = P.h;

  // The loads go in three groups, each issued ahead of the arithmetic that
  // uses it (pinned by scheduling barriers: the scheduler would sink each
  // load to its use): (1) the stage vectors (B-operand tiles; entries
  // past the vector's end read the next stage's or 0: they meet the zero
  // columns of the matrix tiles), x_{t+1} and dx_{t+1}
  const rsrc_t rX = make_rsrc(it.X + (size_t)b * N1 * n, N1 * n);
  const rsrc_t rdX = make_rsrc(ws.dx + (size_t)b * N1 * n, N1 * n);
  const rsrc_t rU = make_rsrc(it.U + (size_t)b * N * m, N * m);
  const rsrc_t rdU = make_rsrc(ws.du + (size_t)b * N * m, N * m);
  f4 X0[NT], X1[NT], U0, U1, Xn[NT], dXn[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    X0[k] = vload4<WIDE>(rX, tc * n + 16 * k + 4 * g, 0, BIG);
    X1[k] = vload4<WIDE>(rdX, tc * n + 16 * k + 4 * g, 0, BIG);
  }
  U0 = vload4<WIDE>(rU, tc * m + 4 * g, 0, BIG);
  U1 = vload4<WIDE>(rdU, tc * m + 4 * g, 0, BIG);
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    Xn[i] = vload4<WIDE>(rX, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
    dXn[i] = vload4<WIDE>(rdX, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
  }
  // (2) x_ref, u_ref
  f4 xr[NT], ur;
  {
    const rsrc_t rxr = view_rsrc(P.xref, b, N1, n), rur = view_rsrc(P.uref, b, N, m);
#pragma unroll
    for (int k = 0; k < NT; ++k)
      xr[k] = vload4<WIDE>(rxr, (int)(tc * P.xref.st) + 16 * k + 4 * g, 16 * k + 4 * g, n);
    ur = vload4<WIDE>(rur, (int)(tc * P.uref.st) + 4 * g, 4 * g, m);
  }
  __builtin_amdgcn_sched_barrier(0);

  // the defect c = x + h (A x + B u) - x_{t+1} at the iterate (0) and along the
  // step (1); rows past n contribute nothing
  f4 C0[NT], C1[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    f4 a0 = zero4(), a1 = zero4();
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      const f4 Mt = lds_aop<L::LD>(smA, i, k, g, c);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a0 = mma(Mt[s], X0[k][s], a0);
        a1 = mma(Mt[s], X1[k][s], a1);
      }
    }
    const f4 Bt = lds_aop<L::LDM>(smB, i, 0, g, c);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      a0 = mma(Bt[s], U0[s], a0);
      a1 = mma(Bt[s], U1[s], a1);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const bool ok = 16 * i + 4 * g + v < n;
      C0[i][v] = ok ? X0[i][v] + hs * a0[v] - Xn[i][v] : 0.f;
      C1[i][v] = ok ? X1[i][v] + hs * a1[v] - dXn[i][v] : 0.f;
    }
    __builtin_amdgcn_sched_barrier(0);   // one row tile's LDS reads at a time
  }
  // W x: the argument of the model's tanh (W's rows past n are 0)
  f4 w0[NT], w1[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    f4 a0 = zero4(), a1 = zero4();
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      const f4 Mt = lds_aop<L::LD>(smW, i, k, g, c);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        a0 = mma(Mt[s], X0[k][s], a0);
        a1 = mma(Mt[s], X1[k][s], a1);
      }
    }
    w0[i] = a0;
    w1[i] = a1;
    __builtin_amdgcn_sched_barrier(0);
  }

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
      for (int k = 0; k < NT; ++k) {
        const f4 Mt = lds_aop<L::LD>(smQ, i, k, g, c);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          q0 = mma(Mt[s], D0[k][s], q0);
          q1 = mma(Mt[s], X1[k][s], q1);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        pq0 += D0[i][v] * q0[v];
        pq1 += D0[i][v] * q1[v] + X1[i][v] * q0[v];
        pq2 += X1[i][v] * q1[v];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    const f4 Rt = lds_aop<L::LDM>(smR, 0, 0, g, c);
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

  // (3) V_{t+1} and its step, the box
  f4 vn0[NT], vn1[NT], lo, hi;
  {
    const rsrc_t rV = make_rsrc(it.V + (size_t)b * N1 * n, N1 * n);
    const rsrc_t rdV = make_rsrc(ws.dv + (size_t)b * N1 * n, N1 * n);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      vn0[i] = vload4<WIDE>(rV, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
      vn1[i] = vload4<WIDE>(rdV, (tc + 1) * n + 16 * i + 4 * g, 0, BIG);
    }
    const int mb = P.box ? m : 0;   // a problem without boxes: the loads go nowhere
    const rsrc_t rlo = view_rsrc(P.ulo, b, N, m), rhi = view_rsrc(P.uhi, b, N, m);
    lo = vload4<WIDE>(rlo, (int)(tc * P.ulo.st) + 4 * g, 4 * g, mb);
    hi = vload4<WIDE>(rhi, (int)(tc * P.uhi.st) + 4 * g, 4 * g, mb);
  }
  __builtin_amdgcn_sched_barrier(0);

  // equality rows h = e + D x + E u at the iterate and along the step, and
  // their part of phi, mu h + rho_c / 2 h^2, as a polynomial in a (every lane
  // of a column holds the stage's coefficients); RQ rows per pass, the rows
  // past p of the last pass read other rows' data (or 0) and are dropped
  float pe0 = 0.f, pe1 = 0.f, pe2 = 0.f;
  if (HQ && p > 0) {   // rows 4 g + s of the lane's stage, summed over the column
    const rsrc_t rh = make_rsrc(ws.hq + (size_t)b * N * 2 * p, N * 2 * p);
    const rsrc_t rmu = make_rsrc(it.Yeq + (size_t)b * N * p, N * p);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int r = 4 * g + s;
      const bool in = r < p;
      const float h0 = rload(rh, in ? (tc * 2 * p + r) * 4 : kOOB);
      const float h1 = rload(rh, in ? (tc * 2 * p + p + r) * 4 : kOOB);
      const float mu = rload(rmu, in ? (tc * p + r) * 4 : kOOB);
      pe0 += mu * h0 + 0.5f * prm.rho_c * h0 * h0;
      pe1 += mu * h1 + prm.rho_c * h0 * h1;
      pe2 += 0.5f * prm.rho_c * h1 * h1;
    }
    pe0 = col_sum4(pe0);
    pe1 = col_sum4(pe1);
    pe2 = col_sum4(pe2);
  } else if (p > 0) {
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

  const bool box = P.box != 0;
  f4 bu0, bu1, blo, bhi;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const bool ok = 4 * g + v < m;
    bu0[v] = ok ? U0[v] : 0.f;
    bu1[v] = ok ? U1[v] : 0.f;
    blo[v] = ok ? lo[v] : -1.f;
    bhi[v] = ok ? hi[v] : 1.f;
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
