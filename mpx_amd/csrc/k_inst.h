// Instance-per-block kernels for the linearisation (K1) and the merit line
// search (K4): one block of four waves per instance.  The instance's constant
// matrices (A, W, B of the tanh model; Q, R when they do not vary over the
// stages) and its trajectory sit in LDS, rows padded to ld + 1 so that the
// lanes of a wave, one per row, read distinct banks while the vector operand
// is a broadcast.  Each wave takes one stage at a time; lane i owns state row
// i, control i and equality row i (n, m, p <= 64).
//
// stage_terms evaluates, at one stage of a trajectory (x, u) and its next
// state, the augmented-Lagrangian stage cost, the defect c = f(x, u) - x' and
// the merit pieces V'^T c, |c|^2 -- the quantities of k_linearize.h and of
// merit_item in k_linesearch.h -- and in linearisation mode also writes the
// LQR inputs: q (with D'(mu + rho_c h)), r (with E'(mu + rho_c h) and the box
// barrier gradient), the box Hessian diagonal, the defect and the Jacobian
// f_x = I + h (A + kappa diag(sech^2(W x)) W) of the tanh model.  (Forming f_x
// inside the LQR recursion instead measured slower: the recursion is latency
// bound, and the extra loads and products sit on its critical path.)
#pragma once
#include "common.h"
#include "k_lqr.h"

namespace mpxk {

constexpr int kInstWaves = 4;
// per wave: x - xr, u - ur, mu + rho_c h, sech^2(W x), then the stage's
// equality rows [D_t | E_t] when they fit in kEqStage elements
constexpr int kEqStage = 512;
constexpr int kInstScratch = 4 * 64 + kEqStage;

// Offsets (elements) of the instance LDS plan.
struct InstLds {
  int A, W, B, Q, R;      // -1 when not staged
  int X, U, V, scratch, red, total;
};

__host__ __device__ inline InstLds inst_lds(int n, int m, int N, bool tanh_model, bool q_const, bool r_const,
                                            bool with_v) {
  InstLds L;
  int o = 0;
  const int N1 = N + 1;
  L.A = tanh_model ? o : -1; o += tanh_model ? n * (n + 1) : 0;
  L.W = tanh_model ? o : -1; o += tanh_model ? n * (n + 1) : 0;
  L.B = tanh_model ? o : -1; o += tanh_model ? n * (m + 1) : 0;
  L.Q = q_const ? o : -1; o += q_const ? n * (n + 1) : 0;
  L.R = r_const ? o : -1; o += r_const ? m * (m + 1) : 0;
  L.X = o; o += N1 * n;
  L.U = o; o += N * m;
  L.V = with_v ? o : -1; o += with_v ? N1 * n : 0;
  L.scratch = o; o += kInstWaves * kInstScratch;
  L.red = o; o += kInstWaves * 8;
  L.total = o;
  return L;
}

template <typename T>
__device__ void inst_stage_constants(const Prob<T>& P, int b, T* sm, const InstLds& L) {
  const int n = P.n, m = P.m;
  if (L.A >= 0) {
    lds_copy_padded<T, 8>(sm + L.A, P.A.at(b, 0), n, n);
    lds_copy_padded<T, 8>(sm + L.W, P.W.at(b, 0), n, n);
    lds_copy_padded<T, 4>(sm + L.B, P.Bm.at(b, 0), n, m);
  }
  if (L.Q >= 0) lds_copy_padded<T, 8>(sm + L.Q, P.Q.at(b, 0), n, n);
  if (L.R >= 0) lds_copy_padded<T, 4>(sm + L.R, P.R.at(b, 0), m, m);
}

// Stage t < N of the trajectory (x = X[t], u = U[t], x' = X[t+1] in LDS; v' =
// the multiplier of c_t, from vnext).  Returns (cost, v'^T c, |c|^2) summed
// over the wave.  LIN: write q, r, rd, c, the Jacobian and replace the box
// multipliers that are not positive.
template <typename T, bool LIN>
__device__ void stage_terms(const Prob<T>& P, const Prm<T>& prm, const Iter<T>& it, const Ws<T>& ws, int b, int t,
                            const T* sm, const InstLds& L, const T* x, const T* u, const T* xn, const T* vnext,
                            T* scr, T out[3]) {
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const int lane = threadIdx.x & 63;
  T* d = scr;          // x - xr
  T* e = scr + 64;     // u - ur
  T* w = scr + 128;    // mu + rho_c h
  T* sd = scr + 192;   // sech^2(W x)
  T cost = 0, vc = 0, cc = 0;
  // The stage's per-lane entries in global memory are requested together up
  // front: the LDS syncs below keep the compiler from moving a load across
  // them, and each load left behind one costs the wave a memory round trip.
  const T xr_l = lane < n ? P.xref.at(b, t)[lane] : T(0);
  const T ur_l = lane < m ? P.uref.at(b, t)[lane] : T(0);
  const T e_l = lane < p ? P.e.at(b, t)[lane] : T(0);
  const T mu_l = lane < p ? it.Yeq[((size_t)b * N + t) * p + lane] : T(0);
  const T vn_l = lane < n ? vnext[lane] : T(0);
  T lo_l = 0, hi_l = 0, nhi_l = 0, nlo_l = 0;
  if (P.box && lane < m) {
    lo_l = P.ulo.at(b, t)[lane];
    hi_l = P.uhi.at(b, t)[lane];
    if (LIN) {
      const T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
      nhi_l = nu[lane];
      nlo_l = nu[m + lane];
    }
  }
  if (lane < n) d[lane] = x[lane] - xr_l;
  if (lane < m) e[lane] = u[lane] - ur_l;
  // The stage's equality rows [D_t | E_t] go to the wave's scratch with every
  // load in flight at once: the dot-product loops below would otherwise wait
  // on global memory once per group of unrolled iterations.
  const bool eq_lds = p > 0 && p * (n + m) <= kEqStage;
  // FP32 with n % 4 == 0 and x, d 16-byte aligned: the LDS vectors x and d are
  // read four elements at a time
  const bool vec4 = IsF32<T>::value && (n & 3) == 0 && ((reinterpret_cast<size_t>(x) | reinterpret_cast<size_t>(d)) & 15) == 0;
  T* Ds = scr + 256;
  T* Es = Ds + p * n;
  if (eq_lds) {
    const T* D = P.D.at(b, t);
    const T* E = P.E.at(b, t);
    const int pn = p * n, pt = p * (n + m);
    T r[kEqStage / 64];
#pragma unroll
    for (int k = 0; k < kEqStage / 64; ++k) {
      const int idx = lane + 64 * k;
      r[k] = idx < pn ? D[idx] : (idx < pt ? E[idx - pn] : T(0));
    }
#pragma unroll
    for (int k = 0; k < kEqStage / 64; ++k) {
      const int idx = lane + 64 * k;
      if (idx < pt) Ds[idx] = r[k];
    }
    wave_lds_sync();
  }
  if (lane < p) {
    T h = e_l;
    if (eq_lds) {
      for (int j = 0; j < n; ++j) h += Ds[lane * n + j] * x[j];
      for (int j = 0; j < m; ++j) h += Es[lane * m + j] * u[j];
    } else {
      const T* D = P.D.at(b, t);
      const T* E = P.E.at(b, t);
      for (int j = 0; j < n; ++j) h += D[lane * n + j] * x[j];
      for (int j = 0; j < m; ++j) h += E[lane * m + j] * u[j];
    }
    const T mu = mu_l;
    w[lane] = mu + prm.rho_c * h;
    cost += mu * h + T(0.5) * prm.rho_c * h * h;
  }
  wave_lds_sync();
  if (lane < m) {
    const int i = lane;
    T Ru = 0;
    if (L.R >= 0) {
      const T* Rr = sm + L.R + i * (m + 1);
      for (int j = 0; j < m; ++j) Ru += Rr[j] * e[j];
    } else {
      const T* R = P.R.at(b, t);
      for (int j = 0; j < m; ++j) Ru += R[i * m + j] * e[j];
    }
    cost += T(0.5) * e[i] * Ru;
    T r = Ru, dg = 0;
    if (eq_lds) {
      for (int k = 0; k < p; ++k) r += Es[k * m + i] * w[k];
    } else if (p > 0) {
      const T* E = P.E.at(b, t);
      for (int k = 0; k < p; ++k) r += E[k * m + i] * w[k];
    }
    if (P.box) {
      const T lo = lo_l, hi = hi_l;
      const T shi = hi - u[i], slo = u[i] - lo;
      cost -= prm.mu_b * (dev_log(shi) + dev_log(slo));
      if (LIN) {
        T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
        T nhi = nhi_l, nlo = nlo_l;
        if (!(nhi > T(0))) { nhi = prm.mu_b / shi; nu[i] = nhi; }
        if (!(nlo > T(0))) { nlo = prm.mu_b / slo; nu[m + i] = nlo; }
        r += prm.mu_b / shi - prm.mu_b / slo;
        dg = nhi / shi + nlo / slo;
      }
    }
    if (LIN) {
      ws.rh[((size_t)b * N + t) * m + i] = r;
      ws.rd[((size_t)b * N + t) * m + i] = dg;
    }
  }
  if (lane < n) {
    const int i = lane;
    T Qx = 0;
    if (L.Q >= 0) {
      const T* Qr = sm + L.Q + i * (n + 1);
      if (vec4) {   // d four at a time, the same order of sums
        const float4* d4 = reinterpret_cast<const float4*>(d);
        for (int q = 0; q < (n >> 2); ++q) {
          const float4 v = d4[q];
          const T* r = Qr + 4 * q;
          Qx += r[0] * v.x;
          Qx += r[1] * v.y;
          Qx += r[2] * v.z;
          Qx += r[3] * v.w;
        }
      } else {
        for (int j = 0; j < n; ++j) Qx += Qr[j] * d[j];
      }
    } else {
      const T* Q = P.Q.at(b, t);
      for (int j = 0; j < n; ++j) Qx += Q[i * n + j] * d[j];
    }
    cost += T(0.5) * d[i] * Qx;
    T f;
    if (L.A >= 0) {
      const T* Ar = sm + L.A + i * (n + 1);
      const T* Wr = sm + L.W + i * (n + 1);
      const T* Br = sm + L.B + i * (m + 1);
      T ax = 0, wx = 0, bu = 0;
      if (vec4) {   // x four at a time, the same order of sums
        const float4* x4 = reinterpret_cast<const float4*>(x);
        for (int q = 0; q < (n >> 2); ++q) {
          const float4 v = x4[q];
          const T* ra = Ar + 4 * q;
          const T* rw = Wr + 4 * q;
          ax += ra[0] * v.x; wx += rw[0] * v.x;
          ax += ra[1] * v.y; wx += rw[1] * v.y;
          ax += ra[2] * v.z; wx += rw[2] * v.z;
          ax += ra[3] * v.w; wx += rw[3] * v.w;
        }
      } else {
        for (int j = 0; j < n; ++j) { ax += Ar[j] * x[j]; wx += Wr[j] * x[j]; }
      }
      for (int j = 0; j < m; ++j) bu += Br[j] * u[j];
      const T th = dev_tanh(wx);
      if (LIN) sd[i] = T(1) - th * th;
      f = x[i] + P.h * (ax + bu + P.kappa * th);
    } else {
      const T* A = P.A.at(b, t);
      const T* Bm = P.Bm.at(b, t);
      T ax = 0, bu = 0;
      for (int j = 0; j < n; ++j) ax += A[i * n + j] * x[j];
      for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * u[j];
      f = ax + bu;
    }
    const T c = f - xn[i];
    vc += vn_l * c;
    cc += c * c;
    if (LIN) {
      T q = Qx;
      if (eq_lds) {
        for (int k = 0; k < p; ++k) q += Ds[k * n + i] * w[k];
      } else if (p > 0) {
        const T* D = P.D.at(b, t);
        for (int k = 0; k < p; ++k) q += D[k * n + i] * w[k];
      }
      ws.qh[((size_t)b * N1 + t) * n + i] = q;
      ws.cdef[((size_t)b * N1 + t + 1) * n + i] = c;
    }
  }
  if (LIN && L.A >= 0) {
    wave_lds_sync();
    T* Al = ws.Alin + ((size_t)b * N + t) * n * n;
    // element e2 = i n + j sits at e2 + i in the padded LDS rows; (i, j) advance
    // by (64 / n, 64 % n) with a carry instead of a division per element
    const int di = 64 / n, dj = 64 - di * n;
    int i = lane / n, j = lane - (lane / n) * n;
    for (int e2 = lane; e2 < n * n; e2 += 64) {
      Al[e2] = (i == j ? T(1) : T(0)) + P.h * (sm[L.A + e2 + i] + P.kappa * sd[i] * sm[L.W + e2 + i]);
      i += di;
      j += dj;
      if (j >= n) {
        j -= n;
        ++i;
      }
    }
  }
  out[0] = wave_sum(cost);
  out[1] = wave_sum(vc);
  out[2] = wave_sum(cc);
  wave_lds_sync();   // the scratch is rewritten by the next stage
}

// Terminal stage (cost of x_N, q_N) and the initial-state row c_init = x0 - X_0.
template <typename T, bool LIN>
__device__ void terminal_terms(const Prob<T>& P, const Ws<T>& ws, int b, const T* xN, const T* x0row,
                               const T* v0, T* scr, T out[3]) {
  const int n = P.n, N = P.N, N1 = N + 1;
  const int lane = threadIdx.x & 63;
  T* d = scr;
  const T* xr = P.xref.at(b, N);
  if (lane < n) d[lane] = xN[lane] - xr[lane];
  wave_lds_sync();
  T cost = 0, vc = 0, cc = 0;
  if (lane < n) {
    const T* QN = P.QN.at(b, 0);
    T Qx = 0;
    for (int j = 0; j < n; ++j) Qx += QN[lane * n + j] * d[j];
    cost += T(0.5) * d[lane] * Qx;
    const T c = P.x0.at(b, 0)[lane] - x0row[lane];
    vc += v0[lane] * c;
    cc += c * c;
    if (LIN) {
      ws.qh[((size_t)b * N1 + N) * n + lane] = Qx;
      ws.cdef[(size_t)b * N1 * n + lane] = c;
    }
  }
  out[0] = wave_sum(cost);
  out[1] = wave_sum(vc);
  out[2] = wave_sum(cc);
  wave_lds_sync();
}

// (four waves per SIMD in FP32, three in FP64: the stage loop is latency
// bound, and these are the largest register budgets that do not spill)
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 3))) k_linearize_inst(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws) {
  const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  if (skip_instance(ws, b)) return;
  const int lane = tid & 63, wv = tid >> 6;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const InstLds L = inst_lds(n, m, N, P.model == 1, P.Q.st == 0, P.R.st == 0, false);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  inst_stage_constants(P, b, sm, L);
  const T* Xb = it.X + (size_t)b * N1 * n;
  const T* Vb = it.V + (size_t)b * N1 * n;
  lds_copy<T, 8>(sm + L.X, Xb, N1 * n);
  T* Ub = it.U + (size_t)b * N * m;
  // U with the box projection into the interior (in place), four loads of a
  // thread in flight at a time
  for (int base = 0; base < N * m; base += 4 * nt) {
    T uv[4], lov[4], hiv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = base + tid + k * nt;
      uv[k] = lov[k] = hiv[k] = T(0);
      if (e < N * m) {
        uv[k] = Ub[e];
        if (P.box) {
          const int t = e / m, i = e - t * m;
          lov[k] = P.ulo.at(b, t)[i];
          hiv[k] = P.uhi.at(b, t)[i];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = base + tid + k * nt;
      if (e < N * m) {
        T u = uv[k];
        if (P.box) {
          const T lo = lov[k], hi = hiv[k];
          const T w = T(1e-3) * (hi - lo);
          if (u >= hi) u = hi - w;
          if (u <= lo) u = lo + w;
          Ub[e] = u;
        }
        sm[L.U + e] = u;
      }
    }
  }
  __syncthreads();
  if (P.model == 1)   // B of the tanh model's f_u = h B, from its LDS copy
    for (int e = tid; e < n * m; e += nt) ws.Blin[(size_t)b * n * m + e] = P.h * sm[L.B + e + e / m];
  T* scr = sm + L.scratch + wv * kInstScratch;
  for (int t = wv; t < N1; t += kInstWaves) {
    T o[3];
    if (t < N)
      stage_terms<T, true>(P, prm, it, ws, b, t, sm, L, sm + L.X + t * n, sm + L.U + t * m,
                           sm + L.X + (t + 1) * n, Vb + (size_t)(t + 1) * n, scr, o);
    else
      terminal_terms<T, true>(P, ws, b, sm + L.X + N * n, sm + L.X, Vb, scr, o);
    if (lane == 0) {
      T* mp = ws.mpart + ((size_t)b * N1 + t) * NMP;
      mp[MP_COST] = o[0];
      mp[MP_VC] = o[1];
      mp[MP_CC] = o[2];
    }
  }
}

// K1 parallel over stages: the block's 8 waves take consecutive stages of one
// instance and share its constant matrices, staged once in LDS as in
// k_linearize_inst; each wave stages its x_t, u_t (projected into the box
// interior and written back) and x_{t+1} and runs stage_terms on them.  The
// per-stage arithmetic is that of k_linearize_inst, so both write the same
// values; this one keeps more waves in flight per instance.
constexpr int kLinStageWaves = 8;

__host__ __device__ inline int lin_stage_wave_elems(int n, int m) {
  return ((2 * n + m + 3) & ~3) + kInstScratch;
}

template <typename T>
__global__ void __launch_bounds__(512) k_linearize_stage(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws) {
  const int b = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  if (skip_instance(ws, b)) return;
  const int lane = tid & 63, wv = tid >> 6, wpb = nt >> 6;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  const InstLds L = inst_lds(n, m, N, P.model == 1, P.Q.st == 0, P.R.st == 0, false);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  inst_stage_constants(P, b, sm, L);
  if (blockIdx.x == 0 && P.model == 1)
    for (int e = tid; e < n * m; e += nt) ws.Blin[(size_t)b * n * m + e] = P.h * P.Bm.at(b, 0)[e];
  __syncthreads();
  const int t = blockIdx.x * wpb + wv;
  if (t > N) return;                               // whole wave; no barrier follows
  T* x = sm + ((L.X + 3) & ~3) + (size_t)wv * lin_stage_wave_elems(n, m);
  T* xn = x + n;
  T* u = xn + n;
  T* scr = x + ((2 * n + m + 3) & ~3);
  const T* Xb = it.X + (size_t)b * N1 * n;
  const T* Vb = it.V + (size_t)b * N1 * n;
  // terminal wave: x = X_N, xn = X_0 (the initial-state row)
  for (int i = lane; i < n; i += 64) {
    x[i] = Xb[(size_t)t * n + i];
    xn[i] = t < N ? Xb[(size_t)(t + 1) * n + i] : Xb[i];
  }
  if (t < N) {
    T* Ut = it.U + ((size_t)b * N + t) * m;
    for (int i = lane; i < m; i += 64) {
      T v = Ut[i];
      if (P.box) {   // box projection into the interior (in place)
        const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
        const T w = T(1e-3) * (hi - lo);
        if (v >= hi) v = hi - w;
        if (v <= lo) v = lo + w;
        Ut[i] = v;
      }
      u[i] = v;
    }
  }
  wave_lds_sync();
  T o[3];
  if (t < N)
    stage_terms<T, true>(P, prm, it, ws, b, t, sm, L, x, u, xn, Vb + (size_t)(t + 1) * n, scr, o);
  else
    terminal_terms<T, true>(P, ws, b, x, xn, Vb, scr, o);
  if (lane == 0) {
    T* mp = ws.mpart + ((size_t)b * N1 + t) * NMP;
    mp[MP_COST] = o[0];
    mp[MP_VC] = o[1];
    mp[MP_CC] = o[2];
  }
}

// Sum of cnt values over the block's waves (every lane gets the result).
template <typename T, int CNT>
__device__ void inst_block_sum(T (&v)[CNT], T* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < CNT; ++j) v[j] = wave_sum(v[j]);
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < CNT; ++j) red[wv * 8 + j] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < CNT; ++j) {
    T s = 0;
    for (int k = 0; k < kInstWaves; ++k) s += red[k * 8 + j];
    v[j] = s;
  }
  __syncthreads();
}

template <typename T>
__device__ T inst_block_min(T v, T* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const T w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  if (lane == 0) red[wv * 8] = v;
  __syncthreads();
  T r = red[0];
  for (int k = 1; k < kInstWaves; ++k) r = red[k * 8] < r ? red[k * 8] : r;
  __syncthreads();
  return r;
}

// K4 with the trial trajectories staged in LDS (same rules as k_linesearch).
template <typename T>
__global__ void __launch_bounds__(256) k_linesearch_inst(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws, T* o_alpha,
                                                         T* o_rho, T* o_phi0, T* o_dphi, int* o_status,
                                                         int* o_trials) {
  const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  if (skip_instance(ws, b)) return;
  const int lane = tid & 63, wv = tid >> 6;
  const int n = P.n, m = P.m, p = P.p, N = P.N, N1 = N + 1;
  const InstLds L = inst_lds(n, m, N, P.model == 1, P.Q.st == 0, P.R.st == 0, true);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  T* red = sm + L.red;
  // 1) merit and slope partials
  T v6[6] = {0, 0, 0, 0, 0, 0};
  for (int t = tid; t < N1; t += nt) {
    const T* mp = ws.mpart + ((size_t)b * N1 + t) * NMP;
    v6[0] += mp[MP_COST];
    v6[1] += mp[MP_VC];
    v6[2] += mp[MP_CC];
    const T* fp = ws.fpart + ((size_t)b * N1 + t) * NFP;
    v6[3] += fp[FP_DF];
    v6[4] += fp[FP_DC];
    v6[5] += fp[FP_LC];
  }
  inst_block_sum<T, 6>(v6, red);
  const T cost = v6[0], vc = v6[1], cc = v6[2], df = v6[3], dc = v6[4], kap = v6[5] - v6[3];
  T rho = prm.rho0;
  if (cc > T(0)) {
    const T r2 = T(2) * (df + dc + T(0.5) * kap) / cc;
    rho = r2 > rho ? r2 : rho;
  }
  rho = rho < prm.rho_max ? rho : prm.rho_max;
  const T dphi = df + dc - rho * cc;
  const T phi0 = cost + vc + T(0.5) * rho * cc;
  // 2) fraction-to-boundary for slacks and box multipliers
  T amax = 1, adual = 1;
  if (P.box) {
    T ap = 1, ad = 1;
    for (int e = tid; e < N * m; e += nt) {
      const int t = e / m, i = e - t * m;
      const T u = it.U[((size_t)b * N + t) * m + i];
      const T du = ws.du[((size_t)b * N + t) * m + i];
      const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
      const T shi = hi - u, slo = u - lo;
      const T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
      const T nhi = nu[i], nlo = nu[m + i];
      if (du > T(0)) { const T r = prm.tau * shi / du; ap = r < ap ? r : ap; }
      if (du < T(0)) { const T r = prm.tau * slo / (-du); ap = r < ap ? r : ap; }
      const T dhi = prm.mu_b / shi - nhi + nhi * du / shi;
      const T dlo = prm.mu_b / slo - nlo - nlo * du / slo;
      if (dhi < T(0)) { const T r = -prm.tau * nhi / dhi; ad = r < ad ? r : ad; }
      if (dlo < T(0)) { const T r = -prm.tau * nlo / dlo; ad = r < ad ? r : ad; }
    }
    amax = inst_block_min(ap, red);
    adual = inst_block_min(ad, red);
  }
  // 3) trials, evaluated `group` at a time (in order) on LDS copies of the trial point
  inst_stage_constants(P, b, sm, L);
  const T slack = prm.slack * (T(1) + (phi0 < 0 ? -phi0 : phi0));
  const T* Xb = it.X + (size_t)b * N1 * n;
  const T* Ub = it.U + (size_t)b * N * m;
  const T* Vb = it.V + (size_t)b * N1 * n;
  const T* dXb = ws.dx + (size_t)b * N1 * n;
  const T* dUb = ws.du + (size_t)b * N * m;
  const T* dVb = ws.dv + (size_t)b * N1 * n;
  T* scr = sm + L.scratch + wv * kInstScratch;
  T alpha = -1;
  int trials = 0, j0 = 0;
  while (j0 < prm.max_trials && alpha < T(0)) {
    int g = prm.group;
    if (g > prm.max_trials - j0) g = prm.max_trials - j0;
    for (int j = 0; j < g; ++j) {
      T a = amax;
      for (int q = 0; q < j0 + j; ++q) a *= prm.beta;
      __syncthreads();   // previous trial's reads of the trajectory are done
      for (int e = tid; e < N1 * n; e += nt) {
        sm[L.X + e] = Xb[e] + a * dXb[e];
        sm[L.V + e] = Vb[e] + a * dVb[e];
      }
      for (int e = tid; e < N * m; e += nt) sm[L.U + e] = Ub[e] + a * dUb[e];
      __syncthreads();
      T acc[3] = {0, 0, 0};
      for (int t = wv; t < N1; t += kInstWaves) {
        T o[3];
        if (t < N)
          stage_terms<T, false>(P, prm, it, ws, b, t, sm, L, sm + L.X + t * n, sm + L.U + t * m,
                                sm + L.X + (t + 1) * n, sm + L.V + (t + 1) * n, scr, o);
        else
          terminal_terms<T, false>(P, ws, b, sm + L.X + N * n, sm + L.X, sm + L.V, scr, o);
        acc[0] += o[0];
        acc[1] += o[1];
        acc[2] += o[2];
      }
      if (lane != 0) acc[0] = acc[1] = acc[2] = 0;   // wave sums sit on every lane
      inst_block_sum<T, 3>(acc, red);
      const T phi = acc[0] + acc[1] + T(0.5) * rho * acc[2];
      if (alpha < T(0) && phi <= phi0 + prm.c1 * a * dphi + slack) alpha = a;
    }
    trials += g;
    j0 += g;
  }
  int status = 0;
  if (alpha < T(0)) {
    alpha = amax;
    for (int q = 0; q < prm.max_trials - 1; ++q) alpha *= prm.beta;
    status = 2;
  }
  if (!isfinite(phi0) || !isfinite(dphi)) status = 1;
  // 4) multiplier updates (read the old iterate), then the primal-dual step
  if (p > 0) {
    for (int e = tid; e < N * p; e += nt) {
      const int t = e / p, k = e - t * p;
      const T* D = P.D.at(b, t);
      const T* E = P.E.at(b, t);
      const T* X = Xb + (size_t)t * n;
      const T* dX = dXb + (size_t)t * n;
      const T* U = Ub + (size_t)t * m;
      const T* dU = dUb + (size_t)t * m;
      T h = P.e.at(b, t)[k];
      for (int j = 0; j < n; ++j) h += D[k * n + j] * (X[j] + alpha * dX[j]);
      for (int j = 0; j < m; ++j) h += E[k * m + j] * (U[j] + alpha * dU[j]);
      it.Yeq[((size_t)b * N + t) * p + k] += prm.rho_c * h;
    }
  }
  if (P.box) {
    for (int e = tid; e < N * m; e += nt) {
      const int t = e / m, i = e - t * m;
      const T u = Ub[e];
      const T du = dUb[e];
      const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
      const T shi = hi - u, slo = u - lo;
      T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
      const T nhi = nu[i], nlo = nu[m + i];
      nu[i] = nhi + adual * (prm.mu_b / shi - nhi + nhi * du / shi);
      nu[m + i] = nlo + adual * (prm.mu_b / slo - nlo - nlo * du / slo);
    }
  }
  __syncthreads();
  for (int e = tid; e < N1 * n; e += nt) {
    const size_t o = (size_t)b * N1 * n + e;
    it.X[o] += alpha * ws.dx[o];
    it.V[o] += alpha * ws.dv[o];
  }
  for (int e = tid; e < N * m; e += nt) {
    const size_t o = (size_t)b * N * m + e;
    it.U[o] += alpha * ws.du[o];
  }
  if (tid == 0) {
    if (o_alpha) o_alpha[b] = alpha;
    if (o_rho) o_rho[b] = rho;
    if (o_phi0) o_phi0[b] = phi0;
    if (o_dphi) o_dphi[b] = dphi;
    if (o_status) o_status[b] = status;
    if (o_trials) o_trials[b] = trials;
  }
}

}  // namespace mpxk
