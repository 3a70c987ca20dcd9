// KKT residual, model evaluation (g, c) and the warm-start shift.
#pragma once
#include "common.h"

namespace mpxk {

// Squared KKT residual of stage t (block (t, b)):
//   t < N: grad_x L = Q(x - xr) + f_x' V_{t+1} - V_t + D' mu, grad_u L = R(u - ur)
//          + f_u' V_{t+1} + E' mu + nu_hi - nu_lo, h, nu*s - mu_b, c_{t+1}
//   t = N: Q_N (x - xr) - V_N and c_init = x0 - X_0.
template <typename T>
__global__ void k_kkt_stage(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws) {
  const int t = blockIdx.x, b = blockIdx.y;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = P.n, m = P.m, N = P.N, p = P.p, N1 = N + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  T* sx = sm;          // n
  T* su = sx + n;      // m
  T* sv = su + m;      // n   V_{t+1} (or weighted)
  T* sdt = sv + n;     // n   (1 - tanh^2) * V_{t+1}
  T* sred = sdt + n;   // nt
  T acc[1] = {0};
  const T* Xb = it.X + (size_t)b * N1 * n;
  const T* Vb = it.V + (size_t)b * N1 * n;
  if (t < N) {
    const T* Ub = it.U + ((size_t)b * N + t) * m;
    for (int i = tid; i < n; i += nt) { sx[i] = Xb[(size_t)t * n + i]; sv[i] = Vb[(size_t)(t + 1) * n + i]; }
    for (int i = tid; i < m; i += nt) su[i] = Ub[i];
    __syncthreads();
    if (P.model == 1) {
      const T* W = P.W.at(b, 0);
      for (int i = tid; i < n; i += nt) {
        T wx = 0;
        for (int j = 0; j < n; ++j) wx += W[i * n + j] * sx[j];
        const T th = dev_tanh(wx);
        sdt[i] = (T(1) - th * th) * sv[i];
      }
      __syncthreads();
    }
    const T* mu = p > 0 ? it.Yeq + ((size_t)b * N + t) * p : nullptr;
    // state rows: gradient and defect
    for (int i = tid; i < n; i += nt) {
      const T* Q = P.Q.at(b, t);
      const T* xr = P.xref.at(b, t);
      T g = -Vb[(size_t)t * n + i];
      if (P.model == MODEL_CALLER) g += P.q.at(b, t)[i];
      else
        for (int j = 0; j < n; ++j) g += Q[i * n + j] * (sx[j] - xr[j]);
      T f;
      if (P.model == MODEL_CALLER) {
        const T* A = P.A.at(b, t);
        T atv = 0;
        for (int k = 0; k < n; ++k) atv += A[k * n + i] * sv[k];
        g += atv;
        f = P.fbar.at(b, t)[i];
      } else if (P.model == 1) {
        const T* A = P.A.at(b, 0);
        const T* Bm = P.Bm.at(b, 0);
        const T* W = P.W.at(b, 0);
        T atv = 0, wtv = 0, ax = 0, wx = 0, bu = 0;
        for (int k = 0; k < n; ++k) {
          atv += A[k * n + i] * sv[k];
          wtv += W[k * n + i] * sdt[k];
          ax += A[i * n + k] * sx[k];
          wx += W[i * n + k] * sx[k];
        }
        for (int k = 0; k < m; ++k) bu += Bm[i * m + k] * su[k];
        g += sv[i] + P.h * (atv + P.kappa * wtv);
        f = sx[i] + P.h * (ax + bu + P.kappa * dev_tanh(wx));
      } else {
        const T* A = P.A.at(b, t);
        const T* Bm = P.Bm.at(b, t);
        T atv = 0, ax = 0, bu = 0;
        for (int k = 0; k < n; ++k) { atv += A[k * n + i] * sv[k]; ax += A[i * n + k] * sx[k]; }
        for (int k = 0; k < m; ++k) bu += Bm[i * m + k] * su[k];
        g += atv;
        f = ax + bu;
      }
      if (p > 0) {
        const T* D = P.D.at(b, t);
        for (int k = 0; k < p; ++k) g += D[k * n + i] * mu[k];
      }
      const T c = f - Xb[(size_t)(t + 1) * n + i];
      acc[0] += g * g + c * c;
    }
    for (int i = tid; i < m; i += nt) {
      const T* R = P.R.at(b, t);
      const T* ur = P.uref.at(b, t);
      T g = 0;
      if (P.model == MODEL_CALLER) g = P.r.at(b, t)[i];
      else
        for (int j = 0; j < m; ++j) g += R[i * m + j] * (su[j] - ur[j]);
      const T* Bm = P.model == 1 ? P.Bm.at(b, 0) : P.Bm.at(b, t);
      const T hs = P.model == 1 ? P.h : T(1);
      T btv = 0;
      for (int k = 0; k < n; ++k) btv += Bm[k * m + i] * sv[k];
      g += hs * btv;
      if (p > 0) {
        const T* E = P.E.at(b, t);
        for (int k = 0; k < p; ++k) g += E[k * m + i] * mu[k];
      }
      if (P.box) {
        const T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
        const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
        g += nu[i] - nu[m + i];
        const T c1 = nu[i] * (hi - su[i]) - prm.mu_b;
        const T c2 = nu[m + i] * (su[i] - lo) - prm.mu_b;
        acc[0] += c1 * c1 + c2 * c2;
      }
      acc[0] += g * g;
    }
    for (int k = tid; k < p; k += nt) {
      const T* D = P.D.at(b, t);
      const T* E = P.E.at(b, t);
      T h = P.e.at(b, t)[k];
      for (int j = 0; j < n; ++j) h += D[k * n + j] * sx[j];
      for (int j = 0; j < m; ++j) h += E[k * m + j] * su[j];
      acc[0] += h * h;
    }
  } else {
    const T* QN = P.QN.at(b, 0);
    const T* xr = P.xref.at(b, N);
    const T* x0 = P.x0.at(b, 0);
    for (int i = tid; i < n; i += nt) {
      T g = -Vb[(size_t)N * n + i];
      if (P.model == MODEL_CALLER) g += P.q.at(b, N)[i];
      else
        for (int j = 0; j < n; ++j) g += QN[i * n + j] * (Xb[(size_t)N * n + j] - xr[j]);
      const T c = x0[i] - Xb[i];
      acc[0] += g * g + c * c;
    }
  }
  block_sum<T, 1>(acc, 1, sred);
  if (tid == 0) ws.kpart[(size_t)b * N1 + t] = acc[0];
}

template <typename T>
__global__ void k_kkt_reduce(int B, int N1, const T* kpart, T* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  T s = 0;
  for (int t = 0; t < N1; ++t) s += kpart[(size_t)b * N1 + t];
  out[b] = dev_sqrt(s);
}

// Row i of f(x_t, u_t) for the device model families (x, u of instance b).
template <typename T>
__device__ __forceinline__ T model_f_row(const Prob<T>& P, int b, int t, int i, const T* x, const T* u) {
  const int n = P.n, m = P.m;
  if (P.model == MODEL_CALLER) return P.fbar.at(b, t)[i];
  if (P.model == 1) {
    const T* A = P.A.at(b, 0);
    const T* Bm = P.Bm.at(b, 0);
    const T* W = P.W.at(b, 0);
    T ax = 0, wx = 0, bu = 0;
    for (int j = 0; j < n; ++j) { ax += A[i * n + j] * x[j]; wx += W[i * n + j] * x[j]; }
    for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * u[j];
    return x[i] + P.h * (ax + bu + P.kappa * dev_tanh(wx));
  }
  const T* A = P.A.at(b, t);
  const T* Bm = P.Bm.at(b, t);
  T ax = 0, bu = 0;
  for (int j = 0; j < n; ++j) ax += A[i * n + j] * x[j];
  for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * u[j];
  return ax + bu;
}

// This thread's part of the tracking cost g_t at (x, u) (t == N: terminal form).
template <typename T>
__device__ __forceinline__ T stage_cost_part(const Prob<T>& P, int b, int t, const T* x, const T* u) {
  const int n = P.n, m = P.m, N = P.N, tid = threadIdx.x, nt = blockDim.x;
  T acc = 0;
  if (P.model == MODEL_CALLER) return tid == 0 ? P.g.at(b, t)[0] : T(0);
  const T* Q = t == N ? P.QN.at(b, 0) : P.Q.at(b, t);
  const T* xr = P.xref.at(b, t);
  for (int i = tid; i < n; i += nt) {
    T qx = 0;
    for (int j = 0; j < n; ++j) qx += Q[i * n + j] * (x[j] - xr[j]);
    acc += T(0.5) * (x[i] - xr[i]) * qx;
  }
  if (t < N) {
    const T* R = P.R.at(b, t);
    const T* ur = P.uref.at(b, t);
    for (int i = tid; i < m; i += nt) {
      T ru = 0;
      for (int j = 0; j < m; ++j) ru += R[i * m + j] * (u[j] - ur[j]);
      acc += T(0.5) * (u[i] - ur[i]) * ru;
    }
  }
  return acc;
}

// (g, c) of the reference's model evaluator: per-stage cost and constraint rows.
template <typename T>
__global__ void k_model_eval(Prob<T> P, const T* X, const T* U, T* g, T* c) {
  const int t = blockIdx.x, b = blockIdx.y;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sred = reinterpret_cast<T*>(smem_raw);
  const T* x = X + ((size_t)b * N1 + t) * n;
  const T* u = t < N ? U + ((size_t)b * N + t) * m : nullptr;
  T acc[1] = {stage_cost_part(P, b, t, x, u)};
  if (t < N) {
    // defect row t+1
    T* cn = c + ((size_t)b * N1 + t + 1) * n;
    for (int i = tid; i < n; i += nt) cn[i] = model_f_row(P, b, t, i, x, u) - x[n + i];
  } else {
    const T* x0 = P.x0.at(b, 0);
    T* c0 = c + (size_t)b * N1 * n;
    const T* xs = X + (size_t)b * N1 * n;
    for (int i = tid; i < n; i += nt) c0[i] = x0[i] - xs[i];
  }
  block_sum<T, 1>(acc, 1, sred);
  if (tid == 0) g[(size_t)b * N1 + t] = acc[0];
}

// Stop test of the multi-iteration solve, per stage: g_t of the model
// evaluator and the squared norm of one constraint row (block t < N: c_{t+1} =
// f(x_t, u_t) - x_{t+1}; block N: c_init = x0 - X_0) into gpart[b][t].
template <typename T>
__global__ void k_eval_partials(Prob<T> P, Iter<T> it, Ws<T> ws) {
  const int t = blockIdx.x, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sred = reinterpret_cast<T*>(smem_raw);
  const T* x = it.X + ((size_t)b * N1 + t) * n;
  const T* u = t < N ? it.U + ((size_t)b * N + t) * m : nullptr;
  T acc[2] = {stage_cost_part(P, b, t, x, u), T(0)};
  if (t < N) {
    for (int i = tid; i < n; i += nt) {
      const T ci = model_f_row(P, b, t, i, x, u) - x[n + i];
      acc[1] += ci * ci;
    }
  } else {
    const T* x0 = P.x0.at(b, 0);
    const T* xs = it.X + (size_t)b * N1 * n;
    for (int i = tid; i < n; i += nt) {
      const T ci = x0[i] - xs[i];
      acc[1] += ci * ci;
    }
  }
  block_sum<T, 2>(acc, 2, sred);
  if (tid == 0) {
    ws.gpart[((size_t)b * N1 + t) * 2] = acc[0];
    ws.gpart[((size_t)b * N1 + t) * 2 + 1] = acc[1];
  }
}

// Start of a solve: every instance runs, last cost 1e10 (the reference's
// initial last_cost), no iteration counted.
template <typename T>
__global__ void k_solve_init(int B, Ws<T> ws, int* iters) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  ws.run[b] = 1;
  ws.last[b] = T(1e10);
  iters[b] = 0;
}

// After each iteration, one thread per running instance: l2 = sum g_t^2,
// cc = sum |c|^2; i = iterations so far; stop when i > max_iter or when
// last - l2 < dcost_tol and cc < cc_tol, else last = l2.
template <typename T>
__global__ void k_stop(int B, int N1, Ws<T> ws, int max_iter, T dcost_tol, T cc_tol, int* iters, T* o_cost,
                       T* o_cc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || ws.run[b] == 0) return;
  const T* gp = ws.gpart + (size_t)b * N1 * 2;
  T l2 = 0, cc = 0;
  for (int t = 0; t < N1; ++t) {
    const T g = gp[2 * t];
    l2 += g * g;
    cc += gp[2 * t + 1];
  }
  const int i = iters[b] + 1;
  iters[b] = i;
  if (i > max_iter || (ws.last[b] - l2 < dcost_tol && cc < cc_tol)) ws.run[b] = 0;
  else ws.last[b] = l2;
  if (o_cost) o_cost[b] = l2;
  if (o_cc) o_cc[b] = cc;
}

// Number of running instances into ws.nrun[0] (one block).
template <typename T>
__global__ void k_count_running(int B, Ws<T> ws) {
  __shared__ int red[256];
  int c = 0;
  for (int b = threadIdx.x; b < B; b += blockDim.x) c += ws.run[b] != 0 ? 1 : 0;
  red[threadIdx.x] = c;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) ws.nrun[0] = red[0];
}

// Warm-start shift (one block per instance; each thread owns columns, so the
// in-place row shift is race free) with the NaN fallback.
template <typename T>
__device__ void shift_rows(T* Z, int rows, int cols, int shift) {
  // in place, in chunks of 8 entries per thread taken in increasing order:
  // every thread reads its chunk entries' sources (rows r + shift >= r, none
  // written yet) before the barrier and stores them after it
  const int tot = rows * cols, tid = threadIdx.x, nt = blockDim.x;
  const float inv = 1.f / (float)cols;
  for (int base = 0; base < tot; base += 8 * nt) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = base + u * nt + tid;
      v[u] = T(0);
      if (e < tot) {
        int r = (int)((float)e * inv);   // e / cols (e < 2^24), corrected by one
        if (r * cols > e) --r;
        else if ((r + 1) * cols <= e) ++r;
        const int src = r + shift < rows ? r + shift : rows - 1;
        v[u] = Z[(size_t)src * cols + (e - r * cols)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = base + u * nt + tid;
      if (e < tot) Z[e] = v[u];
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void k_shift(Prob<T> P, int shift, Iter<T> it, const T* u_reset, int* reset) {
  const int b = blockIdx.x;
  const int n = P.n, m = P.m, N = P.N, p = P.p, N1 = N + 1;
  T* X = it.X + (size_t)b * N1 * n;
  T* U = it.U + (size_t)b * N * m;
  T* V = it.V + (size_t)b * N1 * n;
  T* Ye = p > 0 && it.Yeq ? it.Yeq + (size_t)b * N * p : nullptr;
  T* Yb = P.box && it.Ybox ? it.Ybox + (size_t)b * N * 2 * m : nullptr;
  const bool bad = isnan(U[0]);
  __syncthreads();
  if (bad) {
    const T* x0 = P.x0.at(b, 0);
    for (int e = threadIdx.x; e < N1 * n; e += blockDim.x) { X[e] = x0[e % n]; V[e] = 0; }
    for (int e = threadIdx.x; e < N * m; e += blockDim.x) U[e] = u_reset ? u_reset[(size_t)b * m + e % m] : T(0);
    if (Ye) for (int e = threadIdx.x; e < N * p; e += blockDim.x) Ye[e] = 0;
    if (Yb) for (int e = threadIdx.x; e < N * 2 * m; e += blockDim.x) Yb[e] = 0;
  } else {
    shift_rows(X, N1, n, shift);
    shift_rows(U, N, m, shift);
    shift_rows(V, N1, n, shift);
    if (Ye) shift_rows(Ye, N, p, shift);
    if (Yb) shift_rows(Yb, N, 2 * m, shift);
  }
  if (threadIdx.x == 0 && reset) reset[b] = bad ? 1 : 0;
}

}  // namespace mpxk
