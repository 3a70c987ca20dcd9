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
      for (int j = 0; j < n; ++j) g += Q[i * n + j] * (sx[j] - xr[j]);
      T f;
      if (P.model == 1) {
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

// (g, c) of the reference's model evaluator: per-stage cost and constraint rows.
template <typename T>
__global__ void k_model_eval(Prob<T> P, const T* X, const T* U, T* g, T* c) {
  const int t = blockIdx.x, b = blockIdx.y;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = P.n, m = P.m, N = P.N, N1 = N + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sred = reinterpret_cast<T*>(smem_raw);
  const T* x = X + ((size_t)b * N1 + t) * n;
  T acc[1] = {0};
  T* crow = c + ((size_t)b * N1 + t) * n;
  if (t == N) {
    const T* QN = P.QN.at(b, 0);
    const T* xr = P.xref.at(b, N);
    for (int i = tid; i < n; i += nt) {
      T qx = 0;
      for (int j = 0; j < n; ++j) qx += QN[i * n + j] * (x[j] - xr[j]);
      acc[0] += T(0.5) * (x[i] - xr[i]) * qx;
    }
  } else {
    const T* u = U + ((size_t)b * N + t) * m;
    const T* Q = P.Q.at(b, t);
    const T* R = P.R.at(b, t);
    const T* xr = P.xref.at(b, t);
    const T* ur = P.uref.at(b, t);
    for (int i = tid; i < n; i += nt) {
      T qx = 0;
      for (int j = 0; j < n; ++j) qx += Q[i * n + j] * (x[j] - xr[j]);
      acc[0] += T(0.5) * (x[i] - xr[i]) * qx;
    }
    for (int i = tid; i < m; i += nt) {
      T ru = 0;
      for (int j = 0; j < m; ++j) ru += R[i * m + j] * (u[j] - ur[j]);
      acc[0] += T(0.5) * (u[i] - ur[i]) * ru;
    }
    // defect row t+1
    T* cn = c + ((size_t)b * N1 + t + 1) * n;
    const T* xn = x + n;
    for (int i = tid; i < n; i += nt) {
      T f;
      if (P.model == 1) {
        const T* A = P.A.at(b, 0);
        const T* Bm = P.Bm.at(b, 0);
        const T* W = P.W.at(b, 0);
        T ax = 0, wx = 0, bu = 0;
        for (int j = 0; j < n; ++j) { ax += A[i * n + j] * x[j]; wx += W[i * n + j] * x[j]; }
        for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * u[j];
        f = x[i] + P.h * (ax + bu + P.kappa * dev_tanh(wx));
      } else {
        const T* A = P.A.at(b, t);
        const T* Bm = P.Bm.at(b, t);
        T ax = 0, bu = 0;
        for (int j = 0; j < n; ++j) ax += A[i * n + j] * x[j];
        for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * u[j];
        f = ax + bu;
      }
      cn[i] = f - xn[i];
    }
  }
  if (t == 0) {
    const T* x0 = P.x0.at(b, 0);
    for (int i = tid; i < n; i += nt) crow[i] = x0[i] - x[i];
  }
  block_sum<T, 1>(acc, 1, sred);
  if (tid == 0) g[(size_t)b * N1 + t] = acc[0];
}

// Warm-start shift (one block per instance; each thread owns columns, so the
// in-place row shift is race free) with the NaN fallback.
template <typename T>
__device__ void shift_rows(T* Z, int rows, int cols, int shift) {
  for (int j = threadIdx.x; j < cols; j += blockDim.x) {
    for (int r = 0; r < rows; ++r) {
      const int src = r + shift < rows ? r + shift : rows - 1;
      Z[(size_t)r * cols + j] = Z[(size_t)src * cols + j];
    }
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
