// K1: linearisation, augmentation and merit partials at the current iterate.
//
// One block per (stage t, instance b).  For t < N it evaluates the dynamics
// f(x_t, u_t), the Jacobian f_x (tanh model; written to Alin), the defect
// c_{t+1} = f - x_{t+1}, the gradients q_t, r_t of the tracking cost with the
// augmented-Lagrangian terms of the equality rows and the box barrier, the box
// Hessian diagonal nu/s, and the stage's merit terms (AL cost, V^T c, |c|^2).
// The block t == N handles the terminal cost and the initial-state row
// c_init = x0 - X_0.  The box projection of U and the replacement of
// non-positive box multipliers by mu_b / s happen here, in place.
// Model 2 (the caller's linearisation) takes f(x_t, u_t), the gradients and the
// stage costs from the caller's views instead of evaluating them.
#pragma once
#include "common.h"

namespace mpxk {

template <typename T>
__global__ void k_linearize(Prob<T> P, Prm<T> prm, Iter<T> it, Ws<T> ws) {
  const int t = blockIdx.x, b = blockIdx.y;
  if (skip_instance(ws, b)) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = P.n, m = P.m, N = P.N, p = P.p, N1 = N + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  T* sx = sm;           // n
  T* su = sx + n;       // m
  T* sw = su + m;       // p   mu + rho_c h
  T* sdt = sw + p;      // n   1 - tanh^2(Wx)
  T* sred = sdt + n;    // NMP * nt
  T acc[NMP] = {T(0), T(0), T(0)};
  const T* Xb = it.X + (size_t)b * N1 * n;
  const T* Vb = it.V + (size_t)b * N1 * n;
  if (t < N) {
    T* Ub = it.U + ((size_t)b * N + t) * m;
    for (int i = tid; i < n; i += nt) sx[i] = Xb[(size_t)t * n + i];
    for (int i = tid; i < m; i += nt) {
      T u = Ub[i];
      if (P.box) {
        const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
        const T w = T(1e-3) * (hi - lo);
        if (u >= hi) u = hi - w;
        if (u <= lo) u = lo + w;
        Ub[i] = u;
      }
      su[i] = u;
    }
    __syncthreads();
    // equality rows h = D x + E u + e, weights w = mu + rho_c h
    if (p > 0) {
      const T* D = P.D.at(b, t);
      const T* E = P.E.at(b, t);
      const T* e = P.e.at(b, t);
      const T* mu = it.Yeq + ((size_t)b * N + t) * p;
      for (int r = tid; r < p; r += nt) {
        T h = e[r];
        for (int j = 0; j < n; ++j) h += D[r * n + j] * sx[j];
        for (int j = 0; j < m; ++j) h += E[r * m + j] * su[j];
        sw[r] = mu[r] + prm.rho_c * h;
        acc[MP_COST] += mu[r] * h + T(0.5) * prm.rho_c * h * h;
      }
    }
    __syncthreads();
    // control gradient with AL and barrier terms; box Hessian diagonal
    {
      const T* R = P.R.at(b, t);
      const T* ur = P.uref.at(b, t);
      const T* E = p > 0 ? P.E.at(b, t) : nullptr;
      T* rh = ws.rh + ((size_t)b * N + t) * m;
      T* rd = ws.rd + ((size_t)b * N + t) * m;
      const bool caller = P.model == MODEL_CALLER;
      const T* rc = caller ? P.r.at(b, t) : nullptr;
      if (caller && tid == 0) acc[MP_COST] += P.g.at(b, t)[0];
      for (int i = tid; i < m; i += nt) {
        T r;
        if (caller) {
          r = rc[i];
        } else {
          T Ru = 0;
          for (int j = 0; j < m; ++j) Ru += R[i * m + j] * (su[j] - ur[j]);
          r = Ru;
          acc[MP_COST] += T(0.5) * (su[i] - ur[i]) * Ru;
        }
        for (int k = 0; k < p; ++k) r += E[k * m + i] * sw[k];
        T dg = 0;
        if (P.box) {
          const T lo = P.ulo.at(b, t)[i], hi = P.uhi.at(b, t)[i];
          const T shi = hi - su[i], slo = su[i] - lo;
          T* nu = it.Ybox + ((size_t)b * N + t) * 2 * m;
          T nhi = nu[i], nlo = nu[m + i];
          if (!(nhi > T(0))) { nhi = prm.mu_b / shi; nu[i] = nhi; }
          if (!(nlo > T(0))) { nlo = prm.mu_b / slo; nu[m + i] = nlo; }
          r += prm.mu_b / shi - prm.mu_b / slo;
          dg = nhi / shi + nlo / slo;
          acc[MP_COST] -= prm.mu_b * (dev_log(shi) + dev_log(slo));
        }
        rh[i] = r;
        rd[i] = dg;
      }
    }
    // state gradient, dynamics, defect
    {
      const T* Q = P.Q.at(b, t);
      const T* xr = P.xref.at(b, t);
      const T* D = p > 0 ? P.D.at(b, t) : nullptr;
      T* qh = ws.qh + ((size_t)b * N1 + t) * n;
      T* cd = ws.cdef + ((size_t)b * N1 + t + 1) * n;
      const T* xn = Xb + (size_t)(t + 1) * n;
      const T* vn = Vb + (size_t)(t + 1) * n;
      const bool caller = P.model == MODEL_CALLER;
      for (int i = tid; i < n; i += nt) {
        T q;
        if (caller) {
          q = P.q.at(b, t)[i];
        } else {
          T Qx = 0;
          for (int j = 0; j < n; ++j) Qx += Q[i * n + j] * (sx[j] - xr[j]);
          q = Qx;
          acc[MP_COST] += T(0.5) * (sx[i] - xr[i]) * Qx;
        }
        for (int k = 0; k < p; ++k) q += D[k * n + i] * sw[k];
        qh[i] = q;
        T f;
        if (caller) {
          f = P.fbar.at(b, t)[i];
        } else if (P.model == 1) {
          const T* A = P.A.at(b, 0);
          const T* Bm = P.Bm.at(b, 0);
          const T* W = P.W.at(b, 0);
          T ax = 0, bu = 0, wx = 0;
          for (int j = 0; j < n; ++j) { ax += A[i * n + j] * sx[j]; wx += W[i * n + j] * sx[j]; }
          for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * su[j];
          const T th = dev_tanh(wx);
          sdt[i] = T(1) - th * th;
          f = sx[i] + P.h * (ax + bu + P.kappa * th);
        } else {
          const T* A = P.A.at(b, t);
          const T* Bm = P.Bm.at(b, t);
          T ax = 0, bu = 0;
          for (int j = 0; j < n; ++j) ax += A[i * n + j] * sx[j];
          for (int j = 0; j < m; ++j) bu += Bm[i * m + j] * su[j];
          f = ax + bu;
        }
        const T c = f - xn[i];
        cd[i] = c;
        acc[MP_VC] += vn[i] * c;
        acc[MP_CC] += c * c;
      }
    }
    if (P.model == 1) {
      __syncthreads();
      const T* A = P.A.at(b, 0);
      const T* W = P.W.at(b, 0);
      T* Al = ws.Alin + ((size_t)b * N + t) * n * n;
      for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e - i * n;
        Al[e] = (i == j ? T(1) : T(0)) + P.h * (A[e] + P.kappa * sdt[i] * W[e]);
      }
      if (t == 0) {
        const T* Bm = P.Bm.at(b, 0);
        T* Bl = ws.Blin + (size_t)b * n * m;
        for (int e = tid; e < n * m; e += nt) Bl[e] = P.h * Bm[e];
      }
    }
  } else {
    // terminal stage and the initial-state row
    const T* QN = P.QN.at(b, 0);
    const T* xr = P.xref.at(b, N);
    const T* x0 = P.x0.at(b, 0);
    const T* xN = Xb + (size_t)N * n;
    for (int i = tid; i < n; i += nt) sx[i] = xN[i];
    __syncthreads();
    T* qh = ws.qh + ((size_t)b * N1 + N) * n;
    T* c0 = ws.cdef + (size_t)b * N1 * n;
    const bool caller = P.model == MODEL_CALLER;
    if (caller && tid == 0) acc[MP_COST] += P.g.at(b, N)[0];
    for (int i = tid; i < n; i += nt) {
      if (caller) {
        qh[i] = P.q.at(b, N)[i];
      } else {
        T Qx = 0;
        for (int j = 0; j < n; ++j) Qx += QN[i * n + j] * (sx[j] - xr[j]);
        qh[i] = Qx;
        acc[MP_COST] += T(0.5) * (sx[i] - xr[i]) * Qx;
      }
      const T c = x0[i] - Xb[i];
      c0[i] = c;
      acc[MP_VC] += Vb[i] * c;
      acc[MP_CC] += c * c;
    }
  }
  block_sum<T, NMP>(acc, NMP, sred);
  if (tid == 0) {
    T* mp = ws.mpart + ((size_t)b * N1 + t) * NMP;
    mp[MP_COST] = acc[MP_COST];
    mp[MP_VC] = acc[MP_VC];
    mp[MP_CC] = acc[MP_CC];
  }
}

}  // namespace mpxk
