// Single-rigid-body quadruped on the device: the dynamics of
// quadruped_srbd_dynamics (mpx/utils/models.py:6-37) with analytic Jacobians,
// and the objective quadruped_srbd_obj with its Gauss-Newton Hessian
// quadruped_srbd_hessian_gn (mpx/utils/objectives.py:7-75), evaluated per
// (instance, stage) into the caller-linearised problem form, and at the line
// search's trial points.  With these kernels the whole SQP iteration of the
// SRBD controller (mpx/utils/mpc_wrapper_srbd.py:58-68) runs on the device.
//
// State x = [p (3), quat (4, w first), dp (3), omega (3)], control u = ground
// reaction forces (3 per foot, 4 feet).  Per stage t: reference row
// [p_ref, quat_ref, dp_ref, omega_ref, contact (4)] (17), parameter row
// [contact (4), p_legs (12)] (16).  W: the 24 x 24 weight of the residual
// [p - p_ref, quat_sub(quat, quat_ref), dp - dp_ref, omega - omega_ref, grf].
//
// The quaternion helpers follow mujoco-mjx's math (quat_to_mat, quat_mul,
// quat_integrate as the axis-angle exponential map, quat_sub as the
// axis-angle logarithm with the angle wrapped into (-pi, pi]).
#pragma once
#include "common.h"
#include "k_trials.h"

namespace mpxk {

constexpr int SRBD_N = 13, SRBD_M = 12, SRBD_REF = 17, SRBD_PAR = 16, SRBD_RES = 24;

template <typename T>
struct Srbd {
  T mass, dt;
  T I[9], Iinv[9];
  View<T> W, ref, par;   // W: 24 x 24 per instance (stage stride 0)
};

template <typename T>
__device__ __forceinline__ T dev_atan2(T y, T x) { return atan2(y, x); }
template <>
__device__ __forceinline__ float dev_atan2<float>(float y, float x) { return atan2f(y, x); }
template <typename T>
__device__ __forceinline__ T dev_sin(T x) { return sin(x); }
template <>
__device__ __forceinline__ float dev_sin<float>(float x) { return sinf(x); }
template <typename T>
__device__ __forceinline__ T dev_cos(T x) { return cos(x); }
template <>
__device__ __forceinline__ float dev_cos<float>(float x) { return cosf(x); }

// rotation matrix of a (not necessarily unit) quaternion, quadratic in q
template <typename T>
__device__ __forceinline__ void quat_mat(const T* q, T* R) {
  const T w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = w * w + x * x - y * y - z * z; R[1] = 2 * (x * y - w * z); R[2] = 2 * (x * z + w * y);
  R[3] = 2 * (x * y + w * z); R[4] = w * w - x * x + y * y - z * z; R[5] = 2 * (y * z - w * x);
  R[6] = 2 * (x * z - w * y); R[7] = 2 * (y * z + w * x); R[8] = w * w - x * x - y * y + z * z;
}

// R' a and (dR/dq_k)' a, k = 0..3 (4 x 3, row k = (dR/dq_k)' a)
template <typename T>
__device__ __forceinline__ void quat_mat_t_apply(const T* q, const T* a, T* Rta, T* dRta) {
  const T w = q[0], x = q[1], y = q[2], z = q[3];
  T R[9];
  quat_mat(q, R);
  for (int i = 0; i < 3; ++i) Rta[i] = R[i] * a[0] + R[3 + i] * a[1] + R[6 + i] * a[2];
  // dR/dw = 2 [[w,-z,y],[z,w,-x],[-y,x,w]], dR/dx = 2 [[x,y,z],[y,-x,-w],[z,w,-x]],
  // dR/dy = 2 [[-y,x,w],[x,y,z],[-w,z,-y]], dR/dz = 2 [[-z,-w,x],[w,-z,y],[x,y,z]]
  const T D[4][9] = {{w, -z, y, z, w, -x, -y, x, w},
                     {x, y, z, y, -x, -w, z, w, -x},
                     {-y, x, w, x, y, z, -w, z, -y},
                     {-z, -w, x, w, -z, y, x, y, z}};
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 3; ++i)
      dRta[k * 3 + i] = T(2) * (D[k][i] * a[0] + D[k][3 + i] * a[1] + D[k][6 + i] * a[2]);
}

template <typename T>
__device__ __forceinline__ void qmul(const T* u, const T* v, T* o) {
  o[0] = u[0] * v[0] - u[1] * v[1] - u[2] * v[2] - u[3] * v[3];
  o[1] = u[0] * v[1] + u[1] * v[0] + u[2] * v[3] - u[3] * v[2];
  o[2] = u[0] * v[2] - u[1] * v[3] + u[2] * v[0] + u[3] * v[1];
  o[3] = u[0] * v[3] + u[1] * v[2] - u[2] * v[1] + u[3] * v[0];
}

// sin(n dt / 2) / n and its derivative in n (series near n = 0)
template <typename T>
__device__ __forceinline__ void sinc_half(T n, T dt, T& s, T& ds) {
  const T a = T(0.5) * dt;
  if (n < T(1e-4)) {
    s = a - a * a * a * n * n / T(6);
    ds = -a * a * a * n / T(3);
  } else {
    const T sn = dev_sin(a * n), cn = dev_cos(a * n);
    s = sn / n;
    ds = (a * cn * n - sn) / (n * n);
  }
}

// mjx's norm treats a vector as zero when every |v_i| <= 1e-8 (jnp.allclose(v,
// 0) with its default atol): value 0 and, under JAX's autodiff, gradient 0.
template <typename T>
__device__ __forceinline__ bool mjx_zero3(const T* v) {
  return fabs(v[0]) <= T(1e-8) && fabs(v[1]) <= T(1e-8) && fabs(v[2]) <= T(1e-8);
}

// qn = quat_integrate(q, v, dt); with J: Jq (4 x 4) = dqn/dq, Jv (4 x 3) = dqn/dv
// (Jv = 0 for a zero v in mjx's sense, as JAX differentiates its safe norm)
template <typename T, bool J>
__device__ void quat_integrate(const T* q, const T* v, T dt, T* qn, T* Jq, T* Jv) {
  const bool vz = mjx_zero3(v);
  const T n = vz ? T(0) : dev_sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const T a = T(0.5) * dt;
  T s, ds;
  sinc_half(n, dt, s, ds);
  T qr[4] = {dev_cos(a * n), vz ? T(0) : v[0] * s, vz ? T(0) : v[1] * s, vz ? T(0) : v[2] * s};
  T raw[4];
  qmul(q, qr, raw);
  const T nn = dev_sqrt(raw[0] * raw[0] + raw[1] * raw[1] + raw[2] * raw[2] + raw[3] * raw[3]);
  const T inv = T(1) / (nn + (nn == T(0) ? T(1e-6) : T(0)));
  for (int k = 0; k < 4; ++k) qn[k] = raw[k] * inv;
  if constexpr (J) {
    // dnorm = (I - qn qn') / nn
    T dn[16];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) dn[i * 4 + j] = ((i == j ? T(1) : T(0)) - qn[i] * qn[j]) * inv;
    // d raw / dq = Rmul(qr): raw = R(qr) q
    const T Rr[16] = {qr[0], -qr[1], -qr[2], -qr[3], qr[1], qr[0], qr[3], -qr[2],
                      qr[2], -qr[3], qr[0], qr[1], qr[3], qr[2], -qr[1], qr[0]};
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        T acc = 0;
        for (int k = 0; k < 4; ++k) acc += dn[i * 4 + k] * Rr[k * 4 + j];
        Jq[i * 4 + j] = acc;
      }
    // d raw / dv = L(q) dqr/dv; dqr/dv: row 0 = -sin(a n) a vhat', rows 1..3 = s I + v (ds vhat)'
    const T vh0 = n > T(0) ? v[0] / n : T(0), vh1 = n > T(0) ? v[1] / n : T(0), vh2 = n > T(0) ? v[2] / n : T(0);
    const T vh[3] = {vh0, vh1, vh2};
    T dqr[12];
    const T sa = dev_sin(a * n) * a;
    for (int j = 0; j < 3; ++j) dqr[j] = -sa * vh[j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dqr[(1 + i) * 3 + j] = (i == j ? s : T(0)) + v[i] * ds * vh[j];
    const T Lq[16] = {q[0], -q[1], -q[2], -q[3], q[1], q[0], -q[3], q[2],
                      q[2], q[3], q[0], -q[1], q[3], -q[2], q[1], q[0]};
    T M[12];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j) {
        T acc = 0;
        for (int k = 0; k < 4; ++k) acc += Lq[i * 4 + k] * dqr[k * 3 + j];
        M[i * 3 + j] = acc;
      }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 3; ++j) {
        T acc = 0;
        for (int k = 0; k < 4; ++k) acc += dn[i * 4 + k] * M[k * 3 + j];
        Jv[i * 3 + j] = vz ? T(0) : acc;
      }
  }
}

// r = quat_sub(u, v) (3); with J: Ju (3 x 4) = dr/du (0 when the vector part
// of inv(v) u is zero in mjx's sense)
template <typename T, bool J>
__device__ void quat_sub(const T* u, const T* v, T* r, T* Ju) {
  const T vi[4] = {v[0], -v[1], -v[2], -v[3]};
  T q[4];
  qmul(vi, u, q);
  const bool qz = mjx_zero3(q + 1);
  const T s = qz ? T(0) : dev_sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  T ang = T(2) * dev_atan2(s, q[0]);
  if (ang > T(3.14159265358979323846)) ang -= T(6.28318530717958647692);
  const T inv = T(1) / (s + (s == T(0) ? T(1e-6) : T(0)));
  for (int i = 0; i < 3; ++i) r[i] = q[1 + i] * inv * ang;
  if constexpr (J) {
    if (qz) {
      for (int k = 0; k < 12; ++k) Ju[k] = T(0);
      return;
    }
    const T w = q[0], d2 = s * s + w * w;
    T Jq[12];   // d r / d q (3 x 4)
    const T c1 = ang / s, c2 = (T(2) * w / d2) / (s * s) - ang / (s * s * s);
    for (int i = 0; i < 3; ++i) {
      Jq[i * 4] = -T(2) * q[1 + i] / d2;
      for (int j = 0; j < 3; ++j) Jq[i * 4 + 1 + j] = (i == j ? c1 : T(0)) + q[1 + i] * q[1 + j] * c2;
    }
    // dq/du = L(inv(v))
    const T Lv[16] = {vi[0], -vi[1], -vi[2], -vi[3], vi[1], vi[0], -vi[3], vi[2],
                      vi[2], vi[3], vi[0], -vi[1], vi[3], -vi[2], vi[1], vi[0]};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 4; ++j) {
        T acc = 0;
        for (int k = 0; k < 4; ++k) acc += Jq[i * 4 + k] * Lv[k * 4 + j];
        Ju[i * 4 + j] = acc;
      }
  }
}

template <typename T>
__device__ __forceinline__ void cross3(const T* a, const T* b, T* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// x' = f(x, u) of models.py:6-37 into xn (13); om_next also returned (for the Jacobian)
template <typename T>
__device__ void srbd_step(const Srbd<T>& S, const T* x, const T* u, const T* par, T* xn) {
  const T dt = S.dt;
  const T* p = x;
  const T* quat = x + 3;
  T force[3] = {0, 0, 0}, wr[3] = {0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    const T c = par[i];
    const T* f = u + 3 * i;
    const T l[3] = {par[4 + 3 * i] - p[0], par[5 + 3 * i] - p[1], par[6 + 3 * i] - p[2]};
    T cr[3];
    cross3(l, f, cr);
    for (int k = 0; k < 3; ++k) { force[k] += f[k] * c; wr[k] += cr[k] * c; }
  }
  const T g[3] = {T(0), T(0), T(-9.81)};
  T dpn[3];
  for (int k = 0; k < 3; ++k) dpn[k] = x[7 + k] + (g[k] + force[k] / S.mass) * dt;
  T Rtw[3], dRtw[12];
  quat_mat_t_apply(quat, wr, Rtw, dRtw);
  const T* om = x + 10;
  T Iom[3], gyro[3];
  for (int i = 0; i < 3; ++i) Iom[i] = S.I[i * 3] * om[0] + S.I[i * 3 + 1] * om[1] + S.I[i * 3 + 2] * om[2];
  cross3(om, Iom, gyro);
  T omn[3];
  for (int i = 0; i < 3; ++i) {
    T acc = 0;
    for (int k = 0; k < 3; ++k) acc += S.Iinv[i * 3 + k] * (Rtw[k] - gyro[k]);
    omn[i] = om[i] + acc * dt;
  }
  for (int k = 0; k < 3; ++k) xn[k] = p[k] + dpn[k] * dt;
  quat_integrate<T, false>(quat, omn, dt, xn + 3, nullptr, nullptr);
  for (int k = 0; k < 3; ++k) { xn[7 + k] = dpn[k]; xn[10 + k] = omn[k]; }
}

// relaxed log barrier of objectives.py:7-13 (alpha 0.1, sigma 5) and its first
// two derivatives, zero where one of its clips is active
template <typename T>
__device__ __forceinline__ void penalty(T cv, T& val, T& d1, T& d2) {
  const T alpha = T(0.1), sigma = T(5);
  const T cl = cv < T(1e-10) ? T(1e-10) : (cv > T(1e6) ? T(1e6) : cv);
  const bool inlog = cv > T(1e-10) && cv < T(1e6);
  const T lg = -alpha * dev_log(cl);
  const T dl = inlog ? T(1) / cv : T(0), d2l = inlog ? -T(1) / (cv * cv) : T(0);
  T v, a1, a2;
  if (cv > sigma) {
    v = lg; a1 = -alpha * dl; a2 = -alpha * d2l;
  } else {
    const T z = (cv - T(2) * sigma) / sigma;
    v = alpha / T(2) * (z * z - T(1)) + lg;
    a1 = alpha * (cv - T(2) * sigma) / (sigma * sigma) - alpha * dl;
    a2 = alpha / (sigma * sigma) - alpha * d2l;
  }
  const bool live = v > T(0) && v < T(1e8);
  val = v < T(0) ? T(0) : (v > T(1e8) ? T(1e8) : v);
  d1 = live ? a1 : T(0);
  d2 = live ? a2 : T(0);
}

// stage cost l_t of quadruped_srbd_obj (t == N: terminal form)
template <typename T>
__device__ T srbd_cost(const T* W, const T* ref, const T* x, const T* u, bool term) {
  T r[SRBD_RES];
  for (int k = 0; k < 3; ++k) { r[k] = x[k] - ref[k]; r[6 + k] = x[7 + k] - ref[7 + k]; r[9 + k] = x[10 + k] - ref[10 + k]; }
  quat_sub<T, false>(x + 3, ref + 3, r + 3, nullptr);
  const int nr = term ? 12 : SRBD_RES;
  if (!term)
    for (int k = 0; k < 12; ++k) r[12 + k] = u[k];
  T acc = 0;
  for (int i = 0; i < nr; ++i) {
    T wr = 0;
    for (int j = 0; j < nr; ++j) wr += W[i * SRBD_RES + j] * r[j];
    acc += r[i] * wr;
  }
  if (!term) {
    for (int i = 0; i < 4; ++i) {
      const T fx = u[3 * i], fy = u[3 * i + 1], fz = u[3 * i + 2];
      const T cv = T(0.7) * fz - dev_sqrt(fx * fx + fy * fy + T(1e-2));
      T v, d1, d2;
      penalty(cv, v, d1, d2);
      acc += v < T(1e-6) ? T(1e-6) : (v > T(1e6) ? T(1e6) : v);
    }
  }
  return T(0.5) * acc;
}

// One thread per (stage t, instance b): the caller-form linearisation of the
// SRBD problem at (X_t, U_t): A = f_x, B = f_u, fbar, Q, R, M (if given), q, r,
// g (t < N); q_N, g_N, QN (t = N).
template <typename T>
__global__ void __launch_bounds__(64) k_srbd_linearize(Prob<T> P, Srbd<T> S, Iter<T> it) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = P.N, N1 = N + 1;
  if (gid >= P.B * N1) return;
  const int b = gid / N1, t = gid - b * N1;
  const bool term = t == N;
  const T* x = it.X + ((size_t)b * N1 + t) * SRBD_N;
  T u[SRBD_M];
  for (int k = 0; k < SRBD_M; ++k) u[k] = term ? T(0) : it.U[((size_t)b * N + t) * SRBD_M + k];
  const T* W = S.W.at(b, 0);
  const T* ref = S.ref.at(b, t);
  // residual and its state Jacobian (the quaternion block Jr, 3 x 4)
  T r[SRBD_RES], Jr[12];
  for (int k = 0; k < 3; ++k) { r[k] = x[k] - ref[k]; r[6 + k] = x[7 + k] - ref[7 + k]; r[9 + k] = x[10 + k] - ref[10 + k]; }
  quat_sub<T, true>(x + 3, ref + 3, r + 3, Jr);
  for (int k = 0; k < 12; ++k) r[12 + k] = u[k];
  const int nr = term ? 12 : SRBD_RES;
  // Wr = W r (rows 0..nr)
  T Wr[SRBD_RES];
  for (int i = 0; i < nr; ++i) {
    T acc = 0;
    for (int j = 0; j < nr; ++j) acc += W[i * SRBD_RES + j] * r[j];
    Wr[i] = acc;
  }
  // J_x columns: residual row of state column c: p,dp,om identity; quat via Jr
  // grad_x = J_x' W r
  T gx[SRBD_N];
  for (int c = 0; c < 3; ++c) { gx[c] = Wr[c]; gx[7 + c] = Wr[6 + c]; gx[10 + c] = Wr[9 + c]; }
  for (int c = 0; c < 4; ++c) gx[3 + c] = Jr[c] * Wr[3] + Jr[4 + c] * Wr[4] + Jr[8 + c] * Wr[5];
  T* qo = const_cast<T*>(P.q.at(b, t));
  for (int c = 0; c < SRBD_N; ++c) qo[c] = gx[c];
  // Q = J_x' W J_x: state column c maps to residual rows rowc (with weights)
  // residual row index of state column c for the identity blocks
  auto jx = [&](int rrow, int c) -> T {     // J_x[rrow][c]
    if (c < 3) return rrow == c ? T(1) : T(0);
    if (c < 7) return (rrow >= 3 && rrow < 6) ? Jr[(rrow - 3) * 4 + (c - 3)] : T(0);
    if (c < 10) return rrow == c - 1 ? T(1) : T(0);
    return rrow == c - 1 ? T(1) : T(0);
  };
  // WJ[i][c] = sum_r W[i][r] J_x[r][c] over the 12 state residual rows
  T* Qo = const_cast<T*>(term ? P.QN.at(b, 0) : P.Q.at(b, t));
  for (int c = 0; c < SRBD_N; ++c) {
    T wj[12];
    for (int i = 0; i < 12; ++i) {
      T acc = 0;
      for (int rr = 0; rr < 12; ++rr) acc += W[i * SRBD_RES + rr] * jx(rr, c);
      wj[i] = acc;
    }
    for (int c2 = 0; c2 < SRBD_N; ++c2) {
      T acc = 0;
      for (int i = 0; i < 12; ++i) acc += jx(i, c2) * wj[i];
      Qo[c2 * SRBD_N + c] = acc;
    }
    if (!term && P.M.p) {
      // M[c2][j] = (J_x' W J_u)[c2][j] = sum_i J_x[i][c2] W[i][12 + j]
      T* Mo = const_cast<T*>(P.M.at(b, t));
      for (int j = 0; j < SRBD_M; ++j) {
        T acc = 0;
        for (int i = 0; i < 12; ++i) acc += jx(i, c) * W[i * SRBD_RES + 12 + j];
        Mo[c * SRBD_M + j] = acc;
      }
    }
  }
  T* go = const_cast<T*>(P.g.at(b, t));
  if (term) {
    T acc = 0;
    for (int i = 0; i < 12; ++i) acc += r[i] * Wr[i];
    go[0] = T(0.5) * acc;
    return;
  }
  // friction cone values, Jacobian rows (3 nonzeros each), penalty derivatives
  T pen = 0, d1[4], h[4], jf[4][3];
  const T* contact = ref + 13;
  for (int i = 0; i < 4; ++i) {
    const T fx = u[3 * i], fy = u[3 * i + 1], fz = u[3 * i + 2];
    const T rt = dev_sqrt(fx * fx + fy * fy + T(1e-2));
    const T cv = T(0.7) * fz - rt;
    T v, a1, a2;
    penalty(cv, v, a1, a2);
    const bool live = v > T(1e-6) && v < T(1e6);      // the outer clip of the objective
    pen += v < T(1e-6) ? T(1e-6) : (v > T(1e6) ? T(1e6) : v);
    d1[i] = live ? a1 : T(0);
    const T hc = a2 < T(1e-6) ? T(1e-6) : (a2 > T(1e6) ? T(1e6) : a2);
    h[i] = hc * contact[i];
    jf[i][0] = -fx / rt;
    jf[i][1] = -fy / rt;
    jf[i][2] = T(0.7);
  }
  {
    T acc = 0;
    for (int i = 0; i < SRBD_RES; ++i) acc += r[i] * Wr[i];
    go[0] = T(0.5) * (acc + pen);
  }
  // grad_u = (W r)[12:] + 1/2 J_f' d1 ; R = W[12:,12:] + J_f' diag(h) J_f
  T* ro = const_cast<T*>(P.r.at(b, t));
  T* Ro = const_cast<T*>(P.R.at(b, t));
  for (int j = 0; j < SRBD_M; ++j) {
    const int fi = j / 3, fk = j - 3 * fi;
    ro[j] = Wr[12 + j] + T(0.5) * jf[fi][fk] * d1[fi];
    for (int k = 0; k < SRBD_M; ++k) {
      const int gi = k / 3, gk = k - 3 * gi;
      T v = W[(12 + j) * SRBD_RES + 12 + k];
      if (fi == gi) v += jf[fi][fk] * h[fi] * jf[gi][gk];
      Ro[j * SRBD_M + k] = v;
    }
  }
  // dynamics and Jacobians
  const T* par = S.par.at(b, t);
  T* fo = const_cast<T*>(P.fbar.at(b, t));
  T xn[SRBD_N];
  srbd_step(S, x, u, par, xn);
  for (int k = 0; k < SRBD_N; ++k) fo[k] = xn[k];
  const T dt = S.dt;
  const T* p = x;
  const T* quat = x + 3;
  const T* om = x + 10;
  T wr[3] = {0, 0, 0}, skf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    const T c = par[i];
    const T* f = u + 3 * i;
    const T l[3] = {par[4 + 3 * i] - p[0], par[5 + 3 * i] - p[1], par[6 + 3 * i] - p[2]};
    T cr[3];
    cross3(l, f, cr);
    for (int k = 0; k < 3; ++k) wr[k] += cr[k] * c;
    // [f]x * c
    skf[1] += -f[2] * c; skf[2] += f[1] * c; skf[3] += f[2] * c;
    skf[5] += -f[0] * c; skf[6] += -f[1] * c; skf[7] += f[0] * c;
  }
  T Rm[9], Rtw[3], dRtw[12];
  quat_mat(quat, Rm);
  quat_mat_t_apply(quat, wr, Rtw, dRtw);
  // Ji = Iinv dt
  T Ji[9];
  for (int k = 0; k < 9; ++k) Ji[k] = S.Iinv[k] * dt;
  // fx rows of omega_next (10..12): [JiR'[f]x | Ji dR'w | 0 | I + Ji(-[om]x I + [I om]x)]
  T Fx[SRBD_N * SRBD_N];
  T Fu[SRBD_N * SRBD_M];
  for (int k = 0; k < SRBD_N * SRBD_N; ++k) Fx[k] = 0;
  for (int k = 0; k < SRBD_N * SRBD_M; ++k) Fu[k] = 0;
  T JiRt[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T acc = 0;
      for (int k = 0; k < 3; ++k) acc += Ji[i * 3 + k] * Rm[j * 3 + k];
      JiRt[i * 3 + j] = acc;
    }
  T Iom[3];
  for (int i = 0; i < 3; ++i) Iom[i] = S.I[i * 3] * om[0] + S.I[i * 3 + 1] * om[1] + S.I[i * 3 + 2] * om[2];
  const T sko[9] = {0, -om[2], om[1], om[2], 0, -om[0], -om[1], om[0], 0};
  const T skI[9] = {0, -Iom[2], Iom[1], Iom[2], 0, -Iom[0], -Iom[1], Iom[0], 0};
  T G[9];   // -[om]x I + [I om]x
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T acc = 0;
      for (int k = 0; k < 3; ++k) acc += sko[i * 3 + k] * S.I[k * 3 + j];
      G[i * 3 + j] = -acc + skI[i * 3 + j];
    }
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      T a0 = 0, a1 = 0;
      for (int k = 0; k < 3; ++k) { a0 += JiRt[i * 3 + k] * skf[k * 3 + j]; a1 += Ji[i * 3 + k] * G[k * 3 + j]; }
      Fx[(10 + i) * SRBD_N + j] = a0;
      Fx[(10 + i) * SRBD_N + 10 + j] = (i == j ? T(1) : T(0)) + a1;
    }
    for (int k = 0; k < 4; ++k) {
      T acc = 0;
      for (int j = 0; j < 3; ++j) acc += Ji[i * 3 + j] * dRtw[k * 3 + j];
      Fx[(10 + i) * SRBD_N + 3 + k] = acc;
    }
    for (int f = 0; f < 4; ++f) {
      const T c = par[f];
      const T l[3] = {par[4 + 3 * f] - p[0], par[5 + 3 * f] - p[1], par[6 + 3 * f] - p[2]};
      const T skl[9] = {0, -l[2], l[1], l[2], 0, -l[0], -l[1], l[0], 0};
      for (int j = 0; j < 3; ++j) {
        T acc = 0;
        for (int k = 0; k < 3; ++k) acc += JiRt[i * 3 + k] * skl[k * 3 + j];
        Fu[(10 + i) * SRBD_M + 3 * f + j] = acc * c;
      }
    }
  }
  // dp_next rows 7..9
  for (int i = 0; i < 3; ++i) {
    Fx[(7 + i) * SRBD_N + 7 + i] = T(1);
    for (int f = 0; f < 4; ++f) Fu[(7 + i) * SRBD_M + 3 * f + i] = par[f] * dt / S.mass;
  }
  // p_new rows 0..2 = I (p) + dt * rows 7..9
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < SRBD_N; ++j) Fx[i * SRBD_N + j] = (i == j ? T(1) : T(0)) + dt * Fx[(7 + i) * SRBD_N + j];
    for (int j = 0; j < SRBD_M; ++j) Fu[i * SRBD_M + j] = dt * Fu[(7 + i) * SRBD_M + j];
  }
  // quat rows 3..6 = Jq (cols 3..6) + Jv * rows 10..12
  T Jq[16], Jv[12], qtmp[4];
  quat_integrate<T, true>(quat, xn + 10, dt, qtmp, Jq, Jv);
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < SRBD_N; ++j) {
      T acc = (j >= 3 && j < 7) ? Jq[i * 4 + (j - 3)] : T(0);
      for (int k = 0; k < 3; ++k) acc += Jv[i * 3 + k] * Fx[(10 + k) * SRBD_N + j];
      Fx[(3 + i) * SRBD_N + j] = acc;
    }
    for (int j = 0; j < SRBD_M; ++j) {
      T acc = 0;
      for (int k = 0; k < 3; ++k) acc += Jv[i * 3 + k] * Fu[(10 + k) * SRBD_M + j];
      Fu[(3 + i) * SRBD_M + j] = acc;
    }
  }
  T* Ao = const_cast<T*>(P.A.at(b, t));
  T* Bo = const_cast<T*>(P.Bm.at(b, t));
  for (int k = 0; k < SRBD_N * SRBD_N; ++k) Ao[k] = Fx[k];
  for (int k = 0; k < SRBD_N * SRBD_M; ++k) Bo[k] = Fu[k];
}

// One thread per (stage t, trial j, instance b): l_t and f at
// x = X_t + a_j dX_t, u = U_t + a_j dU_t, a_j = a_max beta^j.
template <typename T>
__global__ void __launch_bounds__(64) k_srbd_trials(Prob<T> P, Srbd<T> S, Iter<T> it, Ws<T> ws, T beta, int L) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = P.N, N1 = N + 1;
  if (gid >= P.B * L * N1) return;
  const int t = gid % N1, bj = gid / N1, j = bj % L, b = bj / L;
  if (skip_instance(ws, b)) return;
  T a = ws.lsp[(size_t)b * NLS + LS_AMAX];
  for (int q = 0; q < j; ++q) a *= beta;
  const bool term = t == N;
  T x[SRBD_N], u[SRBD_M];
  const size_t ox = ((size_t)b * N1 + t) * SRBD_N;
  for (int k = 0; k < SRBD_N; ++k) x[k] = it.X[ox + k] + a * ws.dx[ox + k];
  const size_t ou = ((size_t)b * N + t) * SRBD_M;
  for (int k = 0; k < SRBD_M; ++k) u[k] = term ? T(0) : it.U[ou + k] + a * ws.du[ou + k];
  ws.trial_g[((size_t)b * L + j) * N1 + t] = srbd_cost(S.W.at(b, 0), S.ref.at(b, t), x, u, term);
  if (!term) {
    T xn[SRBD_N];
    srbd_step(S, x, u, S.par.at(b, t), xn);
    T* fo = ws.trial_f + (((size_t)b * L + j) * N + t) * SRBD_N;
    for (int k = 0; k < SRBD_N; ++k) fo[k] = xn[k];
  }
}

}  // namespace mpxk

namespace mpxk {

// Gait timer and reference generator of the SRBD controller
// (mpx/utils/mpc_utils.py:7-14 timer_run, :146-268 reference_generator_srbd,
// without the terrain estimator as in mpx/config/config_srbd.py), one thread
// per instance.  Per instance in: t_timer (4), x (13), foot (12), input (7:
// v, omega, base height), contact (4), liftoff (12, updated in place); out:
// reference [N+1][17], parameter [N+1][16], foot_ref_dot [N+1][12] (may be
// null).  With advance_dt > 0 the instance's t_timer is first advanced by
// timer_run(duty, step_freq, t_timer, advance_dt) in place (the controller's
// per-tick timer, mpx/utils/mpc_wrapper_srbd.py:106).
template <typename T>
struct SrbdGait {
  T dt, mass, duty, step_freq, step_height, clearance, advance_dt;
  T foot0[12];
  int N, B;
  T *t_timer, *liftoff;
  const T *x, *foot, *input, *contact;
  T *ref, *par, *foot_dot;
};

template <typename T>
__device__ __forceinline__ void timer_step(T duty, T freq, T dt, T* tm, T* c) {
  for (int i = 0; i < 4; ++i) {
    T v = tm[i] + dt * freq;
    if (v > T(1)) v -= T(1);
    tm[i] = v;
    c[i] = v < duty ? T(1) : T(0);
  }
}

template <typename T>
__global__ void k_srbd_reference(SrbdGait<T> G) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= G.B) return;
  const int N = G.N, N1 = N + 1;
  T tm[4], lo[12], foot[12], cin[4], in[7];
  for (int i = 0; i < 4; ++i) { tm[i] = G.t_timer[b * 4 + i]; cin[i] = G.contact[b * 4 + i]; }
  for (int i = 0; i < 12; ++i) { lo[i] = G.liftoff[b * 12 + i]; foot[i] = G.foot[b * 12 + i]; }
  for (int i = 0; i < 7; ++i) in[i] = G.input[b * 7 + i];
  if (G.advance_dt > T(0)) {
    T c[4];
    timer_step(G.duty, G.step_freq, G.advance_dt, tm, c);
    for (int i = 0; i < 4; ++i) G.t_timer[b * 4 + i] = tm[i];
  }
  const T* x = G.x + (size_t)b * 13;
  const T q0 = x[3], q1 = x[4], q2 = x[5], q3 = x[6];
  const T yaw = dev_atan2(T(2) * (q0 * q3 + q1 * q2), T(1) - T(2) * (q2 * q2 + q3 * q3));
  const T cy = dev_cos(yaw), sy = dev_sin(yaw);
  const T height = in[6] + (lo[2] + lo[5] + lo[8] + lo[11]) / T(4);
  const T p[3] = {x[0], x[1], height};
  const T v[3] = {cy * in[0] - sy * in[1], sy * in[0] + cy * in[1], in[2]};
  const T vn = dev_sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const T v0[2] = {-v[0] / (vn + T(1e-6)) * G.clearance, -v[1] / (vn + T(1e-6)) * G.clearance};
  // footholds of the swing legs: raibert-style offset from the projected nominal feet
  const T gain = dev_sqrt(in[6] / T(9.81));
  T fh[2][4];
  for (int i = 0; i < 4; ++i) {
    const T f0x = G.foot0[3 * i], f0y = G.foot0[3 * i + 1];
    const T proj[2] = {p[0] + cy * f0x - sy * f0y, p[1] + sy * f0x + cy * f0y};
    for (int d = 0; d < 2; ++d)
      fh[d][i] = T(0.5) * v[d] * G.duty / G.step_freq + gain * (x[7 + d] - v[d]) + proj[d];
  }
  // early contact: a leg the timer lifts that still touches down late in its swing
  T des[4], ct[4] = {tm[0], tm[1], tm[2], tm[3]}, early[4];
  timer_step(G.duty, G.step_freq, G.dt, ct, des);
  for (int i = 0; i < 4; ++i)
    early[i] = (des[i] == T(0) && cin[i] == T(1) && ct[i] > T(0.5) + T(0.5) * G.duty) ? T(1) : T(0);
  T* ref = G.ref + (size_t)b * N1 * 17;
  T* par = G.par + (size_t)b * N1 * 16;
  T* fd = G.foot_dot ? G.foot_dot + (size_t)b * N1 * 12 : nullptr;
  // the recursion's previous foot row starts as the current feet; the
  // previous contact of stage 0 is 0 (the reference reads the still-zero last
  // row there, JAX indexing -1)
  T prev[12], cprev[4] = {0, 0, 0, 0};
  for (int i = 0; i < 12; ++i) prev[i] = foot[i];
  T timer[4] = {tm[0], tm[1], tm[2], tm[3]};
  for (int t = 0; t < N1; ++t) {
    T c[4];
    timer_step(G.duty, G.step_freq, G.dt, timer, c);
    const T ts = T(t) * G.dt;
    T* r = ref + (size_t)t * 17;
    r[0] = ts * v[0] + p[0];
    r[1] = ts * v[1] + p[1];
    r[2] = height;
    r[3] = T(1); r[4] = T(0); r[5] = T(0); r[6] = T(0);
    for (int k = 0; k < 3; ++k) { r[7 + k] = v[k]; r[10 + k] = in[3 + k]; }
    T row[12], drow[12];
    for (int i = 0; i < 4; ++i) {
      if (c[i] == T(0) && cprev[i] != T(0)) {
        lo[3 * i] = prev[3 * i]; lo[3 * i + 1] = prev[3 * i + 1]; lo[3 * i + 2] = prev[3 * i + 2];
      }
      const T s = (timer[i] - G.duty) / (T(1) - G.duty);
      const bool stance = c[i] > T(0) || early[i] == T(1);
      for (int d = 0; d < 2; ++d) {
        const T cur = lo[3 * i + d], h = fh[d][i], w = v0[d];
        const T a2 = T(3) * (h - cur) - T(2) * w, a3 = w - T(2) * (h - cur);
        row[3 * i + d] = stance ? prev[3 * i + d] : cur + w * s + a2 * s * s + a3 * s * s * s;
        drow[3 * i + d] = stance ? T(0) : T(2) * a2 * s + T(3) * a3 * s * s + w;
      }
      {
        const T cur = lo[3 * i + 2], h = cur, top = cur + G.step_height, w = T(0.7);
        const T a = T(16) * top - T(8) * h - T(8) * cur - T(2) * w;
        const T bb = T(5) * w + T(14) * h + T(18) * cur - T(32) * top;
        const T cc = T(16) * top - T(5) * h - T(11) * cur - T(4) * w;
        row[3 * i + 2] = stance ? prev[3 * i + 2] : (((a * s + bb) * s + cc) * s + w) * s + cur;
        drow[3 * i + 2] = stance ? T(0) : ((T(4) * a * s + T(3) * bb) * s + T(2) * cc) * s + w;
      }
      r[13 + i] = c[i];
      par[(size_t)t * 16 + i] = c[i];
      cprev[i] = c[i];
    }
    for (int k = 0; k < 12; ++k) {
      par[(size_t)t * 16 + 4 + k] = row[k];
      if (fd) fd[(size_t)t * 12 + k] = drow[k];
      prev[k] = row[k];
    }
  }
  for (int i = 0; i < 12; ++i) G.liftoff[b * 12 + i] = lo[i];
}

}  // namespace mpxk
