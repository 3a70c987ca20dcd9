"""FP64 restatement of the reference's single-rigid-body quadruped model and
its Gauss-Newton objective -- the caller functions the SRBD controller hands
the solver (mpx/utils/mpc_wrapper_srbd.py:58-68).

TEST INFRASTRUCTURE ONLY (the checker of the device SRBD kernels): nothing in
the product imports it.

Restated from
* ``quadruped_srbd_dynamics`` (mpx/utils/models.py:6-37): state x = [p (3),
  quat (4, w first), dp (3), omega (3)], control u = ground reaction forces
  (3 per contact), stage parameter [contact (4), p_legs (12)];
* ``penalty``, ``quadruped_srbd_obj``, ``quadruped_srbd_hessian_gn``
  (mpx/utils/objectives.py:7-75);
* the quaternion helpers they call from mujoco-mjx 3.3.0 (pinned at
  pyproject.toml:17-18, absent here): ``quat_to_mat``, ``quat_mul``,
  ``quat_inv``, ``quat_integrate`` (axis-angle exponential map), ``quat_sub``
  (axis-angle logarithm of inv(v) u, angle wrapped into (-pi, pi]),
  ``normalize_with_norm`` (x / (|x| + 1e-6 [|x| == 0])).  These follow the
  published mjx/_src/math.py; restated from the mathematics, not copied.

Reference layout used here (per stage t = 0..N): reference row
[p_ref (3), quat_ref (4), dp_ref (3), omega_ref (3), contact (4)] (the
contact columns 13.. are what quadruped_srbd_hessian_gn reads,
objectives.py:70); parameter row [contact (4), p_legs (12)] (models.py:16-17).
Derivatives are analytic; tests/test_srbd.py pins them by central differences.
"""
from __future__ import annotations

import numpy as np

GRAV = np.array([0.0, 0.0, -9.81])
MU = 0.7
NC = 4
N_STATE = 13
N_CTRL = 3 * NC


# --------------------------------------------------------------------------
# quaternion helpers (mujoco-mjx math, restated)
# --------------------------------------------------------------------------

def quat_mul(u, v):
    return np.array([
        u[0] * v[0] - u[1] * v[1] - u[2] * v[2] - u[3] * v[3],
        u[0] * v[1] + u[1] * v[0] + u[2] * v[3] - u[3] * v[2],
        u[0] * v[2] - u[1] * v[3] + u[2] * v[0] + u[3] * v[1],
        u[0] * v[3] + u[1] * v[2] - u[2] * v[1] + u[3] * v[0],
    ])


def quat_mul_left(u):
    """L(u) with quat_mul(u, v) = L(u) v."""
    return np.array([[u[0], -u[1], -u[2], -u[3]],
                     [u[1], u[0], -u[3], u[2]],
                     [u[2], u[3], u[0], -u[1]],
                     [u[3], -u[2], u[1], u[0]]])


def quat_mul_right(v):
    """R(v) with quat_mul(u, v) = R(v) u."""
    return np.array([[v[0], -v[1], -v[2], -v[3]],
                     [v[1], v[0], v[3], -v[2]],
                     [v[2], -v[3], v[0], v[1]],
                     [v[3], v[2], -v[1], v[0]]])


def quat_inv(q):
    return q * np.array([1.0, -1.0, -1.0, -1.0])


def quat_to_mat(q):
    """Rotation matrix of a (not necessarily unit) quaternion, quadratic in q."""
    w, x, y, z = q
    return np.array([
        [w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z],
    ])


def quat_to_mat_grad(q):
    """dR/dq_k, k = 0..3 (each 3 x 3)."""
    w, x, y, z = q
    dw = 2 * np.array([[w, -z, y], [z, w, -x], [-y, x, w]])
    dx = 2 * np.array([[x, y, z], [y, -x, -w], [z, w, -x]])
    dy = 2 * np.array([[-y, x, w], [x, y, z], [-w, z, -y]])
    dz = 2 * np.array([[-z, -w, x], [w, -z, y], [x, y, z]])
    return [dw, dx, dy, dz]


def _is_zero(v):
    """mjx's norm treats a vector as zero when jnp.allclose(v, 0) holds (every
    |v_i| <= 1e-8, allclose's default atol); its value and gradient are then 0."""
    return bool(np.all(np.abs(v) <= 1e-8))


def _norm_safe(v):
    n = 0.0 if _is_zero(v) else float(np.sqrt(v @ v))
    return n, v / (n + (1e-6 if n == 0.0 else 0.0))


def quat_integrate(q, v, dt):
    """q * exp(v dt / 2), normalised."""
    n, axis = _norm_safe(v)
    ang = n * dt
    qr = np.concatenate([[np.cos(ang / 2)], axis * np.sin(ang / 2)])
    qn = quat_mul(q, qr)
    return _norm_safe(qn)[1]


def _sinc_half(n, dt):
    """sin(n dt / 2) / n and its derivative in n (series near n = 0)."""
    a = 0.5 * dt
    if n < 1e-4:
        s = a - (a ** 3) * n * n / 6.0
        ds = -(a ** 3) * n / 3.0
    else:
        s = np.sin(a * n) / n
        ds = (a * np.cos(a * n) * n - np.sin(a * n)) / (n * n)
    return s, ds


def quat_integrate_jac(q, v, dt):
    """(d quat_integrate / dq (4 x 4), d / dv (4 x 3)).  For a zero v in mjx's
    sense (_is_zero) the rotation is the identity and, as JAX's autodiff of
    the safe norm gives, d / dv = 0."""
    if _is_zero(v):
        qr = np.array([1.0, 0.0, 0.0, 0.0])
        qn = quat_mul(q, qr)
        nn = float(np.sqrt(qn @ qn))
        dnorm = (np.eye(4) - np.outer(qn, qn) / (nn * nn)) / nn
        return dnorm @ quat_mul_right(qr), np.zeros((4, 3))
    n = float(np.sqrt(v @ v))
    a = 0.5 * dt
    s, ds = _sinc_half(n, dt)
    qr = np.concatenate([[np.cos(a * n)], v * s])
    vhat = v / n if n > 0 else np.zeros(3)
    dqr = np.zeros((4, 3))
    dqr[0] = -np.sin(a * n) * a * vhat
    dqr[1:] = s * np.eye(3) + np.outer(v, ds * vhat)
    qn = quat_mul(q, qr)
    nn = float(np.sqrt(qn @ qn))
    dnorm = (np.eye(4) - np.outer(qn, qn) / (nn * nn)) / nn
    return dnorm @ quat_mul_right(qr), dnorm @ quat_mul_left(q) @ dqr


def quat_sub(u, v):
    """Axis-angle of inv(v) u, angle wrapped into (-pi, pi]."""
    q = quat_mul(quat_inv(v), u)
    s, axis = _norm_safe(q[1:])
    ang = 2.0 * np.arctan2(s, q[0])
    if ang > np.pi:
        ang -= 2.0 * np.pi
    return axis * ang


def quat_sub_jac(u, v):
    """d quat_sub(u, v) / du (3 x 4); 0 when the vector part of inv(v) u is
    zero in mjx's sense (JAX's autodiff through the safe norm)."""
    q = quat_mul(quat_inv(v), u)
    dq = quat_mul_left(quat_inv(v))
    qv, w = q[1:], q[0]
    if _is_zero(qv):
        return np.zeros((3, 4))
    s = float(np.sqrt(qv @ qv))
    ang = 2.0 * np.arctan2(s, w)
    if ang > np.pi:
        ang -= 2.0 * np.pi
    d2 = s * s + w * w
    J = np.zeros((3, 4))
    J[:, 0] = -2.0 * qv / d2                       # d(axis ang)/dw = axis * (-2 s / d2)
    J[:, 1:] = (ang / s) * np.eye(3) + np.outer(qv, qv) * ((2.0 * w / d2) / (s * s) - ang / s ** 3)
    return J @ dq


# --------------------------------------------------------------------------
# dynamics (mpx/utils/models.py:6-37)
# --------------------------------------------------------------------------

def _skew(a):
    return np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])


def dynamics(c, x, u, t, parameter):
    """x_{t+1}; ``c`` holds mass, inertia, inertia_inv, dt."""
    mass, I, Iinv, dt = c["mass"], c["inertia"], c["inertia_inv"], c["dt"]
    p, quat, dp, om = x[:3], x[3:7], x[7:10], x[10:13]
    contact, legs = parameter[t, :4], parameter[t, 4:]
    f = [u[3 * i:3 * i + 3] for i in range(NC)]
    force = sum(f[i] * contact[i] for i in range(NC))
    dp_next = dp + (GRAV + force / mass) * dt
    wrench = sum(np.cross(legs[3 * i:3 * i + 3] - p, f[i]) * contact[i] for i in range(NC))
    Rm = quat_to_mat(quat)
    om_next = om + Iinv @ (Rm.T @ wrench - np.cross(om, I @ om)) * dt
    p_new = p + dp_next * dt
    quat_new = quat_integrate(quat, om_next, dt)
    return np.concatenate([p_new, quat_new, dp_next, om_next])


def dynamics_jac(c, x, u, t, parameter):
    """(f_x 13 x 13, f_u 13 x 12), analytic."""
    mass, I, Iinv, dt = c["mass"], c["inertia"], c["inertia_inv"], c["dt"]
    p, quat, om = x[:3], x[3:7], x[10:13]
    contact, legs = parameter[t, :4], parameter[t, 4:]
    f = [u[3 * i:3 * i + 3] for i in range(NC)]
    wrench = sum(np.cross(legs[3 * i:3 * i + 3] - p, f[i]) * contact[i] for i in range(NC))
    Rm = quat_to_mat(quat)
    dR = quat_to_mat_grad(quat)
    fx = np.zeros((13, 13))
    fu = np.zeros((13, 12))
    # dp_next
    fx[7:10, 7:10] = np.eye(3)
    for i in range(NC):
        fu[7:10, 3 * i:3 * i + 3] = np.eye(3) * contact[i] * dt / mass
    # omega_next = om + Iinv (R' w - om x I om) dt
    dw_dp = sum(_skew(f[i]) * contact[i] for i in range(NC))          # d((l - p) x f)/dp = [f]x
    Ji = Iinv * dt
    fx[10:13, 0:3] = Ji @ Rm.T @ dw_dp
    for k in range(4):
        fx[10:13, 3 + k] = Ji @ (dR[k].T @ wrench)
    fx[10:13, 10:13] = np.eye(3) + Ji @ (-_skew(om) @ I + _skew(I @ om))
    for i in range(NC):
        fu[10:13, 3 * i:3 * i + 3] = Ji @ Rm.T @ _skew(legs[3 * i:3 * i + 3] - p) * contact[i]
    # p_new = p + dp_next dt
    fx[0:3, 0:3] = np.eye(3)
    fx[0:3, :] += dt * fx[7:10, :]
    fu[0:3, :] = dt * fu[7:10, :]
    # quat_new = quat_integrate(quat, omega_next, dt)
    om_next = dynamics(c, x, u, t, parameter)[10:13]
    Jq, Jv = quat_integrate_jac(quat, om_next, dt)
    fx[3:7, 3:7] = Jq
    fx[3:7, :] += Jv @ fx[10:13, :]
    fu[3:7, :] = Jv @ fu[10:13, :]
    return fx, fu


# --------------------------------------------------------------------------
# objective (mpx/utils/objectives.py:7-75)
# --------------------------------------------------------------------------

ALPHA_P, SIGMA_P = 0.1, 5.0


def _safe_log(x):
    return np.log(np.clip(x, 1e-10, 1e6))


def penalty(cv, alpha=ALPHA_P, sigma=SIGMA_P):
    """Relaxed log barrier of objectives.py:7-13, elementwise."""
    quad = alpha / 2 * (np.square((cv - 2 * sigma) / sigma) - 1.0)
    logb = -alpha * _safe_log(cv)
    return np.clip(np.where(cv > sigma, logb, quad + logb), 0.0, 1e8)


def penalty_d(cv, alpha=ALPHA_P, sigma=SIGMA_P):
    """(penalty', penalty'') elementwise, with the derivative conventions of
    the clips: zero where a clip is active (value outside its range), the
    log's own derivative inside."""
    cv = np.asarray(cv, dtype=float)
    val = penalty(cv, alpha, sigma)
    inlog = (cv > 1e-10) & (cv < 1e6)
    dlog = np.where(inlog, 1.0 / cv, 0.0)
    d2log = np.where(inlog, -1.0 / (cv * cv), 0.0)
    d1 = np.where(cv > sigma, -alpha * dlog, alpha * (cv - 2 * sigma) / sigma ** 2 - alpha * dlog)
    d2 = np.where(cv > sigma, -alpha * d2log, alpha / sigma ** 2 - alpha * d2log)
    live = (val > 0.0) & (val < 1e8)
    return np.where(live, d1, 0.0), np.where(live, d2, 0.0)


def friction_cone(u):
    fx, fy, fz = u[0::3], u[1::3], u[2::3]
    return MU * fz - np.sqrt(fx * fx + fy * fy + 1e-2)


def friction_cone_jac(u):
    J = np.zeros((NC, N_CTRL))
    for i in range(NC):
        fx, fy = u[3 * i], u[3 * i + 1]
        r = np.sqrt(fx * fx + fy * fy + 1e-2)
        J[i, 3 * i] = -fx / r
        J[i, 3 * i + 1] = -fy / r
        J[i, 3 * i + 2] = MU
    return J


def residual(x, u, ref_t):
    return np.concatenate([x[:3] - ref_t[:3], quat_sub(x[3:7], ref_t[3:7]), x[7:10] - ref_t[7:10],
                           x[10:13] - ref_t[10:13], u])


def cost(W, reference, x, u, t, N):
    """l_t of quadruped_srbd_obj (0.5 x the weighted residual, stage penalties)."""
    r = residual(x, u, reference[t])
    if t == N:
        rx = r[:12]
        return 0.5 * rx @ W[:12, :12] @ rx
    pen = np.clip(penalty(friction_cone(u)), 1e-6, 1e6)
    return 0.5 * (r @ W @ r + np.sum(pen))


def cost_grad(W, reference, x, u, t, N):
    """(grad_x l_t, grad_u l_t), analytic."""
    r = residual(x, u, reference[t])
    Jx = np.zeros((24, 13))
    Jx[0:3, 0:3] = np.eye(3)
    Jx[3:6, 3:7] = quat_sub_jac(x[3:7], reference[t][3:7])
    Jx[6:9, 7:10] = np.eye(3)
    Jx[9:12, 10:13] = np.eye(3)
    if t == N:
        return Jx[:12].T @ (W[:12, :12] @ r[:12]), np.zeros(N_CTRL)
    gx = Jx.T @ (W @ r)
    fc = friction_cone(u)
    pv = penalty(fc)
    d1, _ = penalty_d(fc)
    live = (pv > 1e-6) & (pv < 1e6)          # the outer clip of objectives.py:26
    gu = (W @ r)[12:] + 0.5 * friction_cone_jac(u).T @ np.where(live, d1, 0.0)
    return gx, gu


def hessian_gn(W, reference, x, u, t):
    """quadruped_srbd_hessian_gn: (J_x'W J_x, J_u'W J_u + H_c, J_x'W J_u)."""
    Jx = np.zeros((24, 13))
    Jx[0:3, 0:3] = np.eye(3)
    Jx[3:6, 3:7] = quat_sub_jac(x[3:7], reference[t][3:7])
    Jx[6:9, 7:10] = np.eye(3)
    Jx[9:12, 10:13] = np.eye(3)
    Ju = np.zeros((24, 12))
    Ju[12:] = np.eye(12)
    contact = reference[t][13:13 + NC]
    _, d2 = penalty_d(friction_cone(u))
    h = np.clip(d2, 1e-6, 1e6) * contact
    Jf = friction_cone_jac(u)
    Hc = Jf.T @ np.diag(h) @ Jf
    return Jx.T @ W @ Jx, Ju.T @ W @ Ju + Hc, Jx.T @ W @ Ju


def caller_instance(consts, W, reference, parameter, x0, N, p=0, D=None, E=None, e=None):
    """The SRBD problem as an oracle caller instance (pdilqr ``model == "caller"``)."""
    inst = dict(n=N_STATE, m=N_CTRL, N=N, p=p, model="caller", x0=np.asarray(x0, float),
                ulo=None, uhi=None)
    inst["f"] = lambda x, u, t: dynamics(consts, x, u, t, parameter)
    inst["fjac"] = lambda x, u, t: dynamics_jac(consts, x, u, t, parameter)
    inst["cost"] = lambda x, u, t: cost(W, reference, x, u, t, N)
    inst["grad"] = lambda x, u, t: cost_grad(W, reference, x, u, t, N)

    def hess(x, u, t):
        Qh, Rh, Mh = hessian_gn(W, reference, x, u, t)
        return Qh, Rh, Mh
    inst["hess"] = hess
    inst["D"] = np.zeros((N, p, N_STATE)) if D is None else D
    inst["E"] = np.zeros((N, p, N_CTRL)) if E is None else E
    inst["e"] = np.zeros((N, p)) if e is None else e
    return inst


# --------------------------------------------------------------------------
# gait timer and reference generator (mpx/utils/mpc_utils.py:7-14, :146-268)
# --------------------------------------------------------------------------

def timer_run(duty_factor, step_freq, leg_time, dt):
    """(contact, leg_time) after advancing each leg's gait phase by dt step_freq,
    wrapped back by 1 past 1; a leg is in stance while its phase < duty_factor
    (mpx/utils/mpc_utils.py:7-14)."""
    leg_time = leg_time + dt * step_freq
    leg_time = np.where(leg_time > 1, leg_time - 1, leg_time)
    return np.where(leg_time < duty_factor, 1.0, 0.0), leg_time


def spline_xy(cur, fh, v0, s):
    """Cubic from cur (slope v0) to fh over s in [0, 1] (mpc_utils.py:203-208)."""
    a2 = 3 * (fh - cur) - 2 * v0
    a3 = v0 - 2 * (fh - cur)
    return cur + v0 * s + a2 * s ** 2 + a3 * s ** 3


def spline_xy_dot(cur, fh, v0, s):
    a2 = 3 * (fh - cur) - 2 * v0
    a3 = v0 - 2 * (fh - cur)
    return 2 * a2 * s + 3 * a3 * s ** 2 + v0


def spline_z(cur, fh, top, s, v0=0.7):
    """Quartic through cur (s = 0, slope v0), top (s = 1/2) and fh (s = 1)
    (mpc_utils.py:210-219)."""
    a = 16 * top - 8 * fh - 8 * cur - 2 * v0
    b = 5 * v0 + 14 * fh + 18 * cur - 32 * top
    c = 16 * top - 5 * fh - 11 * cur - 4 * v0
    return a * s ** 4 + b * s ** 3 + c * s ** 2 + v0 * s + cur


def spline_z_dot(cur, fh, top, s, v0=0.7):
    a = 16 * top - 8 * fh - 8 * cur - 2 * v0
    b = 5 * v0 + 14 * fh + 18 * cur - 32 * top
    c = 16 * top - 5 * fh - 11 * cur - 4 * v0
    return 4 * a * s ** 3 + 3 * b * s ** 2 + 2 * c * s + v0


def reference_generator(N, dt, mass, foot0, t_timer, x, foot, inp, duty_factor, step_freq, step_height,
                        liftoff, contact, clearance_speed):
    """reference_generator_srbd with use_terrain_estimator = False
    (mpx/utils/mpc_utils.py:146-268; the SRBD config's setting,
    mpx/config/config_srbd.py): returns reference (N+1, 17) = [p_ref, quat_ref,
    dp_ref, omega_ref, contact], parameter (N+1, 16) = [contact, foot_ref], the
    updated liftoff (12) and foot_ref_dot (N+1, 12).  ``inp`` = [v (3),
    omega (3), base height]."""
    nc = NC
    p, quat, dp = x[:3], x[3:7], x[7:10]
    yaw = np.arctan2(2 * (quat[0] * quat[3] + quat[1] * quat[2]), 1 - 2 * (quat[2] ** 2 + quat[3] ** 2))
    Ryaw = np.array([[np.cos(yaw), -np.sin(yaw), 0.0], [np.sin(yaw), np.cos(yaw), 0.0], [0.0, 0.0, 1.0]])
    height = inp[6] + np.sum(liftoff[2::3]) / nc
    p = np.array([p[0], p[1], height])
    quat_ref = np.array([1.0, 0.0, 0.0, 0.0])
    # pitch of quat_ref is 0 without the terrain estimator: R_pitch = I
    v_ref = Ryaw @ inp[:3]
    w_ref = inp[3:6]
    ts = np.arange(N + 1) * dt
    ref = np.zeros((N + 1, 17))
    ref[:, 0] = ts * v_ref[0] + p[0]
    ref[:, 1] = ts * v_ref[1] + p[1]
    ref[:, 2] = height
    ref[:, 3:7] = quat_ref
    ref[:, 7:10] = v_ref
    ref[:, 10:13] = w_ref
    foot_ref = np.tile(foot, (N + 1, 1))
    foot_dot = np.zeros((N + 1, 3 * nc))
    foot0_proj = np.tile(p, nc) + np.concatenate([Ryaw @ foot0[3 * i:3 * i + 3] for i in range(nc)])
    des, cur_timer = timer_run(duty_factor, step_freq, t_timer, dt)
    early = ((des == 0) & (contact == 1) & (cur_timer > 0.5 + 0.5 * duty_factor)).astype(float)
    v0 = -v_ref / (np.linalg.norm(v_ref) + 1e-6) * clearance_speed
    fh = [0.5 * v_ref[d] * duty_factor / step_freq + np.sqrt(inp[6] / 9.81) * (dp[d] - v_ref[d])
          + foot0_proj[d::3] for d in (0, 1)]
    lx, ly, lz = liftoff[0::3].copy(), liftoff[1::3].copy(), liftoff[2::3].copy()
    cs = np.zeros((N + 1, nc))
    timer = np.array(t_timer, dtype=float)
    for t in range(N + 1):
        prev = foot_ref[t - 1]            # t = 0 reads row N (still the current feet), as JAX indexes -1
        c_new, timer = timer_run(duty_factor, step_freq, timer, dt)
        cs[t] = c_new
        lifting = (cs[t] == 0) & (cs[t - 1] != 0)
        lx = np.where(lifting, prev[0::3], lx)
        ly = np.where(lifting, prev[1::3], ly)
        lz = np.where(lifting, prev[2::3], lz)
        s = (timer - duty_factor) / (1 - duty_factor)
        stance = (c_new > 0) | (early == 1)
        row = np.zeros(3 * nc)
        row[0::3] = np.where(stance, prev[0::3], spline_xy(lx, fh[0], v0[0], s))
        row[1::3] = np.where(stance, prev[1::3], spline_xy(ly, fh[1], v0[1], s))
        row[2::3] = np.where(stance, prev[2::3], spline_z(lz, lz, lz + step_height, s))
        foot_ref[t] = row
        drow = np.zeros(3 * nc)
        drow[0::3] = np.where(stance, 0.0, spline_xy_dot(lx, fh[0], v0[0], s))
        drow[1::3] = np.where(stance, 0.0, spline_xy_dot(ly, fh[1], v0[1], s))
        drow[2::3] = np.where(stance, 0.0, spline_z_dot(lz, lz, lz + step_height, s))
        foot_dot[t] = drow
    ref[:, 13:17] = cs
    par = np.concatenate([cs, foot_ref], axis=1)
    lo = np.zeros(3 * nc)
    lo[0::3], lo[1::3], lo[2::3] = lx, ly, lz
    return ref, par, lo, foot_dot
