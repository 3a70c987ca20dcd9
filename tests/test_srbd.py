"""The SRBD oracle (oracle/srbd.py) against what the mathematics fixes:
central differences of its analytic Jacobians and gradients, rotation
properties of the quaternion helpers, and closed forms of the rigid-body
step.  CPU only."""
import numpy as np
import pytest

from oracle import srbd


def consts():
    I = np.array([[0.2310941359705289, -0.0014987128245817424, -0.021400468992761768],
                  [-0.0014987128245817424, 1.4485084687476608, 0.0004641447134275615],
                  [-0.021400468992761768, 0.0004641447134275615, 1.503217877350808]])
    return dict(mass=24.637, inertia=I, inertia_inv=np.linalg.inv(I), dt=0.02)


def _rand_quat(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    return q if q[0] > 0 else -q


def _state(rng):
    return np.concatenate([rng.standard_normal(3) * 0.1 + [0, 0, 0.33], _rand_quat(rng),
                           rng.standard_normal(3) * 0.3, rng.standard_normal(3) * 0.5])


def _param(rng, N):
    par = np.zeros((N + 1, 16))
    par[:, :4] = rng.integers(0, 2, size=(N + 1, 4))
    par[:, 4:] = rng.standard_normal((N + 1, 12)) * 0.2
    return par


def _fd(fn, z, h=1e-6):
    cols = []
    for j in range(z.size):
        e = np.zeros_like(z)
        e[j] = h
        cols.append((fn(z + e) - fn(z - e)) / (2 * h))
    return np.stack(cols, axis=-1)


def test_quaternion_helpers():
    rng = np.random.default_rng(0)
    q = _rand_quat(rng)
    R = srbd.quat_to_mat(q)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    assert abs(np.linalg.det(R) - 1) < 1e-12
    v = rng.standard_normal(3)
    # q * (0, v) * q^-1 rotates v by R
    rot = srbd.quat_mul(srbd.quat_mul(q, np.concatenate([[0.0], v])), srbd.quat_inv(q))[1:]
    np.testing.assert_allclose(rot, R @ v, atol=1e-12)
    # integrate then subtract recovers omega dt
    om = rng.standard_normal(3)
    q2 = srbd.quat_integrate(q, om, 0.02)
    np.testing.assert_allclose(srbd.quat_sub(q2, q), om * 0.02, atol=1e-12)
    for k, dR in enumerate(srbd.quat_to_mat_grad(q)):
        e = np.zeros(4); e[k] = 1e-6
        np.testing.assert_allclose(dR, (srbd.quat_to_mat(q + e) - srbd.quat_to_mat(q - e)) / 2e-6, atol=1e-8)


def test_quaternion_jacobians_fd():
    rng = np.random.default_rng(1)
    for _ in range(5):
        q, v, r = _rand_quat(rng), rng.standard_normal(3), _rand_quat(rng)
        Jq, Jv = srbd.quat_integrate_jac(q, v, 0.02)
        np.testing.assert_allclose(Jq, _fd(lambda z: srbd.quat_integrate(z, v, 0.02), q), atol=1e-8)
        np.testing.assert_allclose(Jv, _fd(lambda z: srbd.quat_integrate(q, z, 0.02), v), atol=1e-8)
        np.testing.assert_allclose(srbd.quat_sub_jac(q, r), _fd(lambda z: srbd.quat_sub(z, r), q), atol=1e-7)
    # small rotation rate: the series branch
    q, v = _rand_quat(rng), 1e-6 * rng.standard_normal(3)
    Jq, Jv = srbd.quat_integrate_jac(q, v, 0.02)
    np.testing.assert_allclose(Jv, _fd(lambda z: srbd.quat_integrate(q, z, 0.02), v, h=1e-9), atol=1e-6)


def test_rigid_body_step_closed_forms():
    """No contact: free fall with constant spin of a symmetric inertia; the
    position update is semi-implicit Euler."""
    c = consts()
    c["inertia"] = np.eye(3) * 0.5
    c["inertia_inv"] = np.eye(3) * 2.0
    rng = np.random.default_rng(2)
    x = _state(rng)
    par = np.zeros((2, 16))
    xn = srbd.dynamics(c, x, rng.standard_normal(12), 0, par)
    np.testing.assert_allclose(xn[7:10], x[7:10] + srbd.GRAV * 0.02, atol=1e-15)
    np.testing.assert_allclose(xn[10:13], x[10:13], atol=1e-15)        # om x I om = 0 for I = cI
    np.testing.assert_allclose(xn[0:3], x[0:3] + xn[7:10] * 0.02, atol=1e-15)
    # one foot in contact directly below the CoM pushing up: pure vertical force, no torque
    par[0, 0] = 1.0
    par[0, 4:7] = x[:3] - [0, 0, 0.3]
    u = np.zeros(12); u[2] = 24.637 * 9.81
    xn = srbd.dynamics(c, x, u, 0, par)
    np.testing.assert_allclose(xn[7:10], x[7:10], atol=1e-12)


def test_dynamics_jacobians_fd():
    c = consts()
    rng = np.random.default_rng(3)
    par = _param(rng, 3)
    for t in range(3):
        x, u = _state(rng), rng.standard_normal(12) * 50
        fx, fu = srbd.dynamics_jac(c, x, u, t, par)
        np.testing.assert_allclose(fx, _fd(lambda z: srbd.dynamics(c, z, u, t, par), x), atol=1e-7)
        np.testing.assert_allclose(fu, _fd(lambda z: srbd.dynamics(c, x, z, t, par), u), atol=1e-9)


def _weights():
    return np.diag(np.concatenate([[0, 0, 1e4], [1e3, 1e3, 0], [1e3] * 3, [1e1] * 3, [1e-2] * 12]))


def test_cost_gradient_fd():
    rng = np.random.default_rng(4)
    N = 3
    W = _weights()
    ref = np.zeros((N + 1, 17))
    ref[:, 3:7] = [_rand_quat(rng) for _ in range(N + 1)]
    ref[:, 13:] = rng.integers(0, 2, size=(N + 1, 4))
    for t in range(N + 1):
        x = _state(rng)
        u = rng.standard_normal(12) * 0.5
        u[2::3] = np.abs(u[2::3]) * 0.5          # friction cone values in the barrier's live range
        gx, gu = srbd.cost_grad(W, ref, x, u, t, N)
        np.testing.assert_allclose(gx, _fd(lambda z: srbd.cost(W, ref, z, u, t, N), x), rtol=1e-6, atol=1e-4)
        np.testing.assert_allclose(gu, _fd(lambda z: srbd.cost(W, ref, x, z, t, N), u), rtol=1e-6, atol=1e-6)


def test_penalty_second_derivative_fd():
    cv = np.array([0.05, 0.3, 1.0, 2.0, 4.0, 6.0])
    d1, d2 = srbd.penalty_d(cv)
    h = 1e-6
    np.testing.assert_allclose(d1, (srbd.penalty(cv + h) - srbd.penalty(cv - h)) / (2 * h), atol=1e-6)
    np.testing.assert_allclose(d2, (srbd.penalty_d(cv + h)[0] - srbd.penalty_d(cv - h)[0]) / (2 * h), atol=1e-5)


def test_gauss_newton_hessian_structure():
    """J_x'W J_x with the quaternion block of quat_sub's Jacobian; no cross term
    (the residual splits state and force); the friction part only for
    contacts in stance."""
    rng = np.random.default_rng(5)
    W = _weights()
    ref = np.zeros((2, 17))
    ref[:, 3:7] = _rand_quat(rng)
    ref[0, 13:] = [1, 0, 1, 0]
    x, u = _state(rng), np.abs(rng.standard_normal(12))
    Q, R, M = srbd.hessian_gn(W, ref, x, u, 0)
    assert np.all(M == 0)
    np.testing.assert_allclose(Q, Q.T, atol=1e-9)
    assert Q[2, 2] == 1e4 and Q[7, 7] == 1e3
    Jq = srbd.quat_sub_jac(x[3:7], ref[0, 3:7])
    np.testing.assert_allclose(Q[3:7, 3:7], Jq.T @ W[3:6, 3:6] @ Jq, atol=1e-9)
    # swing legs (contact 0) add nothing beyond W_grf
    np.testing.assert_allclose(R[3:6, 3:6], W[15:18, 15:18], atol=1e-15)


def test_zero_vector_conventions():
    """At a rotation rate or orientation error that is zero in mjx's sense
    (every component <= 1e-8) the value is the identity rotation / zero error
    and the derivative through mjx's safe norm is 0, as JAX's autodiff gives."""
    q = np.array([1.0, 0.0, 0.0, 0.0])
    Jq, Jv = srbd.quat_integrate_jac(q, np.zeros(3), 0.02)
    assert np.all(Jv == 0)
    np.testing.assert_allclose(srbd.quat_integrate(q, np.full(3, 1e-9), 0.02), q, atol=0)
    assert np.all(srbd.quat_sub_jac(q, q) == 0)
    assert np.all(srbd.quat_sub(q, q) == 0)
    # just above the threshold the analytic derivative holds
    v = np.array([2e-8, 0.0, 0.0])
    _, Jv = srbd.quat_integrate_jac(q, v, 0.02)
    np.testing.assert_allclose(Jv[1:], 0.01 * np.eye(3), atol=1e-12)


# --------------------------------------------------------------------------
# gait timer and reference generator
# --------------------------------------------------------------------------

def test_gait_splines_boundary_conditions():
    """The swing splines start at the liftoff point with the stated slope, end
    at the foothold, and the height spline passes its apex at mid swing."""
    cur, fh, v0, top = 0.3, 0.55, -0.4, 0.065
    assert abs(srbd.spline_xy(cur, fh, v0, 0.0) - cur) < 1e-15
    assert abs(srbd.spline_xy(cur, fh, v0, 1.0) - fh) < 1e-14
    assert abs(srbd.spline_xy_dot(cur, fh, v0, 0.0) - v0) < 1e-15
    h = 1e-6
    for s in (0.2, 0.7):
        fd = (srbd.spline_xy(cur, fh, v0, s + h) - srbd.spline_xy(cur, fh, v0, s - h)) / (2 * h)
        assert abs(fd - srbd.spline_xy_dot(cur, fh, v0, s)) < 1e-8
        fd = (srbd.spline_z(0.0, 0.0, top, s + h) - srbd.spline_z(0.0, 0.0, top, s - h)) / (2 * h)
        assert abs(fd - srbd.spline_z_dot(0.0, 0.0, top, s)) < 1e-8
    assert abs(srbd.spline_z(0.02, 0.02, top, 0.0) - 0.02) < 1e-15
    assert abs(srbd.spline_z(0.02, 0.02, top, 1.0) - 0.02) < 1e-14
    assert abs(srbd.spline_z(0.02, 0.02, top, 0.5) - top) < 1e-14
    assert abs(srbd.spline_z_dot(0.02, 0.02, top, 0.0) - 0.7) < 1e-15


def test_timer_and_reference_generator_invariants():
    """Timer: phases advance by dt f and wrap; stance while phase < duty.
    Generator: p_ref moves at the yaw-rotated commanded velocity from the
    measured position at the commanded height, contacts follow the timer
    stage by stage, stance feet stay put, and the liftoff point is recorded
    when a foot leaves the ground."""
    c, t = srbd.timer_run(0.65, 1.35, np.array([0.5, 0.0, 0.99, 0.64]), 0.02)
    np.testing.assert_allclose(t, [0.527, 0.027, 0.017, 0.667])
    np.testing.assert_array_equal(c, [1, 1, 1, 0])
    foot0 = np.array([0.27, 0.19, 0.0, 0.27, -0.19, 0.0, -0.21, 0.19, 0.0, -0.21, -0.19, 0.0])
    yaw = 0.3
    x = np.concatenate([[0.1, 0.2, 0.33], [np.cos(yaw / 2), 0, 0, np.sin(yaw / 2)], [0.3, 0.0, 0.0], np.zeros(3)])
    foot = np.tile([0.1, 0.2, 0.0], 4) + foot0
    inp = np.array([0.4, 0.1, 0.0, 0.0, 0.0, 0.2, 0.33])
    N, dt = 30, 0.02
    t0 = np.array([0.5, 0.0, 0.0, 0.5])
    ref, par, lo, fd = srbd.reference_generator(N, dt, 24.6, foot0, t0, x, foot, inp, 0.65, 1.35, 0.065, foot,
                                                np.ones(4), 0.4)
    v = np.array([np.cos(yaw) * 0.4 - np.sin(yaw) * 0.1, np.sin(yaw) * 0.4 + np.cos(yaw) * 0.1])
    np.testing.assert_allclose(ref[:, 0], 0.1 + np.arange(N + 1) * dt * v[0], atol=1e-12)
    np.testing.assert_allclose(ref[:, 1], 0.2 + np.arange(N + 1) * dt * v[1], atol=1e-12)
    np.testing.assert_allclose(ref[:, 2], 0.33, atol=1e-15)
    timer = t0.copy()
    for k in range(N + 1):
        cs, timer = srbd.timer_run(0.65, 1.35, timer, dt)
        np.testing.assert_array_equal(ref[k, 13:17], cs)
        np.testing.assert_array_equal(par[k, 0:4], cs)
    stance_all = np.all(par[:, 0:4] == 1, axis=0)
    for i in np.where(stance_all)[0]:
        np.testing.assert_allclose(par[:, 4 + 3 * i:7 + 3 * i], np.tile(foot[3 * i:3 * i + 3], (N + 1, 1)))
    lifted = [i for i in range(4) if np.any(par[:, i] == 0)]
    assert lifted
    for i in lifted:
        np.testing.assert_allclose(lo[3 * i:3 * i + 3], foot[3 * i:3 * i + 3], atol=1e-15)
        assert par[:, 6 + 3 * i].max() > 0.03          # the swing foot lifts
