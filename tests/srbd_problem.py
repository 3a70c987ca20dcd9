"""Seeded SRBD quadruped problems in the layout of the reference's SRBD
controller (mpx/config/config_srbd.py, mpx/utils/mpc_wrapper_srbd.py): the
Aliengo constants, the block-diagonal residual weight W, a trotting contact
sequence, a forward-walking reference and foot positions.  Test
infrastructure only."""
from __future__ import annotations

import numpy as np

MASS = 24.637
INERTIA = np.array([[0.2310941359705289, -0.0014987128245817424, -0.021400468992761768],
                    [-0.0014987128245817424, 1.4485084687476608, 0.0004641447134275615],
                    [-0.021400468992761768, 0.0004641447134275615, 1.503217877350808]])
DT = 0.02
P0 = np.array([0.0, 0.0, 0.33])
LEGS0 = np.array([0.27092872, 0.193, 0.0, 0.27092872, -0.193, 0.0,
                  -0.20887128, 0.193, 0.0, -0.20887128, -0.193, 0.0])


def weight():
    """block_diag(Qp, Qrot, Qdp, Qomega, Qgrf) of config_srbd.py."""
    return np.diag(np.concatenate([[0, 0, 1e4], [1e3, 1e3, 0], [1e3] * 3, [1e1] * 3, [1e-2] * 12]))


def consts():
    return dict(mass=MASS, inertia=INERTIA, inertia_inv=np.linalg.inv(INERTIA), dt=DT)


def make(i, N=25, perturb=True, seed=0):
    """(x0, reference (N+1, 17), parameter (N+1, 16)) of instance i."""
    rng = np.random.default_rng([91, int(i), int(seed)])
    v = np.array([0.3 + 0.2 * rng.uniform(), 0.1 * rng.standard_normal(), 0.0])
    phase = int(rng.integers(0, 10))
    ref = np.zeros((N + 1, 17))
    par = np.zeros((N + 1, 16))
    for t in range(N + 1):
        ref[t, 0:3] = P0 + v * t * DT
        ref[t, 3] = 1.0
        ref[t, 7:10] = v
        trot = ((t + phase) // 5) % 2
        contact = np.array([1, 0, 0, 1]) if trot == 0 else np.array([0, 1, 1, 0])
        if (t + phase) % 5 == 0:
            contact = np.ones(4)                       # all feet down at the switch
        ref[t, 13:17] = contact
        par[t, 0:4] = contact
        par[t, 4:] = LEGS0 + np.tile(ref[t, 0:3] * np.array([1, 1, 0]), 4)
    x0 = np.concatenate([P0, [1.0, 0, 0, 0], np.zeros(6)])
    if perturb:
        x0[0:3] += 0.01 * rng.standard_normal(3)
        q = np.array([1.0, *(0.02 * rng.standard_normal(3))])
        x0[3:7] = q / np.linalg.norm(q)
        x0[7:13] += 0.05 * rng.standard_normal(6)
    return x0, ref, par


def start(x0, N):
    """The controller's reset warm start: X = x0 tiled, U = u_ref = 0, V = 0
    (mpx/utils/mpc_wrapper_srbd.py:49-55)."""
    return np.tile(x0, (N + 1, 1)), np.zeros((N, 12)), np.zeros((N + 1, 13))
