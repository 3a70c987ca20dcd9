"""Error model for FP32 results against the FP64 oracle.

An FP32 computation is at best backward stable: it returns the exact answer
for inputs perturbed at the level of FP32 rounding.  The FP64 oracle's answer
on inputs perturbed by random relative amounts of u32 = 2^-24 therefore shows
how far an ideal FP32 solve may land from the FP64 answer on this instance:
the problem's own sensitivity, measured rather than guessed.  A test then
asserts that the device's FP32 error stays within a stated multiple of that
spread (rounding accumulates over the O(n) terms of each product and the N
stages of a pass, so the multiple is a few hundred ulps of the input
perturbation, not one).
"""
from __future__ import annotations

import numpy as np

U32 = 2.0 ** -24
_INPUT_KEYS = ("A", "B", "W", "Q", "R", "QN", "D", "E", "e", "x0", "xref", "uref")


def perturb_instance(inst, rng, rel=U32):
    """A copy of ``inst`` whose float arrays are multiplied elementwise by
    (1 + rel * xi), xi uniform in [-1, 1]."""
    d = dict(inst)
    for k in _INPUT_KEYS:
        a = inst.get(k)
        if isinstance(a, np.ndarray) and a.size:
            d[k] = a * (1.0 + rel * rng.uniform(-1.0, 1.0, size=a.shape))
    return d


def perturb_start(start, rng, rel=U32):
    return tuple(a * (1.0 + rel * rng.uniform(-1.0, 1.0, size=a.shape)) for a in start)


def rel_err(a, b):
    """max over instances of |a - b|_inf / max(1, |b|_inf) (leading axis = instance)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.size == 0:
        return 0.0
    B = a.shape[0]
    d = np.abs(a - b).reshape(B, -1).max(axis=1)
    s = np.maximum(1.0, np.abs(b).reshape(B, -1).max(axis=1))
    return float(np.max(d / s))
