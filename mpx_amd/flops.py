"""Algorithmic work of one SQP iteration per instance (1 FMA = 2 flops).

These are the counts the roofline figures use.  The sequential-Riccati LQR
count is the standard F_seq = N (4n^3 + 8n^2 m + 4n m^2 + m^3/3) plus the
equality-row augmentation F_eq = 2p(n+m)^2 + 2p(n+m) and the forward pass
4n^2 + 4nm per stage; the chunked scan adds its leaf combines, suffix
combines and segment maps on top.
"""
from __future__ import annotations


def lqr_flops(n, m, N, p, segments=1):
    f_seq = 4 * n**3 + 8 * n**2 * m + 4 * n * m**2 + m**3 / 3.0
    f_eq = 2 * p * (n + m) ** 2 + 2 * p * (n + m)
    f_fwd = 4 * n**2 + 4 * n * m
    total = N * (f_seq + f_eq + f_fwd)
    P = max(1, segments)
    if P > 1:
        leaves = (N + 1) * (P - 1) / P                       # up-sweep leaf combines
        total += leaves * (5 * n**3 + 12 * n**2 * m + f_eq)
        total += (P - 1) * (8 * n**3)                        # suffix value combines
        total += N * (P - 2) / P * (2 * n**3 + 2 * n**2 * m)  # forward segment maps
    return total


def riccati_bwd_flops(n, m, N, p):
    """The backward Riccati recursion alone (the k_lqr_wave kernel), per instance:
    per stage S A and A'(S A) (4n^3), S B, B'(S A) and G'K (6n^2 m), B'(S B) and
    H^{-1} G (4n m^2), the m x m factorisation (m^3/3), the equality-row terms
    D'D, E'D, E'E (2p(n+m)^2) and the affine vector terms (S c + s, A'(.),
    B'(.), G'k, H^{-1} g: 4n^2 + 4nm + 2m^2 + 2p(n+m))."""
    per = (4 * n**3 + 6 * n**2 * m + 4 * n * m**2 + m**3 / 3.0 + 2 * p * (n + m) ** 2
           + 4 * n**2 + 4 * n * m + 2 * m**2 + 2 * p * (n + m))
    return N * per


def linearize_flops(n, m, N, p):
    return N * (9 * n**2 + 2 * n * m + 2 * m**2 + 4 * p * (n + m)) + 2 * n**2


def trial_flops(n, m, N, p):
    """One merit evaluation at a trial point (dynamics, cost, AL rows)."""
    return (N + 1) * (6 * n**2 + 2 * n * m + 2 * m**2 + 2 * p * (n + m) + 10 * n + 10 * m)


def lqr_bytes(n, m, N, p, dsize=4, tanh_model=True):
    """Compulsory HBM bytes of the LQR phase: stage data read once in the backward
    pass and again in the forward pass, S/s/K/k written and read back, steps written."""
    per_stage_in = (n * n if tanh_model else n * n + n * m) + p * (n + m) + 2 * n + 2 * m
    per_stage_out = n * n + n + m * n + m
    fwd_in = n * n + n * m + m * n + m + n * n + n + 2 * n + n * n + m * m
    fwd_out = 2 * n + m
    return dsize * N * (per_stage_in + 2 * per_stage_out + fwd_in + fwd_out)


def sqp_flops(n, m, N, p, segments=1, trials=1):
    return lqr_flops(n, m, N, p, segments) + linearize_flops(n, m, N, p) + trials * trial_flops(n, m, N, p)
