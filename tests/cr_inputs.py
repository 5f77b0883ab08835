"""Seeded inputs of the constitutive-law tests (CPU host build and GPU
batch): stresses scaled to the yield stress and strain increments to
sigma_y / E, so that about half of the plastic cases pass the yield surface."""
import numpy as np

N = 2000


def inputs(seed, n=N):
    rng = np.random.default_rng(seed)
    E = 10 ** rng.uniform(8.5, 11.5, n)
    nu = rng.uniform(0.05, 0.45, n)
    sy = 10 ** rng.uniform(7.5, 9.0, n)
    s = rng.normal(size=(n, 6)) * (sy[:, None] * rng.uniform(0.01, 0.25, n)[:, None])
    de = rng.normal(size=(n, 6)) * (sy / E)[:, None] * rng.uniform(0.0, 1.5, n)[:, None]
    return E, nu, sy, s, de


def hardening(seed, sy, n=N):
    """alpha, C0, C1 of the J2-hardening cases."""
    rng = np.random.default_rng(seed)
    alpha = rng.uniform(0.0, 0.3, n) * sy
    C1 = 10 ** rng.uniform(8, 15, n)
    C0 = 7.0e8 / C1 * rng.uniform(0.5, 2.0, n)
    return alpha, C0, C1


def neohookean_inputs(seed, n=N):
    rng = np.random.default_rng(seed)
    E = 10 ** rng.uniform(6, 10, n)
    nu = rng.uniform(0.05, 0.45, n)
    s = rng.normal(size=(n, 6)) * 1e6
    Fm = np.eye(3)[None] + rng.normal(size=(n, 3, 3)) * 0.1   # [row][col]
    return E, nu, s, Fm
