"""Charged hard-sphere configurations for the Monte Carlo model tests."""
import numpy as np


def hard_sphere_system(n, packing, sigma=1.0, seed=0):
    """n non-overlapping spheres of diameter sigma at the given packing
    fraction in a periodic cube (random sequential addition), charges
    alternating +1/-1.  Returns (r, q, L)."""
    L = (n * np.pi / 6.0 * sigma ** 3 / packing) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 3))
    k = 0
    while k < n:
        x = rng.uniform(0, L, 3)
        if k:
            dv = r[:k] - x
            dv -= L * np.rint(dv / L)
            if (dv * dv).sum(1).min() < sigma * sigma:
                continue
        r[k] = x
        k += 1
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return r, q, L


def min_distance_to_others(r, i, x, L):
    dv = np.delete(r, i, axis=0) - x
    dv -= L * np.rint(dv / L)
    return np.sqrt((dv * dv).sum(1).min())
