"""Charged hard-sphere configurations for the Monte Carlo model tests."""
import numpy as np


def hard_sphere_system(n, packing, sigma=1.0, seed=0):
    """n non-overlapping spheres of diameter sigma at the given packing
    fraction in a periodic cube (random sequential addition), charges
    alternating +1/-1.  Returns (r, q, L)."""
    L = (n * np.pi / 6.0 * sigma ** 3 / packing) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 3))
    # candidates are tested against the particles of the 27 cells around
    # them (cells of side >= sigma): the same test as against all particles
    nc = max(1, int(L // sigma))
    cell = L / nc
    grid = {}
    k = 0
    while k < n:
        x = rng.uniform(0, L, 3)
        c = (x // cell).astype(int) % nc
        ids = [j for a in (-1, 0, 1) for b in (-1, 0, 1) for e in (-1, 0, 1)
               for j in grid.get(((c[0] + a) % nc, (c[1] + b) % nc, (c[2] + e) % nc), ())]
        if ids:
            dv = r[np.unique(ids)] - x
            dv -= L * np.rint(dv / L)
            if (dv * dv).sum(1).min() < sigma * sigma:
                continue
        r[k] = x
        grid.setdefault(tuple(c), []).append(k)
        k += 1
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return r, q, L


def min_distance_to_others(r, i, x, L):
    dv = np.delete(r, i, axis=0) - x
    dv -= L * np.rint(dv / L)
    return np.sqrt((dv * dv).sum(1).min())
