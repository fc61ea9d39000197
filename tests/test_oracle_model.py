"""The oracle's Monte Carlo model (Bjerrum scaling, hard-sphere rejection),
pinned by identities with the plain Metropolis chain and by brute-force
geometry.  The model has no reference counterpart; the scaling is the
reference Ewald's eps = 1/l_B (src/ewald.cpp:187)."""
import numpy as np

import oracle_lib as O
from hs_system import hard_sphere_system, min_distance_to_others


def _moves(n, nm, seed, step=1.0):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, nm), rng.uniform(-step, step, (nm, 3)), rng.uniform(0, 1, nm)


def _tree(r, q, L, leaf=40, digits=6):
    t = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=digits)
    t.form_pw()
    return t


def test_model_without_core_is_the_plain_chain():
    rng = np.random.default_rng(1)
    L = 20.0
    r = rng.uniform(0, L, (600, 3))
    q = np.where(np.arange(600) % 2 == 0, 1.0, -1.0)
    idx, d, u = _moves(600, 80, 2)
    dE1, acc1, _ = _tree(r, q, L).mc_chain(idx, d, u, 1.0)
    dE2, acc2 = _tree(r, q, L).mc_chain_model(idx, d, u, 1.0, 1.0, 0.0)
    assert (acc1 == acc2).all() and np.array_equal(dE1, dE2)
    # Bjerrum length lB at beta is the plain chain at beta * lB, energies scaled by lB
    dE3, acc3, _ = _tree(r, q, L).mc_chain(idx, d, u, 2.5)
    dE4, acc4 = _tree(r, q, L).mc_chain_model(idx, d, u, 1.0, 2.5, 0.0)
    assert (acc3 == acc4).all()
    assert np.allclose(dE4, 2.5 * dE3, rtol=1e-14, atol=0)
    assert 0.05 < acc4.mean() < 0.95


def test_hard_core_rejections_are_exactly_the_overlaps():
    r, q, L = hard_sphere_system(500, 0.2, seed=3)
    idx, d, u = _moves(500, 300, 4)
    t = _tree(r, q, L)
    dE, acc = t.mc_chain_model(idx, d, u, 1.0, 2.0, 1.0)
    pos = r.copy()
    n_ovl = 0
    for k in range(len(idx)):
        i = idx[k]
        x = np.mod(pos[i] + d[k], L)
        ovl = min_distance_to_others(pos, i, x, L) < 1.0
        n_ovl += ovl
        assert np.isinf(dE[k]) == ovl, k
        if ovl:
            assert acc[k] == 0
        if acc[k]:
            pos[i] = x
    assert 10 < n_ovl < 290 and acc.sum() > 10
    assert np.abs(t.positions() - pos).max() < 1e-12
    for i in range(len(pos)):  # the chain never creates an overlap
        assert min_distance_to_others(pos, i, pos[i], L) >= 1.0
