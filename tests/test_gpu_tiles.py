"""Launch shapes and the tile autotuner: the chain does not depend on the
shape, and tuning leaves the system untouched.

Bars (FP64): a batched chain with a non-default shape takes the plain
chain's decisions with dE within 1e-9 and positions within 1e-12; after
autotuning, the system's positions and energy are unchanged and the tuned
batched chain (batch 0) again equals the plain chain.
"""
import sys

import numpy as np
import pytest

import pdmk4mc_amd as P

pytestmark = pytest.mark.gpu
L = 25.0


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def system(n=4000, seed=3, leaf=40):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, L, (n, 3))
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    g = P.HPDMKPtTree(P.HPDMKParams(leaf, 6, L, init=P.DIRECT), r, q, P.HPDMK_FP64)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    return r, q, g


def moves(n, M, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, M), rng.uniform(-1, 1, (M, 3)), rng.uniform(0, 1, M)


def test_chain_independent_of_launch_shape():
    log("tiles: non-default shape")
    _, _, g1 = system()
    _, _, g2 = system()
    assert g1.max_depth >= 4
    g2.set_tiles(win=1, diff=3, near=5, apply=300, batch=40)
    assert g2.tiles() == {"win": 1, "diff": 3, "near": 5, "apply": 300, "batch": 40}
    idx, d, u = moves(4000, 600, 4)
    dE1, a1, _ = g1.mc_chain(idx, d, u, mode=0)
    dE2, a2, _ = g2.mc_chain(idx, d, u, mode=1, batch=0)
    assert (a1 == a2).all() and np.abs(dE1 - dE2).max() < 1e-9
    assert np.abs(g1.positions() - g2.positions()).max() < 1e-12


def test_autotune_keeps_state_and_chain():
    log("tiles: autotune")
    _, _, g = system()
    _, _, ref = system()
    x0, E0 = g.positions(), g.eval_energy()
    sidx, sd, su = moves(4000, 512, 9)
    tiles, ms = g.autotune(sidx, sd, su)
    log("tuned", tiles, ms)
    assert ms > 0 and 1 <= tiles["batch"] <= 256 and 1 <= tiles["near"] <= 32
    assert g.tiles() == tiles
    assert np.array_equal(g.positions(), x0) and abs(g.eval_energy() - E0) < 1e-12 * abs(E0)  # FP64 atomics: summation order
    idx, d, u = moves(4000, 500, 10)
    dE1, a1, _ = ref.mc_chain(idx, d, u, mode=0)
    dE2, a2, _ = g.mc_chain(idx, d, u, mode=1, batch=0)
    assert (a1 == a2).all() and np.abs(dE1 - dE2).max() < 1e-9


def test_bad_tiles_rejected():
    _, _, g = system(n=1000)
    with pytest.raises(P.HPDMKError):
        g.set_tiles(win=0)
    with pytest.raises(P.HPDMKError):
        g.set_tiles(near=33)
    with pytest.raises(P.HPDMKError):
        g.set_tiles(batch=300)
