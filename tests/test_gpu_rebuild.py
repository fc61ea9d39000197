"""Tree rebuild from the current positions (hpdmk_tree_rebuild; no
counterpart in the reference, which keeps the first topology).

Bars (written in each test):
* after a long FP64 chain the incremental energy equals the rebuilt system's
  energy to 1e-10 of |E| (particle lists are exact on every level, so the
  drift check a rebuild provides reads zero), and both equal the oracle's
  energy of a fresh tree on the same positions to 1e-10;
* the chain continued on the rebuilt system takes the oracle's decisions;
* a system driven into HPDMK_ECAPACITY is recovered by a rebuild.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
import pdmk4mc_amd as P

pytestmark = pytest.mark.gpu
L = 20.0


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_rebuild_energy_and_chain():
    log("rebuild after a chain")
    n = 3000
    r, q = O.reference_test_system(n, L)
    g = P.HPDMKPtTree(P.HPDMKParams(20, 6, L, init=P.DIRECT), r, q)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    rng = np.random.default_rng(4)
    m = 4000
    idx = rng.integers(0, n, m)
    d = rng.uniform(-1.5, 1.5, (m, 3))
    u = rng.uniform(0, 1, m)
    g.mc_chain(idx, d, u, beta=1.0, mode=1, batch=32)
    E_inc = g.eval_energy()
    g2 = g.rebuild()
    E_new = g2.eval_energy()
    assert abs(E_new - E_inc) < 1e-10 * abs(E_inc), (E_new, E_inc)
    pos = g2.positions()
    assert np.abs(pos - g.positions()).max() == 0.0
    t = O.OracleTree(pos, q, L, n_per_leaf=20, digits=6)
    t.form_pw()
    Eo = t.eval_energy().sum()
    assert abs(E_new - Eo) < 1e-10 * abs(Eo)
    assert g2.n_nodes == t.n_nodes
    # the chain goes on from the rebuilt system exactly like the oracle's
    m2 = 300
    i2 = rng.integers(0, n, m2)
    d2 = rng.uniform(-1.0, 1.0, (m2, 3))
    u2 = rng.uniform(0, 1, m2)
    dE_g, acc_g, _ = g2.mc_chain(i2, d2, u2, beta=1.0, mode=0)
    dE_o, acc_o, _ = t.mc_chain(i2, d2, u2, 1.0)
    assert (acc_g == acc_o).all()
    assert np.abs(dE_g - dE_o).max() < 1e-9
    g.close()
    g2.close()


def test_rebuild_recovers_capacity_overflow():
    log("rebuild after capacity overflow")
    n = 2000
    r, q = O.reference_test_system(n, L)
    g = P.HPDMKPtTree(P.HPDMKParams(64, 3, L, init=P.DIRECT), r, q)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    # drive particles into one corner with forced updates until a leaf is full
    target = np.array([1.0, 1.0, 1.0])
    pos = g.positions()
    overflow = False
    for i in range(n):
        d = target + 0.3 * np.random.default_rng(i).uniform(-1, 1, 3) - pos[i]
        try:
            g.update_shift(i, *d)
        except P.HPDMKError as e:
            assert e.code == -4, e  # HPDMK_ECAPACITY
            overflow = True
            break
    assert overflow
    g.clear_error()
    g2 = g.rebuild()
    E2 = g2.eval_energy()
    t = O.OracleTree(g2.positions(), q, L, n_per_leaf=64, digits=3)
    t.form_pw()
    assert abs(E2 - t.eval_energy().sum()) < 1e-10 * abs(E2)
    g2.update_shift(0, 3.0, 0.0, 0.0)  # the rebuilt system takes moves again
    g.close()
    g2.close()


def test_rebuild_refuses_deeper_tree_under_hard_core():
    """A crowded leaf makes the rebuilt tree deeper; its finest box is then
    smaller than a hard-sphere diameter that the first tree allowed, and the
    overlap search (the target leaf's near list) would miss contacts: the
    rebuild must refuse instead of clipping silently."""
    log("rebuild refused under a hard core wider than the new finest box")
    n = 2000
    r, q = O.reference_test_system(n, L)
    g = P.HPDMKPtTree(P.HPDMKParams(64, 3, L, init=P.DIRECT), r, q)
    finest = L / 2 ** (g.max_depth - 1)
    g.mc_set_model(1.0, 0.9 * finest)
    pos = g.positions()
    target = np.array([1.0, 1.0, 1.0])
    for i in range(60):  # 60 arrivals fit the leaf's slots (count + 64) and exceed the leaf cap
        d = target + 0.2 * np.random.default_rng(i).uniform(-1, 1, 3) - pos[i]
        g.update_shift(i, *d)
    with pytest.raises(P.HPDMKError) as e:
        g.rebuild()
    assert e.value.code == -1, e.value  # HPDMK_EINVAL
    g.mc_set_model(1.0, 0.0)
    g2 = g.rebuild()  # without the hard core the deeper tree is fine
    assert g2.max_depth > g.max_depth
    g.close()
    g2.close()


def test_chain_rejects_bad_beta():
    log("beta must be finite and > 0")
    n = 1000
    r, q = O.reference_test_system(n, L)
    g = P.HPDMKPtTree(P.HPDMKParams(64, 3, L, init=P.DIRECT), r, q)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    idx = np.arange(8)
    d = np.full((8, 3), 0.1)
    u = np.full(8, 0.5)
    for beta in (0.0, -1.0, float("nan"), float("inf")):
        for mode in (0, 1):
            with pytest.raises(P.HPDMKError) as e:
                g.mc_chain(idx, d, u, beta=beta, mode=mode)
            assert e.value.code == -1
    g.close()
