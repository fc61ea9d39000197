"""GPU Monte Carlo model (Bjerrum scaling, hard-sphere rejection) against the
oracle's model chain, and the batched chain against the plain chain.

Bars: FP64 plain chain: decisions identical to the oracle's, the same trials
flagged as overlaps (dE = +inf), finite dE within 1e-9.  Batched chains
(K = 8, 64): decisions and overlap flags identical to the plain chain, final
positions to 1e-12 (FP64); at least 99% equal decisions in mixed precision.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from chain_check import MIXED_DE_BOUND, assert_same_chain, thresholds
import pdmk4mc_amd as P
from hs_system import hard_sphere_system

pytestmark = pytest.mark.gpu
LB, SIGMA = 2.0, 1.0


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _moves(n, nm, seed, step=1.0):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, nm), rng.uniform(-step, step, (nm, 3)), rng.uniform(0, 1, nm)


def _gpu(r, q, L, leaf, precision=P.HPDMK_FP64, digits=6):
    g = P.HPDMKPtTree(P.HPDMKParams(leaf, digits, L, init=P.DIRECT), r, q, precision)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    g.mc_set_model(LB, SIGMA)
    return g


def _same_energies(a, b, tol):
    ia, ib = np.isinf(a), np.isinf(b)
    assert (ia == ib).all()
    assert np.abs(a[~ia] - b[~ib]).max() < tol


@pytest.mark.parametrize("leaf", [64, 16])
def test_model_chain_fp64_matches_oracle(leaf):
    log("model chain fp64 leaf", leaf)
    r, q, L = hard_sphere_system(2000, 0.2, SIGMA, seed=5)
    g = _gpu(r, q, L, leaf)
    t = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=6)
    t.form_pw()
    idx, d, u = _moves(2000, 300, 6)
    dEg, accg, _ = g.mc_chain(idx, d, u, beta=1.0, mode=0)
    dEo, acco = t.mc_chain_model(idx, d, u, 1.0, LB, SIGMA)
    assert (accg == acco).all()
    _same_energies(dEg, dEo, 1e-9)
    assert np.isinf(dEg).sum() > 10 and 0.1 < accg.mean() < 0.9
    assert np.abs(g.positions() - t.positions()).max() < 1e-12


@pytest.mark.parametrize("precision", [P.HPDMK_FP64, P.HPDMK_MIXED])
@pytest.mark.parametrize("batch", [8, 64])
def test_model_batched_equals_plain(precision, batch):
    log("model batched vs plain", precision, batch)
    r, q, L = hard_sphere_system(2000, 0.2, SIGMA, seed=8)
    g1, g2 = _gpu(r, q, L, 16, precision), _gpu(r, q, L, 16, precision)
    idx, d, u = _moves(2000, 1000, 9)
    dE1, acc1, _ = g1.mc_chain(idx, d, u, beta=1.0, mode=0)
    dE2, acc2, _ = g2.mc_chain(idx, d, u, beta=1.0, mode=1, batch=batch)
    if precision == P.HPDMK_FP64:
        assert (acc1 == acc2).all()
        _same_energies(dE1, dE2, 1e-9)
        assert np.abs(g1.positions() - g2.positions()).max() < 1e-12
    else:
        assert_same_chain(dE1, acc1, dE2, acc2, thresholds(u), MIXED_DE_BOUND)


def test_model_rejects_core_larger_than_finest_box():
    r, q, L = hard_sphere_system(2000, 0.2, SIGMA, seed=5)
    g = _gpu(r, q, L, 16)
    with pytest.raises(P.HPDMKError):
        g.mc_set_model(1.0, L)
    with pytest.raises(P.HPDMKError):
        g.mc_set_model(0.0, 0.5)
