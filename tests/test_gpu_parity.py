"""GPU parity of the HIP path against the CPU oracle (same inputs, same tree).

Bars (written in each test):
* FP64 mode: energies and trial dE agree with the oracle to 1e-10 relative to
  |E| (the reference normalises dE errors by the total energy,
  test/hpdmk_test.cpp:465) and to 1e-9 absolute on each dE;
* chained eval+update and the Metropolis chain reproduce the oracle's chain;
* the batched chain takes the same decisions as the plain chain;
* mixed mode meets the reference's Ewald contracts at digits 3 and 6.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from chain_check import MIXED_DE_BOUND, assert_same_chain, thresholds
import pdmk4mc_amd as P

pytestmark = pytest.mark.gpu
L = 20.0


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(n, leaf, digits, precision=P.HPDMK_FP64, seed=None, L=L):
    if seed is None:
        r, q = O.reference_test_system(n, L)
    else:
        rng = np.random.default_rng(seed)
        r = rng.uniform(0, L, (n, 3))
        q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    g = P.HPDMKPtTree(P.HPDMKParams(leaf, digits, L, init=P.DIRECT), r, q, precision)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    return r, q, g


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_energy_fp64_matches_oracle(digits):
    log("energy fp64 digits", digits)
    r, q, g = make(1000, 10, digits)
    t = O.OracleTree(r, q, L, n_per_leaf=10, digits=digits)
    t.form_pw()
    Eo = t.eval_energy()
    Eg = g.energy_parts()
    assert g.n_nodes == t.n_nodes
    assert np.abs(Eg - Eo).max() < 1e-10 * abs(Eo.sum()), (Eg, Eo)


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_shift_energy_fp64_matches_oracle(digits):
    log("shift fp64 digits", digits)
    r, q, g = make(1000, 10, digits)
    t = O.OracleTree(r, q, L, n_per_leaf=10, digits=digits)
    t.form_pw()
    E = abs(t.eval_energy().sum())
    idx, d = O.reference_test_moves(1000, L, 10)
    for k in range(10):
        eg, pg = g.eval_shift_energy(idx[k], *d[k], parts=True)
        eo, po = t.eval_shift_energy(idx[k], d[k], parts=True)
        assert abs(eg - eo) < 1e-10 * E and abs(eg - eo) < 1e-9, (k, eg, eo, pg, po)


@pytest.mark.parametrize("digits", [3, 6, 12])
def test_update_chain_fp64_matches_oracle(digits):
    log("update chain fp64 digits", digits)
    r, q, g = make(1000, 10, digits)
    t = O.OracleTree(r, q, L, n_per_leaf=10, digits=digits)
    t.form_pw()
    idx, d = O.reference_test_moves(1000, L, 10)
    for k in range(10):
        eg = g.eval_shift_energy(idx[k], *d[k])
        eo = t.eval_shift_energy(idx[k], d[k])
        assert abs(eg - eo) < 1e-9, (k, eg, eo)
        g.update_shift(idx[k], *d[k])
        t.update_shift(idx[k], d[k])
    assert np.abs(g.positions() - t.positions()).max() < 1e-12
    Eg, Eo = g.energy_parts(), t.eval_energy()
    assert np.abs(Eg - Eo).max() < 1e-10 * abs(Eo.sum())


def test_update_chain_vs_ewald_mixed():
    """test/hpdmk_test.cpp:470-533 contract in mixed precision (digits 3 and 6)."""
    for digits in (3, 6):
        log("update chain mixed vs ewald digits", digits)
        r, q, g = make(1000, 10, digits, P.HPDMK_MIXED)
        idx, d = O.reference_test_moves(1000, L, 10)
        dE_ew = O.ewald_shift_chain(r, q, L, 5.0, 1.0, idx, d, np.ones(10))
        for k in range(10):
            eg = g.eval_shift_energy(idx[k], *d[k])
            g.update_shift(idx[k], *d[k])
            assert abs(eg - dE_ew[k]) < 10.0 ** (1 - digits), (digits, k, eg, dE_ew[k])


def _moves(n, nm, seed, step=1.0, L=L):
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, n, nm)
    d = rng.uniform(-step, step, (nm, 3))
    u = rng.uniform(0, 1, nm)
    return idx, d, u


@pytest.mark.parametrize("digits", [3, 6])
def test_metropolis_chain_fp64_matches_oracle(digits):
    log("metropolis plain fp64 digits", digits)
    r, q, g = make(2000, 40, digits, seed=7)
    t = O.OracleTree(r, q, L, n_per_leaf=40, digits=digits)
    t.form_pw()
    idx, d, u = _moves(2000, 200, 11)
    dEg, accg, _ = g.mc_chain(idx, d, u, beta=1.0, mode=0)
    dEo, acco, _ = t.mc_chain(idx, d, u, beta=1.0)
    assert (accg == acco).all()
    assert np.abs(dEg - dEo).max() < 1e-9
    assert 0.2 < accg.mean() < 0.95


@pytest.mark.parametrize("precision", [P.HPDMK_FP64, P.HPDMK_MIXED])
@pytest.mark.parametrize("batch", [8, 40, 64, 128])  # mixed, 2..64: window pair terms on the matrix cores
def test_batched_chain_equals_plain(precision, batch):
    log("batched vs plain", precision, batch)
    r, q, g1 = make(2000, 40, 6, precision, seed=3)
    _, _, g2 = make(2000, 40, 6, precision, seed=3)
    idx, d, u = _moves(2000, 400, 5)
    dE1, acc1, _ = g1.mc_chain(idx, d, u, beta=1.0, mode=0)
    dE2, acc2, _ = g2.mc_chain(idx, d, u, beta=1.0, mode=1, batch=batch)
    if precision == P.HPDMK_FP64:
        assert (acc1 == acc2).all()
        assert np.abs(dE1 - dE2).max() < 1e-9
        assert np.abs(g1.positions() - g2.positions()).max() < 1e-12
        assert np.abs(g1.energy_parts() - g2.energy_parts()).max() < 1e-9
    else:
        assert_same_chain(dE1, acc1, dE2, acc2, thresholds(u), MIXED_DE_BOUND)


def test_batched_chain_mixed_across_compactions():
    """Mixed precision over 5000 moves (the matrix-core update, with and
    without the next batch's set-up in its launch, across compactions):
    the plain FP64 chain's decisions except at a near tie."""
    log("batched mixed vs plain FP64 across compactions")
    r, q, g1 = make(2000, 40, 6, P.HPDMK_FP64, seed=9)
    _, _, g2 = make(2000, 40, 6, P.HPDMK_MIXED, seed=9)
    idx, d, u = _moves(2000, 5000, 13)
    dE1, acc1, _ = g1.mc_chain(idx, d, u, beta=1.0, mode=0)
    dE2, acc2, _ = g2.mc_chain(idx, d, u, beta=1.0, mode=1, batch=64)
    k = assert_same_chain(dE1, acc1, dE2, acc2, thresholds(u), MIXED_DE_BOUND)
    log(f"first differing decision at move {k} of {len(acc1)}")
    g1.close()
    g2.close()


@pytest.mark.parametrize("batch", [64, 128])
def test_batched_chain_equals_plain_across_compactions(batch):
    """5000 moves: the slot compactions (every 2048 moves) fall between
    batches, so the run alternates between updates that set up the next batch
    in the same launch and updates followed by a compaction and a set-up of
    its own; batches over 64 move their particles in the plan."""
    log("batched vs plain across compactions", batch)
    r, q, g1 = make(2000, 40, 6, P.HPDMK_FP64, seed=9)
    _, _, g2 = make(2000, 40, 6, P.HPDMK_FP64, seed=9)
    idx, d, u = _moves(2000, 5000, 13)
    dE1, acc1, _ = g1.mc_chain(idx, d, u, beta=1.0, mode=0)
    dE2, acc2, _ = g2.mc_chain(idx, d, u, beta=1.0, mode=1, batch=batch)
    assert (acc1 == acc2).all()
    assert np.abs(dE1 - dE2).max() < 1e-9
    assert np.abs(g1.positions() - g2.positions()).max() < 1e-12
    assert np.abs(g1.energy_parts() - g2.energy_parts()).max() < 1e-9
    assert 0.2 < acc2.mean() < 0.95
    g1.close()
    g2.close()


@pytest.mark.parametrize("precision", [P.HPDMK_FP64, P.HPDMK_MIXED])
def test_batched_chain_equals_plain_digits3(precision):
    """Digits 3: smaller window and difference grids (the Gram blocks' mode
    count, the update's matrix-core tiles) on the same chain test."""
    log("batched vs plain, digits 3", precision)
    r, q, g1 = make(2000, 40, 3, precision, seed=5)
    _, _, g2 = make(2000, 40, 3, precision, seed=5)
    idx, d, u = _moves(2000, 400, 7)
    dE1, acc1, _ = g1.mc_chain(idx, d, u, beta=1.0, mode=0)
    dE2, acc2, _ = g2.mc_chain(idx, d, u, beta=1.0, mode=1, batch=64)
    if precision == P.HPDMK_FP64:
        assert (acc1 == acc2).all()
        assert np.abs(dE1 - dE2).max() < 1e-9
        assert np.abs(g1.positions() - g2.positions()).max() < 1e-12
    else:
        assert_same_chain(dE1, acc1, dE2, acc2, thresholds(u), MIXED_DE_BOUND)


@pytest.mark.parametrize("digits", [9, 12])
def test_batched_chain_equals_plain_wide_grids(digits):
    """Digits 9 and 12 (FP64): grids wider than 28 points per axis take the
    wide instantiation of the update (and of the update that sets up the
    next batch in its launch)."""
    log("batched vs plain, wide grids", digits)
    r, q, g1 = make(2000, 40, digits, P.HPDMK_FP64, seed=21)
    _, _, g2 = make(2000, 40, digits, P.HPDMK_FP64, seed=21)
    idx, d, u = _moves(2000, 400, 23)
    dE1, acc1, _ = g1.mc_chain(idx, d, u, beta=1.0, mode=0)
    dE2, acc2, _ = g2.mc_chain(idx, d, u, beta=1.0, mode=1, batch=64)
    assert (acc1 == acc2).all()
    assert np.abs(dE1 - dE2).max() < 1e-9
    assert np.abs(g1.positions() - g2.positions()).max() < 1e-12
    assert np.abs(g1.energy_parts() - g2.energy_parts()).max() < 1e-9
    g1.close()
    g2.close()


@pytest.mark.parametrize("precision", [P.HPDMK_FP64, P.HPDMK_MIXED])
def test_batched_chain_single_trial_batches(precision):
    """Moves of one particle in a row cut the batch at every repeat: batches
    of one trial (and a batch size of 1) take the plain chain's decisions."""
    log("batched single-trial batches", precision)
    r, q, g1 = make(2000, 40, 6, precision, seed=4)
    _, _, g2 = make(2000, 40, 6, precision, seed=4)
    _, _, g3 = make(2000, 40, 6, precision, seed=4)
    idx, d, u = _moves(2000, 240, 6)
    idx[1::2] = idx[0::2]  # every particle moves twice in a row
    idx[100:110] = idx[100]
    dE1, acc1, _ = g1.mc_chain(idx, d, u, beta=1.0, mode=0)
    dE2, acc2, _ = g2.mc_chain(idx, d, u, beta=1.0, mode=1, batch=16)
    dE3, acc3, _ = g3.mc_chain(idx, d, u, beta=1.0, mode=1, batch=1)
    for dE, acc in ((dE2, acc2), (dE3, acc3)):
        if precision == P.HPDMK_FP64:
            assert (acc1 == acc).all()
            assert np.abs(dE1 - dE).max() < 1e-9
        else:
            assert_same_chain(dE1, acc1, dE, acc, thresholds(u), MIXED_DE_BOUND)
    if precision == P.HPDMK_FP64:
        assert np.abs(g1.positions() - g2.positions()).max() < 1e-12
        assert np.abs(g1.positions() - g3.positions()).max() < 1e-12


def test_batch_eval_equals_single_eval():
    r, q, g = make(1000, 10, 6)
    idx, d = O.reference_test_moves(1000, L, 10)
    batch = g.eval_shift_energy_batch(idx, d)
    single = np.array([g.eval_shift_energy(idx[k], *d[k]) for k in range(10)])
    assert np.abs(batch - single).max() < 1e-12
