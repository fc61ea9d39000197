"""CPU checks of the chain comparison used by the mixed-precision GPU tests:
mistakes planted in a copy of a reference chain must fail it."""
import numpy as np
import pytest

from chain_check import assert_same_chain, thresholds


def _chain(seed=0, n=200):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, 1, n)
    dE = rng.normal(0, 1, n)
    thr = thresholds(u)
    return dE, (dE < thr).astype(np.int32), thr


def test_identical_chains_pass():
    dE, acc, thr = _chain()
    assert assert_same_chain(dE, acc, dE + 1e-7, acc, thr, 1e-5) == len(acc)


def test_near_tie_divergence_passes():
    dE, acc, thr = _chain()
    k = 50
    dE[k] = thr[k] - 1e-7  # the reference accepts by a hair
    acc[k] = 1
    got_acc = acc.copy()
    got_acc[k] = 0
    got_acc[k + 1:] = 1 - got_acc[k + 1:]  # anything may follow a legitimate divergence
    got_dE = dE.copy()
    got_dE[k + 1:] += 5.0
    assert assert_same_chain(dE, acc, got_dE, got_acc, thr, 1e-5) == k


def test_divergence_away_from_threshold_fails():
    dE, acc, thr = _chain()
    k = int(np.argmax(np.abs(dE - thr) > 0.5))
    got = acc.copy()
    got[k] = 1 - got[k]
    with pytest.raises(AssertionError):
        assert_same_chain(dE, acc, dE, got, thr, 1e-5)


def test_energy_error_before_divergence_fails():
    dE, acc, thr = _chain()
    got = dE.copy()
    got[10] += 1e-3
    with pytest.raises(AssertionError):
        assert_same_chain(dE, acc, got, acc, thr, 1e-5)


def test_overlap_mismatch_fails():
    dE, acc, thr = _chain()
    dE[5], acc[5] = np.inf, 0
    got = dE.copy()
    got[5] = 0.3
    with pytest.raises(AssertionError):
        assert_same_chain(dE, acc, got, acc, thr, 1e-5)
