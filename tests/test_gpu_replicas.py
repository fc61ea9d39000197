"""Replica sets: ragged independent systems advanced together, one move of
every replica per step (HIP-graph replays).  Each replica's chain must be the
plain chain of that system on its own.

Bars: FP64: decisions identical and dE within 1e-9 of the replica's plain
chain run through its own hpdmk_tree, final positions to 1e-12, and one
replica against the oracle's Metropolis chain; mixed: at least 99% equal
decisions.  Graph replays and direct launches give identical results.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from chain_check import MIXED_DE_BOUND, assert_same_chain, thresholds
import pdmk4mc_amd as P

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ragged(n_rep, seed, lo=300, hi=3000, density=0.1, leaf=40, digits=6):
    rng = np.random.default_rng(seed)
    systems, params = [], []
    for _ in range(n_rep):
        n = int(rng.integers(lo, hi + 1))
        L = (n / density) ** (1.0 / 3.0)
        r = rng.uniform(0, L, (n, 3))
        q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
        systems.append((r, q))
        params.append(P.HPDMKParams(leaf, digits, L, init=P.DIRECT))
    return params, systems


def moves(systems, M, seed):
    rng = np.random.default_rng(seed)
    idx = np.stack([rng.integers(0, len(q), M) for _, q in systems])
    d = rng.uniform(-1.0, 1.0, (len(systems), M, 3))
    u = rng.uniform(0, 1, (len(systems), M))
    return idx, d, u


def single(params, r, q, precision):
    g = P.HPDMKPtTree(params, r, q, precision)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    return g


@pytest.mark.parametrize("precision", [P.HPDMK_FP64, P.HPDMK_MIXED])
def test_replica_chains_equal_single_chains(precision):
    log("replicas vs single chains", precision)
    params, systems = ragged(12, 1)
    rs = P.HPDMKReplicas(params, systems, precision)
    depths = {rs.tree(k).max_depth for k in range(rs.n)}
    assert depths >= {3, 4}, depths  # ragged depths: with and without difference levels
    rs.form_pw()
    idx, d, u = moves(systems, 100, 2)
    rs.mc_load(idx, d, u)
    rs.mc_run(1.0, graph_steps=16)
    dE, acc = rs.mc_results()
    for k, (r, q) in enumerate(systems):
        g = single(params[k], r, q, precision)
        dEs, accs, _ = g.mc_chain(idx[k], d[k], u[k], beta=1.0, mode=0)
        if precision == P.HPDMK_FP64:
            assert (acc[k] == accs).all(), k
            assert np.abs(dE[k] - dEs).max() < 1e-9, k
            assert np.abs(rs.tree(k).positions() - g.positions()).max() < 1e-12
        else:
            assert_same_chain(dEs, accs, dE[k], acc[k], thresholds(u[k]), MIXED_DE_BOUND)
        g.close()
    assert 0.2 < acc.mean() < 0.95
    rs.close()


def test_replica_matches_oracle_and_graph_equals_direct():
    log("replicas: oracle, graph vs direct vs whole-run graph, compaction")
    params, systems = ragged(4, 7, lo=500, hi=1500)
    idx, d, u = moves(systems, 700, 8)  # > 512 steps: one slot compaction inside the run
    out = []
    for gs in (64, 0, 700):  # 700: the whole run, compaction included, as one graph launch
        rs = P.HPDMKReplicas(params, systems, P.HPDMK_FP64)
        rs.form_pw()
        rs.mc_load(idx, d, u)
        rs.mc_run(1.0, graph_steps=gs)
        out.append(rs.mc_results() + ([rs.tree(k).positions() for k in range(rs.n)],))
        E = [rs.tree(k).eval_energy() for k in range(rs.n)]
        rs.close()
    (dE1, acc1, x1), (dE2, acc2, x2), (dE3, acc3, x3) = out
    for dEk, acck, xk in ((dE2, acc2, x2), (dE3, acc3, x3)):
        assert (acc1 == acck).all() and np.abs(dE1 - dEk).max() < 1e-9
        assert max(np.abs(a - b).max() for a, b in zip(x1, xk)) < 1e-12
    r, q = systems[0]
    t = O.OracleTree(r, q, params[0].L, n_per_leaf=40, digits=6)
    t.form_pw()
    dEo, acco, _ = t.mc_chain(idx[0][:200], d[0][:200], u[0][:200], 1.0)
    assert (acc1[0][:200] == acco).all() and np.abs(dE1[0][:200] - dEo).max() < 1e-9
    # energy of the final state: incremental device state vs a fresh build
    for k in range(len(systems)):
        g = single(params[k], x2[k], systems[k][1], P.HPDMK_FP64)
        assert abs(g.eval_energy() - E[k]) < 1e-9 * max(1.0, abs(E[k]))
        g.close()
