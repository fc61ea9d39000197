"""The mixed-precision DIRECT build of the outgoing expansions (FP32 MFMA
contractions with FP64 sums) against the all-FP64 build of the same system.

Bars (written in each test): every expansion element within 2e-6 of the
largest element of its box (FP32 phases and products, FP64 sums), and the
total energy within 1e-6 relative, the digits-6 tolerance (measured: 2e-8 to
4e-7; the dense L = 20 systems lose digits to the cancellation between the
window sum and its self term).  The FP64 build itself is pinned to the
oracle to 1e-10 (test_gpu_parity.py).  A rebuild after accepted moves (holes
and appended particles in the leaf slots) is checked the same way.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
import pdmk4mc_amd as P

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(n, L, leaf, digits, seed=5):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, L, (n, 3))
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    gs = []
    for prec in (P.HPDMK_FP64, P.HPDMK_MIXED):
        g = P.HPDMKPtTree(P.HPDMKParams(leaf, digits, L, init=P.DIRECT), r, q, prec)
        g.form_outgoing_pw()
        g.form_incoming_pw()
        gs.append(g)
    return r, q, gs[0], gs[1]


def compare(g64, g32, nodes):
    D = g64.max_depth
    a, b = g64.get_expansion(0), g32.get_expansion(0)
    worst = np.abs(a - b).max() / np.abs(a).max()
    for nd in range(g64.n_nodes):
        if 2 <= nodes[nd, 0] <= D - 2:
            a, b = g64.get_expansion(1, nd), g32.get_expansion(1, nd)
            worst = max(worst, np.abs(a - b).max() / max(np.abs(a).max(), 1.0))
    return worst


@pytest.mark.parametrize("n,L,leaf,digits", [(20000, 20.0, 64, 3), (20000, 20.0, 64, 6), (100000, 100.0, 200, 6)])
def test_mfma_build_matches_fp64(n, L, leaf, digits):
    log("mfma build", n, L, leaf, digits)
    r, q, g64, g32 = pair(n, L, leaf, digits)
    nodes = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=digits).nodes()
    worst = compare(g64, g32, nodes)
    E64, E32 = g64.eval_energy(), g32.eval_energy()
    p64, p32 = g64.energy_parts(), g32.energy_parts()
    log(f"  expansions rel {worst:.3e}, energy rel {abs(E32 - E64) / abs(E64):.3e}, parts64 {p64}, "
        f"part errors rel to E {(p32 - p64) / abs(E64)}")
    assert worst < 2e-6, worst
    assert abs(E32 - E64) < 1e-6 * abs(E64), (E32, E64)


def test_mfma_rebuild_after_moves():
    log("mfma rebuild after moves")
    n, L, leaf = 20000, 20.0, 64
    r, q, g64, g32 = pair(n, L, leaf, 6)
    rng = np.random.default_rng(11)
    for _ in range(300):  # forced updates: identical positions in both systems
        i = int(rng.integers(0, n))
        d = rng.uniform(-2.0, 2.0, 3)
        g64.update_shift(i, *d)
        g32.update_shift(i, *d)
    assert np.abs(g64.positions() - g32.positions()).max() < 1e-12
    for g in (g64, g32):
        g.form_outgoing_pw()
        g.form_incoming_pw()
    nodes = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=6).nodes()
    assert compare(g64, g32, nodes) < 2e-6
    E64, E32 = g64.eval_energy(), g32.eval_energy()
    assert abs(E32 - E64) < 1e-6 * abs(E64)
