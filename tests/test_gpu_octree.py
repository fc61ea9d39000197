"""The octree built on the GPU (hpdmk_plan_tree_device, the build
hpdmk_tree_create uses) against the host build (hpdmk_plan_tree, itself
checked node for node against the oracle's tree in tests/test_library.py).

Bar: identical, node for node in Morton pre-order (level, leaf flag, parent,
cell, particles in the subtree, existing colleagues, near-list length) and
particle for particle (leaf of every particle), for uniform systems from 1e3
to 1e6 particles (the test systems and the benchmark workloads), a clustered
system whose refinement needs many 2:1 balance splits, ragged replicas, and
the leaf caps the tests and benchmarks use; and the energy of a system built
with the device tree equals the oracle's."""
import sys

import numpy as np
import pytest

import oracle_lib as O
import pdmk4mc_amd as P

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_tree(r, L, leaf):
    params = P.HPDMKParams(leaf, 3, L, init=P.DIRECT)
    nh, lh = P.plan_tree(params, r)
    nd, ld, ms = P.plan_tree_device(params, r)
    assert nd.shape == nh.shape, (nd.shape, nh.shape)
    assert np.array_equal(nd, nh)
    assert np.array_equal(ld, lh)
    return nh, ms


def clustered(n, L, seed):
    """Half the particles in three tight Gaussian clumps, half uniform: deep,
    uneven refinement next to shallow regions (the balance's work)."""
    rng = np.random.default_rng(seed)
    m = n // 2
    centres = rng.uniform(0, L, (3, 3))
    r = [np.mod(centres[k] + rng.normal(0, 0.02 * L, (m // 3, 3)), L) for k in range(3)]
    r.append(rng.uniform(0, L, (n - 3 * (m // 3), 3)))
    return np.clip(np.concatenate(r), 0.0, np.nextafter(L, 0.0))


@pytest.mark.parametrize("n,L,leaf", [(1000, 20.0, 64), (3000, 20.0, 20), (20000, 20.0, 64), (100000, 100.0, 200),
                                      (1000000, 215.443469, 200), (100000, 100.0, 8)])
def test_uniform_systems(n, L, leaf):
    log("device octree, uniform", n, L, leaf)
    r = np.random.default_rng(n + leaf).uniform(0, L, (n, 3))
    nodes, ms = same_tree(r, L, leaf)
    log(f"  {len(nodes)} nodes, {nodes[:, 0].max() + 1} levels, device build {ms:.1f} ms")


@pytest.mark.parametrize("n,leaf", [(20000, 16), (200000, 64)])
def test_clustered_systems(n, leaf):
    log("device octree, clustered", n, leaf)
    L = 50.0
    r = clustered(n, L, seed=n)
    nodes, ms = same_tree(r, L, leaf)
    depth = nodes[:, 0].max() + 1
    assert depth >= 7  # deep refinement beside shallow boxes
    log(f"  {len(nodes)} nodes, {depth} levels, device build {ms:.1f} ms")


def test_reference_test_system_and_replicas():
    log("device octree, reference test systems and ragged replicas")
    r, q = O.reference_test_system(2000, 20.0)
    same_tree(r, 20.0, 64)
    rng = np.random.default_rng(5)
    for _ in range(8):
        n_r = int(rng.integers(2000, 20001))
        L_r = (n_r / 0.1) ** (1.0 / 3.0)
        same_tree(rng.uniform(0, L_r, (n_r, 3)), L_r, 200)


def test_shallow_tree_refused_like_the_host():
    r = np.random.default_rng(1).uniform(0, 10.0, (50, 3))
    params = P.HPDMKParams(64, 3, 10.0, init=P.DIRECT)
    with pytest.raises(P.HPDMKError) as e1:
        P.plan_tree(params, r)
    with pytest.raises(P.HPDMKError) as e2:
        P.plan_tree_device(params, r)
    assert e1.value.code == e2.value.code == -2
    bad = r.copy()
    bad[3, 1] = 10.0  # outside [0, L)
    with pytest.raises(P.HPDMKError) as e3:
        P.plan_tree_device(P.HPDMKParams(8, 3, 10.0, init=P.DIRECT), bad)
    assert e3.value.code == -1


def test_energy_with_device_tree_matches_oracle():
    log("energy of a system built on the device tree")
    L = 30.0
    r = clustered(6000, L, seed=3)
    q = np.where(np.arange(6000) % 2 == 0, 1.0, -1.0)
    g = P.HPDMKPtTree(P.HPDMKParams(32, 6, L, init=P.DIRECT), r, q)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    t = O.OracleTree(r, q, L, n_per_leaf=32, digits=6)
    t.form_pw()
    assert g.n_nodes == t.n_nodes
    Eo = t.eval_energy().sum()
    assert abs(g.eval_energy() - Eo) < 1e-10 * abs(Eo)
    g.close()
