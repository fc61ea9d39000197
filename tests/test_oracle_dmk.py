"""Pins the oracle's DMK restatement with the reference's own contracts
(test/hpdmk_test.cpp), run on the CPU.

* tree invariants (SCTL PtTree semantics; topology identity is unpinned);
* PlanewaveDirect / PlanewaveSingle: expansions vs brute-force sums
  (test/hpdmk_test.cpp:17-315);
* Energy: plane-wave energies vs the real-space direct sums of
  src/energy_direct.cpp, relative to E_Ewald (test/hpdmk_test.cpp:317-372);
* ShiftEnergy: dE vs Ewald dE relative to E_Ewald (:374-467);
* Update: chained eval+update vs Ewald, absolute (:470-533).
"""
import numpy as np
import pytest

import oracle_lib as O

L = 20.0


def _adjacent(ca, la, cb, lb, m):
    """periodic touching test of two boxes given integer cells at levels la, lb (finest level m)"""
    n = 1 << m
    for a in range(3):
        lo_a, hi_a = ca[a] << (m - la), (ca[a] + 1) << (m - la)
        lo_b, hi_b = cb[a] << (m - lb), (cb[a] + 1) << (m - lb)
        ok = False
        for s in (-n, 0, n):
            if lo_b + s <= hi_a and lo_a <= hi_b + s:
                ok = True
        if not ok:
            return False
    return True


@pytest.mark.parametrize("n,leaf", [(1000, 10), (1000, 64), (1100, 5)])
def test_tree_invariants(n, leaf):
    r, q = O.reference_test_system(n, L)
    t = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=3)
    nodes = t.nodes()  # level, leaf, parent, cx, cy, cz, npts, ncol, ncg
    leaves = np.where(nodes[:, 1] == 1)[0]
    assert nodes[leaves, 6].sum() == n
    assert (nodes[leaves, 6] <= leaf).all()
    nonleaf = (nodes[:, 1] == 0) & (nodes[:, 0] >= 2)
    assert (nodes[nonleaf, 7] == 26).all()
    # every level-0/1 box refined, leaves only from level 2 on
    assert (nodes[leaves, 0] >= 2).all()
    # particle -> leaf consistent with positions
    pl = t.particle_leaf()
    pos = t.positions()
    for i in range(0, n, 37):
        lv = nodes[pl[i], 0]
        cell = np.floor(pos[i] / L * (1 << lv)).astype(int)
        assert (cell == nodes[pl[i], 3:6]).all()
    # 2:1 balance over the 26 neighbours (periodic)
    m = nodes[:, 0].max()
    for ia in leaves:
        for ib in leaves:
            if ib <= ia or abs(int(nodes[ia, 0]) - int(nodes[ib, 0])) <= 1:
                continue
            assert not _adjacent(nodes[ia, 3:6], nodes[ia, 0], nodes[ib, 3:6], nodes[ib, 0], m)


def test_tree_too_shallow_raises():
    r, q = O.reference_test_system(50, L)
    with pytest.raises(RuntimeError, match="max depth"):
        O.OracleTree(r, q, L, n_per_leaf=200, digits=3)


def test_planewave_direct():
    n = 1000
    r, q = O.reference_test_system(n, L)
    t = O.OracleTree(r, q, L, n_per_leaf=5, digits=3)
    t.form_pw()
    nw = t.n_window
    dw = 2 * nw + 1
    dk0 = 2 * np.pi / L
    root = t.get_pw(0)
    rng = np.random.default_rng(1)
    for _ in range(10):
        ix, iy, iz = rng.integers(0, dw), rng.integers(0, dw), rng.integers(0, nw + 1)
        k = np.array([ix - nw, iy - nw, iz - nw]) * dk0
        direct = np.sum(q * np.exp(-1j * (r @ k)))
        assert abs(root[ix + iy * dw + iz * dw * dw] - direct) < 1e-10 * max(1, abs(direct))
    nodes = t.nodes()
    pos = t.positions()
    pl = t.particle_leaf()
    nd = t.n_diff
    d = 2 * nd + 1
    for l in range(2, t.max_depth - 1):
        dk = (2 * np.pi / 3) / (L / 2 ** l)
        ids = [i for i in np.where(nodes[:, 0] == l)[0] if nodes[i, 6] > 0]
        node = ids[0]
        cell = nodes[node, 3:6]
        members = [i for i in range(n) if (np.floor(pos[i] / L * (1 << l)).astype(int) == cell).all()]
        out = t.get_pw(1, node)
        for _ in range(10):
            ix, iy, iz = rng.integers(0, d), rng.integers(0, d), rng.integers(0, nd + 1)
            k = np.array([ix - nd, iy - nd, iz - nd]) * dk
            direct = np.sum(q[members] * np.exp(-1j * (pos[members] @ k)))
            assert abs(out[ix + iy * d + iz * d * d] - direct) < 1e-9
        # incoming = conj(self) + colleagues' images (test/hpdmk_test.cpp:135-223)
        nonleaf = [i for i in np.where(nodes[:, 0] == l)[0] if nodes[i, 1] == 0]
        node = nonleaf[0]
        inc = t.get_pw(2, node) - np.conj(t.get_pw(1, node))
        c = (nodes[node, 3:6] + 0.5) * (L / 2 ** l)
        nb = []
        for i in range(n):
            dx = pos[i] - c
            dx -= L * np.round(dx / L)
            if (np.abs(dx) < 1.5 * L / 2 ** l).all() and not (np.abs(dx) < 0.5 * L / 2 ** l).all():
                nb.append(c + dx)
        nb_idx = None
        for _ in range(10):
            ix, iy, iz = rng.integers(0, d), rng.integers(0, d), rng.integers(0, nd + 1)
            k = np.array([ix - nd, iy - nd, iz - nd]) * dk
            # charges of the neighbour particles, matched through their wrapped positions
            direct = 0
            for i in range(n):
                dx = pos[i] - c
                dx -= L * np.round(dx / L)
                if (np.abs(dx) < 1.5 * L / 2 ** l).all() and not (np.abs(dx) < 0.5 * L / 2 ** l).all():
                    direct += q[i] * np.exp(1j * ((c + dx) @ k))
            assert abs(inc[ix + iy * d + iz * d * d] - direct) < 1e-9


def test_planewave_single():
    n = 1000
    r, q = O.reference_test_system(n, L)
    t = O.OracleTree(r, q, L, n_per_leaf=5, digits=3)
    rng = np.random.default_rng(2)
    for _ in range(5):
        x = rng.uniform(0, L, 3)
        path, pws = t.single_pw(x)
        nw = t.n_window
        dw = 2 * nw + 1
        g = np.arange(dw) - nw
        kx, ky, kz = np.meshgrid(g, g, np.arange(nw + 1) - nw, indexing="ij")
        ph = np.exp(-1j * (2 * np.pi / L) * (kx * x[0] + ky * x[1] + kz * x[2]))
        assert np.abs(pws[0] - ph.transpose(2, 1, 0).reshape(-1)).max() < 1e-12
        nd = t.n_diff
        d = 2 * nd + 1
        g = np.arange(d) - nd
        kx, ky, kz = np.meshgrid(g, g, np.arange(nd + 1) - nd, indexing="ij")
        for l in range(2, min(t.max_depth - 1, len(path))):
            dk = (2 * np.pi / 3) / (L / 2 ** l)
            ph = np.exp(-1j * dk * (kx * x[0] + ky * x[1] + kz * x[2]))
            assert np.abs(pws[l] - ph.transpose(2, 1, 0).reshape(-1)).max() < 1e-12


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_energy_decomposition(digits):
    for n in (1000, 1100):
        r, q = O.reference_test_system(n, L)
        E_ew = O.ewald_energy(r, q, L, 4.0, 1.0)
        t = O.OracleTree(r, q, L, n_per_leaf=10, digits=digits)
        t.form_pw()
        E = t.eval_energy()
        Ed = t.eval_energy_direct()
        tol = 5 * 10.0 ** (1 - digits)
        assert (np.abs(E - Ed) / abs(E_ew) < tol).all()
        assert abs(E.sum() - Ed.sum()) / abs(E_ew) < tol
        if digits <= 6:  # Ewald at s = 4 is itself only good to ~erfc(4) = 1.5e-8
            assert abs(E.sum() - E_ew) / abs(E_ew) < tol


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_shift_energy_vs_ewald(digits):
    n = 1000
    r, q = O.reference_test_system(n, L)
    idx, d = O.reference_test_moves(n, L, 10)
    E_old = O.ewald_energy(r, q, L, 5.0, 1.0)
    dE_ew = O.ewald_shift_chain(r, q, L, 5.0, 1.0, idx, d, np.zeros(10))
    t = O.OracleTree(r, q, L, n_per_leaf=10, digits=digits)
    t.form_pw()
    for k in range(10):
        dE = t.eval_shift_energy(idx[k], d[k])
        assert abs(dE - dE_ew[k]) / abs(E_old) < 5 * 10.0 ** (1 - digits)


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_update_chain_vs_ewald(digits):
    n = 1000
    r, q = O.reference_test_system(n, L)
    idx, d = O.reference_test_moves(n, L, 10)
    dE_ew = O.ewald_shift_chain(r, q, L, 5.0, 1.0, idx, d, np.ones(10))
    t = O.OracleTree(r, q, L, n_per_leaf=10, digits=digits)
    t.form_pw()
    for k in range(10):
        dE = t.eval_shift_energy(idx[k], d[k])
        t.update_shift(idx[k], d[k])
        assert abs(dE - dE_ew[k]) < 10.0 ** (1 - digits)


def test_update_matches_rebuild():
    """Incremental expansions after accepted moves equal a fresh build at the
    moved positions when the leaf topology is unchanged (exact in FP64)."""
    n = 1000
    r, q = O.reference_test_system(n, L)
    t = O.OracleTree(r, q, L, n_per_leaf=10, digits=6)
    t.form_pw()
    rng = np.random.default_rng(5)
    leaf0 = t.particle_leaf()
    moved = 0
    for _ in range(40):
        i = int(rng.integers(0, n))
        dd = rng.uniform(-0.3, 0.3, 3)
        t.eval_shift_energy(i, dd)
        t.update_shift(i, dd)
        moved += 1
    E_inc = t.eval_energy()
    pos = t.positions()
    t2 = O.OracleTree(pos, q, L, n_per_leaf=10, digits=6)
    # same topology is required for the identity; small moves may still cross
    # a leaf that overflows, in which case the rebuilt tree differs
    if t2.n_nodes == t.n_nodes:
        t2.form_pw()
        E_new = t2.eval_energy()
        assert np.abs(E_inc - E_new).max() < 1e-9
