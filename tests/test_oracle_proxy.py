"""Pins the oracle's restatement of the PROXY initialisation
(src/upward_pass.cpp, include/chebychev.hpp), on the CPU.

* the Vandermonde inverse against the closed form of Chebyshev discrete
  orthogonality, V^-1[a, i] = (2 - delta_a0) / p T_a(x_i);
* the child-to-parent matrices against the polynomial identity
  T_l(x/2 -+ 1/2) = sum_a A[l, a] T_a(x) at random points (both halves);
* the coefficient-to-plane-wave matrix against e^{-i k (pi/3) x} at random x
  (Chebyshev interpolation error of the order used);
* transferred coefficients of every non-leaf box equal the coefficients its
  own particles give (the transfer is exact), so child order and axes are
  right;
* PlanewaveProxy (test/hpdmk_test.cpp:17-127, 547-549): the PROXY expansions
  against brute-force sums, root to 1e-3 relative and boxes to 1e-3
  absolute, plus agreement with DIRECT;
* the digits-6 and digits-9 extensions (orders 28 and 44): expansions and
  energies agree with DIRECT.
"""
import numpy as np
import pytest

import oracle_lib as O

L = 20.0


def cheb_T(p, x):
    x = np.asarray(x, dtype=float)
    T = np.zeros((p,) + x.shape)
    T[0] = 1.0
    if p > 1:
        T[1] = x
    for k in range(2, p):
        T[k] = 2 * x * T[k - 1] - T[k - 2]
    return T


@pytest.mark.parametrize("p", [18, 28, 44])
def test_vandermonde_inverse_closed_form(p):
    x = np.cos(np.pi * ((p - np.arange(p) - 1) + 0.5) / p)
    closed = (2.0 - (np.arange(p) == 0)) [:, None] / p * cheb_T(p, x)  # [a, i]
    Vi = O.cheb_matrix(0, p)
    assert np.abs(Vi - closed).max() < 1e-12


@pytest.mark.parametrize("p", [18, 28, 44])
def test_child_to_parent_identity(p):
    x = np.random.default_rng(1).uniform(-1, 1, 50)
    Tx = cheb_T(p, x)
    for half, shift in ((0, -0.5), (1, 0.5)):
        A = O.cheb_matrix(1 + half, p)
        lhs = cheb_T(p, 0.5 * x + shift)
        assert np.abs(lhs - A @ Tx).max() < 1e-9 * p


@pytest.mark.parametrize("p,n,tol", [(18, 6, 1e-6), (28, 12, 1e-6), (44, 19, 1e-10)])
def test_poly2pw_interpolates_plane_waves(p, n, tol):
    P = O.cheb_matrix(3, p, n)
    x = np.random.default_rng(2).uniform(-1, 1, 400)
    k = np.arange(-n, n + 1)
    exact = np.exp(-1j * np.outer(k, x) * np.pi / 3)
    assert np.abs(P @ cheb_T(p, x) - exact).max() < tol


def test_proxy_order_table():
    assert O.proxy_order(3) == 18  # src/upward_pass.cpp:9-14
    assert O.proxy_order(6) == 28
    assert O.proxy_order(9) == 44
    with pytest.raises(RuntimeError):
        O.proxy_order(12)


@pytest.mark.parametrize("n,leaf,digits", [(1000, 5, 3), (3000, 6, 6), (1000, 5, 9)])
def test_transfer_equals_direct_coefficients(n, leaf, digits):
    r, q = O.reference_test_system(n, L)
    t = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=digits)
    t.form_pw_proxy()
    nodes = t.nodes()
    D = t.max_depth
    checked = 0
    for nd in range(t.n_nodes):
        lv, isleaf, npts = nodes[nd, 0], nodes[nd, 1], nodes[nd, 6]
        if lv < 2 or lv > D - 3 or isleaf or npts == 0:
            continue
        a = t.proxy_coeffs(nd)
        b = t.proxy_coeffs(nd, from_particles=True)
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
        checked += 1
    assert checked > 0


def _brute(pos, q, nodes, t, nd, k):
    lv = nodes[nd, 0]
    cell = nodes[nd, 3:6]
    inside = np.all(np.floor(pos / L * (1 << lv)).astype(int) == cell, axis=1)
    return np.sum(q[inside] * np.exp(-1j * (pos[inside] @ k)))


def test_planewave_proxy_contract():
    """test/hpdmk_test.cpp:17-127 with init = PROXY (n_per_leaf 5, digits 3,
    L 20, 1000 charges): root within 1e-3 relative, boxes within 1e-3."""
    r, q = O.reference_test_system(1000, L)
    t = O.OracleTree(r, q, L, n_per_leaf=5, digits=3)
    t.form_pw_proxy()
    pos = t.positions()
    nodes = t.nodes()
    nw, nd_ = t.n_window, t.n_diff
    dw, dd = 2 * nw + 1, 2 * nd_ + 1
    rng = np.random.default_rng(3)
    root = t.get_pw(0)
    dk0 = 2 * np.pi / L
    for _ in range(10):
        ix, iy, iz = rng.integers(0, dw), rng.integers(0, dw), rng.integers(0, nw + 1)
        kv = np.array([ix - nw, iy - nw, iz - nw]) * dk0
        ref = np.sum(q * np.exp(-1j * (r @ kv)))
        got = root[ix + iy * dw + iz * dw * dw]
        assert abs(got.real / ref.real - 1) < 1e-3 and abs(got.imag / ref.imag - 1) < 1e-3
    D = t.max_depth
    nboxes = 0
    for nd in range(t.n_nodes):
        lv = nodes[nd, 0]
        if lv < 2 or lv > D - 2 or nodes[nd, 6] == 0:
            continue
        pw = t.get_pw(1, nd)
        dk = (2 * np.pi / 3) / (L / 2 ** lv)
        for _ in range(3):
            ix, iy, iz = rng.integers(0, dd), rng.integers(0, dd), rng.integers(0, nd_ + 1)
            kv = np.array([ix - nd_, iy - nd_, iz - nd_]) * dk
            ref = _brute(pos, q, nodes, t, nd, kv)
            assert abs(pw[ix + iy * dd + iz * dd * dd] - ref) < 1e-3
        nboxes += 1
    assert nboxes > 0


@pytest.mark.parametrize("n,leaf,digits,tol", [(1000, 5, 3, 1e-5), (3000, 6, 6, 1e-7), (1000, 5, 9, 1e-10)])
def test_proxy_matches_direct(n, leaf, digits, tol):
    """PROXY expansions against DIRECT ones (relative to the box's total |q|),
    and the energies of the two builds."""
    r, q = O.reference_test_system(n, L)
    td = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=digits)
    td.form_pw()
    tp = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=digits)
    tp.form_pw_proxy()
    nodes = td.nodes()
    D = td.max_depth
    qabs = np.abs(q).sum()
    worst = 0.0
    for nd in range(td.n_nodes):
        if 2 <= nodes[nd, 0] <= D - 2 and nodes[nd, 6] > 0:
            a, b = td.get_pw(1, nd), tp.get_pw(1, nd)
            worst = max(worst, np.abs(a - b).max())
    assert worst < tol * qabs
    Ed, Ep = td.eval_energy(), tp.eval_energy()
    assert abs(Ed.sum() - Ep.sum()) < 10 * tol * abs(Ed.sum())
