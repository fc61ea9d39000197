"""Pins the oracle's PSWF constants (CPU only).

* residual polynomial vs 1 - int_0^x psi0 / C0 (test/pswf_test.cpp:14-35);
* lambda from the 200-point quadrature at x = 0.5 (src/pswf.cpp:1208-1219)
  against the closed form lambda = 2 a_0 / psi0(0) of the integral equation at x = 0;
* psi0 shape against scipy's prolate angular function S_00(c, x) (normalisation
  cancels in every ratio the method uses);
* the direct_eval contract (test/direct_eval_test.cpp:55-111).
"""
import numpy as np
import pytest

import oracle_lib as O


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_residual_poly_matches_prolate_integral(digits):
    p = O.pswf_params(digits)
    for x in np.arange(0.01, 1.0, 0.01):
        val = 1 - O.prolate0_int(p["c"], x) / p["C0"]
        assert abs(val - O.residual_poly(digits, x)) < 10.0 ** (-digits)


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_lambda_two_routes(digits):
    p = O.pswf_params(digits)
    a = O.prolate0_coefs(p["c"])
    lam_closed = 2 * a[0] / O.prolate0_eval(p["c"], 0.0)
    assert abs(p["lam"] - lam_closed) < 1e-13


@pytest.mark.parametrize("digits", [3, 6, 9])
def test_psi0_shape_against_scipy(digits):
    from scipy.special import pro_ang1
    c = O.pswf_params(digits)["c"]
    xs = np.linspace(-0.95, 0.95, 9)
    ratio = np.array([O.prolate0_eval(c, x) / pro_ang1(0, 0, c, x)[0] for x in xs])
    assert np.ptp(ratio) / abs(ratio.mean()) < 1e-9


def test_psi0_zero_outside_unit_interval():
    # src/pswf.cpp:980-985: psi0 = 0 for |x| > 1 when c < 45
    c = O.pswf_params(6)["c"]
    assert O.prolate0_eval(c, 1.0001) == 0.0 and O.prolate0_eval(c, -3.0) == 0.0


def test_window_sizes_table():
    # n_window = floor(2c/pi) (src/tree.cpp:58) and n_diff (include/utils.hpp:20-45)
    expect = {3: (4, 6), 6: (8, 12), 9: (13, 19), 12: (17, 26)}
    for d, (nw, nd) in expect.items():
        p = O.pswf_params(d)
        assert (p["n_window"], p["n_diff"]) == (nw, nd)


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
@pytest.mark.parametrize("n_trg", [10, 99, 100])
def test_direct_eval_contract(digits, n_trg):
    rng = np.random.default_rng(digits * 1000 + n_trg)
    r_src = np.array([0.1, 0.2, 0.3])
    r_trg = rng.uniform(-1, 1, (n_trg, 3))
    q_trg = rng.uniform(-1, 1, n_trg)
    cutoff = 0.8
    p = O.pswf_params(digits)
    ref = 0.0
    for t in range(n_trg):
        rr = np.linalg.norm(r_trg[t] - r_src)
        if 0 < rr < cutoff:
            ref += (1 - O.prolate0_int(p["c"], rr / cutoff) / p["C0"]) / rr * q_trg[t] * 0.5
    got = O.direct_eval(r_src, 0.5, r_trg, q_trg, cutoff, digits)
    assert abs(got - ref) < 10.0 ** (-digits)
