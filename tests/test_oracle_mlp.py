"""Pins the learned-proposal oracle (oracle/mlp_oracle.py): the gradient by
finite differences, the Gaussian log-density by scipy, the features by a
two-particle closed form and by translation and axis-permutation symmetry,
the BF16 rounding by hand-made bit patterns, and the BF16 inference against
the FP64 forward pass (equal on BF16-exact inputs up to FP32 rounding, close
to BF16 precision otherwise)."""
import pathlib
import sys

import numpy as np
from scipy.stats import norm

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as M  # noqa: E402

CFG = (0.5, 0.05, 1.0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.4, M.N_PARAMS)
    phi = rng.normal(0, 0.7, (25, 8))
    delta = rng.normal(0, 0.5, (25, 3))
    wt = rng.integers(0, 2, 25).astype(float)
    loss, g = M.nll_and_grad(w, phi, delta, wt, *CFG)
    h = 1e-6
    for p in rng.choice(M.N_PARAMS, 60, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (M.nll_and_grad(wp, phi, delta, wt, *CFG)[0] - M.nll_and_grad(wm, phi, delta, wt, *CFG)[0]) / (2 * h)
        assert abs(fd - g[p]) < 1e-6 * max(1.0, abs(fd)), (p, fd, g[p])
    assert loss > 0


def test_log_density_is_the_gaussian():
    rng = np.random.default_rng(1)
    mu, sig, d = rng.normal(size=3), rng.uniform(0.1, 1, 3), rng.normal(size=3)
    assert abs(M.gauss_logq(d, mu, sig) - norm.logpdf(d, mu, sig).sum()) < 1e-12


def test_two_particle_features_closed_form():
    L, rf = 10.0, 2.0
    r = np.array([[1.0, 1.0, 1.0], [9.7, 1.0, 1.0]])  # neighbour 1.3 away across the x boundary
    q = np.array([1.0, -1.0])
    s = M.feature_sums(r, q, L, 0, r[0], rf)
    v = np.array([1.3, 0.0, 0.0])
    assert np.allclose(s[:3], -v / (1.3 ** 2 + 0.25) ** 1.5, atol=1e-14)
    assert s[3] == 1.0  # one neighbour
    assert np.allclose(s[4:], (1 - 1.3 / rf) ** 2 * np.array([1.0, 0, 0]), atol=1e-14)
    phi = M.features(r, q, L, 0, r[0], rf)
    assert phi[0] == 1.0 and np.allclose(phi[1:4], np.tanh(s[:3]))


def test_features_symmetries():
    rng = np.random.default_rng(2)
    L, rf = 12.0, 2.5
    r = rng.uniform(0, L, (300, 3))
    q = np.where(np.arange(300) % 2 == 0, 1.0, -1.0)
    x = rng.uniform(0, L, 3)
    f0 = M.features(r, q, L, 7, x, rf)
    shift = np.array([3.3, -5.1, 7.7])
    f1 = M.features(np.mod(r + shift, L), q, L, 7, np.mod(x + shift, L), rf)
    assert np.allclose(f0, f1, atol=1e-12)
    perm = [2, 0, 1]
    f2 = M.features(r[:, perm], q, L, 7, x[perm], rf)
    assert np.allclose(f2[1:4], f0[1:4][perm], atol=1e-12) and np.allclose(f2[5:8], f0[5:8][perm], atol=1e-12)
    assert abs(f2[4] - f0[4]) < 1e-12


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7 + 2.0 ** -8, -2.5, 0.0], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3 2^-9 rounds up;
    # 1 + 2^-7 + 2^-8 is a tie between 1 + 2^-7 (odd) and 1 + 2^-6: to 1 + 2^-6
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -2.5, 0.0], dtype=np.float32)
    assert np.array_equal(M.bf16(x), want)
    y = np.random.default_rng(3).normal(size=1000).astype(np.float32)
    b = M.bf16(y)
    assert np.array_equal(M.bf16(b), b)  # idempotent
    assert np.all(np.abs(b - y) <= np.abs(y) * 2.0 ** -8 * 1.0001)


def test_bf16_inference_matches_fp64_forward():
    rng = np.random.default_rng(4)
    # BF16-exact weights and features, activations kept away from rounding
    # effects by a single active path: the two forward passes agree to FP32
    w = np.zeros(M.N_PARAMS)
    W = M.unpack(np.arange(M.N_PARAMS, dtype=float))
    offs = np.cumsum([0] + [int(np.prod(sh)) for sh in M.SHAPES])
    phi = np.array([[1.0, 0.5, -0.25, 0.125, 0.75, 0.0, 0.5, -1.0]])
    w[offs[0]:offs[1]] = rng.integers(-4, 5, offs[1] - offs[0]) / 8.0
    for k in (2, 4, 6):
        w[offs[k]:offs[k + 1]] = rng.integers(-4, 5, offs[k + 1] - offs[k]) / 64.0
    mu64, sig64, _ = M.forward(w, phi, 0.5, 0.05, 1.0)
    mub, sigb = M.forward_bf16(w, phi, 0.5, 0.05, 1.0)
    assert np.allclose(mub, mu64, atol=0.02) and np.allclose(sigb, sig64, atol=0.02)
    # random weights: within BF16 precision of the FP64 pass
    w = rng.normal(0, 0.3, M.N_PARAMS)
    phis = rng.normal(0, 0.7, (50, 8))
    mu64, sig64, _ = M.forward(w, phis, 0.5, 0.05, 1.0)
    mub, sigb = M.forward_bf16(w, phis, 0.5, 0.05, 1.0)
    assert np.abs(mub - mu64).max() < 0.05 and np.abs(sigb - sig64).max() < 0.05
    assert np.abs(mub - mu64).max() > 1e-7  # the rounding is really there
