"""The policy-gradient samples of the proposal training
(pdmk4mc_amd/proposal.py) against finite differences of the objective it
climbs, on a toy chain where the objective is a plain function of the
weights: J(w) = E_z[alpha |jump|^2] with delta = mu(phi_o) + sigma(phi_o) z
(z fixed: common random numbers), target features phi_t(delta), a smooth
energy landscape dE(delta) and alpha = min(1, exp(-dE + H)).

The weighted score sum of the samples, divided by the move count, is an
unbiased estimate of the directional derivative of J; it must agree with
the central difference of J within 4 standard errors, the standard error
being small against the derivative (the check has power), and the estimate
without the reverse-density term (plain REINFORCE on the origin density)
must miss by more than 5 standard errors (a dropped term fails).  The MLP
is the oracle's FP64 forward pass (oracle/mlp_oracle.py).
"""
import pathlib
import sys

import numpy as np
import pytest

from pdmk4mc_amd import proposal as PR

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as M  # noqa: E402

CFG = (0.5, 0.05, 1.5)  # max_mean, sigma_min, sigma_max


def init(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for s in M.SHAPES:
        parts.append(rng.normal(0, scale / np.sqrt(s[0]), s).reshape(-1) if len(s) == 2 else rng.normal(0, 0.3, s))
    return np.concatenate(parts)


class Toy:
    def __init__(self, n, seed):
        rng = np.random.default_rng(seed)
        base = np.concatenate([np.sign(rng.normal(size=(4, 1))), np.tanh(rng.normal(0, 1.0, (4, 7)))], 1)
        self.phio = base[rng.integers(0, 4, n)]
        self.z = rng.normal(size=(n, 3))
        self.A = rng.normal(0, 0.8, (3, 8))
        self.A[:, 0] = 0.0  # the charge feature does not move
        self.g = rng.normal(0, 1.0, 3)

    def phit(self, d):
        pt = np.tanh(self.phio + d @ self.A)
        pt[:, 0] = self.phio[:, 0]
        return pt

    def state(self, w):
        mu, sg, _ = M.forward(w, self.phio, *CFG)
        d = mu + sg * self.z
        lq = M.gauss_logq(d, mu, sg)
        pt = self.phit(d)
        mu2, sg2, _ = M.forward(w, pt, *CFG)
        H = M.gauss_logq(-d, mu2, sg2) - lq
        dE = d @ self.g + 0.8 * (d * d).sum(1)
        return d, pt, dE, H

    def J(self, w, objective):
        d, _, dE, H = self.state(w)
        return PR.esjd(dE, H, d, 1.0, objective)


def dlogq(w, v, feat, delta, h=1e-5):
    """per-sample directional derivative of log q(delta | feat) along v"""
    lp = M.gauss_logq(delta, *M.forward(w + h * v, feat, *CFG)[:2])
    lm = M.gauss_logq(delta, *M.forward(w - h * v, feat, *CFG)[:2])
    return (lp - lm) / (2 * h)


@pytest.mark.parametrize("objective", ["distance", "energy"])
def test_policy_gradient_matches_finite_difference(objective):
    n = 200000
    toy = Toy(n, seed=3)
    w = init(4)
    v = np.random.default_rng(5).normal(size=w.size)
    v /= np.linalg.norm(v)
    h = 1e-4
    fd = (toy.J(w + h * v, objective) - toy.J(w - h * v, objective)) / (2 * h)

    d, pt, dE, H = toy.state(w)
    feat, dl, wt, _ = PR.policy_gradient_samples(toy.phio, pt, d, dE, H, 1.0, objective)
    alpha = PR.acceptance_probability(dE, H, 1.0)
    reward = alpha * PR.jump_size(d, dE, 1.0, objective)
    wt = wt * reward.std()  # undo the unit-deviation scaling
    s = dlogq(w, v, feat, dl)
    contrib = wt[:n] * s[:n]  # per move: the origin sample, plus its target sample if it has one
    part = np.flatnonzero((alpha < 1.0) & (alpha > 0.0))
    assert len(part) == len(feat) - n
    contrib[part] += wt[n:] * s[n:]
    est, se = contrib.mean(), contrib.std() / np.sqrt(n)
    # plain REINFORCE on the origin density alone (drops the reverse-density term)
    naive = ((reward - reward.mean()) * s[:n]).mean()
    print(f"{objective}: fd {fd:.5f} estimate {est:.5f} +- {se:.5f}, without reverse term {naive:.5f}; "
          f"moves with 0 < alpha < 1: {len(part) / n:.3f}")
    assert abs(est - fd) < 4 * se
    assert se < 0.1 * abs(fd)
    assert abs(naive - fd) > 5 * se


def test_acceptance_probability_and_jumps():
    dE = np.array([-1.0, 0.5, np.inf, 2.0])
    H = np.array([0.0, 0.2, 0.0, 3.0])
    a = PR.acceptance_probability(dE, H, 2.0)
    assert np.allclose(a, [1.0, np.exp(-0.8), 0.0, np.exp(-1.0)])
    d = np.array([[1.0, 2.0, 2.0]] * 4)
    assert np.allclose(PR.jump_size(d, dE, 2.0, "distance"), 9.0)
    assert np.allclose(PR.jump_size(d, dE, 2.0, "energy"), [4.0, 1.0, 0.0, 16.0])
    with pytest.raises(ValueError):
        PR.jump_size(d, dE, 2.0, "other")
