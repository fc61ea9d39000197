"""CPU oracle -- TEST INFRASTRUCTURE ONLY -- of the learned proposal (no
reference counterpart; the energy it is combined with is the DMK oracle's).

FP64 numpy restatement of the neighbourhood features, the MLP forward pass,
the weighted Gaussian NLL with its gradient (hand-written backpropagation),
and the Metropolis-Hastings chain with the learned proposal.  Definitions:

features of point x for particle i, over j != i with minimum image v = x - x_j,
0 < r = |v| < rf:
    s1 = sum q_j v / (r^2 + 0.25)^{3/2},  s2 = sum (1 - r/rf)^2,
    s3 = sum (1 - r/rf)^2 v / r,
    phi = [q_i, tanh(q_i s1), tanh(s2 / 4), tanh(s3 / 2)]
MLP: h1 = tanh(phi W1 + b1), h2 = tanh(h1 W2 + b2), o = h2 W3 + b3,
    mu = M tanh(o[:3]), sigma = smin + (smax - smin) sigmoid(o[3:])
NLL(delta) = sum_a log sigma_a + (delta_a - mu_a)^2 / (2 sigma_a^2)
log q(delta) = -NLL(delta) - 1.5 log(2 pi)
H = log q(-delta | phi(t)) - log q(delta | phi(o)); accept iff no overlap and
    lB dE - H / beta < -ln(u) / beta.
"""
import numpy as np

N_FEAT, N_HID, N_OUT = 8, 32, 6
SHAPES = [(N_FEAT, N_HID), (N_HID,), (N_HID, N_HID), (N_HID,), (N_HID, N_OUT), (N_OUT,)]
N_PARAMS = sum(int(np.prod(s)) for s in SHAPES)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    out, k = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(w[k:k + n].reshape(s))
        k += n
    return out


def feature_sums(r, q, L, i, x, rf):
    v = x[None, :] - r
    v -= L * np.rint(v / L)
    r2 = (v * v).sum(1)
    keep = (r2 > 0) & (r2 < rf * rf)
    keep[i] = False
    v, r2, qj = v[keep], r2[keep], q[keep]
    rr = np.sqrt(r2)
    s1 = (qj[:, None] * v / ((r2 + 0.25) ** 1.5)[:, None]).sum(0)
    wgt = (1 - rr / rf) ** 2
    return np.concatenate([s1, [wgt.sum()], (wgt[:, None] * v / rr[:, None]).sum(0)])


def features(r, q, L, i, x, rf):
    s = feature_sums(r, q, L, i, x, rf)
    qi = q[i]
    return np.concatenate([[qi], np.tanh(qi * s[:3]), [np.tanh(s[3] / 4.0)], np.tanh(s[4:] / 2.0)])


def forward(w, phi, max_mean, smin, smax):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    phi = np.atleast_2d(phi)
    h1 = np.tanh(phi @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    o = h2 @ W3 + b3
    mu = max_mean * np.tanh(o[:, :3])
    sig = smin + (smax - smin) / (1 + np.exp(-o[:, 3:]))
    return mu, sig, (phi, h1, h2, o)


def gauss_logq(delta, mu, sig):
    y = (delta - mu) / sig
    return -(np.log(sig) + 0.5 * y * y).sum(-1) - 1.5 * np.log(2 * np.pi)


def nll_and_grad(w, phi, delta, wt, max_mean, smin, smax):
    """sum_s wt_s NLL_s and its gradient with respect to the flat weights."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    mu, sig, (phi, h1, h2, o) = forward(w, phi, max_mean, smin, smax)
    wt = np.asarray(wt, dtype=np.float64)[:, None]
    y = (delta - mu) / sig
    loss = float((wt * (np.log(sig) + 0.5 * y * y)).sum())
    dmu = -wt * y / sig
    dsig = wt * (1 - y * y) / sig
    th = np.tanh(o[:, :3])
    sg = 1 / (1 + np.exp(-o[:, 3:]))
    do = np.concatenate([dmu * max_mean * (1 - th * th), dsig * (smax - smin) * sg * (1 - sg)], axis=1)
    gW3, gb3 = h2.T @ do, do.sum(0)
    g2 = (do @ W3.T) * (1 - h2 * h2)
    gW2, gb2 = h1.T @ g2, g2.sum(0)
    g1 = (g2 @ W2.T) * (1 - h1 * h1)
    gW1, gb1 = phi.T @ g1, g1.sum(0)
    return loss, np.concatenate([g.reshape(-1) for g in (gW1, gb1, gW2, gb2, gW3, gb3)])


def learned_chain(tree, r, q, L, idx, z, u, beta, w, rf, max_mean, smin, smax, bjerrum=1.0, sigma=0.0):
    """Plain Metropolis-Hastings chain with the learned proposal on the DMK
    oracle tree `tree` (eval_shift_energy / update_shift); returns
    (dE reported, accept, delta, H)."""
    pos = np.array(r, dtype=np.float64, copy=True)
    n = len(idx)
    dE, acc, dl, Hs = np.zeros(n), np.zeros(n, dtype=np.int32), np.zeros((n, 3)), np.zeros(n)
    for k in range(n):
        i = int(idx[k])
        o = pos[i].copy()
        mu, sg, _ = forward(w, features(pos, q, L, i, o, rf), max_mean, smin, smax)
        delta = mu[0] + sg[0] * z[k]
        t = np.mod(o + delta, L)
        mu2, sg2, _ = forward(w, features(pos, q, L, i, t, rf), max_mean, smin, smax)
        H = gauss_logq(-delta, mu2[0], sg2[0]) - gauss_logq(delta, mu[0], sg[0])
        overlap = False
        if sigma > 0:
            v = t[None, :] - pos
            v -= L * np.rint(v / L)
            d2 = (v * v).sum(1)
            d2[i] = np.inf
            overlap = bool((d2 < sigma * sigma).any())
        e = tree.eval_shift_energy(i, delta)
        dE[k] = np.inf if overlap else bjerrum * e
        ok = (not overlap) and (bjerrum * e - H / beta < -np.log(u[k]) / beta)
        acc[k], dl[k], Hs[k] = ok, delta, H
        if ok:
            tree.update_shift(i, delta)
            pos[i] = t
    return dE, acc, dl, Hs
