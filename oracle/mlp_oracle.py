"""CPU oracle -- TEST INFRASTRUCTURE ONLY -- of the learned proposal (no
reference counterpart; the energy it is combined with is the DMK oracle's).

FP64 numpy restatement of the neighbourhood features, the MLP forward pass,
the weighted Gaussian NLL with its gradient (hand-written backpropagation),
and the Metropolis-Hastings chain with the learned proposal.  Definitions:

features of point x for particle i, over j != i with minimum image v = x - x_j,
0 < r = |v| < rf:
    s1 = sum q_j v / (r^2 + 0.25)^{3/2},  s2 = number of neighbours,
    s3 = sum (1 - r/rf)^2 v / r,
    phi = [q_i, tanh(q_i s1), tanh(s2 / 8), tanh(s3 / 2)]
MLP: h1 = tanh(phi W1 + b1), h2 = tanh(h1 W2 + b2), h3 = tanh(h2 W3 + b3),
    o = h3 W4 + b4,  mu = M tanh(o[:3]), sigma = smin + (smax - smin) sigmoid(o[3:])
(forward: FP64; forward_bf16: the library's deployed inference, BF16
operands -- features, weights, hidden activations rounded to nearest even --
with FP32 accumulation, FP32 biases and tanh; the products of BF16 values are
exact, so the only freedom left is the FP32 rounding of the sums)
NLL(delta) = sum_a log sigma_a + (delta_a - mu_a)^2 / (2 sigma_a^2)
log q(delta) = -NLL(delta) - 1.5 log(2 pi)
H = log q(-delta | phi(t)) - log q(delta | phi(o)); accept iff no overlap and
    lB dE - H / beta < -ln(u) / beta.
"""
import numpy as np

N_FEAT, N_HID, N_OUT = 8, 64, 6
SHAPES = [(N_FEAT, N_HID), (N_HID,), (N_HID, N_HID), (N_HID,), (N_HID, N_HID), (N_HID,), (N_HID, N_OUT), (N_OUT,)]
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
    return np.concatenate([s1, [float(len(rr))], (wgt[:, None] * v / rr[:, None]).sum(0)])


def features(r, q, L, i, x, rf):
    s = feature_sums(r, q, L, i, x, rf)
    qi = q[i]
    return np.concatenate([[qi], np.tanh(qi * s[:3]), [np.tanh(s[3] / 8.0)], np.tanh(s[4:] / 2.0)])


def forward(w, phi, max_mean, smin, smax):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(w)
    phi = np.atleast_2d(phi)
    h1 = np.tanh(phi @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    h3 = np.tanh(h2 @ W3 + b3)
    o = h3 @ W4 + b4
    mu = max_mean * np.tanh(o[:, :3])
    sig = smin + (smax - smin) / (1 + np.exp(-o[:, 3:]))
    return mu, sig, (phi, h1, h2, h3, o)


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    u = ((u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)).astype(np.uint32)
    return u.view(np.float32)


def forward_bf16(w, phi, max_mean, smin, smax):
    """The deployed inference: BF16 operands, sums rounded to FP32, FP32
    bias and tanh; mu and sigma formed in FP64 from the FP32 outputs."""
    Ws = unpack(w)
    x = bf16(np.atleast_2d(phi).astype(np.float32))
    for k in range(4):
        W = bf16(Ws[2 * k].astype(np.float32)).astype(np.float64)
        b = Ws[2 * k + 1].astype(np.float32)
        acc = (x.astype(np.float64) @ W).astype(np.float32) + b
        x = bf16(np.tanh(acc)) if k < 3 else acc
    o = x.astype(np.float64)
    mu = max_mean * np.tanh(o[:, :3])
    sig = smin + (smax - smin) / (1 + np.exp(-o[:, 3:]))
    return mu, sig


def gauss_logq(delta, mu, sig):
    y = (delta - mu) / sig
    return -(np.log(sig) + 0.5 * y * y).sum(-1) - 1.5 * np.log(2 * np.pi)


def nll_and_grad(w, phi, delta, wt, max_mean, smin, smax):
    """sum_s wt_s NLL_s and its gradient with respect to the flat weights."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(w)
    mu, sig, (phi, h1, h2, h3, o) = forward(w, phi, max_mean, smin, smax)
    wt = np.asarray(wt, dtype=np.float64)[:, None]
    y = (delta - mu) / sig
    loss = float((wt * (np.log(sig) + 0.5 * y * y)).sum())
    dmu = -wt * y / sig
    dsig = wt * (1 - y * y) / sig
    th = np.tanh(o[:, :3])
    sg = 1 / (1 + np.exp(-o[:, 3:]))
    do = np.concatenate([dmu * max_mean * (1 - th * th), dsig * (smax - smin) * sg * (1 - sg)], axis=1)
    gW4, gb4 = h3.T @ do, do.sum(0)
    g3 = (do @ W4.T) * (1 - h3 * h3)
    gW3, gb3 = h2.T @ g3, g3.sum(0)
    g2 = (g3 @ W3.T) * (1 - h2 * h2)
    gW2, gb2 = h1.T @ g2, g2.sum(0)
    g1 = (g2 @ W2.T) * (1 - h1 * h1)
    gW1, gb1 = phi.T @ g1, g1.sum(0)
    return loss, np.concatenate([g.reshape(-1) for g in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])


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
        mu, sg = forward_bf16(w, features(pos, q, L, i, o, rf), max_mean, smin, smax)
        delta = mu[0] + sg[0] * z[k]
        t = np.mod(o + delta, L)
        mu2, sg2 = forward_bf16(w, features(pos, q, L, i, t, rf), max_mean, smin, smax)
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


def replay_learned_chain(tree, r, q, L, idx, z, delta_dev, acc_dev, beta, w, rf, max_mean, smin, smax, bjerrum=1.0,
                         sigma=0.0):
    """Replays a learned chain run elsewhere (the GPU's): at every move the
    oracle forms the features, the proposal (BF16 inference emulation) and
    the DMK energy change at the displacement the GPU used, from the state
    the GPU's earlier decisions left, and applies the GPU's decision.
    Returns (delta, H, dE reported at the GPU's displacement, features at the
    target) as the oracle computes them; the caller checks the GPU's
    decisions against them."""
    pos = np.array(r, dtype=np.float64, copy=True)
    n = len(idx)
    dl, Hs, dE, ftt = np.zeros((n, 3)), np.zeros(n), np.zeros(n), np.zeros((n, N_FEAT))
    for k in range(n):
        i = int(idx[k])
        o = pos[i].copy()
        mu, sg = forward_bf16(w, features(pos, q, L, i, o, rf), max_mean, smin, smax)
        dl[k] = mu[0] + sg[0] * z[k]
        d = np.asarray(delta_dev[k], dtype=np.float64)
        t = np.mod(o + d, L)
        ftt[k] = features(pos, q, L, i, t, rf)
        mu2, sg2 = forward_bf16(w, ftt[k], max_mean, smin, smax)
        Hs[k] = gauss_logq(-d, mu2[0], sg2[0]) - gauss_logq(d, mu[0], sg[0])
        overlap = False
        if sigma > 0:
            v = t[None, :] - pos
            v -= L * np.rint(v / L)
            d2 = (v * v).sum(1)
            d2[i] = np.inf
            overlap = bool((d2 < sigma * sigma).any())
        e = tree.eval_shift_energy(i, d)
        dE[k] = np.inf if overlap else bjerrum * e
        if acc_dev[k]:
            tree.update_shift(i, d)
            pos[i] = t
    return dl, Hs, dE, ftt
