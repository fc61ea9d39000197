"""Learned proposal on the GPU against the oracle (oracle/mlp_oracle.py and the
DMK oracle).

Bars: recorded origin features equal the oracle's to 1e-12; the deployed
BF16 matrix-core inference equals the oracle's emulation of it (BF16
operands, FP32 sums) to 1e-5 on at least 90% of samples and to 0.02 on all
(the sums' FP32 rounding order is the hardware's, so a sum can fall on the
other side of a BF16 rounding boundary), and the FP64 forward pass to BF16
precision (0.05); the FP32 weighted-NLL gradient equals the FP64 oracle's to
1e-4 relative; training lowers the NLL of data with a learnable mean; the
FP64 learned chain, replayed move by move by the oracle on the GPU's own
trajectory: the oracle's displacement (from the same z) within 1e-5 on at
least 90% of moves and 0.02 on all, H within 2e-3 (0.1 where a flip moved
the proposal), dE at the GPU's displacement to 1e-9, and every decision the
oracle's unless the decision variable lies within 1e-3 of its threshold,
and the recorded target features the oracle's to 1e-12.  Policy-gradient
training (pdmk4mc_amd/proposal.py) on a hard-sphere chain raises the
expected squared jump of its warm start.
"""
import pathlib
import sys

import numpy as np
import pytest

import oracle_lib as O
from chain_check import MIXED_DE_BOUND, assert_same_chain, thresholds
import pdmk4mc_amd as P
from pdmk4mc_amd import proposal as PR
from hs_system import hard_sphere_system

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu
CFG = dict(max_mean=0.4, sigma_min=0.05, sigma_max=0.8)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_tree(r, q, L, leaf):
    g = P.HPDMKPtTree(P.HPDMKParams(leaf, 6, L, init=P.DIRECT), r, q, P.HPDMK_FP64)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    return g


@pytest.mark.parametrize("leaf,rf", [(64, 2.5), (16, 2.0)])
def test_recorded_features_match_oracle(leaf, rf):
    log("features", leaf, rf)
    r, q, L = hard_sphere_system(2000, 0.2, 1.0, seed=4)
    g = gpu_tree(r, q, L, leaf)
    g.mc_set_proposal(learned=False, rf=rf)
    rng = np.random.default_rng(5)
    idx, d, u = rng.integers(0, 2000, 120), rng.uniform(-1, 1, (120, 3)), rng.uniform(0, 1, 120)
    _, acc, _ = g.mc_chain(idx, d, u, mode=0)
    _, ft, _ = g.mc_proposal_results()
    pos = r.copy()
    for k in range(120):
        i = idx[k]
        assert np.abs(ft[k] - M.features(pos, q, L, i, pos[i], rf)).max() < 1e-12, k
        if acc[k]:
            pos[i] = np.mod(pos[i] + d[k], L)
    assert np.abs(g.positions() - pos).max() < 1e-12


def test_mlp_forward_and_gradient_match_oracle():
    rng = np.random.default_rng(6)
    w = P.mlp_init(3, scale=1.5)
    phi = np.concatenate([np.sign(rng.normal(size=(1000, 1))), np.tanh(rng.normal(0, 1.5, (1000, 7)))], 1)
    delta = rng.normal(0, 0.4, (1000, 3))
    wt = rng.integers(0, 2, 1000).astype(float)
    ms, g, loss = P.mlp_eval(w, phi, delta, wt, **CFG)
    mub, sigb = M.forward_bf16(w, phi, CFG["max_mean"], CFG["sigma_min"], CFG["sigma_max"])
    err = np.maximum(np.abs(ms[:, :3] - mub).max(1), np.abs(ms[:, 3:] - sigb).max(1))
    log(f"BF16 inference vs emulation: median {np.median(err):.2e}, max {err.max():.2e}, "
        f"beyond 1e-5: {(err > 1e-5).mean():.3f}")
    assert (err <= 1e-5).mean() >= 0.9 and err.max() < 0.02
    mu, sig, _ = M.forward(w.astype(np.float64), phi, CFG["max_mean"], CFG["sigma_min"], CFG["sigma_max"])
    assert np.abs(ms[:, :3] - mu).max() < 0.05 and np.abs(ms[:, 3:] - sig).max() < 0.05
    lo, go = M.nll_and_grad(w.astype(np.float64), phi, delta, wt, CFG["max_mean"], CFG["sigma_min"], CFG["sigma_max"])
    assert abs(loss - lo) < 1e-5 * abs(lo)
    assert np.abs(g - go).max() < 1e-4 * np.abs(go).max()


def test_training_lowers_nll():
    rng = np.random.default_rng(7)
    n = 20000
    phi = np.concatenate([np.sign(rng.normal(size=(n, 1))), np.tanh(rng.normal(0, 1.0, (n, 7)))], 1)
    delta = 0.3 * phi[:, 1:4] + 0.1 * rng.normal(size=(n, 3))
    wt = np.ones(n)
    w0 = P.mlp_init(1)
    _, _, l0 = P.mlp_eval(w0, phi, delta, wt, **CFG)
    w, losses = P.mlp_train(phi, delta, wt, w0, lr=3e-3, epochs=15, batch=1024, **CFG)
    _, _, l1 = P.mlp_eval(w, phi, delta, wt, **CFG)
    log("nll per sample", l0 / n, "->", l1 / n, losses[[0, -1]])
    assert l1 < l0 - 0.5 * n and losses[-1] < losses[0]
    ms, _, _ = P.mlp_eval(w, phi, delta, wt, **CFG)
    assert np.corrcoef(ms[:, 0], 0.3 * phi[:, 1])[0, 1] > 0.95


def test_learned_chain_matches_oracle():
    log("learned chain")
    r, q, L = hard_sphere_system(2000, 0.2, 1.0, seed=9)
    leaf, rf = 16, 2.0
    w = P.mlp_init(5, scale=1.0)
    g = gpu_tree(r, q, L, leaf)
    g.mc_set_model(2.0, 1.0)
    g.mc_set_proposal(learned=True, weights=w, rf=rf, **CFG)
    rng = np.random.default_rng(10)
    idx, z, u = rng.integers(0, 2000, 150), rng.normal(size=(150, 3)), rng.uniform(0, 1, 150)
    dEg, accg, _ = g.mc_chain(idx, z, u, beta=1.0, mode=0)
    dlg, _, Hg = g.mc_proposal_results()
    fttg = g.mc_target_features()
    t = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=6)
    t.form_pw()
    dlo, Ho, dEo, ftto = M.replay_learned_chain(t, r, q, L, idx, z, dlg, accg, 1.0, w, rf, CFG["max_mean"],
                                          CFG["sigma_min"], CFG["sigma_max"], bjerrum=2.0, sigma=1.0)
    assert np.abs(fttg - ftto).max() < 1e-12  # features at the target: the reverse density's input
    ed = np.abs(dlg - dlo).max(1)
    eh = np.abs(Hg - Ho)
    log(f"replay: delta beyond 1e-5 {(ed > 1e-5).mean():.3f} max {ed.max():.2e}; H median {np.median(eh):.2e} "
        f"max {eh.max():.2e}")
    assert (ed <= 1e-5).mean() >= 0.9 and ed.max() < 0.02
    flip = ed > 1e-5
    assert eh[~flip].max(initial=0.0) < 2e-3 and eh.max() < 0.1
    fin = np.isfinite(dEo)
    assert (np.isfinite(dEg) == fin).all() and np.abs(dEg[fin] - dEo[fin]).max() < 1e-9
    thr = -np.log(u)
    x = np.where(fin, dEo, np.inf) - Ho
    clear = np.abs(x - thr) > 1e-3 + eh
    assert (accg[clear] == (x < thr)[clear]).all()
    log("moves decided within the tie band:", int((~clear).sum()))
    assert 5 < accg.sum() < 145


@pytest.mark.parametrize("precision", [P.HPDMK_FP64, P.HPDMK_MIXED])
@pytest.mark.parametrize("batch", [8, 64])
@pytest.mark.parametrize("spheres", [True, False])  # the decision scan has a variant for each
def test_batched_learned_chain_equals_plain(precision, batch, spheres):
    log("batched learned", precision, batch, spheres)
    r, q, L = hard_sphere_system(2000, 0.2, 1.0, seed=11)
    w = P.mlp_init(7, scale=1.0)
    runs = []
    for mode in (0, 1):
        g = P.HPDMKPtTree(P.HPDMKParams(16, 6, L, init=P.DIRECT), r, q, precision)
        g.form_outgoing_pw()
        g.form_incoming_pw()
        if spheres:
            g.mc_set_model(2.0, 1.0)
        g.mc_set_proposal(learned=True, weights=w, rf=2.0, **CFG)
        rng = np.random.default_rng(12)
        idx, z, u = rng.integers(0, 2000, 600), rng.normal(size=(600, 3)), rng.uniform(0, 1, 600)
        dE, acc, _ = g.mc_chain(idx, z, u, beta=1.0, mode=mode, batch=batch)
        dl, _, H = g.mc_proposal_results()
        runs.append((dE, acc, dl, H, g.positions()))
        g.close()
    (dE1, a1, d1, H1, x1), (dE2, a2, d2, H2, x2) = runs
    if precision == P.HPDMK_FP64:
        assert (a1 == a2).all()
        assert np.abs(d1 - d2).max() < 1e-12 and np.abs(H1 - H2).max() < 1e-9
        fin = np.isfinite(dE1)
        assert (np.isfinite(dE2) == fin).all() and np.abs(dE1[fin] - dE2[fin]).max() < 1e-9
        assert np.abs(x1 - x2).max() < 1e-12
    else:
        # the decision variable of a learned proposal is dE - H / beta (beta = 1);
        # before the first near tie the proposals and log-ratios agree too
        k = assert_same_chain(dE1, a1, dE2, a2, thresholds(u), MIXED_DE_BOUND, score_ref=dE1 - H1)
        assert np.abs(d1[:k] - d2[:k]).max(initial=0.0) < 1e-5
        assert np.abs(H1[:k] - H2[:k]).max(initial=0.0) < MIXED_DE_BOUND


def test_batched_learned_chain_equals_plain_across_compactions():
    """5000 learned moves on hard spheres (FP64): slot compactions between
    batches and several host rounds of steps (the stale-proposal cuts leave
    moves for later rounds, whose first step was set up by the last update
    of the round before)."""
    log("batched learned across compactions")
    r, q, L = hard_sphere_system(2000, 0.2, 1.0, seed=11)
    w = P.mlp_init(7, scale=1.0)
    rng = np.random.default_rng(14)
    m = 5000
    idx, z, u = rng.integers(0, 2000, m), rng.normal(size=(m, 3)), rng.uniform(0, 1, m)
    runs = []
    for mode in (0, 1):
        g = P.HPDMKPtTree(P.HPDMKParams(16, 6, L, init=P.DIRECT), r, q, P.HPDMK_FP64)
        g.form_outgoing_pw()
        g.form_incoming_pw()
        g.mc_set_model(2.0, 1.0)
        g.mc_set_proposal(learned=True, weights=w, rf=2.0, **CFG)
        dE, acc, _ = g.mc_chain(idx, z, u, beta=1.0, mode=mode, batch=64)
        dl, _, H = g.mc_proposal_results()
        runs.append((dE, acc, dl, H, g.positions()))
        g.close()
    (dE1, a1, d1, H1, x1), (dE2, a2, d2, H2, x2) = runs
    assert (a1 == a2).all()
    assert np.abs(d1 - d2).max() < 1e-12 and np.abs(H1 - H2).max() < 1e-9
    fin = np.isfinite(dE1)
    assert (np.isfinite(dE2) == fin).all() and np.abs(dE1[fin] - dE2[fin]).max() < 1e-9
    assert np.abs(x1 - x2).max() < 1e-12
    assert 0.05 < a1.mean() < 0.95


def test_batched_learned_chain_mixed_across_compactions():
    """The mixed-precision learned chain over 5000 moves on hard spheres
    (BF16 network, matrix-core update with the next proposals in its launch,
    compactions between batches): the plain FP64 learned chain's decisions
    except at a near tie of dE - H / beta."""
    log("batched learned mixed across compactions")
    r, q, L = hard_sphere_system(2000, 0.2, 1.0, seed=11)
    w = P.mlp_init(7, scale=1.0)
    rng = np.random.default_rng(15)
    m = 5000
    idx, z, u = rng.integers(0, 2000, m), rng.normal(size=(m, 3)), rng.uniform(0, 1, m)
    runs = []
    for mode, prec in ((0, P.HPDMK_FP64), (1, P.HPDMK_MIXED)):
        g = P.HPDMKPtTree(P.HPDMKParams(16, 6, L, init=P.DIRECT), r, q, prec)
        g.form_outgoing_pw()
        g.form_incoming_pw()
        g.mc_set_model(2.0, 1.0)
        g.mc_set_proposal(learned=True, weights=w, rf=2.0, **CFG)
        dE, acc, _ = g.mc_chain(idx, z, u, beta=1.0, mode=mode, batch=64)
        dl, _, H = g.mc_proposal_results()
        runs.append((dE, acc, dl, H))
        g.close()
    (dE1, a1, d1, H1), (dE2, a2, d2, H2) = runs
    k = assert_same_chain(dE1, a1, dE2, a2, thresholds(u), MIXED_DE_BOUND, score_ref=dE1 - H1)
    log(f"first differing decision at move {k} of {m}")
    assert np.abs(d1[:k] - d2[:k]).max(initial=0.0) < 1e-5
    assert np.abs(H1[:k] - H2[:k]).max(initial=0.0) < MIXED_DE_BOUND


def test_policy_gradient_training_raises_expected_jump():
    log("policy gradient")
    r, q, L = hard_sphere_system(3000, 0.2, 1.0, seed=13)
    n = len(q)
    g = P.HPDMKPtTree(P.HPDMKParams(32, 6, L, init=P.DIRECT), r, q, P.HPDMK_MIXED)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    g.mc_set_model(2.0, 1.0)
    cfg = dict(rf=2.0, max_mean=0.5, sigma_min=0.02, sigma_max=2.0)
    rng = np.random.default_rng(14)
    PR.run_uniform(g, n, 0.5, 30 * n, 1.0, 64, rng)  # equilibration
    w, hist = PR.train_proposal(g, n, cfg, warm_step=1.0, warm_moves=30000, rounds=8, round_moves=30000,
                                lr=3e-4, pg_batch=8192, seed=15)
    for h in hist:
        log({k: (round(v, 4) if isinstance(v, float) else v) for k, v in h.items()})
    w_warm, _ = PR.train_proposal(g, n, cfg, warm_step=1.0, warm_moves=30000, rounds=0, seed=15)
    js = {}
    for name, ww in (("warm", w_warm), ("trained", w)):
        out = PR.run_learned(g, n, ww, cfg, 60000, 1.0, 64, rng)
        js[name] = PR.esjd(out[0], out[3], out[2], 1.0)
    log("expected squared jump per move", js)
    assert js["trained"] > 1.05 * js["warm"]
