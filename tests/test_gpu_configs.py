"""The benchmark workloads at full size, on the GPU, against the oracle and
the FP64 Ewald reporter.

* N = 1e5 (L = 100, +-1 alternating, digits 6, leaf 200): FP64 trial dE equal
  to the oracle's to 1e-9 (absolute); mixed trial dE within 2e-6 of it; the
  FP64 plain chain takes the oracle's decisions; the FP64 batched chain
  equals the FP64 plain chain (decisions identical, dE 1e-9, positions
  1e-12); the mixed batched chain equals the FP64 plain chain up to near ties
  (tests/chain_check.py); E and dE of both precisions meet the reference's
  Ewald contract 5e-5 of |E| (test/hpdmk_test.cpp:465) and the measured
  digits-6 level 1e-7 of |E| for dE.
* N = 1e6 (L = 215.4, 90% +-1 and 10% +-2, neutral, digits 6, leaf 200): the
  mixed build against the FP64 build (energy 1e-6 relative; root and sampled
  box expansions within 2e-6 of their largest element); 1000 trial moves in
  both precisions against FP64 Ewald energy differences (the reference's
  contract 5e-5 of |E|, and 1e-7 measured level); a mixed batched chain
  against the FP64 plain chain.
* 256 ragged replicas (N_r ~ U{2000..20000}, density 0.1, digits 6, leaf
  200): FP64 replica chains equal their own FP64 plain chains (four replicas:
  smallest, largest, two more), the smallest equals the oracle's chain, and
  every mixed replica chain equals its FP64 replica chain up to near ties.
* N = 2e4 charged hard spheres (sigma 1, packing 0.2, Bjerrum length 2,
  digits 6, leaf 200) with the learned proposal: the FP64 plain chain
  replayed move by move by the oracle (DMK oracle energies, the oracle's
  features and BF16 inference emulation) with the bars of
  tests/test_gpu_proposal.py, and the FP64 batched chain equal to it
  (decisions, displacements, H, dE, positions).
"""
import sys

import numpy as np
import pytest

import pathlib

import oracle_lib as O
import pdmk4mc_amd as P
from chain_check import MIXED_DE_BOUND, assert_same_chain, thresholds
from hs_system import hard_sphere_system

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as MO  # noqa: E402

pytestmark = pytest.mark.gpu
CONTRACT_6 = 5e-5  # 5 * 10^(1 - digits), test/hpdmk_test.cpp:465


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def built(params, r, q, prec):
    g = P.HPDMKPtTree(params, r, q, prec)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    return g


def system_1e5(seed=2026):
    rng = np.random.default_rng(seed)
    n, L = 100000, 100.0
    return rng.uniform(0, L, (n, 3)), np.where(np.arange(n) % 2 == 0, 1.0, -1.0), L


def system_1e6(seed=2067):
    n, L = 1000000, 215.443469
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, L, (n, 3))
    k = np.arange(n)
    q = np.where(k % 2 == 0, 1.0, -1.0)
    q[k % 20 < 2] *= 2.0  # one +2 and one -2 in every 20: 10% divalent, neutral
    return r, q, L


def test_n1e5_trials_and_chains_vs_oracle():
    log("N=1e5: oracle build")
    r, q, L = system_1e5()
    n = len(q)
    params = P.HPDMKParams(200, 6, L, init=P.DIRECT)
    t = O.OracleTree(r, q, L, n_per_leaf=200, digits=6, precision=64, fast=True)
    t.form_pw()
    rng = np.random.default_rng(5)
    m = 64
    idx = rng.integers(0, n, m)
    d = rng.uniform(-1.0, 1.0, (m, 3))
    eo = np.array([t.eval_shift_energy(int(i), dd) for i, dd in zip(idx, d)])
    log("N=1e5: trial dE, FP64 and mixed")
    g64 = built(params, r, q, P.HPDMK_FP64)
    gmx = built(params, r, q, P.HPDMK_MIXED)
    e64 = g64.eval_shift_energy_batch(idx, d)
    emx = gmx.eval_shift_energy_batch(idx, d)
    assert np.abs(e64 - eo).max() < 1e-9, np.abs(e64 - eo).max()
    assert np.abs(emx - eo).max() < 2e-6, np.abs(emx - eo).max()
    E64, Emx = g64.eval_energy(), gmx.eval_energy()
    Eo = t.eval_energy().sum()
    assert abs(E64 - Eo) < 1e-10 * abs(Eo)
    assert abs(Emx - Eo) < 1e-6 * abs(Eo)
    log("N=1e5: Ewald")
    ew = P.Ewald(L, 5.0, 0.5, 1.0, q, r)
    Ee = ew.compute_energy()
    de = ew.shift_chain(idx, d)
    ew.close()
    for E, dE in ((E64, e64), (Emx, emx)):
        assert abs(E - Ee) < CONTRACT_6 * abs(Ee)
        assert np.abs(dE - de).max() < 1e-7 * abs(Ee), np.abs(dE - de).max() / abs(Ee)
    log("N=1e5: chains")
    mc = 400
    ci = rng.integers(0, n, mc)
    cd = rng.uniform(-1.0, 1.0, (mc, 3))
    cu = rng.uniform(0.0, 1.0, mc)
    dEo, acco, _ = t.mc_chain(ci, cd, cu, 1.0)
    dEp, accp, _ = g64.mc_chain(ci, cd, cu, beta=1.0, mode=0)
    assert (accp == acco).all()
    assert np.abs(dEp - dEo).max() < 1e-9
    gb = built(params, r, q, P.HPDMK_FP64)
    dEb, accb, _ = gb.mc_chain(ci, cd, cu, beta=1.0, mode=1, batch=64)
    assert (accb == accp).all() and np.abs(dEb - dEp).max() < 1e-9
    assert np.abs(gb.positions() - g64.positions()).max() < 1e-12
    dEm, accm, _ = gmx.mc_chain(ci, cd, cu, beta=1.0, mode=1, batch=64)
    assert_same_chain(dEp, accp, dEm, accm, thresholds(cu), MIXED_DE_BOUND)
    assert 0.5 < accp.mean() < 0.95
    for g in (g64, gmx, gb):
        g.close()


def test_n1e6_build_trials_and_chain_vs_ewald():
    log("N=1e6: builds")
    r, q, L = system_1e6()
    n = len(q)
    assert q.sum() == 0 and (np.abs(q) == 2).mean() == pytest.approx(0.1)
    params = P.HPDMKParams(200, 6, L, init=P.DIRECT)
    g64 = built(params, r, q, P.HPDMK_FP64)
    gmx = built(params, r, q, P.HPDMK_MIXED)
    assert g64.max_depth == 6
    E64, Emx = g64.eval_energy(), gmx.eval_energy()
    assert abs(Emx - E64) < 1e-6 * abs(E64), (Emx, E64)
    a, b = g64.get_expansion(0), gmx.get_expansion(0)
    assert np.abs(a - b).max() < 2e-6 * np.abs(a).max()
    plan, _ = P.plan_tree(params, r)  # per node: level, is_leaf, ...
    D = g64.max_depth
    boxes = np.nonzero((plan[:, 0] >= 2) & (plan[:, 0] <= D - 2))[0]
    for node in np.random.default_rng(3).choice(boxes, 12, replace=False):
        for which in (1, 2):
            if which == 2 and plan[node, 1]:
                continue
            a, b = g64.get_expansion(which, int(node)), gmx.get_expansion(which, int(node))
            assert np.abs(a - b).max() < 2e-6 * max(np.abs(a).max(), 1.0), (which, node)
    log("N=1e6: 1000 trial moves vs FP64 Ewald")
    rng = np.random.default_rng(41)
    m = 1000
    idx = rng.integers(0, n, m)
    d = rng.uniform(-1.0, 1.0, (m, 3))
    e64 = g64.eval_shift_energy_batch(idx, d)
    emx = gmx.eval_shift_energy_batch(idx, d)
    ew = P.Ewald(L, 4.0, 0.1, 1.0, q, r)
    Ee = ew.compute_energy()
    de = ew.shift_chain(idx, d)
    ew.close()
    log(f"N=1e6: E64 {E64} Ewald {Ee}; dE err fp64 {np.abs(e64 - de).max() / abs(Ee):.3e}, "
        f"mixed {np.abs(emx - de).max() / abs(Ee):.3e} of |E|")
    assert abs(E64 - Ee) < CONTRACT_6 * abs(Ee)
    for dE in (e64, emx):
        err = np.abs(dE - de).max() / abs(Ee)
        assert err < CONTRACT_6 and err < 1e-7, err
    log("N=1e6: mixed batched chain vs FP64 plain chain")
    mc = 2000
    ci = rng.integers(0, n, mc)
    cd = rng.uniform(-1.0, 1.0, (mc, 3))
    cu = rng.uniform(0.0, 1.0, mc)
    dEp, accp, _ = g64.mc_chain(ci, cd, cu, beta=1.0, mode=0)
    dEm, accm, _ = gmx.mc_chain(ci, cd, cu, beta=1.0, mode=1, batch=64)
    assert_same_chain(dEp, accp, dEm, accm, thresholds(cu), MIXED_DE_BOUND)
    g64.close()
    gmx.close()


def ragged_set(n_rep=256, seed=2033):
    rng = np.random.default_rng(seed)
    params, systems = [], []
    for _ in range(n_rep):
        n_r = int(rng.integers(2000, 20001))
        L_r = (n_r / 0.1) ** (1.0 / 3.0)
        systems.append((rng.uniform(0, L_r, (n_r, 3)), np.where(np.arange(n_r) % 2 == 0, 1.0, -1.0)))
        params.append(P.HPDMKParams(200, 6, L_r, init=P.DIRECT))
    return params, systems


def test_256_ragged_replicas():
    log("256 replicas: build")
    params, systems = ragged_set()
    M = 300
    rng = np.random.default_rng(9)
    idx = np.stack([rng.integers(0, len(q), M) for _, q in systems])
    d = rng.uniform(-1.0, 1.0, (len(systems), M, 3))
    u = rng.uniform(0, 1, (len(systems), M))
    res = {}
    for prec in (P.HPDMK_FP64, P.HPDMK_MIXED):
        rs = P.HPDMKReplicas(params, systems, prec)
        rs.form_pw()
        rs.mc_load(idx, d, u)
        rs.mc_run(1.0, graph_steps=64)
        dE, acc = rs.mc_results()
        res[prec] = (dE, acc, [rs.tree(k).positions() for k in (0, 1, 2, 3)] if prec == P.HPDMK_FP64 else None)
        if prec == P.HPDMK_FP64:
            sizes = [len(q) for _, q in systems]
            pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 200})
            pos = {k: rs.tree(k).positions() for k in pick}
        rs.close()
    dE64, acc64, _ = res[P.HPDMK_FP64]
    log("256 replicas: single chains and oracle")
    for k in pick:
        r, q = systems[k]
        g = built(params[k], r, q, P.HPDMK_FP64)
        dEs, accs, _ = g.mc_chain(idx[k], d[k], u[k], beta=1.0, mode=0)
        assert (accs == acc64[k]).all(), k
        assert np.abs(dEs - dE64[k]).max() < 1e-9, k
        assert np.abs(g.positions() - pos[k]).max() < 1e-12, k
        g.close()
    k0 = int(np.argmin([len(q) for _, q in systems]))
    r, q = systems[k0]
    t = O.OracleTree(r, q, params[k0].L, n_per_leaf=200, digits=6)
    t.form_pw()
    dEo, acco, _ = t.mc_chain(idx[k0], d[k0], u[k0], 1.0)
    assert (acco == acc64[k0]).all() and np.abs(dEo - dE64[k0]).max() < 1e-9
    log("256 replicas: mixed vs FP64")
    dEm, accm, _ = res[P.HPDMK_MIXED]
    for k in range(len(systems)):
        assert_same_chain(dE64[k], acc64[k], dEm[k], accm[k], thresholds(u[k]), MIXED_DE_BOUND)
    assert 0.5 < acc64.mean() < 0.95


def test_n2e4_hard_spheres_learned_chain():
    log("N=2e4 hard spheres: build")
    r, q, L = hard_sphere_system(20000, 0.2, 1.0, seed=2031)
    n = len(q)
    params = P.HPDMKParams(200, 6, L, init=P.DIRECT)
    cfg = dict(max_mean=0.5, sigma_min=0.02, sigma_max=2.0)
    rf = 2.0
    w = P.mlp_init(2032, scale=1.0)
    rng = np.random.default_rng(2033)
    m = 300
    idx, z, u = rng.integers(0, n, m), rng.normal(size=(m, 3)), rng.uniform(0, 1, m)
    runs = []
    for mode in (0, 1):
        g = built(params, r, q, P.HPDMK_FP64)
        g.mc_set_model(2.0, 1.0)
        g.mc_set_proposal(learned=True, weights=w, rf=rf, **cfg)
        dE, acc, _ = g.mc_chain(idx, z, u, beta=1.0, mode=mode, batch=64)
        dl, _, H = g.mc_proposal_results()
        runs.append((dE, acc, dl, H, g.mc_target_features(), g.positions()))
        g.close()
    (dEg, accg, dlg, Hg, fttg, xg), (dEb, accb, dlb, Hb, _, xb) = runs
    assert (accb == accg).all()
    assert np.abs(dlb - dlg).max() < 1e-12 and np.abs(Hb - Hg).max() < 1e-9
    fin = np.isfinite(dEg)
    assert (np.isfinite(dEb) == fin).all() and np.abs(dEb[fin] - dEg[fin]).max() < 1e-9
    assert np.abs(xb - xg).max() < 1e-12
    log("N=2e4 hard spheres: oracle replay")
    t = O.OracleTree(r, q, L, n_per_leaf=200, digits=6, precision=64, fast=True)
    t.form_pw()
    dlo, Ho, dEo, ftto = MO.replay_learned_chain(t, r, q, L, idx, z, dlg, accg, 1.0, w, rf, cfg["max_mean"],
                                                 cfg["sigma_min"], cfg["sigma_max"], bjerrum=2.0, sigma=1.0)
    assert np.abs(fttg - ftto).max() < 1e-12
    ed = np.abs(dlg - dlo).max(1)
    eh = np.abs(Hg - Ho)
    log(f"replay: delta beyond 1e-5 {(ed > 1e-5).mean():.3f} max {ed.max():.2e}; H max {eh.max():.2e}")
    assert (ed <= 1e-5).mean() >= 0.9 and ed.max() < 0.02
    assert eh[ed <= 1e-5].max(initial=0.0) < 2e-3 and eh.max() < 0.1
    fin = np.isfinite(dEo)
    assert (np.isfinite(dEg) == fin).all() and np.abs(dEg[fin] - dEo[fin]).max() < 1e-9
    thr = -np.log(u)
    x = np.where(fin, dEo, np.inf) - Ho
    clear = np.abs(x - thr) > 1e-3 + eh
    assert (accg[clear] == (x < thr)[clear]).all()
    log("accepted", int(accg.sum()), "overlaps", int((~fin).sum()), "tie band", int((~clear).sum()))
    assert 5 < accg.sum() < m - 5 and (~fin).sum() > 5
