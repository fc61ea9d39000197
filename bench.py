#!/usr/bin/env python3
"""Throughput of the DMK Monte Carlo hot path on one MI355X.

Workload (N = 1e5 system): N = 100000 point charges uniform in a periodic
cube L = 100, charges alternating +1/-1, PSWF tolerance 1e-6 (digits 6), leaf
cap 200; trial moves pick a particle uniformly and displace it by
U[-1,1]^3; Metropolis at beta = 1.  A step is 1e4 trial moves of one chain
(energy change, decision, accept/update).  Moves, displacements and uniforms
are in HBM before the timed region starts.

Prints one JSON line: trial moves/s of the batched chain (the headline), the
plain one-trial-at-a-time chain beside it, the roofline of the dominant
kernel and the CPU baseline (the oracle, i.e. the CPU restatement of the
reference path, one thread, on a bounded sample of the same chain).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def workload(n, L, n_moves, seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, L, (n, 3))
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    mv = np.random.default_rng(seed + 1)
    idx = mv.integers(0, n, n_moves)
    d = mv.uniform(-1.0, 1.0, (n_moves, 3))
    u = mv.uniform(0.0, 1.0, n_moves)
    return r, q, idx, d, u


def trial_bytes(st, Mw, Md, bc, br, bn):
    """Algorithmic HBM bytes of the trial moves with stats st (per move:
    levels_t, levels_o, nbr_t, nbr_o, arrays_t, arrays_o)."""
    lt, lo, nt, no = st[:, 0], st[:, 1], st[:, 2], st[:, 3]
    return Mw * (bc + br) + (lt + lo) * Md * bc + np.maximum(lt, lo) * Md * br + bn * (nt + no)


def accept_bytes(st, acc, Mw, Md, bc):
    return acc * (2 * Mw * bc + 2 * (st[:, 4] + st[:, 5]) * Md * bc)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--moves-per-step", type=int, default=10000)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--L", type=float, default=100.0)
    ap.add_argument("--digits", type=int, default=6)
    ap.add_argument("--leaf", type=int, default=200)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--plain-moves", type=int, default=3000)
    ap.add_argument("--cpu-moves", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2026)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); independent replicas shard "
              "trivially across GPUs and no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    import pdmk4mc_amd as P

    prec = P.HPDMK_MIXED if args.precision == "mixed" else P.HPDMK_FP64
    bc, br, bn = (8, 4, 16) if prec == P.HPDMK_MIXED else (16, 8, 32)
    M = args.moves_per_step
    n_total = (args.warmup + args.steps) * M
    r, q, idx, d, u = workload(args.n, args.L, n_total + args.plain_moves, args.seed)
    params = P.HPDMKParams(args.leaf, args.digits, args.L, init=P.DIRECT)

    t0 = time.time()
    g = P.HPDMKPtTree(params, r, q, prec)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    E0 = g.eval_energy()
    t_build = time.time() - t0
    print(f"# tree: {g.n_nodes} nodes, {g.max_depth} levels, {g.n_leaves} leaves, max near {g.max_near}; "
          f"build+energy {t_build:.2f}s E={E0:.6f}", file=sys.stderr, flush=True)

    # warmup then timed steps of the batched chain (moves already in HBM)
    if args.warmup:
        g.mc_load(idx[:args.warmup * M], d[:args.warmup * M], u[:args.warmup * M])
        g.mc_run(1.0, 1, args.batch)
    sl = slice(args.warmup * M, n_total)
    g.mc_load(idx[sl], d[sl], u[sl])
    t0 = time.perf_counter()
    gpu_ms = g.mc_run(1.0, 1, args.batch)
    wall = time.perf_counter() - t0
    dE_b, acc_b = g.mc_results()
    stats_b = g.mc_stats()
    moves = args.steps * M
    value = moves / wall
    print(f"# batched: {moves} moves in {wall:.3f}s (gpu {gpu_ms / 1e3:.3f}s) = {value:.0f} moves/s, "
          f"acceptance {acc_b.mean():.3f}", file=sys.stderr, flush=True)

    # per-kernel times on a profiled segment (events around every launch)
    prof_moves = min(M, 4000)
    psl = slice(n_total - prof_moves, n_total)
    gp = P.HPDMKPtTree(params, r, q, prec)
    gp.form_outgoing_pw()
    gp.form_incoming_pw()
    gp.set_profiling(True)
    gp.mc_load(idx[psl], d[psl], u[psl])
    gp.mc_run(1.0, 1, args.batch)
    kt = gp.kernel_times()
    phases = gp.phase_times()
    _, acc_p = gp.mc_results()
    st_p = gp.mc_stats()
    Mw, Md = g.dim_window, g.dim_diff
    kb = {"trial": trial_bytes(st_p, Mw, Md, bc, br, bn).sum(),
          "update": accept_bytes(st_p, acc_p, Mw, Md, bc).sum()}
    dom = max(kt, key=lambda k: kt[k][0])
    ms_tot, cnt = kt[dom]
    avg_ms = ms_tot / max(cnt, 1)
    per_launch = kb.get(dom, 0.0) / max(cnt, 1)
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "bytes_per_launch": int(per_launch), "avg_launch_us": round(avg_ms * 1e3, 3),
                "kernel_ms": {k: [round(v[0], 3), v[1]] for k, v in kt.items()}, "phase_us": phases}
    gp.close()

    # plain chain (one trial kernel + one update kernel per move) on the same state
    gpl = P.HPDMKPtTree(params, r, q, prec)
    gpl.form_outgoing_pw()
    gpl.form_incoming_pw()
    psl = slice(n_total, n_total + args.plain_moves)
    gpl.mc_load(idx[psl][:200], d[psl][:200], u[psl][:200])
    gpl.mc_run(1.0, 0)
    gpl.mc_load(idx[psl], d[psl], u[psl])
    t0 = time.perf_counter()
    gpl.mc_run(1.0, 0)
    plain_wall = time.perf_counter() - t0
    plain = args.plain_moves / plain_wall
    gpl.close()
    gpp = P.HPDMKPtTree(params, r, q, prec)
    gpp.form_outgoing_pw()
    gpp.form_incoming_pw()
    gpp.set_profiling(True)
    gpp.mc_load(idx[psl][:500], d[psl][:500], u[psl][:500])
    gpp.mc_run(1.0, 0)
    plain_prof = {"kernel_ms": {k: [round(v[0], 3), v[1]] for k, v in gpp.kernel_times().items()},
                  "phase_us": gpp.phase_times()}
    gpp.close()
    print(f"# plain: {plain:.0f} moves/s; batched/plain = {value / plain:.2f}", file=sys.stderr, flush=True)

    # accuracy of the arithmetic mode: trial dE of this precision vs the FP64 path on the same state
    acc_info = None
    if prec == P.HPDMK_MIXED:
        g64 = P.HPDMKPtTree(params, r, q, P.HPDMK_FP64)
        g64.form_outgoing_pw()
        g64.form_incoming_pw()
        gm = P.HPDMKPtTree(params, r, q, prec)
        gm.form_outgoing_pw()
        gm.form_incoming_pw()
        sidx, sd = idx[:256], d[:256]
        e64 = g64.eval_shift_energy_batch(sidx, sd)
        em = gm.eval_shift_energy_batch(sidx, sd)
        E64 = g64.eval_energy()
        acc_info = {"dE_vs_fp64_max_abs": float(np.abs(em - e64).max()),
                    "dE_vs_fp64_max_rel_to_E": float(np.abs(em - e64).max() / abs(E64)),
                    "E_mixed_vs_fp64_rel": float(abs(gm.eval_energy() - E64) / abs(E64)), "sample": 256}
        g64.close()
        gm.close()

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O

        S = args.cpu_moves
        t = O.OracleTree(r, q, args.L, n_per_leaf=args.leaf, digits=args.digits, precision=64, fast=True)
        t.form_pw()
        sl = slice(args.warmup * M, args.warmup * M + S)
        dE_c, acc_c, t2 = t.mc_chain(idx[sl], d[sl], u[sl], 1.0)
        cpu = {"value": round(S / t2.sum(), 2), "unit": "trial moves/s", "cores": 1, "kind": "port",
               "sample": f"first {S} moves of the timed chain, FP64, oracle built -O3 -ffast-math, one thread",
               "eval_us": round(t2[0] / S * 1e6, 2), "update_us_per_accept": round(t2[1] / max(1, acc_c.sum()) * 1e6, 2),
               "agree_with_gpu_decisions": float((acc_c == acc_b[:S]).mean())}
        print(f"# cpu: {cpu}", file=sys.stderr, flush=True)

    out = {
        "metric": "dE trial moves/sec at 1e-6 rel err, N=1e5; batched+learned vs plain GPU",
        "value": round(value, 2), "unit": "trial moves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64acc" if prec == P.HPDMK_MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": "N=1e5 periodic cube L=100, +-1 alternating charges, digits 6, leaf 200, "
                               "1e4 shifts/step dx~U[-1,1]^3, beta=1",
                   "n": args.n, "L": args.L, "digits": args.digits, "leaf": args.leaf, "batch": args.batch,
                   "moves_per_step": M, "precision": args.precision},
        "plain_gpu": {"value": round(plain, 2), "unit": "trial moves/s", "moves": args.plain_moves,
                      "profile": plain_prof},
        "batched_over_plain": round(value / plain, 3),
        "acceptance": round(float(acc_b.mean()), 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": acc_info,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
