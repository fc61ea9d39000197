"""The C++ class of include/hpdmk_tree.hpp, driven by a compiled C++ caller
(tests/cpp/tree_check.cpp) exactly like the reference drives HPDMKPtTree in
its shift-energy test (test/hpdmk_test.cpp:374-465), against the oracle.

Bars: Real = double (FP64 device state): total energy, its parts, every dE of
a 10-move eval+update chain and the final energy agree with the oracle to
1e-10 relative to |E|.  Real = float (mixed precision): the same quantities to
the digits-6 Ewald contract of the reference, 5e-5 relative to |E|
(test/hpdmk_test.cpp:465 with 5 * 10^(1-d)).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from pdmk4mc_amd.build import CPP_CHECK

pytestmark = pytest.mark.gpu
L, N, LEAF = 20.0, 1000, 10


def run_cpp(tmp_path, r, q, digits, idx, d, real):
    f = tmp_path / "in.txt"
    with open(f, "w") as fh:
        fh.write(f"{len(q)} {L:.17g} {digits} {LEAF}\n")
        for k in range(len(q)):
            fh.write(f"{r[k, 0]:.17g} {r[k, 1]:.17g} {r[k, 2]:.17g} {q[k]:.17g}\n")
        fh.write(f"{len(idx)}\n")
        for k in range(len(idx)):
            fh.write(f"{int(idx[k])} {d[k, 0]:.17g} {d[k, 1]:.17g} {d[k, 2]:.17g}\n")
    print("tree_check", real, "digits", digits, flush=True)
    sys.stdout.flush()
    out = subprocess.run([str(CPP_CHECK), str(f), real], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    res = {"dE": []}
    for line in out.stdout.splitlines():
        w = line.split()
        if w[0] == "E":
            res["E"] = np.array([float(x) for x in w[1:]])
        elif w[0] == "dE":
            res["dE"].append(float(w[1]))
        elif w[0] == "E_end":
            res["E_end"] = float(w[1])
        else:
            res["flag"] = w[0]
    res["dE"] = np.array(res["dE"])
    return res


def oracle_chain(r, q, digits, idx, d):
    t = O.OracleTree(r, q, L, n_per_leaf=LEAF, digits=digits)
    t.form_pw()
    parts = t.eval_energy()
    dE = []
    for k in range(len(idx)):
        dE.append(t.eval_shift_energy(idx[k], d[k]))
        t.update_shift(idx[k], d[k])
    t2 = O.OracleTree(t.positions(), q, L, n_per_leaf=LEAF, digits=digits)
    t2.form_pw()
    return parts, np.array(dE), t2.eval_energy().sum()


@pytest.mark.parametrize("digits", [3, 6])
@pytest.mark.parametrize("real", ["double", "float"])
def test_cpp_class_matches_oracle(tmp_path, digits, real):
    assert os.path.exists(CPP_CHECK), "tests/cpp/_build/tree_check is built by pdmk4mc_amd/build.py"
    r, q = O.reference_test_system(N, L)
    idx, d = O.reference_test_moves(N, L, 10)
    if real == "float":  # the caller hands float inputs to the class
        r, q, d = (a.astype(np.float32).astype(np.float64) for a in (r, q, d))
    got = run_cpp(tmp_path, r, q, digits, idx, d, real)
    parts, dE, E_end = oracle_chain(r, q, digits, idx, d)
    E = abs(parts.sum())
    tol = 1e-10 * E if real == "double" else 5e-5 * E
    assert abs(got["E"][0] - parts.sum()) < tol
    assert np.abs(got["E"][1:] - parts).max() < tol
    assert np.abs(got["dE"] - dE).max() < tol, (got["dE"], dE)
    # the device keeps exact membership, so the incremental state equals a rebuild
    assert abs(got["E_end"] - E_end) < tol
    assert got["flag"] == "bad_index_rejected"
