"""CPU checks of the product library: it loads, exports every entry point of
include/hpdmk.h, and its host-side set-up (octree, PSWF constants) agrees with
the oracle.  No compute call needs a GPU here."""
import re
import pathlib

import numpy as np
import pytest

import oracle_lib as O
import pdmk4mc_amd as P

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_library_exports_every_header_symbol():
    import ctypes
    lib = ctypes.CDLL(str(P.lib_path()))
    header = (ROOT / "include" / "hpdmk.h").read_text()
    names = set(re.findall(r"^\w[\w\s\*]*?\b(hpdmk_\w+)\s*\(", header, re.M))
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert names <= set(P.hpdmk.SIGNATURES), names - set(P.hpdmk.SIGNATURES)


def test_default_params_match_reference():
    p = P.load_library().hpdmk_default_params()
    assert (p.n_per_leaf, p.digits, p.prolate_order, p.init) == (200, 3, 16.0, P.PROXY)


@pytest.mark.parametrize("digits", [3, 6, 9, 12])
def test_pswf_constants_match_oracle(digits):
    a = P.pswf_constants(digits)
    b = O.pswf_params(digits)
    assert abs(a["lam"] - b["lam"]) < 1e-13
    assert abs(a["diff0"] - b["diff0"]) < 1e-11 * abs(b["diff0"])
    # psi0 normalisation is free; the method only uses psi0(0)/C0
    assert abs(a["psi0_0"] / a["C0"] - b["psi0_0"] / b["C0"]) < 1e-13
    assert (a["n_window"], a["n_diff"]) == (b["n_window"], b["n_diff"])


def test_bad_digits_rejected():
    with pytest.raises(P.HPDMKError):
        P.pswf_constants(5)


@pytest.mark.parametrize("n,L,leaf", [(1000, 20.0, 10), (1000, 20.0, 64), (1100, 20.0, 5), (20000, 58.5, 200)])
def test_octree_identical_to_oracle(n, L, leaf):
    r, q = O.reference_test_system(n, L)
    t = O.OracleTree(r, q, L, n_per_leaf=leaf, digits=3)
    on, ol = t.nodes(), t.particle_leaf()
    pn, pl = P.plan_tree(P.HPDMKParams(leaf, 3, L), r)
    assert on.shape == pn.shape
    assert (on[:, :8] == pn[:, :8]).all()
    assert (ol == pl).all()
    # near list = self + leaves under colleagues + coarse-grain leaves
    for i in np.where(pn[:, 1] == 1)[0][:50]:
        assert pn[i, 8] >= 1 + on[i, 8]


def test_plan_rejects_shallow_tree_and_bad_coordinates():
    r, q = O.reference_test_system(50, 20.0)
    with pytest.raises(P.HPDMKError, match="max depth"):
        P.plan_tree(P.HPDMKParams(200, 3, 20.0), r)
    r[0, 0] = 25.0
    with pytest.raises(P.HPDMKError, match=r"\[0, L\)"):
        P.plan_tree(P.HPDMKParams(5, 3, 20.0), r)


def test_cpp_class_header_compiles_and_caller_is_built():
    import subprocess
    from pdmk4mc_amd.build import CPP_CHECK
    out = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                          "-x", "c++", "-"], input='#include "hpdmk_tree.hpp"\n'
                         "template class hpdmk_amd::HPDMKPtTree<float>;\n"
                         "template class hpdmk_amd::HPDMKPtTree<double>;\n",
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert CPP_CHECK.exists()
    # argument errors are reported before any device call
    res = subprocess.run([str(CPP_CHECK)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2 and "usage" in res.stderr


def test_read_particle_info_follows_reference(tmp_path):
    """src/utils.cpp:14-28: the first line is a header; every other line that
    parses as four numbers is a particle (extra fields ignored, other lines
    skipped)."""
    f = tmp_path / "particles.txt"
    f.write_text("x y z q\n1 2 3 -1\n  4.5\t5.5 6.5 1.0\nnot a particle\n7 8 9\n1e-3 2e-3 3e-3 2 extra\n")
    r, q = P.read_particle_info(f)
    assert r.shape == (3, 3)
    assert np.array_equal(r, [[1, 2, 3], [4.5, 5.5, 6.5], [1e-3, 2e-3, 3e-3]])
    assert np.array_equal(q, [-1.0, 1.0, 2.0])
    with pytest.raises(P.HPDMKError):
        P.read_particle_info(tmp_path / "missing.txt")
