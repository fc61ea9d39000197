"""hpdmk_eval_shift_energy_batch_noalloc (no counterpart in the reference):
the allocation-free variant of hpdmk_eval_shift_energy_batch, called through
the C ABI.  Bar: the same kernels on the same state, so its energies equal
the allocating entry point's bit for bit, on two calls in a row (the second
with fewer moves, so stale staging from the first call would show), in FP64
and mixed precision; the FP64 energies also match the oracle to 1e-9; after
the first call on the device (its set-up looks up the staging array, which
makes the runtime load the module that holds it) a call leaves the device's
free memory unchanged; more moves than the staging array holds is
HPDMK_EINVAL."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
import pdmk4mc_amd as P
from pdmk4mc_amd import hpdmk as PH

pytestmark = pytest.mark.gpu
L = 20.0
_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(g, idx, d):
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1)
    out = np.full(idx.size, np.nan)
    rc = PH.load_library().hpdmk_eval_shift_energy_batch_noalloc(g._h, idx.size, idx.ctypes.data_as(_lp),
                                                               d.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return rc, out


def free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("prec", [P.HPDMK_FP64, P.HPDMK_MIXED])
def test_noalloc_equals_batch_entry(prec):
    log(f"noalloc batch, precision {prec}")
    n = 3000
    r, q = O.reference_test_system(n, L)
    g = P.HPDMKPtTree(P.HPDMKParams(32, 6, L, init=P.DIRECT), r, q, prec)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    rng = np.random.default_rng(11)
    t = None
    if prec == P.HPDMK_FP64:
        t = O.OracleTree(r, q, L, n_per_leaf=32, digits=6)
        t.form_pw()
    rc, _ = noalloc(g, [0], [0.1, 0.0, 0.0])  # first call on the device: set-up
    assert rc == 0
    for m in (700, 130):  # second call smaller
        idx = rng.integers(0, n, m)
        d = rng.uniform(-1.0, 1.0, (m, 3))
        ref = g.eval_shift_energy_batch(idx, d)
        free0 = free_device_bytes()
        rc, got = noalloc(g, idx, d)
        assert rc == 0, PH.load_library().hpdmk_last_error()
        assert free_device_bytes() == free0
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        if t is not None:
            eo = np.array([t.eval_shift_energy(int(i), dd) for i, dd in zip(idx[:40], d[:40])])
            assert np.abs(got[:40] - eo).max() < 1e-9
    g.close()


def test_noalloc_rejects_oversized_batch():
    log("noalloc batch over the staging limit")
    n = 1000
    r, q = O.reference_test_system(n, L)
    g = P.HPDMKPtTree(P.HPDMKParams(64, 3, L, init=P.DIRECT), r, q)
    g.form_outgoing_pw()
    g.form_incoming_pw()
    m = 4097
    rc, _ = noalloc(g, np.zeros(m, dtype=np.int64), np.zeros((m, 3)))
    assert rc == -1
    assert b"at most 4096" in PH.load_library().hpdmk_last_error()
    rc, out = noalloc(g, np.arange(5), np.full((5, 3), 0.25))  # still usable afterwards
    assert rc == 0 and np.array_equal(out, g.eval_shift_energy_batch(np.arange(5), np.full((5, 3), 0.25)))
    g.close()
