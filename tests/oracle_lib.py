"""ctypes binding of the CPU oracle (oracle/_build/liboracle*.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the pdmk4mc_amd package.
"""
import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"

_P = C.POINTER
_dp = _P(C.c_double)
_ip = _P(C.c_int)
_lp = _P(C.c_long)


def build(quiet=True):
    """Compile the oracle with its Makefile (g++ only; no GPU needed)."""
    out = subprocess.run(["make", "-C", str(ORACLE_DIR), "-j2"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _load(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    path = ORACLE_DIR / "_build" / name
    if not path.exists():
        build()
    lib = C.CDLL(str(path))
    sig = {
        "oracle_last_error": (C.c_char_p, []),
        "oracle_pswf_params": (C.c_int, [C.c_int, _dp]),
        "oracle_prolate0_eval": (C.c_double, [C.c_double, C.c_double]),
        "oracle_prolate0_int": (C.c_double, [C.c_double, C.c_double]),
        "oracle_prolate0_intx2": (C.c_double, [C.c_double, C.c_double]),
        "oracle_prolate0_lambda": (C.c_double, [C.c_double]),
        "oracle_prolate0_coefs": (C.c_int, [C.c_double, _dp, C.c_int]),
        "oracle_residual_poly": (C.c_double, [C.c_int, C.c_double]),
        "oracle_direct_eval": (C.c_double, [_dp, C.c_double, C.c_int, _dp, _dp, C.c_double, C.c_int]),
        "oracle_tree_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp]),
        "oracle_tree_destroy": (None, [C.c_void_p]),
        "oracle_tree_info": (C.c_int, [C.c_void_p, _ip]),
        "oracle_tree_nodes": (C.c_int, [C.c_void_p, _ip]),
        "oracle_particle_leaf": (C.c_int, [C.c_void_p, _ip]),
        "oracle_positions": (C.c_int, [C.c_void_p, _dp]),
        "oracle_form_pw": (C.c_int, [C.c_void_p]),
        "oracle_get_pw": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
        "oracle_form_pw_proxy": (C.c_int, [C.c_void_p]),
        "oracle_proxy_coeffs": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
        "oracle_cheb_matrix": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp]),
        "oracle_proxy_order": (C.c_int, [C.c_int]),
        "oracle_single_pw": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                                       C.c_int, _ip]),
        "oracle_eval_energy": (C.c_int, [C.c_void_p, _dp]),
        "oracle_eval_energy_direct": (C.c_int, [C.c_void_p, _dp]),
        "oracle_eval_shift_energy": (C.c_int, [C.c_void_p, C.c_long, C.c_double, C.c_double, C.c_double, _dp,
                                               _dp]),
        "oracle_update_shift": (C.c_int, [C.c_void_p, C.c_long, C.c_double, C.c_double, C.c_double]),
        "oracle_mc_chain": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _dp, C.c_double, _dp, _ip, _dp]),
        "oracle_mc_chain_model": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                            _dp, _ip]),
        "oracle_ewald_energy": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp]),
        "oracle_ewald_potential": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                                _dp, C.c_double, C.c_double, C.c_double]),
        "oracle_ewald_shift_chain": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                               C.c_int, _lp, _dp, _ip, _dp]),
        "oracle_random_init": (None, [_dp, C.c_long, C.c_double, C.c_double]),
        "oracle_test_moves": (None, [C.c_int, C.c_double, C.c_int, _lp, _dp]),
    }
    for fn, (res, args) in sig.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


_LIBS = {}


def lib(fast=False):
    if fast not in _LIBS:
        _LIBS[fast] = _load(fast)
    return _LIBS[fast]


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _check(rc, L):
    if rc != 0:
        raise RuntimeError(L.oracle_last_error().decode())


# ---------------------------------------------------------------- PSWF
def pswf_params(digits):
    out = np.zeros(8)
    L = lib()
    _check(L.oracle_pswf_params(digits, out.ctypes.data_as(_dp)), L)
    return dict(c=out[0], lam=out[1], C0=out[2], diff0=out[3], psi0_0=out[4], n_window=int(out[5]),
                n_diff=int(out[6]), n_poly=int(out[7]))


def prolate0_eval(c, x):
    return lib().oracle_prolate0_eval(c, x)


def prolate0_int(c, r):
    return lib().oracle_prolate0_int(c, r)


def prolate0_coefs(c):
    out = np.zeros(512)
    n = lib().oracle_prolate0_coefs(c, out.ctypes.data_as(_dp), 512)
    return out[:n]


def residual_poly(digits, x):
    return lib().oracle_residual_poly(digits, x)


def direct_eval(r_src, q_src, r_trg, q_trg, cutoff, digits):
    rs, prs = _d(r_src)
    rt, prt = _d(r_trg)
    qt, pqt = _d(q_trg)
    return lib().oracle_direct_eval(prs, q_src, len(qt), prt, pqt, cutoff, digits)


# ---------------------------------------------------------------- inputs
def random_init(n, lo, hi):
    """include/utils.hpp:122-129: default-seeded mt19937, re-seeded per call."""
    v = np.zeros(n)
    lib().oracle_random_init(v.ctypes.data_as(_dp), n, lo, hi)
    return v


def reference_test_system(n_src, L):
    """Inputs of test/hpdmk_test.cpp:389-393: r ~ U[0,L)^3, q = 2u-1 then neutralised."""
    r = random_init(3 * n_src, 0.0, L).reshape(n_src, 3)
    q = random_init(n_src, -1.0, 1.0)
    q = q - q.sum() / n_src
    return r, q


def reference_test_moves(n_src, L, ntrials):
    idx = np.zeros(ntrials, dtype=np.int64)
    d = np.zeros((ntrials, 3))
    lib().oracle_test_moves(n_src, L, ntrials, idx.ctypes.data_as(_lp), d.ctypes.data_as(_dp))
    return idx, d


# ---------------------------------------------------------------- PROXY
def proxy_order(digits):
    p = lib().oracle_proxy_order(int(digits))
    if p < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return p


def cheb_matrix(which, p, n=0):
    """0: V^-1 (p x p); 1, 2: child-to-parent matrices of the lower and upper
    half; 3: coefficient-to-plane-wave matrix P ((2n+1) x p, complex)."""
    size = (2 * n + 1) * p * 2 if which == 3 else p * p
    out = np.zeros(size)
    cnt = lib().oracle_cheb_matrix(int(which), int(p), int(n), out.ctypes.data_as(_dp))
    if cnt < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    if which == 3:
        return (out[0::2] + 1j * out[1::2]).reshape(2 * n + 1, p)
    return out.reshape(p, p)


# ---------------------------------------------------------------- tree
class OracleTree:
    """CPU restatement of HPDMKPtTree<Real> (FP64 by default)."""

    def __init__(self, r, q, L, n_per_leaf=200, digits=3, precision=64, fast=False):
        self._L = lib(fast)
        r, pr = _d(np.asarray(r).reshape(-1))
        q, pq = _d(q)
        self.n = len(q)
        self.h = self._L.oracle_tree_create(precision, n_per_leaf, digits, float(L), self.n, pr, pq)
        if not self.h:
            raise RuntimeError(self._L.oracle_last_error().decode())
        info = np.zeros(7, dtype=np.int32)
        self._L.oracle_tree_info(self.h, info.ctypes.data_as(_ip))
        (self.n_nodes, self.max_depth, self.n_window, self.n_diff, self.dim_window, self.dim_diff,
         _) = [int(v) for v in info]
        self.L_box = float(L)
        self.digits = digits

    def close(self):
        if self.h:
            self._L.oracle_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nodes(self):
        out = np.zeros((self.n_nodes, 9), dtype=np.int32)
        self._L.oracle_tree_nodes(self.h, out.ctypes.data_as(_ip))
        return out

    def particle_leaf(self):
        out = np.zeros(self.n, dtype=np.int32)
        self._L.oracle_particle_leaf(self.h, out.ctypes.data_as(_ip))
        return out

    def positions(self):
        out = np.zeros((self.n, 3))
        self._L.oracle_positions(self.h, out.ctypes.data_as(_dp))
        return out

    def form_pw(self):
        _check(self._L.oracle_form_pw(self.h), self._L)

    def form_pw_proxy(self):
        """PROXY outgoing expansions (Chebyshev upward pass), then incoming."""
        _check(self._L.oracle_form_pw_proxy(self.h), self._L)

    def proxy_coeffs(self, node, from_particles=False):
        p = proxy_order(self.digits)
        out = np.zeros(p ** 3)
        n = self._L.oracle_proxy_coeffs(self.h, int(node), int(from_particles), out.ctypes.data_as(_dp))
        if n < 0:
            return None
        return out.reshape(p, p, p)  # [c, b, a]: a (x) fastest

    def get_pw(self, which, node=0):
        dim = self.dim_window if which == 0 else self.dim_diff
        out = np.zeros(2 * dim)
        n = self._L.oracle_get_pw(self.h, which, node, out.ctypes.data_as(_dp))
        if n < 0:
            return None
        return out[0::2] + 1j * out[1::2]

    def single_pw(self, x, q=1.0):
        stride = 2 * max(self.dim_window, self.dim_diff)
        out = np.zeros(self.max_depth * stride)
        path = np.zeros(self.max_depth, dtype=np.int32)
        n = self._L.oracle_single_pw(self.h, x[0], x[1], x[2], q, out.ctypes.data_as(_dp), stride,
                                     path.ctypes.data_as(_ip))
        pws = []
        for l in range(n):
            seg = out[l * stride:(l + 1) * stride]
            dim = self.dim_window if l == 0 else self.dim_diff
            pws.append(seg[0:2 * dim:2] + 1j * seg[1:2 * dim:2])
        return path[:n].copy(), pws

    def eval_energy(self):
        out = np.zeros(3)
        _check(self._L.oracle_eval_energy(self.h, out.ctypes.data_as(_dp)), self._L)
        return out

    def eval_energy_direct(self):
        out = np.zeros(3)
        _check(self._L.oracle_eval_energy_direct(self.h, out.ctypes.data_as(_dp)), self._L)
        return out

    def eval_shift_energy(self, i, d, parts=False):
        e = np.zeros(1)
        p = np.zeros(5)
        _check(self._L.oracle_eval_shift_energy(self.h, int(i), float(d[0]), float(d[1]), float(d[2]),
                                                e.ctypes.data_as(_dp), p.ctypes.data_as(_dp)), self._L)
        return (e[0], p) if parts else e[0]

    def update_shift(self, i, d):
        _check(self._L.oracle_update_shift(self.h, int(i), float(d[0]), float(d[1]), float(d[2])), self._L)

    def mc_chain(self, idx, d, u, beta):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        d, pd = _d(d)
        u, pu = _d(u)
        n = len(idx)
        dE = np.zeros(n)
        acc = np.zeros(n, dtype=np.int32)
        t2 = np.zeros(2)
        _check(self._L.oracle_mc_chain(self.h, n, idx.ctypes.data_as(_lp), pd, pu, beta, dE.ctypes.data_as(_dp),
                                       acc.ctypes.data_as(_ip), t2.ctypes.data_as(_dp)), self._L)
        return dE, acc, t2


    def mc_chain_model(self, idx, d, u, beta, bjerrum, sigma):
        """Chain with Bjerrum-scaled energies and a hard core of diameter sigma."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        d, pd = _d(d)
        u, pu = _d(u)
        n = len(idx)
        dE = np.zeros(n)
        acc = np.zeros(n, dtype=np.int32)
        _check(self._L.oracle_mc_chain_model(self.h, n, idx.ctypes.data_as(_lp), pd, pu, beta, bjerrum, sigma,
                                             dE.ctypes.data_as(_dp), acc.ctypes.data_as(_ip)), self._L)
        return dE, acc


# ---------------------------------------------------------------- Ewald
def ewald_energy(r, q, L, s, alpha, eps=1.0):
    r, pr = _d(np.asarray(r).reshape(-1))
    q, pq = _d(q)
    return lib().oracle_ewald_energy(L, s, alpha, eps, len(q), pq, pr)


def ewald_potential(r, q, L, s, alpha, x, eps=1.0):
    r, pr = _d(np.asarray(r).reshape(-1))
    q, pq = _d(q)
    return lib().oracle_ewald_potential(L, s, alpha, eps, len(q), pq, pr, x[0], x[1], x[2])


def ewald_shift_chain(r, q, L, s, alpha, idx, d, apply, eps=1.0):
    r, pr = _d(np.asarray(r).reshape(-1))
    q, pq = _d(q)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    d, pd = _d(np.asarray(d).reshape(-1))
    apply = np.ascontiguousarray(apply, dtype=np.int32)
    dE = np.zeros(len(idx))
    L_ = lib()
    _check(L_.oracle_ewald_shift_chain(L, s, alpha, eps, len(q), pq, pr, len(idx), idx.ctypes.data_as(_lp), pd,
                                       apply.ctypes.data_as(_ip), dE.ctypes.data_as(_dp)), L_)
    return dE
