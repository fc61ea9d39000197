"""ctypes mirror of the reference's HPDMKPtTree (include/tree.hpp:28-148) over
the C ABI in include/hpdmk.h.  Fails loudly when libhpdmk_amd.so is missing:
there is no CPU fallback."""
import ctypes as C
import pathlib

import numpy as np

DIRECT, PROXY = 1, 2
HPDMK_FP64, HPDMK_MIXED = 0, 1
_LIB = pathlib.Path(__file__).resolve().parent / "libhpdmk_amd.so"
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)


class HPDMKParams(C.Structure):
    """include/hpdmk.h:11-17 (defaults of the reference)."""
    _fields_ = [("n_per_leaf", C.c_int), ("digits", C.c_int), ("L", C.c_double), ("prolate_order", C.c_double),
                ("init", C.c_int)]

    def __init__(self, n_per_leaf=200, digits=3, L=0.0, prolate_order=16.0, init=PROXY):
        super().__init__(n_per_leaf, digits, L, prolate_order, init)


class HPDMKError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None

# every entry point of include/hpdmk.h with its ctypes signature
SIGNATURES = {
    "hpdmk_default_params": (HPDMKParams, []),
    "hpdmk_last_error": (C.c_char_p, []),
    "hpdmk_tree_create": (C.c_int, [HPDMKParams, C.c_int, _dp, _dp, C.POINTER(C.c_void_p)]),
    "hpdmk_tree_create_ex": (C.c_int, [HPDMKParams, C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_void_p)]),
    "hpdmk_tree_destroy": (C.c_int, [C.c_void_p]),
    "hpdmk_tree_rebuild": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "hpdmk_read_particle_info": (C.c_int, [C.c_char_p, C.c_int, _dp, _dp, _ip]),
    "hpdmk_tree_info": (C.c_int, [C.c_void_p, _ip]),
    "hpdmk_plan_tree_device": (C.c_int, [HPDMKParams, C.c_int, _dp, _ip, _ip, _ip, _dp]),
    "hpdmk_form_outgoing_pw": (C.c_int, [C.c_void_p]),
    "hpdmk_form_incoming_pw": (C.c_int, [C.c_void_p]),
    "hpdmk_eval_energy": (C.c_int, [C.c_void_p, _dp, _dp]),
    "hpdmk_eval_shift_energy": (C.c_int, [C.c_void_p, C.c_long, C.c_double, C.c_double, C.c_double, _dp, _dp]),
    "hpdmk_update_shift": (C.c_int, [C.c_void_p, C.c_long, C.c_double, C.c_double, C.c_double]),
    "hpdmk_get_positions": (C.c_int, [C.c_void_p, _dp]),
    "hpdmk_eval_shift_energy_batch": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _dp]),
    "hpdmk_eval_shift_energy_batch_noalloc": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _dp]),
    "hpdmk_mc_load": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _dp]),
    "hpdmk_mc_run": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int, _dp]),
    "hpdmk_mc_results": (C.c_int, [C.c_void_p, _dp, _ip]),
    "hpdmk_mc_stats": (C.c_int, [C.c_void_p, _ip]),
    "hpdmk_get_expansion": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _ip]),
    "hpdmk_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "hpdmk_kernel_times": (C.c_int, [C.c_void_p, _dp, _ip]),
    "hpdmk_clear_error": (C.c_int, [C.c_void_p]),
    "hpdmk_copy_probe": (C.c_int, [C.c_long, C.c_int, _dp]),
    "hpdmk_mc_set_model": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "hpdmk_ewald_create": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                     C.POINTER(C.c_void_p)]),
    "hpdmk_ewald_destroy": (C.c_int, [C.c_void_p]),
    "hpdmk_ewald_compute_energy": (C.c_int, [C.c_void_p, _dp, _dp]),
    "hpdmk_ewald_compute_potential": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    "hpdmk_ewald_shift_chain": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _ip, _dp]),
    "hpdmk_mc_set_proposal": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_double, C.c_double,
                                        C.c_double, C.c_double]),
    "hpdmk_mc_proposal_results": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "hpdmk_mc_target_features": (C.c_int, [C.c_void_p, _dp]),
    "hpdmk_mlp_eval": (C.c_int, [C.c_int, C.POINTER(C.c_float), _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_float), _dp]),
    "hpdmk_mlp_train": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp, C.POINTER(C.c_float), _dp]),
    "hpdmk_set_tiles": (C.c_int, [C.c_void_p, _ip, C.c_int]),
    "hpdmk_get_tiles": (C.c_int, [C.c_void_p, _ip]),
    "hpdmk_autotune": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _dp, C.c_double, _ip, _dp]),
    "hpdmk_replicas_create": (C.c_int, [C.c_int, C.c_void_p, C.c_int, _ip, _dp, _dp, C.POINTER(C.c_void_p)]),
    "hpdmk_replicas_destroy": (C.c_int, [C.c_void_p]),
    "hpdmk_replicas_count": (C.c_int, [C.c_void_p, _ip]),
    "hpdmk_replicas_tree": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "hpdmk_replicas_form_pw": (C.c_int, [C.c_void_p]),
    "hpdmk_replicas_mc_load": (C.c_int, [C.c_void_p, C.c_int, _lp, _dp, _dp]),
    "hpdmk_replicas_mc_run": (C.c_int, [C.c_void_p, C.c_double, C.c_int, _dp]),
    "hpdmk_replicas_mc_results": (C.c_int, [C.c_void_p, _dp, _ip]),
    "hpdmk_phase_times": (C.c_int, [C.c_void_p, _dp, _ip]),
    "hpdmk_plan_tree": (C.c_int, [HPDMKParams, C.c_int, _dp, _ip, _ip, _ip]),
    "hpdmk_pswf_constants": (C.c_int, [C.c_int, _dp]),
}


def lib_path():
    return _LIB


def load_library():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            raise ImportError(f"{_LIB} is missing: build it with pdmk4mc_amd.build.build() (hipcc, gfx950)")
        lib = C.CDLL(str(_LIB))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise HPDMKError(rc, load_library().hpdmk_last_error().decode())


def _d(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    return a, a.ctypes.data_as(_dp)


def read_particle_info(path):
    """src/utils.cpp:14-28: (r (N x 3), q) from an "x y z q" text file with
    one header line."""
    L = load_library()
    n = C.c_int(0)
    _check(L.hpdmk_read_particle_info(str(path).encode(), 0, None, None, C.byref(n)))
    r = np.zeros((n.value, 3))
    q = np.zeros(n.value)
    _check(L.hpdmk_read_particle_info(str(path).encode(), n.value, r.ctypes.data_as(_dp), q.ctypes.data_as(_dp),
                                      C.byref(n)))
    return r, q


def plan_tree(params, r_src):
    """Host-only octree plan: (nodes[n,9], particle_leaf[n_src]); no device needed."""
    L = load_library()
    r, pr = _d(r_src)
    n = r.size // 3
    nn = C.c_int()
    _check(L.hpdmk_plan_tree(params, n, pr, C.byref(nn), None, None))
    nodes = np.zeros((nn.value, 9), dtype=np.int32)
    leaf = np.zeros(n, dtype=np.int32)
    _check(L.hpdmk_plan_tree(params, n, pr, C.byref(nn), nodes.ctypes.data_as(_ip), leaf.ctypes.data_as(_ip)))
    return nodes, leaf


def plan_tree_device(params, r_src):
    """The same octree plan built on the GPU: (nodes[n,9], particle_leaf[n_src], ms)."""
    L = load_library()
    r, pr = _d(r_src)
    n = r.size // 3
    nn = C.c_int()
    ms = C.c_double()
    _check(L.hpdmk_plan_tree_device(params, n, pr, C.byref(nn), None, None, None))
    nodes = np.zeros((nn.value, 9), dtype=np.int32)
    leaf = np.zeros(n, dtype=np.int32)
    _check(L.hpdmk_plan_tree_device(params, n, pr, C.byref(nn), nodes.ctypes.data_as(_ip), leaf.ctypes.data_as(_ip),
                                    C.byref(ms)))
    return nodes, leaf, ms.value


def pswf_constants(digits):
    out = np.zeros(7)
    _check(load_library().hpdmk_pswf_constants(int(digits), out.ctypes.data_as(_dp)))
    return dict(c=out[0], lam=out[1], C0=out[2], diff0=out[3], psi0_0=out[4], n_window=int(out[5]),
                n_diff=int(out[6]))


MLP_FEATURES = 8
MLP_LAYERS = ((8, 64), (64, 64), (64, 64), (64, 6))
MLP_PARAMS = sum(a * b + b for a, b in MLP_LAYERS)


def mlp_init(seed=0, scale=1.0):
    """Initial proposal-MLP weights (float32): scaled Gaussian layers, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in MLP_LAYERS:
        parts.append(rng.normal(0, scale / np.sqrt(fan_in), (fan_in, fan_out)).reshape(-1))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def mlp_eval(weights, feat, delta, weight, max_mean=0.5, sigma_min=0.05, sigma_max=1.0):
    """GPU inference and gradient: (mu_sigma[n,6] from the deployed BF16 matrix-core
    inference, FP32 grad[MLP_PARAMS] of the weighted NLL sum, that sum)."""
    w = np.ascontiguousarray(weights, dtype=np.float32)
    feat, pf = _d(np.asarray(feat).reshape(-1))
    delta, pdl = _d(np.asarray(delta).reshape(-1))
    weight, pw = _d(weight)
    cfg, pc = _d([max_mean, sigma_min, sigma_max])
    n = weight.size
    ms = np.zeros((n, 6))
    g = np.zeros(MLP_PARAMS, dtype=np.float32)
    loss = C.c_double()
    _check(load_library().hpdmk_mlp_eval(n, w.ctypes.data_as(C.POINTER(C.c_float)), pf, pdl, pw, pc,
                                         ms.ctypes.data_as(_dp), g.ctypes.data_as(C.POINTER(C.c_float)),
                                         C.byref(loss)))
    return ms, g, loss.value


def mlp_train(feat, delta, weight, weights0, max_mean=0.5, sigma_min=0.05, sigma_max=1.0, lr=3e-3, epochs=20,
              batch=1024, seed=0, signed=False):
    """Adam on the mean weighted NLL of the displacements; returns (weights, loss per epoch).
    signed: weights are rewards minus a baseline (normalised by the sample count), so
    the descent is the policy gradient of the mean reward."""
    w = np.array(weights0, dtype=np.float32, copy=True)
    feat, pf = _d(np.asarray(feat).reshape(-1))
    delta, pdl = _d(np.asarray(delta).reshape(-1))
    weight, pw = _d(weight)
    cfg, pc = _d([max_mean, sigma_min, sigma_max, lr, epochs, batch, seed, 1.0 if signed else 0.0])
    losses = np.zeros(int(epochs))
    _check(load_library().hpdmk_mlp_train(weight.size, pf, pdl, pw, pc, w.ctypes.data_as(C.POINTER(C.c_float)),
                                          losses.ctypes.data_as(_dp)))
    return w, losses


def copy_probe(n_bytes=256 << 20, reps=5):
    """Read+write GB/s of a device streaming copy of n_bytes (HBM probe)."""
    out = C.c_double(0.0)
    _check(load_library().hpdmk_copy_probe(int(n_bytes), int(reps), C.byref(out)))
    return out.value


class HPDMKPtTree:
    """Device-resident DMK system; method names follow include/tree.hpp:98-132."""

    def __init__(self, params, r_src, charge, precision=HPDMK_FP64, _borrowed=None):
        L = load_library()
        r, pr = _d(r_src)
        q, pq = _d(charge)
        if r.size != 3 * q.size:
            raise ValueError("r_src must hold 3 coordinates per charge")
        self.params = params
        self.n_src = q.size
        self.precision = precision
        self._owned = _borrowed is None
        if _borrowed is None:
            h = C.c_void_p()
            _check(L.hpdmk_tree_create_ex(params, precision, q.size, pr, pq, C.byref(h)))
        else:
            h = _borrowed
        self._h = h
        info = np.zeros(10, dtype=np.int32)
        _check(L.hpdmk_tree_info(h, info.ctypes.data_as(_ip)))
        (self.n_nodes, self.max_depth, self.n_window, self.n_diff, self.dim_window, self.dim_diff,
         self.n_pw_boxes, self.max_near, self.n_leaves, self.n_slots) = [int(v) for v in info]

    # ---- lifecycle
    def rebuild(self):
        """A new system with the octree rebuilt from the current positions
        (fresh leaf slots; expansions formed as in this one; Monte Carlo
        settings carried over, move lists not).  This system stays usable."""
        h = C.c_void_p()
        _check(load_library().hpdmk_tree_rebuild(self._h, C.byref(h)))
        g = HPDMKPtTree.__new__(HPDMKPtTree)
        g.params, g.n_src, g.precision, g._owned, g._h = self.params, self.n_src, self.precision, True, h
        info = np.zeros(10, dtype=np.int32)
        _check(load_library().hpdmk_tree_info(h, info.ctypes.data_as(_ip)))
        (g.n_nodes, g.max_depth, g.n_window, g.n_diff, g.dim_window, g.dim_diff,
         g.n_pw_boxes, g.max_near, g.n_leaves, g.n_slots) = [int(v) for v in info]
        return g

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                load_library().hpdmk_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- reference interface
    def form_outgoing_pw(self):
        _check(load_library().hpdmk_form_outgoing_pw(self._h))

    def form_incoming_pw(self):
        _check(load_library().hpdmk_form_incoming_pw(self._h))

    def get_expansion(self, which, node=0):
        """which 0: root outgoing; 1: outgoing of `node`; 2: incoming of `node`
        (the reference's tree.outgoing_pw / incoming_pw, half grid)."""
        n = self.dim_window if which == 0 else self.dim_diff
        out = np.zeros(2 * n)
        cnt = C.c_int(0)
        _check(load_library().hpdmk_get_expansion(self._h, int(which), int(node), out.ctypes.data_as(_dp),
                                                  C.byref(cnt)))
        return out[0::2] + 1j * out[1::2]

    def energy_parts(self):
        e = C.c_double()
        parts = np.zeros(3)
        _check(load_library().hpdmk_eval_energy(self._h, C.byref(e), parts.ctypes.data_as(_dp)))
        return parts

    def eval_energy(self):
        return float(self.energy_parts().sum())

    def eval_energy_window(self):
        return float(self.energy_parts()[0])

    def eval_energy_diff(self):
        return float(self.energy_parts()[1])

    def eval_energy_res(self):
        return float(self.energy_parts()[2])

    def eval_shift_energy(self, i, dx, dy, dz, parts=False):
        e = C.c_double()
        p = np.zeros(5)
        _check(load_library().hpdmk_eval_shift_energy(self._h, int(i), float(dx), float(dy), float(dz), C.byref(e),
                                                      p.ctypes.data_as(_dp)))
        return (e.value, p) if parts else e.value

    def update_shift(self, i, dx, dy, dz):
        _check(load_library().hpdmk_update_shift(self._h, int(i), float(dx), float(dy), float(dz)))

    # ---- extensions
    def positions(self):
        r = np.zeros((self.n_src, 3))
        _check(load_library().hpdmk_get_positions(self._h, r.ctypes.data_as(_dp)))
        return r

    def eval_shift_energy_batch(self, idx, d):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        d, pd = _d(d)
        out = np.zeros(idx.size)
        _check(load_library().hpdmk_eval_shift_energy_batch(self._h, idx.size, idx.ctypes.data_as(_lp), pd,
                                                            out.ctypes.data_as(_dp)))
        return out

    def mc_load(self, idx, d, u):
        self._mc_idx = np.ascontiguousarray(idx, dtype=np.int64)
        d, pd = _d(d)
        u, pu = _d(u)
        _check(load_library().hpdmk_mc_load(self._h, self._mc_idx.size, self._mc_idx.ctypes.data_as(_lp), pd, pu))

    def mc_set_model(self, bjerrum=1.0, sigma=0.0):
        """Bjerrum scaling of the chain's energies and hard-sphere diameter (0 = none)."""
        _check(load_library().hpdmk_mc_set_model(self._h, float(bjerrum), float(sigma)))

    def mc_set_proposal(self, learned=False, weights=None, rf=0.0, max_mean=0.5, sigma_min=0.05, sigma_max=1.0):
        """learned: displacements from the MLP (mc_load's d then holds N(0,1) triples);
        rf > 0 records the origin features of every move."""
        wp = None
        if weights is not None:
            self._mlp_w = np.ascontiguousarray(weights, dtype=np.float32)
            if self._mlp_w.size != MLP_PARAMS:
                raise ValueError(f"MLP weights must hold {MLP_PARAMS} values")
            wp = self._mlp_w.ctypes.data_as(C.POINTER(C.c_float))
        _check(load_library().hpdmk_mc_set_proposal(self._h, int(bool(learned)), wp, float(rf), float(max_mean),
                                                    float(sigma_min), float(sigma_max)))

    def mc_proposal_results(self):
        """(displacements used, origin features, proposal log-ratios) of the last run."""
        n = self._mc_idx.size
        dl = np.zeros((n, 3))
        ft = np.zeros((n, MLP_FEATURES))
        H = np.zeros(n)
        _check(load_library().hpdmk_mc_proposal_results(self._h, dl.ctypes.data_as(_dp), ft.ctypes.data_as(_dp),
                                                        H.ctypes.data_as(_dp)))
        return dl, ft, H

    def mc_target_features(self):
        """Features at each move's target of the last learned-proposal run
        (the input of the reverse-move density)."""
        ft = np.zeros((self._mc_idx.size, MLP_FEATURES))
        _check(load_library().hpdmk_mc_target_features(self._h, ft.ctypes.data_as(_dp)))
        return ft

    def set_tiles(self, win=2, diff=8, near=8, apply=1024, batch=64):
        t = np.array([win, diff, near, apply], dtype=np.int32)
        _check(load_library().hpdmk_set_tiles(self._h, t.ctypes.data_as(_ip), int(batch)))

    def tiles(self):
        out = np.zeros(5, dtype=np.int32)
        _check(load_library().hpdmk_get_tiles(self._h, out.ctypes.data_as(_ip)))
        return dict(zip(["win", "diff", "near", "apply", "batch"], out.tolist()))

    def autotune(self, idx, d, u, beta=1.0):
        """Fastest launch shape for the batched chain, timed on a twin of the
        current state over the sample moves; returns (tiles, best ms)."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        d, pd = _d(d)
        u, pu = _d(u)
        out = np.zeros(5, dtype=np.int32)
        ms = C.c_double()
        _check(load_library().hpdmk_autotune(self._h, idx.size, idx.ctypes.data_as(_lp), pd, pu, float(beta),
                                             out.ctypes.data_as(_ip), C.byref(ms)))
        return dict(zip(["win", "diff", "near", "apply", "batch"], out.tolist())), ms.value

    def mc_run(self, beta=1.0, mode=0, batch=64):
        ms = C.c_double()
        _check(load_library().hpdmk_mc_run(self._h, float(beta), int(mode), int(batch), C.byref(ms)))
        return ms.value

    def mc_results(self):
        n = self._mc_idx.size
        dE = np.zeros(n)
        acc = np.zeros(n, dtype=np.int32)
        _check(load_library().hpdmk_mc_results(self._h, dE.ctypes.data_as(_dp), acc.ctypes.data_as(_ip)))
        return dE, acc

    def mc_chain(self, idx, d, u, beta=1.0, mode=0, batch=64):
        self.mc_load(idx, d, u)
        ms = self.mc_run(beta, mode, batch)
        dE, acc = self.mc_results()
        return dE, acc, ms

    def mc_stats(self):
        st = np.zeros((self._mc_idx.size, 8), dtype=np.int32)
        _check(load_library().hpdmk_mc_stats(self._h, st.ctypes.data_as(_ip)))
        return st

    def set_profiling(self, level=1):
        """0 off; 1 kernel times and per-move counts; 2 also device phase timers."""
        _check(load_library().hpdmk_set_profiling(self._h, int(level)))

    def kernel_times(self):
        ms = np.zeros(6)
        cnt = np.zeros(6, dtype=np.int32)
        _check(load_library().hpdmk_kernel_times(self._h, ms.ctypes.data_as(_dp), cnt.ctypes.data_as(_ip)))
        return dict(zip(["trial", "update", "pair_corr", "scan", "setup", "propose"], zip(ms.tolist(), cnt.tolist())))

    def phase_times(self):
        us = np.zeros(12)
        cnt = np.zeros(12, dtype=np.int32)
        _check(load_library().hpdmk_phase_times(self._h, us.ctypes.data_as(_dp), cnt.ctypes.data_as(_ip)))
        names = ["scan_assemble", "scan_decide", "plan_mark", "plan_compact", "plan_relocate", "trial_final",
                 "apply_lists", "apply_accumulate", "apply_writeback", "scan_prep", "scan_walk", "plan_unite"]
        return {n: (round(float(u), 2), int(c)) for n, u, c in zip(names, us, cnt)}

    def clear_error(self):
        _check(load_library().hpdmk_clear_error(self._h))


class HPDMKReplicas:
    """A set of independent systems of different sizes advanced together
    (one Metropolis move of every replica per step, steps replayed from a HIP
    graph).  systems: list of (r_src, charge); params: one HPDMKParams per
    replica.  tree(r) gives replica r as an HPDMKPtTree owned by the set."""

    def __init__(self, params, systems, precision=HPDMK_FP64):
        L = load_library()
        if len(params) != len(systems):
            raise ValueError("one HPDMKParams per replica")
        self.n = len(systems)
        self.n_src = np.array([len(q) for _, q in systems], dtype=np.int32)
        r_all, pr = _d(np.concatenate([np.asarray(r, dtype=np.float64).reshape(-1) for r, _ in systems]))
        q_all, pq = _d(np.concatenate([np.asarray(q, dtype=np.float64) for _, q in systems]))
        arr = (HPDMKParams * self.n)(*params)
        h = C.c_void_p()
        _check(L.hpdmk_replicas_create(self.n, C.cast(arr, C.c_void_p), precision, self.n_src.ctypes.data_as(_ip),
                                       pr, pq, C.byref(h)))
        self._h = h
        self._trees = []
        for k, (r, q) in enumerate(systems):
            th = C.c_void_p()
            _check(L.hpdmk_replicas_tree(h, k, C.byref(th)))
            self._trees.append(HPDMKPtTree(params[k], r, q, precision, _borrowed=th))
        self._n_moves = 0

    def close(self):
        if getattr(self, "_h", None):
            for t in self._trees:
                t.close()
            load_library().hpdmk_replicas_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tree(self, r):
        return self._trees[r]

    def form_pw(self):
        _check(load_library().hpdmk_replicas_form_pw(self._h))

    def mc_load(self, idx, d, u):
        """idx (n_rep, M) replica-local indices, d (n_rep, M, 3), u (n_rep, M)."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        M = idx.shape[1]
        d, pd = _d(np.asarray(d).reshape(-1))
        u, pu = _d(np.asarray(u).reshape(-1))
        if idx.shape[0] != self.n or d.size != 3 * self.n * M or u.size != self.n * M:
            raise ValueError("move arrays must be (n_rep, M), (n_rep, M, 3), (n_rep, M)")
        self._keep = (idx, d, u)
        _check(load_library().hpdmk_replicas_mc_load(self._h, M, idx.ctypes.data_as(_lp), pd, pu))
        self._n_moves = M

    def mc_run(self, beta=1.0, graph_steps=64):
        ms = C.c_double()
        _check(load_library().hpdmk_replicas_mc_run(self._h, float(beta), int(graph_steps), C.byref(ms)))
        return ms.value

    def mc_results(self):
        dE = np.zeros((self.n, self._n_moves))
        acc = np.zeros((self.n, self._n_moves), dtype=np.int32)
        _check(load_library().hpdmk_replicas_mc_results(self._h, dE.ctypes.data_as(_dp), acc.ctypes.data_as(_ip)))
        return dE, acc


class Ewald:
    """FP64 Ewald summation on the GPU; names follow include/ewald.hpp:14-57."""

    def __init__(self, L, s, alpha, eps, q, r):
        q, pq = _d(q)
        r, pr = _d(np.asarray(r).reshape(-1))
        if r.size != 3 * q.size:
            raise ValueError("r must hold 3 coordinates per charge")
        self.n = q.size
        h = C.c_void_p()
        _check(load_library().hpdmk_ewald_create(float(L), float(s), float(alpha), float(eps), q.size, pq, pr,
                                                 C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load_library().hpdmk_ewald_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute_energy(self, parts=False):
        e = C.c_double()
        p = np.zeros(3)
        _check(load_library().hpdmk_ewald_compute_energy(self._h, C.byref(e), p.ctypes.data_as(_dp)))
        return (e.value, p) if parts else e.value

    def compute_potential(self, points):
        pts, pp = _d(np.asarray(points).reshape(-1))
        phi = np.zeros(pts.size // 3)
        _check(load_library().hpdmk_ewald_compute_potential(self._h, phi.size, pp, phi.ctypes.data_as(_dp)))
        return phi

    def shift_chain(self, idx, d, apply=None):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        d, pd = _d(d)
        out = np.zeros(idx.size)
        ap = None
        if apply is not None:
            apply = np.ascontiguousarray(apply, dtype=np.int32)
            ap = apply.ctypes.data_as(_ip)
        _check(load_library().hpdmk_ewald_shift_chain(self._h, idx.size, idx.ctypes.data_as(_lp), pd, ap,
                                                      out.ctypes.data_as(_dp)))
        return out
