"""Python front end of libisg.so, the MI355X MLFMA / BiCGSTAB / DBIM solver.

Mirrors the reference's operator surface (vars.h:14-26): ``mlfma``,
``mlfmah``, ``aggregate``, ``aggregateh`` and ``bicgs`` on complex128 vectors
in tree order, plus batched device-pointer variants and the DBIM driver.
Everything runs in the HIP library; this module only marshals arguments.
Import succeeds without a GPU, but creating a :class:`Solver` raises if the
library is missing or no HIP device is present -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libisg.so"))
HEADER_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "include", "isg.h"))

ISG_OK = 0
SOLVE_STATUS = {0: "converged", 1: "not converged", 2: "rho breakdown", 3: "omega breakdown"}


class IsgError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("level", ctypes.c_int),
        ("box", ctypes.c_int),
        ("res", ctypes.c_double),
        ("scale", ctypes.c_double),
        ("beta", ctypes.c_int),
        ("ninter", ctypes.c_int),
        ("numrx", ctypes.c_int),
        ("numtx", ctypes.c_int),
        ("toli", ctypes.c_double),
        ("iti", ctypes.c_int),
        ("regular", ctypes.c_double),
        ("max_rhs", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("mixed_margin", ctypes.c_double),
        ("graph", ctypes.c_int),
        ("radius", ctypes.c_double),
        ("products", ctypes.c_int),
        ("solve_start", ctypes.c_int),
        ("setup", ctypes.c_int),
    ]


class SolveInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iters", ctypes.c_int), ("relres", ctypes.c_double)]


class DbimStats(ctypes.Structure):
    _fields_ = [
        ("alpha", ctypes.c_double),
        ("misfit", ctypes.c_double),
        ("object_error", ctypes.c_double),
        ("object_cost", ctypes.c_double),
        ("gradient_norm2", ctypes.c_double),
        ("object_norm2", ctypes.c_double),
        ("beta_re", ctypes.c_double),
        ("beta_im", ctypes.c_double),
        ("solves", ctypes.c_longlong),
        ("solve_iters", ctypes.c_longlong),
        ("matvecs", ctypes.c_longlong),
        ("not_converged", ctypes.c_longlong),
        ("matvecs_fp32", ctypes.c_longlong),
        ("seconds", ctypes.c_double),
        ("field_iters", ctypes.c_longlong),
        ("field_seconds", ctypes.c_double),
        ("lambda_abs", ctypes.c_double),
        ("net_seconds", ctypes.c_double),
        ("start_ratio", ctypes.c_double),
    ]


_V = ctypes.c_void_p
_P = ctypes.POINTER
_D = _P(ctypes.c_double)
_I = _P(ctypes.c_int)

_SIGS = {
    "isg_config_default": (None, [_P(Config)]),
    "isg_create": (ctypes.c_int, [_P(Config), _P(_V)]),
    "isg_destroy": (None, [_V]),
    "isg_last_error": (ctypes.c_char_p, [_V]),
    "isg_numunk": (ctypes.c_int, [_V]),
    "isg_level": (ctypes.c_int, [_V]),
    "isg_numsamp": (ctypes.c_int, [_V, ctypes.c_int]),
    "isg_k0": (ctypes.c_double, [_V]),
    "isg_geometry": (ctypes.c_int, [_V, _D, _I]),
    "isg_stream": (_V, [_V]),
    "isg_mlfma": (ctypes.c_int, [_V, _D, _D]),
    "isg_mlfmah": (ctypes.c_int, [_V, _D, _D]),
    "isg_mlfma_many": (ctypes.c_int, [_V, _D, _D, ctypes.c_int]),
    "isg_aggregate": (ctypes.c_int, [_V, _D, _D]),
    "isg_aggregateh": (ctypes.c_int, [_V, _D, _D]),
    "isg_bicgs": (ctypes.c_int, [_V, _D, _D, _D, ctypes.c_int, _P(SolveInfo)]),
    "isg_mlfma_dev": (ctypes.c_int, [_V, _V, _V, ctypes.c_int]),
    "isg_aggregate_dev": (ctypes.c_int, [_V, _V, _V, ctypes.c_int]),
    "isg_mlfma_f32_dev": (ctypes.c_int, [_V, _V, _V, ctypes.c_int]),
    "isg_synchronize": (ctypes.c_int, [_V]),
    "isg_aggregateh_dev": (ctypes.c_int, [_V, _V, _V, ctypes.c_int]),
    "isg_bicgs_dev": (ctypes.c_int, [_V, _V, _V, ctypes.c_long, _V, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _P(SolveInfo)]),
    "isg_bicgs_dev_static": (ctypes.c_int, [_V, _V, _V, ctypes.c_long, _V, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, _P(SolveInfo)]),
    "isg_stage": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]),
    "isg_counters": (ctypes.c_int, [_V, _P(ctypes.c_longlong)]),
    "isg_dbim_setup": (ctypes.c_int, [_V, _D, ctypes.c_int, _P(DbimStats)]),
    "isg_dbim_setup_measured": (ctypes.c_int, [_V, _D, _D, ctypes.c_int, _P(DbimStats)]),
    "isg_profile": (ctypes.c_int, [_V, ctypes.c_int]),
    "isg_profile_read": (ctypes.c_int, [_V, ctypes.c_int, _D, _P(ctypes.c_longlong)]),
    "isg_dbim_iterate": (ctypes.c_int, [_V, _P(DbimStats)]),
    "isg_dbim_restart": (ctypes.c_int, [_V]),
    "isg_set_progress": (ctypes.c_int, [_V, _V, _V]),
    "isg_dbim_get": (ctypes.c_int, [_V, ctypes.c_int, _D]),
    "isg_dbim_set": (ctypes.c_int, [_V, ctypes.c_int, _D]),
    "isg_warmnet_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_ulonglong, _P(_V)]),
    "isg_warmnet_destroy": (None, [_V]),
    "isg_warmnet_last_error": (ctypes.c_char_p, [_V]),
    "isg_warmnet_num_params": (ctypes.c_long, []),
    "isg_warmnet_get_params": (ctypes.c_int, [_V, _P(ctypes.c_float)]),
    "isg_warmnet_set_params": (ctypes.c_int, [_V, _P(ctypes.c_float)]),
    "isg_warmnet_get_grads": (ctypes.c_int, [_V, _P(ctypes.c_float)]),
    "isg_warmnet_set_scales": (ctypes.c_int, [_V, _P(ctypes.c_float)]),
    "isg_warmnet_get_scales": (ctypes.c_int, [_V, _P(ctypes.c_float)]),
    "isg_warmnet_forward_dev": (ctypes.c_int, [_V, _V, _V, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "isg_warmnet_loss_grad_dev": (ctypes.c_int, [_V, _V, _V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]),
    "isg_warmnet_forward": (ctypes.c_int, [_V, _P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int]),
    "isg_warmnet_loss_grad": (ctypes.c_int, [_V, _P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, _D]),
    "isg_warmnet_adam": (ctypes.c_int, [_V, ctypes.c_float]),
    "isg_warmnet_save": (ctypes.c_int, [_V, ctypes.c_char_p]),
    "isg_warmnet_load": (ctypes.c_int, [_V, ctypes.c_char_p]),
    "isg_dbim_set_warmnet": (ctypes.c_int, [_V, _V]),
    "isg_autotune": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int]),
    "isg_tuning_set": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "isg_tuning_get": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I]),
    "isg_tuning_save": (ctypes.c_int, [_V, ctypes.c_char_p]),
    "isg_tuning_load": (ctypes.c_int, [_V, ctypes.c_char_p]),
    "isg_dbim_write_cg": (ctypes.c_int, [_V, ctypes.c_char_p, ctypes.c_int]),
    "isg_dbim_read_cg": (ctypes.c_int, [_V, ctypes.c_char_p]),
    "isg_hist_append": (ctypes.c_int, [ctypes.c_char_p, _P(DbimStats), ctypes.c_int]),
    "isg_phantom_read_txt": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_double, _D]),
    "isg_phantom_write_txt": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_double, _D]),
    "isg_raster_bin_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _D]),
    "isg_raster_bin_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _D]),
    "isg_phantom_blobs": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_ulonglong, _D]),
    "isg_phantom_ellipses": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_ulonglong,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]),
}

_lib = None


def declared_symbols():
    """Names of every function declared in include/isg.h."""
    import re
    txt = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(isg_[a-z0-9_]+)\s*\(", txt)))


def load(path=None):
    """Load libisg.so (build it first with inverse-solver-gpu_amd/build.py)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise IsgError("libisg.so not built at %s; run inverse-solver-gpu_amd/build.py" % p)
    L = ctypes.CDLL(p)
    for name, (res, args) in _SIGS.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(_D)


def _c128(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if n is not None and a.size != n:
        raise ValueError("expected %d complex values, got %d" % (n, a.size))
    return a


def default_config(**kw):
    c = Config()
    load().isg_config_default(ctypes.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Solver:
    """One operator + solver context on the GPU (setup_mlfma + setup_gpu +
    setup_bicgs of the reference, mom.cpp:184-190)."""

    def __init__(self, tuning=True, **kw):
        """tuning: load the committed kernel-variant cache (when present) or a
        given cache path; False keeps the built-in rules."""
        self.lib = load()
        self.cfg = default_config(**kw)
        h = _V()
        rc = self.lib.isg_create(ctypes.byref(self.cfg), ctypes.byref(h))
        self.h = h
        if rc != ISG_OK:
            msg = self.lib.isg_last_error(h).decode() if h else "invalid configuration"
            self.close()
            raise IsgError("isg_create failed (%d): %s" % (rc, msg))
        self.N = self.lib.isg_numunk(h)
        self.level = self.lib.isg_level(h)
        self.numsamp = [self.lib.isg_numsamp(h, i) for i in range(self.level)]
        self.k0 = self.lib.isg_k0(h)
        self.numrx = self.cfg.numrx
        if tuning is True:
            if os.path.exists(self.TUNING_FILE):
                self.tuning_load(self.TUNING_FILE)
        elif tuning:
            self.tuning_load(tuning)

    def close(self):
        if getattr(self, "h", None):
            self.lib.isg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc, what):
        if rc < 0:
            raise IsgError("%s failed (%d): %s" % (what, rc, self.lib.isg_last_error(self.h).decode()))
        return rc

    def geometry(self):
        pos = np.empty(self.N, np.complex128)
        r2t = np.empty(self.N, np.int32)
        self._err(self.lib.isg_geometry(self.h, _dp(pos), r2t.ctypes.data_as(_I)), "isg_geometry")
        return pos, r2t

    def mlfma(self, x):
        x = _c128(x, self.N)
        r = np.empty(self.N, np.complex128)
        self._err(self.lib.isg_mlfma(self.h, _dp(x), _dp(r)), "isg_mlfma")
        return r

    def mlfma_many(self, X):
        """Batched r[k] = G x[k] for an (nrhs, N) array, one pass."""
        X = np.ascontiguousarray(X, dtype=np.complex128)
        nrhs = X.shape[0]
        if X.shape != (nrhs, self.N):
            raise ValueError("expected (nrhs, %d)" % self.N)
        R = np.empty_like(X)
        self._err(self.lib.isg_mlfma_many(self.h, _dp(X), _dp(R), nrhs), "isg_mlfma_many")
        return R

    def mlfmah(self, x):
        x = _c128(x, self.N)
        r = np.empty(self.N, np.complex128)
        self._err(self.lib.isg_mlfmah(self.h, _dp(x), _dp(r)), "isg_mlfmah")
        return r

    def aggregate(self, x):
        x = _c128(x, self.N)
        r = np.empty(self.numrx, np.complex128)
        self._err(self.lib.isg_aggregate(self.h, _dp(x), _dp(r)), "isg_aggregate")
        return r

    def aggregateh(self, r):
        r = _c128(r, self.numrx)
        x = np.empty(self.N, np.complex128)
        self._err(self.lib.isg_aggregateh(self.h, _dp(x), _dp(r)), "isg_aggregateh")
        return x

    def bicgs(self, o, b, x0=None, herm=False):
        """Solve (I - G diag(o)) x = b (or its adjoint); o is contrast * k0^2."""
        o = _c128(o, self.N)
        b = _c128(b, self.N)
        x = np.zeros(self.N, np.complex128) if x0 is None else _c128(x0, self.N).copy()
        info = SolveInfo()
        rc = self._err(self.lib.isg_bicgs(self.h, _dp(x), _dp(o), _dp(b), int(herm), ctypes.byref(info)),
                       "isg_bicgs")
        return x, {"status": info.status, "iters": info.iters, "relres": info.relres, "rc": rc}

    def stage(self, which, level, rhs=0):
        out = np.empty(4 ** level * self.numsamp[level], np.complex128)
        self._err(self.lib.isg_stage(self.h, which, level, rhs, _dp(out)), "isg_stage")
        return out.reshape(4 ** level, self.numsamp[level])

    KERNELS = {"near": 0, "m2l": 1, "bicgs_update": 2, "near_fp32": 3, "m2l_fp32": 4, "p2m": 5, "interp": 6,
               "receivers": 7}

    def profile(self, enable=True):
        """Start (clearing totals) or stop live HIP-event timing of the hot kernels."""
        self._err(self.lib.isg_profile(self.h, int(enable)), "isg_profile")

    def profile_read(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._err(self.lib.isg_profile_read(self.h, self.KERNELS[kernel], ctypes.byref(ms), ctypes.byref(n)),
                  "isg_profile_read")
        return ms.value, n.value

    # ---- autotuner (stage: 0 near, 1 p2m, 2 m2l, 3 m2m, 4 l2l, 5 leaf = fused leaf M2L + last L2L, 0 / 1)
    TUNE_STAGES = {"near": 0, "p2m": 1, "m2l": 2, "m2m": 3, "l2l": 4, "leaf": 5}
    TUNING_FILE = os.path.normpath(os.path.join(_HERE, "..", "tuning", "mi355x.txt"))

    def autotune(self, nrhs=1, reps=3):
        self._err(self.lib.isg_autotune(self.h, int(nrhs), int(reps)), "isg_autotune")

    def tuning_set(self, precision, batch_class, stage, level, value):
        self._err(self.lib.isg_tuning_set(self.h, precision, batch_class, self.TUNE_STAGES[stage], level, value),
                  "isg_tuning_set")

    def tuning_get(self, precision, batch_class, stage, level=0):
        v = ctypes.c_int()
        self._err(self.lib.isg_tuning_get(self.h, precision, batch_class, self.TUNE_STAGES[stage], level,
                                          ctypes.byref(v)), "isg_tuning_get")
        return v.value

    def tuning_save(self, path):
        self._err(self.lib.isg_tuning_save(self.h, str(path).encode()), "isg_tuning_save")

    def tuning_load(self, path=None):
        """Load a tuning cache (default: the committed MI355X cache)."""
        self._err(self.lib.isg_tuning_load(self.h, str(path or self.TUNING_FILE).encode()), "isg_tuning_load")

    _PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int)

    def set_progress(self, fn=None):
        """fn(done, total) after each batch of the measurement synthesis of a
        later Dbim(...) without measurements; None turns it off."""
        self._progress_cb = self._PROGRESS(lambda user, done, total: fn(done, total)) if fn else None
        ptr = ctypes.cast(self._progress_cb, ctypes.c_void_p) if fn else None
        self._err(self.lib.isg_set_progress(self.h, ptr, None), "isg_set_progress")

    def counters(self):
        out = (ctypes.c_longlong * 5)()
        self._err(self.lib.isg_counters(self.h, out), "isg_counters")
        return {"matvecs": out[0], "solves": out[1], "solve_iters": out[2], "matvecs_fp32": out[3],
                "refinements": out[4]}

    # ---- device-pointer surface (pointers are integers, e.g. torch data_ptr())
    def mlfma_dev(self, x_ptr, r_ptr, nrhs=1):
        return self._err(self.lib.isg_mlfma_dev(self.h, _V(x_ptr), _V(r_ptr), nrhs), "isg_mlfma_dev")

    def synchronize(self):
        """Wait for the work queued on the context's stream by the *_dev calls."""
        self._err(self.lib.isg_synchronize(self.h), "isg_synchronize")

    def mlfma_f32_dev(self, x_ptr, r_ptr, nrhs=1):
        """FP32 operator on interleaved complex64 device vectors (precision 1 contexts)."""
        return self._err(self.lib.isg_mlfma_f32_dev(self.h, _V(x_ptr), _V(r_ptr), nrhs), "isg_mlfma_f32_dev")

    def aggregate_dev(self, x_ptr, r_ptr, nrhs=1):
        return self._err(self.lib.isg_aggregate_dev(self.h, _V(x_ptr), _V(r_ptr), nrhs), "isg_aggregate_dev")

    def aggregateh_dev(self, x_ptr, r_ptr, nrhs=1):
        return self._err(self.lib.isg_aggregateh_dev(self.h, _V(x_ptr), _V(r_ptr), nrhs), "isg_aggregateh_dev")

    def bicgs_dev_static(self, x_ptr, o_ptr, b_ptr, nrhs=1, herm=False, o_stride=0, x_zero=False):
        """bicgs_dev with the solver's scratch in the library's static device
        array (no device allocation; calls overlapping on one GPU race)."""
        infos = (SolveInfo * nrhs)()
        rc = self._err(self.lib.isg_bicgs_dev_static(self.h, _V(x_ptr), _V(o_ptr), o_stride, _V(b_ptr),
                                                     int(herm), nrhs, int(x_zero), infos), "isg_bicgs_dev_static")
        return rc, [{"status": i.status, "iters": i.iters, "relres": i.relres} for i in infos]

    def bicgs_dev(self, x_ptr, o_ptr, b_ptr, nrhs=1, herm=False, o_stride=0, x_zero=False):
        infos = (SolveInfo * nrhs)()
        rc = self._err(self.lib.isg_bicgs_dev(self.h, _V(x_ptr), _V(o_ptr), o_stride, _V(b_ptr), int(herm),
                                              nrhs, int(x_zero), infos), "isg_bicgs_dev")
        return rc, [{"status": i.status, "iters": i.iters, "relres": i.relres} for i in infos]


_DBIM_FIELDS = {"o": 0, "od": 1, "ot": 2, "mes": 3, "mesd": 4, "inc": 5, "tot": 6, "dtot": 8}


def _stats_dict(st):
    return {f: getattr(st, f) for f, _ in DbimStats._fields_}


class Dbim:
    """DBIM / nonlinear-CG inversion on the GPU (mom.cpp:222-400).

    ``solver`` fixes the operator, the transmitter count (numtx), the batch
    size (max_rhs) and the BiCGSTAB tolerance (toli, iti)."""

    def __init__(self, solver, o_true_tree, field_start=0, measurements=None):
        self.s = solver
        st = DbimStats()
        if measurements is None:
            o = _c128(o_true_tree, solver.N)
            solver._err(solver.lib.isg_dbim_setup(solver.h, _dp(o), int(field_start), ctypes.byref(st)),
                        "isg_dbim_setup")
        else:
            m = _c128(measurements, solver.cfg.numtx * solver.cfg.numrx)
            optr = None if o_true_tree is None else _dp(_c128(o_true_tree, solver.N))
            solver._err(solver.lib.isg_dbim_setup_measured(solver.h, _dp(m), optr, int(field_start),
                                                           ctypes.byref(st)), "isg_dbim_setup_measured")
        self.setup_stats = _stats_dict(st)

    def iterate(self):
        st = DbimStats()
        self.s._err(self.s.lib.isg_dbim_iterate(self.s.h, ctypes.byref(st)), "isg_dbim_iterate")
        return _stats_dict(st)

    def restart(self):
        """Back to the state after setup: the next iterate() is iteration 1
        again (the first gradient is kept, not recomputed)."""
        self.s._err(self.s.lib.isg_dbim_restart(self.s.h), "isg_dbim_restart")

    def _size(self, name):
        T, R, N = self.s.cfg.numtx, self.s.cfg.numrx, self.s.N
        return {"o": N, "od": N, "ot": N, "mes": T * R, "mesd": T * R, "inc": T * N, "tot": T * N,
                "dtot": T * N}[name]

    def get(self, name):
        out = np.empty(self._size(name), np.complex128)
        self.s._err(self.s.lib.isg_dbim_get(self.s.h, _DBIM_FIELDS[name], _dp(out)), "isg_dbim_get")
        return out

    def set(self, name, value):
        v = _c128(value, self._size(name))
        self.s._err(self.s.lib.isg_dbim_set(self.s.h, _DBIM_FIELDS[name], _dp(v)), "isg_dbim_set")

    def write_cg(self, path, which="o"):
        """cg_<k>.bin (which 'o') or cg_ref.bin (which 'true'): raster complex doubles."""
        code = {"o": 0, "true": 7}[which]
        self.s._err(self.s.lib.isg_dbim_write_cg(self.s.h, str(path).encode(), code), "isg_dbim_write_cg")

    def read_cg(self, path):
        """Set the current contrast from a raster cg_<k>.bin."""
        self.s._err(self.s.lib.isg_dbim_read_cg(self.s.h, str(path).encode()), "isg_dbim_read_cg")

    def set_warmnet(self, net):
        """Network for field_start 3 (kept referenced by this object)."""
        self._net = net
        self.s._err(self.s.lib.isg_dbim_set_warmnet(self.s.h, net.h), "isg_dbim_set_warmnet")


def _fp(a):
    return a.ctypes.data_as(_P(ctypes.c_float))


class WarmNet:
    """Warm-start network on the GPU (csrc/isg_warmnet.hip): a 4-level conv
    encoder-decoder from 8 channels (Re/Im contrast, Re/Im incident field,
    Re/Im first-order start, Re/Im its first-order step; fields divided by the
    transmitter's incident-field rms) to a 2-channel correction of the
    first-order start.  Images are raster NHWC FP32 with H, W multiples of 16
    (the network runs on their space-to-depth-by-2 form)."""

    NUM_SCALES = 9
    IN_CH = 8

    def __init__(self, max_batch=4, max_pixels=512 * 512, seed=1, path=None):
        self.lib = load()
        self.h = _V()
        rc = self.lib.isg_warmnet_create(int(max_batch), int(max_pixels), int(seed), ctypes.byref(self.h))
        if rc != 0:
            msg = self.lib.isg_warmnet_last_error(self.h).decode() if self.h else "isg_warmnet_create failed"
            if self.h:
                self.lib.isg_warmnet_destroy(self.h)
                self.h = None
            raise IsgError(msg)
        self.max_batch, self.max_pixels = max_batch, max_pixels
        if path is not None:
            self.load(path)

    def close(self):
        if self.h:
            self.lib.isg_warmnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc, what):
        if rc != 0:
            raise IsgError("%s: %s" % (what, self.lib.isg_warmnet_last_error(self.h).decode()))

    @staticmethod
    def num_params():
        return int(load().isg_warmnet_num_params())

    @property
    def params(self):
        out = np.empty(self.num_params(), np.float32)
        self._err(self.lib.isg_warmnet_get_params(self.h, _fp(out)), "isg_warmnet_get_params")
        return out

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.num_params():
            raise ValueError("expected %d parameters" % self.num_params())
        self._err(self.lib.isg_warmnet_set_params(self.h, _fp(w)), "isg_warmnet_set_params")

    def grads(self):
        out = np.empty(self.num_params(), np.float32)
        self._err(self.lib.isg_warmnet_get_grads(self.h, _fp(out)), "isg_warmnet_get_grads")
        return out

    @property
    def scales(self):
        out = np.empty(self.NUM_SCALES, np.float32)
        self._err(self.lib.isg_warmnet_get_scales(self.h, _fp(out)), "isg_warmnet_get_scales")
        return out

    @scales.setter
    def scales(self, s):
        s = np.ascontiguousarray(s, dtype=np.float32)
        if s.size != self.NUM_SCALES:
            raise ValueError("expected %d scales" % self.NUM_SCALES)
        self._err(self.lib.isg_warmnet_set_scales(self.h, _fp(s)), "isg_warmnet_set_scales")

    def forward_dev(self, in_ptr, out_ptr, batch, H, W):
        self._err(self.lib.isg_warmnet_forward_dev(self.h, _V(in_ptr), _V(out_ptr), batch, H, W),
                  "isg_warmnet_forward_dev")

    def loss_grad_dev(self, in_ptr, target_ptr, batch, H, W):
        loss = ctypes.c_double()
        self._err(self.lib.isg_warmnet_loss_grad_dev(self.h, _V(in_ptr), _V(target_ptr), batch, H, W,
                                                     ctypes.byref(loss)), "isg_warmnet_loss_grad_dev")
        return loss.value

    def adam(self, lr):
        self._err(self.lib.isg_warmnet_adam(self.h, float(lr)), "isg_warmnet_adam")

    def save(self, path):
        self._err(self.lib.isg_warmnet_save(self.h, str(path).encode()), "isg_warmnet_save")

    def load(self, path):
        self._err(self.lib.isg_warmnet_load(self.h, str(path).encode()), "isg_warmnet_load")

    # host arrays (staged through device buffers the library keeps)
    def forward(self, x):
        """x: (batch, H, W, 8) float32 host array -> (batch, H, W, 2)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty(x.shape[:3] + (2,), np.float32)
        self._err(self.lib.isg_warmnet_forward(self.h, _fp(x), _fp(out), *x.shape[:3]), "isg_warmnet_forward")
        return out

    def loss_grad(self, x, target):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = ctypes.c_double()
        self._err(self.lib.isg_warmnet_loss_grad(self.h, _fp(x), _fp(t), *x.shape[:3], ctypes.byref(loss)),
                  "isg_warmnet_loss_grad")
        return loss.value


def hist_append(path, stats, create=False):
    """Append the hist.txt line of a stats dict (setup_born.cpp:239)."""
    st = DbimStats()
    for f, _ in DbimStats._fields_:
        if f in stats:
            setattr(st, f, stats[f])
    if load().isg_hist_append(str(path).encode(), ctypes.byref(st), 1 if create else 0) != 0:
        raise IsgError("cannot write %s" % path)


def read_phantom_txt(path, side, scale=0.1):
    """Reference phantom text file -> raster (side x side) complex contrast."""
    out = np.empty(side * side, np.complex128)
    if load().isg_phantom_read_txt(str(path).encode(), side * side, scale, _dp(out)) != 0:
        raise IsgError("cannot read phantom %s" % path)
    return out.reshape(side, side)


def write_phantom_txt(path, raster, scale=0.1):
    """Raster contrast -> reference phantom text file (real part / scale)."""
    r = _c128(raster)
    if load().isg_phantom_write_txt(str(path).encode(), r.size, scale, _dp(r)) != 0:
        raise IsgError("cannot write phantom %s" % path)


def read_raster_bin(path, side):
    """cg_<k>.bin / cg_ref.bin -> raster (side x side) complex."""
    out = np.empty(side * side, np.complex128)
    if load().isg_raster_bin_read(str(path).encode(), side * side, _dp(out)) != 0:
        raise IsgError("cannot read %s" % path)
    return out.reshape(side, side)


def phantom_blobs(side, n, seed, res=0.1):
    """Raster (side x side) complex contrast: n seeded Gaussian blobs."""
    out = np.empty(side * side, np.complex128)
    if load().isg_phantom_blobs(side, res, n, seed, _dp(out)) != 0:
        raise ValueError("bad phantom arguments")
    return out.reshape(side, side)


def phantom_ellipses(side, n, seed, res=0.1, vmin=0.05, vmax=0.8, imag=0.02):
    """Raster (side x side) complex contrast: n seeded ellipses."""
    out = np.empty(side * side, np.complex128)
    if load().isg_phantom_ellipses(side, res, n, seed, vmin, vmax, imag, _dp(out)) != 0:
        raise ValueError("bad phantom arguments")
    return out.reshape(side, side)


def raster_to_tree(raster, r2t):
    """Reorder a raster image into the solver's tree order (setup_born.cpp:109)."""
    flat = np.ascontiguousarray(raster).ravel()
    out = np.empty_like(flat)
    out[r2t] = flat
    return out


def tree_to_raster(tree, r2t, side):
    """Inverse of raster_to_tree (setup_born.cpp:117, :214)."""
    return np.asarray(tree)[r2t].reshape(side, side)
