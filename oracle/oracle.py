"""ctypes front end of the CPU oracle (oracle/isg_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the GPU library, never
as the thing measured or shipped.  The product (inverse-solver-gpu_amd/) does
not import this module.

All complex arrays are numpy complex128 (interleaved re/im doubles).  Vectors
over unknowns are in the reference's tree order (Morton boxes, 8x8 pixels
per leaf box, setup_mlfma.cpp:124-153); `unkmap` maps raster index to tree
index as in setup_mlfma.cpp:177-185.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libisg_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_L = ctypes.POINTER(ctypes.c_long)
_V = ctypes.c_void_p


def build(force=False):
    """Compile the oracle with its Makefile (g++ -O3 -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-B" if force else "-s", "-C", _HERE],
                              stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_create": (_V, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int]),
            "orc_destroy": (None, [_V]),
            "orc_set_radius": (None, [_V, ctypes.c_double]),
            "orc_numunk": (ctypes.c_int, [_V]),
            "orc_numsamp": (ctypes.c_int, [_V, ctypes.c_int]),
            "orc_numterm": (ctypes.c_int, [_V, ctypes.c_int]),
            "orc_k0": (ctypes.c_double, [_V]),
            "orc_clusize": (ctypes.c_double, [_V, ctypes.c_int]),
            "orc_pos": (None, [_V, _D]),
            "orc_unkmap": (None, [_V, _I]),
            "orc_centres": (None, [_V, ctypes.c_int, _D]),
            "orc_near_list": (None, [_V, ctypes.c_int, _I]),
            "orc_far_list": (None, [_V, ctypes.c_int, _I]),
            "orc_traid": (None, [_V, ctypes.c_int, _I]),
            "orc_near_blocks": (None, [_V, _D]),
            "orc_coeff": (None, [_V, _D]),
            "orc_basis": (None, [_V, _D]),
            "orc_trans": (None, [_V, ctypes.c_int, _D]),
            "orc_shift": (None, [_V, ctypes.c_int, _D, _D]),
            "orc_interp": (None, [_V, ctypes.c_int, _D, _I]),
            "orc_anterp": (None, [_V, ctypes.c_int, _D, _I]),
            "orc_rx_interp": (None, [_V, _D, _I, _D, _I]),
            "orc_stage": (None, [_V, ctypes.c_int, ctypes.c_int, _D]),
            "orc_green": (None, [_V, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, _D]),
            "orc_hn": (None, [ctypes.c_int, ctypes.c_double, _D]),
            "orc_mlfma": (None, [_V, _D, _D]),
            "orc_mlfmah": (None, [_V, _D, _D]),
            "orc_aggregate": (None, [_V, _D, _D]),
            "orc_aggregateh": (None, [_V, _D, _D]),
            "orc_direct": (None, [_V, _D, _D, ctypes.c_int, ctypes.c_int]),
            "orc_p2p_only": (None, [_V, _D, _D]),
            "orc_bicgs": (ctypes.c_int, [_V, _D, _D, _D, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, _D]),
            "orc_dbim_create": (_V, [_V, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, _D]),
            "orc_dbim_destroy": (None, [_V]),
            "orc_dbim_mutate": (None, [_V, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "orc_dbim_iterate": (ctypes.c_double, [_V, _V]),
            "orc_dbim_hist": (None, [_V, _D]),
            "orc_dbim_get": (None, [_V, _V, ctypes.c_int, _D]),
            "orc_dbim_counts": (None, [_V, _L]),
            "orc_matvecs": (ctypes.c_long, [_V]),
            "orc_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


class Oracle:
    """One MLFMA operator set-up (setup_mlfma.cpp:50-640) on the CPU.

    level counts the root like the reference (mom.cpp:6, 137-138): the grid is
    (8 * 2**(level-1))**2 pixels of `res` wavelengths.  k0 = 2 pi scale.
    """

    def __init__(self, level, numrx=16, res=0.1, scale=1.0, beta=5, ninter=30, radius=0.0):
        """radius > 0: line-source transmitters and receivers on a circle of that
        radius (wavelengths) instead of plane waves and far-field receivers."""
        self.L = lib()
        self.level = level
        self.numrx = numrx
        self.res = res
        self.radius = radius
        self.h = self.L.orc_create(level, res, scale, beta, ninter, numrx)
        if radius > 0:
            self.L.orc_set_radius(self.h, radius)
        self.N = self.L.orc_numunk(self.h)
        self.k0 = self.L.orc_k0(self.h)
        self.numsamp = [self.L.orc_numsamp(self.h, i) for i in range(level)]
        self.numterm = [self.L.orc_numterm(self.h, i) for i in range(level)]
        self.clusize = [self.L.orc_clusize(self.h, i) for i in range(level)]
        self.nclus = [4 ** i for i in range(level)]
        self.ninter = ninter
        self.side = int(round(np.sqrt(self.N)))

    def __del__(self):
        try:
            if self.h:
                self.L.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # ---- geometry and tables
    def pos(self):
        out = np.empty(self.N, np.complex128)
        self.L.orc_pos(self.h, _dp(out))
        return out

    def unkmap(self):
        out = np.empty(self.N, np.int32)
        self.L.orc_unkmap(self.h, _ip(out))
        return out

    def centres(self, i):
        out = np.empty(self.nclus[i], np.complex128)
        self.L.orc_centres(self.h, i, _dp(out))
        return out

    def near_list(self, i):
        out = np.empty(self.nclus[i] * 9, np.int32)
        self.L.orc_near_list(self.h, i, _ip(out))
        return out.reshape(-1, 9)

    def far_list(self, i):
        out = np.empty(self.nclus[i] * 27, np.int32)
        self.L.orc_far_list(self.h, i, _ip(out))
        return out.reshape(-1, 27)

    def traid(self, i):
        out = np.empty(self.nclus[i] * 27, np.int32)
        self.L.orc_traid(self.h, i, _ip(out))
        return out.reshape(-1, 27)

    def near_blocks(self):
        out = np.empty(9 * 64 * 64, np.complex128)
        self.L.orc_near_blocks(self.h, _dp(out))
        return out.reshape(9, 64, 64)

    def coeff(self):
        S = self.numsamp[-1]
        out = np.empty(S * 64, np.complex128)
        self.L.orc_coeff(self.h, _dp(out))
        return out.reshape(S, 64)

    def basis(self):
        S = self.numsamp[-1]
        out = np.empty(64 * S, np.complex128)
        self.L.orc_basis(self.h, _dp(out))
        return out.reshape(64, S)

    def trans(self, i):
        S = self.numsamp[i]
        out = np.empty(49 * S, np.complex128)
        self.L.orc_trans(self.h, i, _dp(out))
        return out.reshape(49, S)

    def shift(self, i):
        S = self.numsamp[i]
        m = np.empty(4 * S, np.complex128)
        l = np.empty(4 * S, np.complex128)
        self.L.orc_shift(self.h, i, _dp(m), _dp(l))
        return m.reshape(4, S), l.reshape(4, S)

    def interp(self, i):
        S = self.numsamp[i]
        w = np.empty(S * self.ninter)
        idx = np.empty(S * self.ninter, np.int32)
        self.L.orc_interp(self.h, i, _dp(w), _ip(idx))
        return w.reshape(S, -1), idx.reshape(S, -1)

    def anterp(self, i):
        S = self.numsamp[i]
        w = np.empty(S * 2 * self.ninter)
        idx = np.empty(S * 2 * self.ninter, np.int32)
        self.L.orc_anterp(self.h, i, _dp(w), _ip(idx))
        return w.reshape(S, -1), idx.reshape(S, -1)

    def rx_interp(self):
        R, S0, ni = self.numrx, self.numsamp[0], self.ninter
        w = np.empty(R * ni)
        idx = np.empty(R * ni, np.int32)
        aw = np.empty(S0 * 2 * ni)
        aidx = np.empty(S0 * 2 * ni, np.int32)
        self.L.orc_rx_interp(self.h, _dp(w), _ip(idx), _dp(aw), _ip(aidx))
        return w.reshape(R, ni), idx.reshape(R, ni), aw.reshape(S0, -1), aidx.reshape(S0, -1)

    def stage(self, which, i):
        """which=0: multipoles, 1: locals of level i after the last operator call."""
        out = np.empty(self.nclus[i] * self.numsamp[i], np.complex128)
        self.L.orc_stage(self.h, which, i, _dp(out))
        return out.reshape(self.nclus[i], self.numsamp[i])

    def green(self, pt, pb):
        out = np.empty(2)
        self.L.orc_green(self.h, pt.real, pt.imag, pb.real, pb.imag, _dp(out))
        return complex(out[0], out[1])

    # ---- operators (tree-ordered vectors)
    def mlfma(self, x):
        x = _c128(x)
        r = np.empty(self.N, np.complex128)
        self.L.orc_mlfma(self.h, _dp(x), _dp(r))
        return r

    def mlfmah(self, x):
        x = _c128(x).copy()
        r = np.empty(self.N, np.complex128)
        self.L.orc_mlfmah(self.h, _dp(x), _dp(r))
        return r

    def aggregate(self, x):
        x = _c128(x)
        r = np.empty(self.numrx, np.complex128)
        self.L.orc_aggregate(self.h, _dp(x), _dp(r))
        return r

    def aggregateh(self, r):
        r = _c128(r)
        x = np.empty(self.N, np.complex128)
        self.L.orc_aggregateh(self.h, _dp(x), _dp(r))
        return x

    def direct(self, x, row0=0, nrow=None):
        x = _c128(x)
        if nrow is None:
            nrow = self.N - row0
        r = np.empty(nrow, np.complex128)
        self.L.orc_direct(self.h, _dp(x), _dp(r), row0, nrow)
        return r

    def p2p_only(self, x):
        x = _c128(x)
        r = np.empty(self.N, np.complex128)
        self.L.orc_p2p_only(self.h, _dp(x), _dp(r))
        return r

    def bicgs(self, g, b, x0=None, herm=False, tol=1e-4, maxit=300):
        g = _c128(g)
        b = _c128(b)
        x = np.zeros(self.N, np.complex128) if x0 is None else _c128(x0).copy()
        info = np.zeros(3)
        st = self.L.orc_bicgs(self.h, _dp(x), _dp(g), _dp(b), int(herm), tol, maxit, _dp(info))
        return x, {"status": st, "iters": int(info[0]), "matvecs": int(info[1]),
                   "relres": float(info[2])}


class Dbim:
    """DBIM / nonlinear-CG loop of mom.cpp:222-400 on the CPU oracle."""

    def __init__(self, orc, numtx, otrue_tree, regular=0.0, tol=1e-4, maxit=300):
        self.o = orc
        self.numtx = numtx
        ot = _c128(otrue_tree)
        self.h = orc.L.orc_dbim_create(orc.h, numtx, regular, tol, maxit, _dp(ot))

    def __del__(self):
        try:
            if self.h:
                self.o.L.orc_dbim_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def iterate(self):
        return self.o.L.orc_dbim_iterate(self.o.h, self.h)

    def mutate(self, alpha=1.0, beta=1.0, reg=1.0):
        """Plant a scale error in the step length, the Polak-Ribiere
        coefficient or the regulariser's gradient term (tests of the pins)."""
        self.o.L.orc_dbim_mutate(self.h, alpha, beta, reg)

    def hist(self):
        out = np.empty(5)
        self.o.L.orc_dbim_hist(self.h, _dp(out))
        return out

    def get(self, name):
        idx = {"o": 0, "od": 1, "ot": 2, "mes": 3, "mesd": 4, "inc": 5, "tot": 6}[name]
        n = {"o": self.o.N, "od": self.o.N, "ot": self.o.N,
             "mes": self.numtx * self.o.numrx, "mesd": self.numtx * self.o.numrx,
             "inc": self.numtx * self.o.N, "tot": self.numtx * self.o.N}[name]
        out = np.empty(n, np.complex128)
        self.o.L.orc_dbim_get(self.o.h, self.h, idx, _dp(out))
        return out

    def counts(self):
        out = np.zeros(2, np.int64)
        self.o.L.orc_dbim_counts(self.h, out.ctypes.data_as(_L))
        return {"solves": int(out[0]), "solve_iters": int(out[1])}
