"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker, never as the product.
The oracle restates code/forcefield0.c and code/forcefield1.c of
hkayabilisim/msmpb (see msm_oracle.c for per-function citations) and, for
the PME comparator, code/pme0.c and code/pme1.c (pme_oracle.c).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile liborc.so with make (gcc, plain C11)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_build.restype = ctypes.c_void_p
        L.orc_build.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _ip,
                                ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                ctypes.c_char_p]
        L.orc_energy.restype = ctypes.c_double
        L.orc_energy.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_info.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_grid_dims.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.orc_stencil_dims.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.orc_copy_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_copy_stencil.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_copy_table.restype = ctypes.c_int
        L.orc_copy_table.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_bspline.restype = ctypes.c_double
        L.orc_bspline.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_top_level.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_restrict.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_prolong.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_grid2grid.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.orc_errest.restype = ctypes.c_double
        L.orc_errest.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_deltaF.restype = ctypes.c_double
        L.orc_deltaF.argtypes = [ctypes.c_void_p]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_build_pairs.restype = ctypes.c_int
        L.orc_build_pairs.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.pmo_build.restype = ctypes.c_void_p
        L.pmo_build.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _ip, ctypes.c_double,
                                ctypes.c_double, ctypes.c_char_p]
        L.pmo_energy.restype = ctypes.c_double
        L.pmo_energy.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.pmo_info.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.pmo_copy_khat.argtypes = [ctypes.c_void_p, _dp]
        L.pmo_copy_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.pmo_bspline.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.pmo_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


class Oracle:
    """FP64 CPU MSM with the reference's FF_build parameter rules.

    nu, L, a0 = -1 and Mtop = None select the reference's defaults.
    compat = 0: corrected algorithm; 1: reference HEAD index layout.
    fft = False: the reference's DFT build (FF_set_FFT(false)).
    """

    def __init__(self, pos, q, edges, nu=-1, L=-1, Mtop=None, a0=-1.0, estErr=-1.0, compat=0,
                 fft=True):
        self.q = np.ascontiguousarray(q, dtype=np.float64)
        self.N = self.q.shape[0]
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.N, 3)
        self.edges = np.ascontiguousarray(edges, dtype=np.float64).reshape(3, 3)
        err = ctypes.create_string_buffer(512)
        if Mtop is not None:
            if np.isscalar(Mtop):
                Mtop = [int(Mtop)] * 3
            mt = (ctypes.c_int * 3)(*[int(m) for m in Mtop])
            mtp = ctypes.cast(mt, _ip)
        else:
            mtp = None
        self._h = lib().orc_build(self.N, _p(self.q), _p(pos), _p(self.edges), int(nu), int(L),
                                  mtp, float(a0), float(estErr), int(compat), 1 if fft else 0,
                                  err)
        if not self._h:
            raise ValueError(err.value.decode())
        info = (ctypes.c_int * 8)()
        dinfo = (ctypes.c_double * 5)()
        lib().orc_info(self._h, info, dinfo)
        self.nu, self.L = info[0], info[1]
        self.Mtop = (info[2], info[3], info[4])
        self.nLim = info[5]
        self.a0, self.relCutoff, self.coeff1, self.coeff2, self.detA = list(dinfo)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            try:
                _lib.orc_free(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def energy(self, pos, q=None, want_phi=True):
        """Returns (E, F[N,3], phi[N] or None, terms[3])."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.N, 3)
        qq = self.q if q is None else np.ascontiguousarray(q, dtype=np.float64)
        F = np.zeros((self.N, 3))
        phi = np.zeros(self.N) if want_phi else None
        terms = np.zeros(3)
        E = lib().orc_energy(self._h, _p(pos), _p(qq), _p(F), _p(phi) if want_phi else None,
                             _p(terms))
        return E, F, phi, terms

    def build_pairs(self, pos, margin=0.0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.N, 3)
        return lib().orc_build_pairs(self._h, _p(pos), float(margin))

    def pair_count(self):
        info = (ctypes.c_int * 8)()
        lib().orc_info(self._h, info, None)
        return info[6]

    def grid_dims(self, l):
        d = (ctypes.c_int * 3)()
        lib().orc_grid_dims(self._h, l, d)
        return tuple(d)

    def stencil_dims(self, l):
        d = (ctypes.c_int * 3)()
        lib().orc_stencil_dims(self._h, l, d)
        return tuple(d)

    def grid(self, l, which):
        out = np.zeros(self.grid_dims(l))
        lib().orc_copy_grid(self._h, l, which, _p(out))
        return out

    def stencil(self, l):
        out = np.zeros(self.stencil_dims(l))
        lib().orc_copy_stencil(self._h, l, _p(out))
        return out

    def table(self, which, l=0, a=0):
        out = np.zeros(4096)
        n = lib().orc_copy_table(self._h, which, l, a, _p(out))
        return out[:n].copy()

    def bspline(self, i, t):
        d = ctypes.c_double()
        v = lib().orc_bspline(self._h, i, t, ctypes.byref(d))
        return v, d.value

    def top_level(self, qL):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        out = np.zeros_like(qL)
        lib().orc_top_level(self._h, _p(qL), _p(out))
        return out

    def restrict(self, l, fine):
        fine = np.ascontiguousarray(fine, dtype=np.float64)
        out = np.zeros(self.grid_dims(l))
        lib().orc_restrict(self._h, l, _p(fine), _p(out))
        return out

    def prolong(self, l, coarse):
        coarse = np.ascontiguousarray(coarse, dtype=np.float64)
        out = np.zeros(self.grid_dims(l - 1))
        lib().orc_prolong(self._h, l, _p(coarse), _p(out))
        return out

    def grid2grid(self, l, ql):
        ql = np.ascontiguousarray(ql, dtype=np.float64)
        out = np.zeros_like(ql)
        lib().orc_grid2grid(self._h, l, _p(ql), _p(out))
        return out

    def errest(self, nu):
        return lib().orc_errest(self._h, nu)

    def deltaF(self):
        return lib().orc_deltaF(self._h)


class PmeOracle:
    """FP64 CPU smooth PME with the reference pme's parameter rules.

    nu, a0, tolDir <= 0 and M = None select the reference's defaults
    (order 6, cutoff 8 A, direct tolerance 1e-5, grid from its spacing
    table).
    """

    def __init__(self, pos, q, edges, nu=-1, M=None, a0=-1.0, tolDir=-1.0):
        self.q = np.ascontiguousarray(q, dtype=np.float64)
        self.N = self.q.shape[0]
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.N, 3)
        self.edges = np.ascontiguousarray(edges, dtype=np.float64).reshape(3, 3)
        err = ctypes.create_string_buffer(512)
        mtp = None
        if M is not None:
            if np.isscalar(M):
                M = [int(M)] * 3
            mt = (ctypes.c_int * 3)(*[int(m) for m in M])
            mtp = ctypes.cast(mt, _ip)
        self._h = lib().pmo_build(self.N, _p(self.q), _p(pos), _p(self.edges), int(nu), mtp,
                                  float(a0), float(tolDir), err)
        if not self._h:
            raise ValueError(err.value.decode())
        info = (ctypes.c_int * 5)()
        dinfo = (ctypes.c_double * 6)()
        lib().pmo_info(self._h, info, dinfo)
        self.nu = info[0]
        self.M = (info[1], info[2], info[3])
        self.pair_count = info[4]
        self.a0, self.beta, self.coeff1, self.coeff2, self.detA, self.tolDir = list(dinfo)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            try:
                _lib.pmo_free(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def energy(self, pos, q=None, want_phi=True):
        """Returns (E, F[N,3], phi[N] or None, terms[3])."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.N, 3)
        qq = self.q if q is None else np.ascontiguousarray(q, dtype=np.float64)
        F = np.zeros((self.N, 3))
        phi = np.zeros(self.N) if want_phi else None
        terms = np.zeros(3)
        E = lib().pmo_energy(self._h, _p(pos), _p(qq), _p(F), _p(phi) if want_phi else None,
                             _p(terms))
        return E, F, phi, terms

    def khat(self):
        out = np.zeros(self.M)
        lib().pmo_copy_khat(self._h, _p(out))
        return out

    def grid(self, which):
        out = np.zeros(self.M)
        lib().pmo_copy_grid(self._h, which, _p(out))
        return out


def pme_bspline(nu, t):
    """(w, dw): M_nu(i + t) and its derivative for i = 0 .. nu-1."""
    w = np.zeros(nu)
    dw = np.zeros(nu)
    lib().pmo_bspline(int(nu), float(t), _p(w), _p(dw))
    return w, dw
