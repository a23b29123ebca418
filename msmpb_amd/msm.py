"""Python interface to the MI355X MSM library.

`MSM` mirrors the reference's library life cycle (FF_new, FF_set_*, FF_build,
FF_energy, FF_delete; code/forcefield.h:70-97 of hkayabilisim/msmpb) and adds
the device-resident entry points used by the autograd binding.
"""
import ctypes

import numpy as np
import torch

from . import _lib

_dp = ctypes.POINTER(ctypes.c_double)

PRECISIONS = {"fp64": _lib.MSM_PREC_FP64, "mixed": _lib.MSM_PREC_MIXED}
LAYOUTS = {"fixed": _lib.MSM_LAYOUT_FIXED, "asis": _lib.MSM_LAYOUT_ASIS}
METHODS = {"msm": _lib.MSM_METHOD_MSM, "pme": _lib.MSM_METHOD_PME}


class MSMError(RuntimeError):
    pass


def _p(a):
    return a.ctypes.data_as(_dp)


def _check(rc, what):
    if rc != 0:
        raise MSMError(f"{what}: {_lib.last_error()}")


class MSM:
    """B-spline multilevel summation for periodic Coulomb energy and forces.

    positions (N, 3) and edges (3, 3: columns are the box vectors, the
    reference's `edges` argument of FF_build) in Angstrom, charges (N,) in e.
    nu, levels, top_grid, cutoff, est_err = -1/None select the reference's
    automatic parameter rules (forcefield0.c:286-348).  precision: "fp64" or
    "mixed"; layout: "fixed" or "asis" (reference HEAD index layout).
    method="pme" evaluates the reference's smooth PME comparator (pme0.c,
    pme1.c) with the same kernels: one grid of top_grid points, the Ewald
    pair term with erfc(beta cutoff) = tol_dir * cutoff (tol_dir <= 0:
    1e-5); levels and est_err do not apply, nu defaults to 6, cutoff to 8.
    fft=False selects the reference's DFT build (FF_set_FFT(false)): its
    default grid and level rules, no rounding to 7-smooth sizes.
    margin widens the neighbour-list statistics (pair_count, nlist_len) to
    pairs within cutoff + margin, as FF_build's margin does in the reference;
    the energy does not depend on it.
    """

    def __init__(self, positions, charges, edges, nu=-1, levels=-1, top_grid=None, cutoff=-1.0,
                 est_err=-1.0, precision="fp64", layout="fixed", stage_timing=False,
                 method="msm", tol_dir=-1.0, margin=0.0, fft=True):
        for name, val, table in (("precision", precision, PRECISIONS), ("layout", layout, LAYOUTS),
                                 ("method", method, METHODS)):
            if val not in table:
                raise ValueError(f"{name} must be one of {sorted(table)} (got {val!r})")
        L = _lib.lib()
        self._q = np.ascontiguousarray(charges, dtype=np.float64).copy()
        self.N = int(self._q.shape[0])
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.N, 3)
        self.edges = np.ascontiguousarray(edges, dtype=np.float64).reshape(3, 3)
        self._ff = L.FF_new(self.N, _p(self._q))
        if nu is not None and nu > 0:
            L.FF_set_orderAcc(self._ff, int(nu))
        if levels is not None and levels > 0:
            L.FF_set_maxLevel(self._ff, int(levels))
        if top_grid is not None:
            tg = [int(top_grid)] * 3 if np.isscalar(top_grid) else [int(m) for m in top_grid]
            arr = (ctypes.c_int * 3)(*tg)
            L.FF_set_topGridDim(self._ff, arr)
        if cutoff is not None and cutoff > 0:
            L.FF_set_cutoff(self._ff, float(cutoff))
        if est_err is not None and est_err > 0:
            L.FF_set_estErr(self._ff, float(est_err))
        L.FF_set_FFT(self._ff, bool(fft))
        _check(L.FF_set_precision(self._ff, PRECISIONS[precision]), "precision")
        _check(L.FF_set_layout_compat(self._ff, LAYOUTS[layout]), "layout")
        _check(L.FF_set_stage_timing(self._ff, 1 if stage_timing else 0), "stage timing")
        _check(L.FF_set_method(self._ff, METHODS[method], float(tol_dir)), "method")
        self.precision = precision
        self.method = method
        L.FF_build(self._ff, self.N, _p(pos), _p(self.edges), float(margin))
        err = _lib.last_error()  # the build's own message, before any other call
        if L.FF_get_maxLevel(self._ff) <= 0 or not self._built():
            self.close()
            raise MSMError(f"FF_build failed: {err}")
        ff = self._ff.contents
        self.nu = ff.orderAcc
        self.levels = ff.maxLevel
        self.top_grid = tuple(ff.topGridDim)
        self.cutoff = ff.cutoff
        self.nlim = ff.nLim
        self.pair_count = ff.nlist_interaction_count
        self.nlist_len = ff.nlist_len
        self.beta = L.FF_get_beta(self._ff)

    def _built(self):
        d = (ctypes.c_int * 3)()
        return _lib.lib().msm_grid_dims(self._ff, 1, d) == 0

    def close(self):
        if getattr(self, "_ff", None):
            try:
                _lib.lib().FF_delete(self._ff)
            except Exception:  # interpreter shutdown
                pass
            self._ff = None

    def __del__(self):
        self.close()

    @property
    def ff(self):
        return self._ff.contents

    def set_precision(self, precision):
        _check(_lib.lib().FF_set_precision(self._ff, PRECISIONS[precision]), "precision")
        self.precision = precision

    def set_stage_timing(self, enable):
        """Record per-stage GPU intervals (read with stage_times())."""
        _check(_lib.lib().FF_set_stage_timing(self._ff, 1 if enable else 0), "stage timing")

    def stage_times(self):
        """Mean seconds per evaluation of each stage since the last call."""
        L = _lib.lib()
        st = (ctypes.c_double * 16)()
        n = L.msm_stage_times(self._ff, st, 16)
        _check(min(n, 0), "stage times")
        return {L.msm_stage_name(k).decode(): st[k] / max(n, 1) for k in range(16)
                if L.msm_stage_name(k)}

    def set_charges(self, charges):
        self._q[:] = np.asarray(charges, dtype=np.float64)

    # -- host path: the drop-in FF_energy ---------------------------------
    def energy(self, positions, potential=False):
        """FF_energy: returns (E, F) or (E, F, phi) with host arrays."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.N, 3)
        F = np.zeros((self.N, 3))
        if potential:
            phi = np.zeros(self.N)
            E = _lib.lib().FF_energy_potential(self._ff, _p(F), _p(pos), _p(phi))
        else:
            E = _lib.lib().FF_energy(self._ff, _p(F), _p(pos), None)
        if not np.isfinite(E):
            raise MSMError(f"FF_energy: {_lib.last_error()}")
        return (E, F, phi) if potential else (E, F)

    def terms(self):
        t = np.zeros(3)
        _check(_lib.lib().FF_get_energy_terms(self._ff, _p(t)), "terms")
        return t

    def grid_dims(self, level):
        d = (ctypes.c_int * 3)()
        _check(_lib.lib().msm_grid_dims(self._ff, level, d), "grid dims")
        return tuple(d)

    def stencil_dims(self, level):
        d = (ctypes.c_int * 3)()
        _check(_lib.lib().msm_stencil_dims(self._ff, level, d), "stencil dims")
        return tuple(d)

    def grid(self, level, which):
        out = np.zeros(self.grid_dims(level))
        _check(_lib.lib().msm_copy_grid(self._ff, level, which, _p(out)), "copy grid")
        return out

    def stencil(self, level):
        out = np.zeros(self.stencil_dims(level))
        _check(_lib.lib().msm_copy_stencil(self._ff, level, _p(out)), "copy stencil")
        return out

    def count_pairs(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.N, 3)
        c = ctypes.c_longlong()
        _check(_lib.lib().msm_count_pairs(self._ff, _p(pos), ctypes.byref(c)), "count pairs")
        return c.value

    # -- device path -------------------------------------------------------
    @staticmethod
    def _dev(t, shape):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64):
            raise MSMError("device entry points take float64 CUDA tensors")
        if tuple(t.shape) != shape or not t.is_contiguous():
            raise MSMError(f"expected a contiguous tensor of shape {shape}, got {tuple(t.shape)}")
        return ctypes.c_void_p(t.data_ptr())

    def eval_device(self, pos, q, forces=True, potential=False, out=None):
        """Stream-ordered evaluation on torch tensors; returns (terms4, F, phi)."""
        p = self._dev(pos, (self.N, 3))
        qq = self._dev(q, (self.N,))
        e = out if out is not None else torch.empty(4, dtype=torch.float64, device=pos.device)
        F = torch.empty((self.N, 3), dtype=torch.float64, device=pos.device) if forces else None
        phi = torch.empty(self.N, dtype=torch.float64, device=pos.device) if potential else None
        s = ctypes.c_void_p(torch.cuda.current_stream(pos.device).cuda_stream)
        rc = _lib.lib().msm_eval_device(
            self._ff, p, qq, ctypes.c_void_p(e.data_ptr()),
            ctypes.c_void_p(F.data_ptr()) if F is not None else None,
            ctypes.c_void_p(phi.data_ptr()) if phi is not None else None, s)
        _check(rc, "msm_eval_device")
        return e, F, phi

    def forward_device(self, pos, q, keep=True):
        """keep: True/1 (grid and short-range parts), 2 (short-range parts
        only: the adjoint recomputes the grid levels), False/0 (nothing)."""
        p = self._dev(pos, (self.N, 3))
        qq = self._dev(q, (self.N,))
        e = torch.empty(4, dtype=torch.float64, device=pos.device)
        s = ctypes.c_void_p(torch.cuda.current_stream(pos.device).cuda_stream)
        _check(_lib.lib().msm_forward_device(self._ff, p, qq, ctypes.c_void_p(e.data_ptr()),
                                             int(keep), s), "msm_forward_device")
        return e

    def adjoint_device(self, pos, q, recompute=False):
        """Returns (force = -dE/dr, phi = dE/dq) as tensors."""
        p = self._dev(pos, (self.N, 3))
        qq = self._dev(q, (self.N,))
        F = torch.empty((self.N, 3), dtype=torch.float64, device=pos.device)
        phi = torch.empty(self.N, dtype=torch.float64, device=pos.device)
        s = ctypes.c_void_p(torch.cuda.current_stream(pos.device).cuda_stream)
        _check(_lib.lib().msm_adjoint_device(self._ff, p, qq, ctypes.c_void_p(F.data_ptr()),
                                             ctypes.c_void_p(phi.data_ptr()),
                                             1 if recompute else 0, s), "msm_adjoint_device")
        return F, phi


class _MSMEnergy(torch.autograd.Function):
    """E(r, q); backward runs the adjoint kernel chain: dE/dr = -F, dE/dq = phi."""

    @staticmethod
    def forward(ctx, pos, q, msm, recompute):
        # the short-range parts are per-atom (kept even when the grid levels
        # are recomputed in backward)
        e = msm.forward_device(pos, q, keep=2 if recompute else 1)
        msm._generation = getattr(msm, "_generation", 0) + 1
        ctx.generation = msm._generation
        ctx.msm = msm
        ctx.recompute = recompute
        ctx.save_for_backward(pos, q)
        return e[0]

    @staticmethod
    def backward(ctx, grad_e):
        pos, q = ctx.saved_tensors
        # the library keeps only the latest forward's state
        stale = ctx.msm._generation != ctx.generation
        F, phi = ctx.msm.adjoint_device(pos, q, recompute=ctx.recompute or stale)
        gpos = -F * grad_e if ctx.needs_input_grad[0] else None
        gq = phi * grad_e if ctx.needs_input_grad[1] else None
        return gpos, gq, None, None


def msm_energy(msm, pos, q, recompute=None):
    """Differentiable MSM energy of float64 CUDA tensors pos (N,3) and q (N,).

    recompute=None recomputes the grid levels in backward for N >= 500k
    atoms instead of keeping the forward's state.
    """
    if recompute is None:
        recompute = msm.N >= 500_000
    return _MSMEnergy.apply(pos, q, msm, bool(recompute))
