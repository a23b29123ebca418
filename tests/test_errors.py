"""Failure handling of the drop-in boundary, forced from the host (no GPU).

The reference has no error reporting: it exits, prints "error" and keeps
writing, or corrupts its heap outside its domain (forcefield0.c:152 and the
small-grid overflow).  This library validates at FF_build, before anything
touches the GPU, reports through FF_last_error() (Python: MSMError with the
same message) and returns NaN from FF_energy when the FF is not built.
"""
import ctypes

import numpy as np
import pytest

from msmpb_amd import MSM, _lib, inputs
from msmpb_amd.msm import MSMError

dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def water():
    return inputs.cfg0_water300()


@pytest.mark.parametrize("kw,msg", [
    (dict(nu=5, levels=2, top_grid=4, cutoff=8.0), "orderAcc must be 4, 6, 8 or 10 (got 5)"),
    (dict(nu=12, levels=2, top_grid=4, cutoff=8.0), "orderAcc must be 4, 6, 8 or 10 (got 12)"),
    # the reference corrupts its heap here (intermediate stencil narrower than nu)
    (dict(nu=6, levels=4, top_grid=2, cutoff=8.0), "level 3 stencil extent 4 is below the order 6"),
    (dict(method="pme", nu=8, cutoff=8.0), "no default PME grid for order 8"),
    (dict(method="pme", nu=6, top_grid=4, cutoff=8.0), "PME grid dimension 4 is below the order 6"),
    (dict(method="pme", nu=6, top_grid=16, cutoff=8.0, tol_dir=0.2), "tolDir * cutoff must be below 1"),
    (dict(method="pme", nu=7, top_grid=16, cutoff=8.0), "orderAcc must be 4, 6, 8 or 10 (got 7)"),
    (dict(method="pme", nu=6, top_grid=16, cutoff=6.0, layout="asis"), "does not apply to the PME method"),
])
def test_build_rejects_parameters_with_reason(kw, msg, water):
    s = water
    with pytest.raises(MSMError) as e:
        MSM(s.pos, s.q, s.edges, **kw)
    assert msg in str(e.value)


def test_build_rejects_skewed_box_beyond_its_domain(water):
    """The reference's own image rules (as-is layout) need a triclinic box's
    cutoff below half its smallest face-to-face width (14.4 A box, cutoff
    8 A); the default layout sums every image and takes it."""
    s = water
    e = s.edges.copy()
    e[0, 1] = 1.0
    with pytest.raises(MSMError, match="triclinic box: the cutoff 8 must be below half"):
        MSM(s.pos, s.q, e, nu=4, levels=2, top_grid=4, cutoff=8.0, layout="asis")


@pytest.mark.parametrize("kw", [dict(precision="fp32"), dict(method="ewald"), dict(layout="padded")])
def test_python_rejects_unknown_modes(kw, water):
    s = water
    with pytest.raises(ValueError):
        MSM(s.pos, s.q, s.edges, **kw)


def test_c_abi_failures_leave_the_ff_unbuilt(water):
    """N mismatch, NULL inputs and invalid switches through the C ABI; the
    FF then refuses to evaluate (NaN and a message) instead of computing."""
    s = water
    L = _lib.lib()
    q = np.ascontiguousarray(s.q)
    pos = np.ascontiguousarray(s.pos)
    edges = np.ascontiguousarray(s.edges)
    ff = L.FF_new(len(q), q.ctypes.data_as(dp))
    try:
        assert L.FF_set_precision(ff, 7) == _lib.MSM_ERR_ARG
        assert "precision" in _lib.last_error()
        assert L.FF_set_method(ff, 9, 0.0) == _lib.MSM_ERR_ARG
        assert "method" in _lib.last_error()
        assert L.FF_set_layout_compat(ff, 3) == _lib.MSM_ERR_ARG
        assert L.FF_set_method(ff, _lib.MSM_METHOD_PME, 0.0) == _lib.MSM_OK
        assert L.FF_set_layout_compat(ff, _lib.MSM_LAYOUT_ASIS) == _lib.MSM_ERR_ARG
        assert "PME" in _lib.last_error()
        assert L.FF_set_method(ff, _lib.MSM_METHOD_MSM, 0.0) == _lib.MSM_OK
        L.FF_build(ff, len(q) + 1, pos.ctypes.data_as(dp), edges.ctypes.data_as(dp), 0.0)
        assert "N differs" in _lib.last_error()
        L.FF_build(ff, len(q), None, edges.ctypes.data_as(dp), 0.0)
        assert "positions are NULL" in _lib.last_error()
        F = np.zeros((len(q), 3))
        E = L.FF_energy(ff, F.ctypes.data_as(dp), pos.ctypes.data_as(dp), None)
        assert np.isnan(E) and "FF_build has not completed" in _lib.last_error()
        d = (ctypes.c_int * 3)()
        assert L.msm_grid_dims(ff, 1, d) == _lib.MSM_ERR_STATE
    finally:
        L.FF_delete(ff)
    ff = L.FF_new(len(q), None)
    try:
        L.FF_build(ff, len(q), pos.ctypes.data_as(dp), edges.ctypes.data_as(dp), 0.0)
        assert "charges are NULL" in _lib.last_error()
    finally:
        L.FF_delete(ff)


def test_pme_c_abi_failed_build_refuses_energy(water, capfd):
    """pme.h: a build outside the domain prints the reason, FF_energy
    returns NaN (the reference would exit or run on)."""
    s = water
    P = _lib.pme_lib()
    q = np.ascontiguousarray(s.q)
    e = ((ctypes.c_double * 3) * 3)()
    for i in range(3):
        for j in range(3):
            e[i][j] = s.edges[i][j]
    ff = P.FF_new(len(q), q.ctypes.data_as(dp), ctypes.byref(e))
    try:
        P.FF_set_orderAcc(ff, 8)     # no default grid for order 8
        P.FF_set_cutoff(ff, 8.0)
        pos = np.ascontiguousarray(s.pos)
        P.FF_build(ff, pos.ctypes.data_as(dp))
        F = np.zeros((len(q), 3))
        assert np.isnan(P.FF_energy(ff, F.ctypes.data_as(dp), pos.ctypes.data_as(dp), None))
        err = capfd.readouterr().err
        assert "no default PME grid" in err and "FF_build has not completed" in err
        assert P.pme_set_precision(ff, 2) == -1
    finally:
        P.FF_delete(ff)
