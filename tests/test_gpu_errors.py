"""Failure handling that needs a built FF (GPU).

Non-finite inputs are detected by the binning kernel itself (it places such
an atom in cell 0 at a finite position and counts it), so every access stays
in bounds and FF_energy reports the failure; the next evaluation with valid
inputs is unaffected.  Level weights other than 1 and a box change beyond
the FF_build buffers are refused on the host.
"""
import ctypes

import numpy as np
import pytest

from helpers import force_err, rel_err
from msmpb_amd import MSM, _lib, inputs

pytestmark = pytest.mark.gpu
dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def built(oracle_mod, torch_cuda):
    s = inputs.cfg0_water300()
    p = s.params
    m = MSM(s.pos, s.q, s.edges, nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"],
            cutoff=p["cutoff"])
    o = oracle_mod.Oracle(s.pos, s.q, s.edges, nu=p["nu"], L=p["levels"], Mtop=p["top_grid"],
                          a0=p["cutoff"])
    Eo, Fo, _, _ = o.energy(s.pos, want_phi=False)
    yield s, m, Eo, Fo
    m.close()


def _energy(m, pos):
    F = np.zeros((m.N, 3))
    x = np.ascontiguousarray(pos)
    return _lib.lib().FF_energy(m._ff, F.ctypes.data_as(dp), x.ctypes.data_as(dp), None), F


@pytest.mark.parametrize("what", ["nan_position", "inf_position", "nan_charge"])
def test_non_finite_input_reported_then_recovers(what, built):
    s, m, Eo, Fo = built
    print(f"[errors] {what}", flush=True)
    pos = s.pos.copy()
    if what == "nan_position":
        pos[17, 1] = np.nan
    elif what == "inf_position":
        pos[3, 0] = np.inf
    else:
        m.set_charges(np.where(np.arange(m.N) == 5, np.nan, s.q))
    E, _ = _energy(m, pos)
    assert np.isnan(E) and "non-finite position or charge" in _lib.last_error()
    m.set_charges(s.q)
    E, F = _energy(m, s.pos)  # the next valid evaluation is exact again
    assert rel_err(E, Eo) < 1e-10 and force_err(F, Fo) < 1e-10


def test_level_weights_other_than_one_refused(built):
    s, m, Eo, _ = built
    w = np.ones(m.levels + 1)
    w[1] = 0.5
    F = np.zeros((m.N, 3))
    x = np.ascontiguousarray(s.pos)
    E = _lib.lib().FF_energy(m._ff, F.ctypes.data_as(dp), x.ctypes.data_as(dp), w.ctypes.data_as(dp))
    assert np.isnan(E) and "weights other than 1" in _lib.last_error()
    E = _lib.lib().FF_energy(m._ff, F.ctypes.data_as(dp), x.ctypes.data_as(dp),
                             np.ones(m.levels + 1).ctypes.data_as(dp))
    assert rel_err(E, Eo) < 1e-10  # all-ones weights are the unweighted evaluation


def test_rebuild_beyond_buffers_refused(torch_cuda):
    """A box change that needs more pair-search cells than FF_build allocated
    is refused (the FF must be built again); the FF reports it is unbuilt."""
    s = inputs.cfg0_water300()
    p = s.params
    m = MSM(s.pos, s.q, s.edges, nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"],
            cutoff=p["cutoff"])
    try:
        e = np.ascontiguousarray(s.edges * 3.0)
        x = np.ascontiguousarray(s.pos * 3.0)
        _lib.lib().FF_rebuild(m._ff, e.ctypes.data_as(dp), x.ctypes.data_as(dp), ctypes.c_double(0.0))
        assert "beyond its buffers" in _lib.last_error()
        assert not m._built()
        E, _ = _energy(m, x)
        assert np.isnan(E)
    finally:
        m.close()
