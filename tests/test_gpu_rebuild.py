"""FF_rebuild (box change, code/forcefield0.c:598-680) on the GPU.

The rebuild generates the level stencils, the top-level kernel and the MFMA
operands on the device; FF_build computes the same stencils on the host.
A rebuild keeps nu, the grids and nLim, like the reference, so the boxes
here are scaled by small factors that leave nLim unchanged (checked) and a
fresh build of the new box is the comparison:
  - the public FF::khat after the rebuild against the fresh build's;
  - energy, forces and potentials after the rebuild against the oracle of
    the new box (FP64 to 1e-10; mixed to the mixed-precision bar).
Positions are scaled with the box, as a barostat moves them.
"""
import ctypes
from fractions import Fraction

import numpy as np
import pytest

from helpers import load_dataset, rel_err, force_err, rms_rel_force
from msmpb_amd import inputs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _spce300(nu, L, M, a0):
    def make():
        edges, q, pos, mass = load_dataset("spceN300")
        return inputs.System("spceN300", pos, q, edges, mass,
                             dict(nu=nu, levels=L, top_grid=M, cutoff=a0))
    return make


CASES = {
    "water300": (inputs.cfg0_water300, 1.003),
    "crystal8192": (inputs.cfg1_crystal8192, 0.998),
    "random3000": (inputs.cfg2_random3000, 1.002),
    "spce300_L1_nu4": (_spce300(4, 1, 24, 9.0), 1.001),  # top level only
    "spce300_L2_nu8": (_spce300(8, 2, 8, 11.0), 0.999),
}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _msm(s, edges, pos, precision="fp64"):
    from msmpb_amd import MSM
    p = s.params
    return MSM(pos, s.q, edges, nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"],
               cutoff=p["cutoff"], precision=precision)


def _rebuild(m, edges, pos):
    from msmpb_amd import _lib
    e = np.ascontiguousarray(edges, dtype=np.float64)
    x = np.ascontiguousarray(pos, dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    _lib.lib().FF_rebuild(m._ff, e.ctypes.data_as(dp), x.ctypes.data_as(dp), ctypes.c_double(0.0))
    assert m._built(), _lib.last_error()


def _cardinal_bspline(nu, x):
    """M_nu(x) at an integer x (the recursion of de Boor), exact."""
    if nu == 1:
        return Fraction(1 if 0 <= x < 1 else 0)
    return (x * _cardinal_bspline(nu - 1, x) + (nu - x) * _cardinal_bspline(nu - 1, x - 1)) / (nu - 1)


def antiblur_gain(nu):
    """Largest gain of the three anti-blurring passes: 1 / Phi^(pi)^2 per axis.

    Phi^(theta) = sum_n M_nu(nu/2 + n) e^{i n theta} is the symbol of the
    grid B-spline; the stencil passes divide by its square, most at the
    checkerboard mode theta = pi (nu = 4: 1/3, 6: 2/15, 8: 17/315)."""
    h = nu // 2
    s = _cardinal_bspline(nu, h) + 2 * sum((-1) ** n * _cardinal_bspline(nu, h + n) for n in range(1, h))
    return float(s) ** -6


def test_antiblur_gain_closed_form():
    assert antiblur_gain(4) == pytest.approx(3.0 ** 6)
    assert antiblur_gain(8) == pytest.approx((315 / 17) ** 6)


def _khat(m, l):
    ff = m.ff
    if l < m.levels:
        n = int(np.prod([min(2 * ff.nLim + 1, d) for d in m.grid_dims(l)]))
    else:
        n = int(np.prod(m.grid_dims(l)))
    return np.ctypeslib.as_array(ff.khat[l], shape=(n,)).copy()


@pytest.mark.parametrize("case", list(CASES))
def test_rebuild_stencils_match_fresh_build(case, torch_cuda):
    make, scale = CASES[case]
    s = make()
    print(f"[rebuild] stencils {case} x{scale}", flush=True)
    e1, x1 = s.edges * scale, s.pos * scale
    m = _msm(s, s.edges, s.pos)
    fresh = _msm(s, e1, x1)
    try:
        _rebuild(m, e1, x1)
        assert m.ff.nLim == fresh.ff.nLim
        assert m.ff.nlist_interaction_count == fresh.pair_count  # recounted in the new box
        # Both are FP64 evaluations of the same sums; the device fuses
        # multiply-adds that the host rounds twice.  The anti-blurring passes
        # multiply the rounding in the checkerboard mode by up to
        # antiblur_gain(nu) (7e2 at nu = 4, 2e5 at nu = 6, 4e7 at nu = 8), so
        # the bar is a few ulps of the peak times that gain.
        tol = 4 * np.finfo(np.float64).eps * antiblur_gain(s.params["nu"])
        for l in range(1, m.levels + 1):
            a, b = _khat(m, l), _khat(fresh, l)
            assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b)), (l, np.max(np.abs(a - b)))
    finally:
        m.close()
        fresh.close()


@pytest.mark.parametrize("case", list(CASES))
def test_rebuild_energy_matches_oracle(case, oracle_mod, torch_cuda):
    make, scale = CASES[case]
    s = make()
    p = s.params
    print(f"[rebuild] energy {case} x{scale}", flush=True)
    e1, x1 = s.edges * scale, s.pos * scale
    o = oracle_mod.Oracle(x1, s.q, e1, nu=p["nu"], L=p["levels"], Mtop=p["top_grid"],
                          a0=p["cutoff"])
    Eo, Fo, phio, _ = o.energy(x1)
    for precision in ("fp64", "mixed"):
        m = _msm(s, s.edges, s.pos, precision)
        try:
            m.energy(s.pos)  # an evaluation in the old box first
            _rebuild(m, e1, x1)
            E, F, phi = m.energy(x1, potential=True)
            if precision == "fp64":
                assert rel_err(E, Eo) < TOL, (E, Eo)
                assert force_err(F, Fo) < TOL
                assert np.max(np.abs(phi - phio)) < TOL * np.max(np.abs(phio))
            else:
                assert rms_rel_force(F, Fo) < 1e-5
                assert rel_err(E, Eo) < 1e-5
        finally:
            m.close()
