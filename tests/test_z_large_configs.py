"""The two full-size workloads at their real shapes on the GPU, against the
oracle (collected after the fast parity tests: the oracle needs ~1.5 min of
one CPU core for the 1M-atom system).

* water96k at order 4 (32,000 rigid waters, 98.6 A box, cutoff 10 A,
  4 levels, 48^3 finest grid): FP64 energy, forces and potentials to 1e-10
  relative; mixed precision rms relative force error <= 1e-5.
* water1m (333,333 waters, 215 A box, order 6, 5 levels, 112^3 finest
  grid): FP64 energy, forces and potentials to 1e-10; mixed rms <= 1e-5;
  the torch.autograd adjoint with the grid levels recomputed in backward
  (the default at this size) against the FP64 forces and potentials; and
  the device footprint of one FF.
"""
import numpy as np
import pytest

from helpers import force_err, rel_err, rms_rel_force
from msmpb_amd import inputs

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _msm(s, precision):
    from msmpb_amd import MSM
    p = s.params
    return MSM(s.pos, s.q, s.edges, nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"],
               cutoff=p["cutoff"], precision=precision)


def _oracle(s, oracle_mod):
    p = s.params
    return oracle_mod.Oracle(s.pos, s.q, s.edges, nu=p["nu"], L=p["levels"], Mtop=p["top_grid"],
                             a0=p["cutoff"])


def _phi_err(phi, phio):
    return np.max(np.abs(phi - phio)) / np.sqrt(np.mean(phio ** 2))


def test_water96k_order4(oracle_mod, torch_cuda):
    print("[large] water96k nu=4: oracle", flush=True)
    s = inputs.cfg3_water96k(4)
    Eo, Fo, phio, _ = _oracle(s, oracle_mod).energy(s.pos)
    print("[large] water96k nu=4: fp64", flush=True)
    m = _msm(s, "fp64")
    try:
        E, F, phi = m.energy(s.pos, potential=True)
        assert rel_err(E, Eo) < TOL, (E, Eo)
        assert force_err(F, Fo) < TOL
        assert _phi_err(phi, phio) < TOL
        print("[large] water96k nu=4: mixed", flush=True)
        m.set_precision("mixed")
        Em, Fm, phim = m.energy(s.pos, potential=True)
        assert rms_rel_force(Fm, Fo) < 1e-5
        assert rel_err(Em, Eo) < 1e-6
        assert np.sqrt(np.mean((phim - phio) ** 2) / np.mean(phio ** 2)) < 1e-5
    finally:
        m.close()


@pytest.fixture(scope="module")
def water1m(oracle_mod, torch_cuda):
    s = inputs.cfg4_water1m()
    print(f"[large] water1m: oracle build ({s.N} atoms)", flush=True)
    o = _oracle(s, oracle_mod)
    print("[large] water1m: oracle evaluation", flush=True)
    Eo, Fo, phio, _ = o.energy(s.pos)
    del o
    return s, Eo, Fo, phio


def test_water1m_fp64_and_footprint(water1m, torch_cuda):
    torch = torch_cuda
    s, Eo, Fo, phio = water1m
    print("[large] water1m: fp64", flush=True)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    m = _msm(s, "fp64")
    torch.cuda.synchronize()
    footprint = free0 - torch.cuda.mem_get_info(0)[0]
    try:
        print(f"[large] water1m: FF footprint {footprint / 2**20:.1f} MiB", flush=True)
        # positions, sorted atoms, forces, potentials, cells and five grid
        # levels; no stored neighbour list (the reference's holds ~640 M ints)
        assert footprint < 512 * 2**20
        E, F, phi = m.energy(s.pos, potential=True)
        assert rel_err(E, Eo) < TOL, (E, Eo)
        assert force_err(F, Fo) < TOL
        assert _phi_err(phi, phio) < TOL
    finally:
        m.close()


def test_water1m_mixed(water1m, torch_cuda):
    s, Eo, Fo, _ = water1m
    print("[large] water1m: mixed", flush=True)
    m = _msm(s, "mixed")
    try:
        E, F = m.energy(s.pos)
        assert rms_rel_force(F, Fo) < 1e-5
        assert rel_err(E, Eo) < 1e-6
    finally:
        m.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_water1m_adjoint_recomputed_levels(precision, water1m, torch_cuda):
    """msm_energy at 1M atoms recomputes the grid levels in backward by
    default; dE/dr and dE/dq against the oracle's forces and potentials."""
    torch = torch_cuda
    from msmpb_amd import msm_energy
    s, Eo, Fo, phio = water1m
    print(f"[large] water1m: autograd adjoint, recomputed levels, {precision}", flush=True)
    m = _msm(s, precision)
    try:
        pos = torch.tensor(s.pos, device="cuda", requires_grad=True)
        q = torch.tensor(s.q, device="cuda", requires_grad=True)
        E = msm_energy(m, pos, q)  # recompute=None -> True for N >= 500k
        E.backward()
        F = -pos.grad.cpu().numpy()
        phi = q.grad.cpu().numpy()
        if precision == "fp64":
            assert rel_err(E.item(), Eo) < TOL
            assert force_err(F, Fo) < TOL
            assert _phi_err(phi, phio) < TOL
        else:
            assert rel_err(E.item(), Eo) < 1e-6
            assert rms_rel_force(F, Fo) < 1e-5
            assert np.sqrt(np.mean((phi - phio) ** 2) / np.mean(phio ** 2)) < 1e-5
    finally:
        m.close()
