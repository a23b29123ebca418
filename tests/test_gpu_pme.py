"""The PME comparator on the GPU against the oracle and the published runs.

The reference's smooth PME (code/pme0.c, code/pme1.c) runs here on the same
kernels as the MSM: the Ewald real-space pairs, B-spline spreading and
interpolation, and the Fourier convolution on one grid.  Checked:
  - FP64 energy, forces, potentials and Fourier kernel against the oracle
    (oracle/pme_oracle.c) to 1e-10; mixed precision to the mixed bar;
  - every distinct run of results/Figure4.json's PME (Box44): energy to
    1e-10, beta, pair count, poterror;
  - the pme.h C ABI of libpme_hip.so (build, energy, rebuild, the public
    struct) and the device-pointer and autograd paths.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from helpers import DATA, force_err, force_scale, load_dataset, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
PUB = json.load(open(os.path.join(os.path.dirname(DATA), "published.json")))["Figure4_pme"]
RUNS = sorted({(t["M"], t["a0"]) for t in PUB})


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _random3000():
    rng = np.random.default_rng(2)
    box = np.array([30.0, 35.0, 40.0])
    pos = rng.uniform(0, 1, (3000, 3)) * box
    q = rng.normal(size=3000)
    q -= q.mean() - 0.5 / 3000  # slightly charged: the neutralising term is exercised
    return np.diag(box), q, pos


def _dataset(name):
    def make():
        edges, q, pos, _ = load_dataset(name)
        return edges, q, pos
    return make


# name: (system, nu, grid, cutoff, tolDir)
CASES = {
    "spce300": (_dataset("spceN300"), 6, 24, 9.0, -1.0),
    "spce300_wide_cutoff": (_dataset("spceN300"), 4, 20, 12.0, -1.0),  # cutoff > box/2: images
    # unit lattice: the cutoff avoids a shell of neighbours (r^2 = 16 would
    # sit exactly on it, where erfc(beta r)/r does not vanish); > box/2
    "nacl512": (_dataset("NaClN512"), 8, 16, 4.1, 1e-8),
    "random3000": (_random3000, 10, (30, 36, 40), 9.0, -1.0),
    # 22 = 2 x 11 and 26 = 2 x 13: no mixed-radix FFT, the direct DFT path
    "spce300_grid_11_13": (_dataset("spceN300"), 6, (22, 25, 26), 9.0, -1.0),
}


def _pme(system, nu, M, a0, tol, precision="fp64"):
    from msmpb_amd import MSM
    edges, q, pos = system
    return MSM(pos, q, edges, nu=nu, top_grid=M, cutoff=a0, precision=precision, method="pme",
               tol_dir=tol)


@pytest.fixture(scope="module")
def reference(oracle_mod):
    cache = {}

    def get(case):
        if case not in cache:
            make, nu, M, a0, tol = CASES[case]
            edges, q, pos = make()
            o = oracle_mod.PmeOracle(pos, q, edges, nu=nu, M=M, a0=a0, tolDir=tol)
            E, F, phi, terms = o.energy(pos)
            cache[case] = ((edges, q, pos), o, E, F, phi, terms)
        return cache[case]
    return get


@pytest.mark.parametrize("case", list(CASES))
def test_pme_fp64_matches_oracle(case, reference, torch_cuda):
    (system, o, Eo, Fo, phio, termso) = reference(case)
    _, nu, M, a0, tol = CASES[case]
    print(f"[pme] fp64 {case}", flush=True)
    m = _pme(system, nu, M, a0, tol)
    try:
        assert m.method == "pme" and m.levels == 1 and m.nu == nu
        assert tuple(m.top_grid) == o.M
        assert rel_err(m.beta, o.beta) < 1e-14
        assert m.pair_count == o.pair_count
        assert rel_err(m.ff.coeff1, o.coeff1) < 1e-13 and rel_err(m.ff.coeff2, o.coeff2) < 1e-13
        kh = np.ctypeslib.as_array(m.ff.khat[1], shape=(int(np.prod(o.M)),))
        assert np.max(np.abs(kh - o.khat().ravel())) <= 1e-13 * np.max(np.abs(kh))
        E, F, phi = m.energy(system[2], potential=True)
        assert rel_err(E, Eo) < TOL, (E, Eo)
        assert force_err(F, Fo, scale=force_scale(system[1], system[0])) < TOL
        assert np.max(np.abs(phi - phio)) < TOL * np.max(np.abs(phio))
        terms = (ctypes.c_double * 3)()
        from msmpb_amd import _lib
        assert _lib.lib().FF_get_energy_terms(m._ff, terms) == 0
        for a, b in zip(terms, termso):
            assert abs(a - b) < TOL * abs(Eo), (list(terms), termso)
    finally:
        m.close()


@pytest.mark.parametrize("case", list(CASES))
def test_pme_mixed_matches_oracle(case, reference, torch_cuda):
    (system, o, Eo, Fo, phio, _) = reference(case)
    _, nu, M, a0, tol = CASES[case]
    print(f"[pme] mixed {case}", flush=True)
    m = _pme(system, nu, M, a0, tol, precision="mixed")
    try:
        E, F, phi = m.energy(system[2], potential=True)
        assert rel_err(E, Eo) < 1e-5, (E, Eo)
        # rms force error relative to the rms force, or to the system's force
        # scale where every force vanishes (a perfect crystal)
        ref = max(np.sqrt(np.mean(np.sum(Fo ** 2, axis=1))), force_scale(system[1], system[0]))
        assert np.sqrt(np.mean(np.sum((F - Fo) ** 2, axis=1))) / ref < 1e-5
        assert np.sqrt(np.mean((phi - phio) ** 2) / np.mean(phio ** 2)) < 1e-5
    finally:
        m.close()


@pytest.mark.parametrize("M,a0", RUNS)
def test_pme_published_box44(M, a0, torch_cuda):
    """Each distinct PME run of results/Figure4.json (order 6, as the
    reference's pme always used): energy, beta, pair count, list capacity,
    poterror."""
    from msmpb_amd import MSM
    t = next(t for t in PUB if t["M"] == M and t["a0"] == a0)
    edges, q, pos, _ = load_dataset("Box44")
    with open(os.path.join(DATA, "Box44.pot")) as f:
        epot = float(f.readline().split()[0])
    print(f"[pme] published Box44 M={M} a0={a0}", flush=True)
    m = MSM(pos, q, edges, nu=t["nu"], top_grid=M, cutoff=a0, method="pme")
    try:
        assert round(m.beta, 6) == t["beta"]
        assert m.pair_count == t["nlist_interaction_count"]
        assert m.nlist_len == t["nlist_len"]
        E, _ = m.energy(pos)
        assert rel_err(E, t["utotal"]) < TOL, (E, t["utotal"])
        assert rel_err(abs(epot - E) / abs(epot), t["poterror"]) < 1e-6
    finally:
        m.close()


def _pme_capi(edges, q, pos, M, a0, nu=None, tol=None):
    from msmpb_amd import _lib
    P = _lib.pme_lib()
    e = ((ctypes.c_double * 3) * 3)()
    for i in range(3):
        for j in range(3):
            e[i][j] = edges[i][j]
    qq = np.ascontiguousarray(q, dtype=np.float64)
    ff = P.FF_new(len(qq), qq.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(e))
    P.FF_set_cutoff(ff, float(a0))
    P.FF_set_topGridDim(ff, (ctypes.c_int * 3)(*([M] * 3 if np.isscalar(M) else M)))
    if nu:
        P.FF_set_orderAcc(ff, nu)
    if tol:
        P.FF_set_tolDir(ff, tol)
    x = np.ascontiguousarray(pos, dtype=np.float64)
    P.FF_build(ff, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return P, ff, qq


def test_pme_capi_published_and_struct(torch_cuda):
    """libpme_hip.so through the reference's pme.h: FF_new / FF_set_* /
    FF_build / FF_energy on Box44 reproduce the published record, and the
    public struct carries what the reference's pme.c prints."""
    t = next(t for t in PUB if t["M"] == 48 and t["a0"] == 9)
    edges, q, pos, _ = load_dataset("Box44")
    print("[pme] C ABI Box44 M=48 a0=9", flush=True)
    P, ff, qq = _pme_capi(edges, q, pos, 48, 9.0)
    try:
        s = ff.contents
        assert s.orderAcc == 6 and list(s.topGridDim) == [48, 48, 48]
        assert s.tolDir == 1e-5 and round(s.beta, 6) == t["beta"]
        assert s.nlist_interaction_count == t["nlist_interaction_count"]
        assert s.nlist_len == t["nlist_len"]  # pme0.c:31
        assert s.detA == pytest.approx(88.0 ** 3, rel=1e-15)
        F = np.zeros((len(qq), 3))
        x = np.ascontiguousarray(pos)
        dp = ctypes.POINTER(ctypes.c_double)
        E = P.FF_energy(ff, F.ctypes.data_as(dp), x.ctypes.data_as(dp), None)
        assert rel_err(E, t["utotal"]) < TOL, (E, t["utotal"])
        assert s.time_partcl2partcl > 0 and s.time_grid2grid > 0
        assert P.FF_get_errEst(ff, len(qq), qq.ctypes.data_as(dp)) > 0
    finally:
        P.FF_delete(ff)


def test_pme_capi_rebuild_matches_oracle(oracle_mod, torch_cuda):
    """FF_rebuild (pme0.c:365-412) for a scaled box: the Fourier kernel is
    generated on the GPU; energy and forces against the oracle of the new
    box."""
    edges, q, pos, _ = load_dataset("spceN300")
    s = 1.004
    e1, x1 = edges * s, pos * s
    o = oracle_mod.PmeOracle(x1, q, e1, nu=6, M=24, a0=9.0)
    Eo, Fo, _, _ = o.energy(x1, want_phi=False)
    print("[pme] C ABI rebuild spceN300 x1.004", flush=True)
    P, ff, qq = _pme_capi(edges, q, pos, 24, 9.0)
    try:
        dp = ctypes.POINTER(ctypes.c_double)
        e = ((ctypes.c_double * 3) * 3)()
        for i in range(3):
            for j in range(3):
                e[i][j] = e1[i][j]
        x = np.ascontiguousarray(x1)
        P.FF_rebuild(ff, ctypes.byref(e), x.ctypes.data_as(dp))
        st = ff.contents
        assert rel_err(st.coeff1, o.coeff1) < 1e-13 and rel_err(st.coeff2, o.coeff2) < 1e-13
        kh = np.ctypeslib.as_array(st.khat, shape=(24 ** 3,))
        assert np.max(np.abs(kh - o.khat().ravel())) <= 1e-12 * np.max(np.abs(kh))
        F = np.zeros((len(qq), 3))
        E = P.FF_energy(ff, F.ctypes.data_as(dp), x.ctypes.data_as(dp), None)
        assert rel_err(E, Eo) < TOL, (E, Eo)
        assert force_err(F, Fo) < TOL
    finally:
        P.FF_delete(ff)


def test_pme_device_and_autograd(reference, torch_cuda):
    """msm_eval_device and the autograd forward/adjoint with method "pme"."""
    torch = torch_cuda
    from msmpb_amd import msm_energy
    (system, o, Eo, Fo, phio, _) = reference("random3000")
    _, nu, M, a0, tol = CASES["random3000"]
    print("[pme] device + autograd random3000", flush=True)
    m = _pme(system, nu, M, a0, tol)
    try:
        pos = torch.tensor(system[2], device="cuda", dtype=torch.float64, requires_grad=True)
        q = torch.tensor(system[1], device="cuda", dtype=torch.float64, requires_grad=True)
        terms, F, phi = m.eval_device(pos.detach(), q.detach(), potential=True)
        torch.cuda.synchronize()
        assert rel_err(terms[0].item(), Eo) < TOL
        assert force_err(F.cpu().numpy(), Fo) < TOL
        E = msm_energy(m, pos, q)
        E.backward()
        assert rel_err(E.item(), Eo) < TOL
        assert force_err(-pos.grad.cpu().numpy(), Fo) < TOL
        assert np.max(np.abs(q.grad.cpu().numpy() - phio)) < TOL * np.max(np.abs(phio))
    finally:
        m.close()
