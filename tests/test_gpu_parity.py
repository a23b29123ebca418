"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Plain FP64 must agree with the oracle to 1e-10 relative in energy and in
forces (max |dF| / rms F), stage by stage on the level grids, for every
layout and for the device-pointer and autograd entry points.
"""
import numpy as np
import pytest

from helpers import force_err, force_scale, load_dataset, rel_err, rms_rel_force
from msmpb_amd import inputs

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _pair(sysm, oracle_mod, precision="fp64", layout="fixed", **over):
    from msmpb_amd import MSM
    p = dict(sysm.params)
    p.update(over)
    m = MSM(sysm.pos, sysm.q, sysm.edges, nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"],
            cutoff=p["cutoff"], precision=precision, layout=layout)
    o = oracle_mod.Oracle(sysm.pos, sysm.q, sysm.edges, nu=p["nu"], L=p["levels"],
                          Mtop=p["top_grid"], a0=p["cutoff"], compat=1 if layout == "asis" else 0)
    return m, o


def _dataset_system(name, nu, L, M, a0):
    edges, q, pos, mass = load_dataset(name)
    return inputs.System(name, pos, q, edges, mass, dict(nu=nu, levels=L, top_grid=M, cutoff=a0))


def _random_system(name, n, box, a0, levels, top, seed):
    """Neutral random charges; small boxes exercise periodic self images
    (box edges below the cutoff) and cells of a0/3 (dense)."""
    r = np.random.default_rng(seed)
    box = np.asarray(box, dtype=float)
    pos = r.uniform(0, 1, (n, 3)) * box
    q = r.uniform(-1, 1, n)
    q -= q.mean()
    return inputs.System(name, pos, q, np.diag(box), None,
                         dict(nu=6, levels=levels, top_grid=top, cutoff=a0))


CASES = {
    "water300": inputs.cfg0_water300,
    "crystal8192": inputs.cfg1_crystal8192,
    "random3000": inputs.cfg2_random3000,
    "spce300_L1_nu4": lambda: _dataset_system("spceN300", 4, 1, 24, 9.0),
    "spce300_L2_nu6": lambda: _dataset_system("spceN300", 6, 2, 6, 9.0),
    "spce300_L1_nu8_a11": lambda: _dataset_system("spceN300", 8, 1, 30, 11.0),
    "spce300_L1_nu6": lambda: _dataset_system("spceN300", 6, 1, 24, 9.0),
    "spce300_L2_nu4_M8": lambda: _dataset_system("spceN300", 4, 2, 8, 9.0),
    "nacl512_L3_nu6": lambda: _dataset_system("NaClN512", 6, 3, 4, 3.0),
    # three levels with a 4^3 top grid: the top level rides on the level-2
    # prolongation (one kernel), at orders 8 and 4
    "spce300_L3_nu8_M4": lambda: _dataset_system("spceN300", 8, 3, 4, 9.0),
    "spce300_L3_nu4_M4": lambda: _dataset_system("spceN300", 4, 3, 4, 9.0),
    "tiny_box_self_images": lambda: _random_system("tiny", 60, (12.0, 13.0, 14.0), 13.0, 1, 8, 7),
    "dense400_L2": lambda: _random_system("dense", 400, (10.0, 11.0, 12.0), 6.0, 2, 5, 7),
}


@pytest.mark.parametrize("case", list(CASES))
def test_fp64_energy_forces_potentials(case, oracle_mod, torch_cuda):
    print(f"[parity] fp64 {case}", flush=True)
    s = CASES[case]()
    m, o = _pair(s, oracle_mod)
    E, F, phi = m.energy(s.pos, potential=True)
    Eo, Fo, phio, to = o.energy(s.pos)
    assert rel_err(E, Eo) < TOL, (E, Eo)
    assert force_err(F, Fo, force_scale(s.q, s.edges)) < TOL
    assert np.max(np.abs(phi - phio)) / np.sqrt(np.mean(phio ** 2)) < TOL
    t = m.terms()
    for k in range(3):
        assert abs(t[k] - to[k]) <= TOL * max(abs(Eo), abs(to[k]))
    # the reference's neighbour statistic: pairs with a nearest image inside a0
    assert m.pair_count == o.pair_count()


@pytest.mark.parametrize("case", ["water300", "crystal8192", "random3000", "spce300_L2_nu6",
                                  "spce300_L3_nu8_M4"])
def test_fp64_stage_grids(case, oracle_mod, torch_cuda):
    print(f"[parity] stages {case}", flush=True)
    s = CASES[case]()
    m, o = _pair(s, oracle_mod)
    m.energy(s.pos)
    o.energy(s.pos)
    for l in range(1, m.levels + 1):
        for which in (0, 1):
            g, go = m.grid(l, which), o.grid(l, which)
            assert g.shape == go.shape
            scale = np.max(np.abs(go))
            assert np.max(np.abs(g - go)) <= TOL * scale, (l, which)
        # stencils are built by independent FP64 code; their rounding grows
        # with the sum of |omega'| per axis (the stencil is omega' (x) omega'
        # (x) omega' applied to the kernel): about 7e2 at order 4, 2e5 at 6
        # and 4e7 at 8 on these grids
        st, sto = m.stencil(l), o.stencil(l)
        amp = 1.0
        for a in range(3):
            w = np.abs(o.table(3, l, a))
            amp *= 2 * np.sum(w) - w[0]
        tol = max(1e-11, 2.2e-16 * amp)
        assert np.max(np.abs(st - sto)) <= tol * np.max(np.abs(sto)), (l, amp)


@pytest.mark.parametrize("case", ["spce300_L1_nu4", "spce300_L1_nu6", "spce300_L2_nu6",
                                  "spce300_L2_nu4_M8", "water300", "nacl512_L3_nu6"])
def test_asis_layout(case, oracle_mod, torch_cuda):
    print(f"[parity] asis {case}", flush=True)
    s = CASES[case]()
    m, o = _pair(s, oracle_mod, layout="asis")
    E, F = m.energy(s.pos)
    Eo, Fo, _, _ = o.energy(s.pos)
    assert rel_err(E, Eo) < TOL
    assert force_err(F, Fo, force_scale(s.q, s.edges)) < TOL


def test_device_path_matches_host(oracle_mod, torch_cuda):
    torch = torch_cuda
    from msmpb_amd.inputs import cfg2_random3000
    s = cfg2_random3000()
    m, o = _pair(s, oracle_mod)
    E, F, phi = m.energy(s.pos, potential=True)
    pos = torch.tensor(s.pos, device="cuda")
    q = torch.tensor(s.q, device="cuda")
    e, Fd, phid = m.eval_device(pos, q, forces=True, potential=True)
    torch.cuda.synchronize()
    assert abs(e[0].item() - E) <= 1e-13 * abs(E)
    assert np.max(np.abs(Fd.cpu().numpy() - F)) <= 1e-13 * np.max(np.abs(F))
    assert np.max(np.abs(phid.cpu().numpy() - phi)) <= 1e-13 * np.max(np.abs(phi))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_caller_side_stream(precision, torch_cuda):
    """A caller on its own (non-default) stream gets the same bits as on the
    default stream, through msm_eval_device and the autograd pair, and the
    results are ordered on its stream (read right after that stream's own
    synchronisation only)."""
    torch = torch_cuda
    from msmpb_amd import MSM, msm_energy
    from msmpb_amd.inputs import cfg1_crystal8192
    s = cfg1_crystal8192()
    p = s.params
    m = MSM(s.pos, s.q, s.edges, nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"],
            cutoff=p["cutoff"], precision=precision)
    try:
        pos = torch.tensor(s.pos, device="cuda")
        q = torch.tensor(s.q, device="cuda")
        e0, F0, phi0 = m.eval_device(pos, q, forces=True, potential=True)
        E0 = msm_energy(m, pos.clone().requires_grad_(True), q)
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        with torch.cuda.stream(side):
            for _ in range(3):  # several in flight on the side stream
                e1, F1, phi1 = m.eval_device(pos, q, forces=True, potential=True)
            pg = pos.clone().requires_grad_(True)
            E1 = msm_energy(m, pg, q)
            E1.backward()
        side.synchronize()
        assert torch.equal(e0, e1) and torch.equal(F0, F1) and torch.equal(phi0, phi1)
        assert E1.item() == E0.item()
        assert torch.equal(-pg.grad, F0)
    finally:
        m.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_capture_replays_exact(precision, torch_cuda):
    """msm_eval_device captured into a graph (torch.cuda.graph, stream
    capture through the library's forked streams): every replay, with eager
    evaluations and synchronisations in between, gives the eager result
    bit for bit.  The evaluation records no memset, so nothing in the graph
    depends on one."""
    torch = torch_cuda
    from msmpb_amd import MSM
    from msmpb_amd.inputs import cfg0_water300
    s = cfg0_water300()
    p = s.params
    m = MSM(s.pos, s.q, s.edges, nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"],
            cutoff=p["cutoff"], precision=precision)
    try:
        pos = torch.tensor(s.pos, device="cuda")
        q = torch.tensor(s.q, device="cuda")
        out = torch.empty(4, dtype=torch.float64, device="cuda")
        _, F0, _ = m.eval_device(pos, q, out=out)
        torch.cuda.synchronize()
        E0, F0 = out.clone(), F0.clone()
        side = torch.cuda.Stream()
        with torch.cuda.stream(side):
            for _ in range(3):
                m.eval_device(pos, q, out=out)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            _, Fg, _ = m.eval_device(pos, q, out=out)
        for _ in range(5):
            out.zero_()
            Fg.zero_()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, E0) and torch.equal(Fg, F0)
            m.eval_device(pos, q)  # eager work between replays
            torch.cuda.synchronize()
    finally:
        m.close()


@pytest.mark.parametrize("recompute", [False, True])
def test_autograd_adjoint(recompute, oracle_mod, torch_cuda):
    torch = torch_cuda
    from msmpb_amd import msm_energy
    from msmpb_amd.inputs import cfg2_random3000
    s = cfg2_random3000()
    m, o = _pair(s, oracle_mod)
    Eo, Fo, phio, _ = o.energy(s.pos)
    pos = torch.tensor(s.pos, device="cuda", requires_grad=True)
    q = torch.tensor(s.q, device="cuda", requires_grad=True)
    E = msm_energy(m, pos, q, recompute=recompute)
    E.backward()
    assert rel_err(E.item(), Eo) < TOL
    assert force_err(-pos.grad.cpu().numpy(), Fo) < TOL
    assert np.max(np.abs(q.grad.cpu().numpy() - phio)) / np.sqrt(np.mean(phio ** 2)) < TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_forward_adjoint_keep_modes(precision, oracle_mod, torch_cuda):
    """msm_forward_device / msm_adjoint_device for every keep mode (all state,
    short-range parts only, nothing) and recompute flag: the adjoint's F and
    phi against the oracle (FP64 to 1e-10, mixed to the mixed bar), and a
    full evaluation in between invalidates the kept state."""
    torch = torch_cuda
    from msmpb_amd.inputs import cfg2_random3000
    s = cfg2_random3000()
    m, o = _pair(s, oracle_mod, precision=precision)
    Eo, Fo, phio, _ = o.energy(s.pos)
    pos = torch.tensor(s.pos, device="cuda")
    q = torch.tensor(s.q, device="cuda")
    tol = TOL if precision == "fp64" else 1e-5
    for keep, recompute, between in ((1, False, False), (1, True, False), (2, True, False),
                                     (2, False, False), (0, False, False), (1, False, True)):
        print(f"[parity] keep={keep} recompute={recompute} eval_between={between} {precision}",
              flush=True)
        e = m.forward_device(pos, q, keep=keep)
        if between:  # overwrites the kept pair parts and grid
            m.eval_device(pos * 1.0, q * 1.0)
        F, phi = m.adjoint_device(pos, q, recompute=recompute)
        assert rel_err(e[0].item(), Eo) < tol
        F, phi = F.cpu().numpy(), phi.cpu().numpy()
        if precision == "fp64":
            assert force_err(F, Fo) < tol, (keep, recompute)
        else:
            assert rms_rel_force(F, Fo) < tol, (keep, recompute)
        assert np.max(np.abs(phi - phio)) / np.sqrt(np.mean(phio ** 2)) < tol, (keep, recompute)


def test_adjoint_central_differences(oracle_mod, torch_cuda):
    """dE/dr and dE/dq from the adjoint chain against central differences of
    the forward energy, FP64 (3000 charges in 30 x 35 x 40 A; steps 1e-5 A
    and 1e-6 e) on a few atoms, to 1e-6 of the force / potential scale (the
    differences' own truncation and rounding error).  The mixed path is not
    differenced: FP32 rounding of the pair terms dominates energy changes at
    these steps."""
    precision = "fp64"
    torch = torch_cuda
    from msmpb_amd import msm_energy
    from msmpb_amd.inputs import cfg2_random3000
    print(f"[parity] central differences {precision}", flush=True)
    s = cfg2_random3000()
    m, _ = _pair(s, oracle_mod, precision=precision)
    pos = torch.tensor(s.pos, device="cuda", requires_grad=True)
    q = torch.tensor(s.q, device="cuda", requires_grad=True)
    msm_energy(m, pos, q).backward()
    gpos, gq = pos.grad.cpu().numpy(), q.grad.cpu().numpy()
    fscale, pscale = np.sqrt(np.mean(gpos ** 2)), np.sqrt(np.mean(gq ** 2))

    def energy(p, qq):
        e, _, _ = m.eval_device(p, qq, forces=False)
        return e[0].item()

    P0 = torch.tensor(s.pos, device="cuda")
    Q0 = torch.tensor(s.q, device="cuda")
    tol = 1e-6
    hr, hq = 1e-5, 1e-6
    for i in (0, 17, 1234, 2999):
        for d in range(3):
            Pp, Pm = P0.clone(), P0.clone()
            Pp[i, d] += hr
            Pm[i, d] -= hr
            fd = (energy(Pp, Q0) - energy(Pm, Q0)) / (2 * hr)
            assert abs(fd - gpos[i, d]) < tol * fscale, (i, d, fd, gpos[i, d])
        Qp, Qm = Q0.clone(), Q0.clone()
        Qp[i] += hq
        Qm[i] -= hq
        fd = (energy(P0, Qp) - energy(P0, Qm)) / (2 * hq)
        assert abs(fd - gq[i]) < tol * pscale, (i, fd, gq[i])


def test_mixed_precision_accuracy(oracle_mod, torch_cuda):
    for case in ("water300", "crystal8192", "random3000", "tiny_box_self_images", "dense400_L2",
                 "spce300_L3_nu8_M4"):
        print(f"[parity] mixed {case}", flush=True)
        s = CASES[case]()
        m, o = _pair(s, oracle_mod, precision="mixed")
        E, F = m.energy(s.pos)
        Eo, Fo, _, _ = o.energy(s.pos)
        assert rms_rel_force(F, Fo) < 1e-5, case
        assert rel_err(E, Eo) < 1e-5, case


def _capacity(c, n, a, edges):
    """The reference's pair-list capacity c N + 2.1 (4/3 pi a^3) N (N-1) / (2 det A),
    truncated (forcefield0.c:32 has c = 200; the published runs c = 3)."""
    return int(c * n + 2.1 * ((4.0 / 3.0) * np.pi * a * a * a * 0.5 * n * (n - 1.0)
                              / abs(np.linalg.det(edges))))


@pytest.mark.parametrize("case", ["water300", "random3000", "tiny_box_self_images", "dense400_L2"])
def test_pair_statistics_with_margin(case, oracle_mod, torch_cuda):
    """FF_build's and FF_rebuild's margin widens the neighbour-list
    statistics to a0 + margin (forcefield0.c:21-36,109-163): the pair count
    against the oracle's list built with the same margin (its own cell
    enumeration), and the list capacity, by the published runs' rule in the
    fixed layout and by the source's in the as-is layout."""
    import ctypes
    from msmpb_amd import MSM, _lib
    print(f"[parity] margin {case}", flush=True)
    s = CASES[case]()
    p = s.params
    o = oracle_mod.Oracle(s.pos, s.q, s.edges, nu=p["nu"], L=p["levels"], Mtop=p["top_grid"],
                          a0=p["cutoff"])
    want0 = o.pair_count()
    for margin in (1.5, 0.25):
        want = o.build_pairs(s.pos, margin)
        assert want >= want0  # equal where every pair is already inside a0
        for layout, c in (("fixed", 3), ("asis", 200)):
            m = MSM(s.pos, s.q, s.edges, nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"],
                    cutoff=p["cutoff"], layout=layout, margin=margin)
            try:
                assert m.pair_count == want, (margin, layout)
                assert m.nlist_len == _capacity(c, s.N, m.cutoff + margin, s.edges)
                # a rebuild in the same box with margin 0 counts within a0
                e = np.ascontiguousarray(s.edges, dtype=np.float64)
                x = np.ascontiguousarray(s.pos, dtype=np.float64)
                dp = ctypes.POINTER(ctypes.c_double)
                _lib.lib().FF_rebuild(m._ff, e.ctypes.data_as(dp), x.ctypes.data_as(dp),
                                      ctypes.c_double(0.0))
                assert m.ff.nlist_interaction_count == want0
                assert m.ff.nlist_len == _capacity(c, s.N, m.cutoff, s.edges)
            finally:
                m.close()


def test_pair_range_table_overflow_dense(oracle_mod, torch_cuda):
    """20,000 random charges in a 30 A box with a0 = 10: every pass of the
    pair kernel has more candidates than its range table holds (2,048
    entries mixed, 1,024 FP64), so both kernels take the walk over the range
    offsets.  Cells are >= a0/3 wide, so |dx|, |dy| <= 3 and the <= 49 cell
    columns form one round, whose candidates include every atom within a0 of
    the pass's first home atom."""
    print("[parity] dense20k", flush=True)
    s = _random_system("dense20k", 20000, (30.0, 30.0, 30.0), 10.0, 2, 6, 11)
    d = s.pos[:100, None, :] - s.pos[None, :, :]
    d -= 30.0 * np.floor(d / 30.0 + 0.5)
    assert np.sum(np.sum(d * d, axis=2) < 100.0, axis=1).min() > 2048
    m, o = _pair(s, oracle_mod)
    E, F, phi = m.energy(s.pos, potential=True)
    Eo, Fo, phio, _ = o.energy(s.pos)
    assert rel_err(E, Eo) < TOL, (E, Eo)
    assert force_err(F, Fo, force_scale(s.q, s.edges)) < TOL
    assert np.max(np.abs(phi - phio)) / np.sqrt(np.mean(phio ** 2)) < TOL
    assert m.pair_count == o.pair_count()
    m.set_precision("mixed")
    Em, Fm = m.energy(s.pos)
    assert rms_rel_force(Fm, Fo) < 1e-5
    assert rel_err(Em, Eo) < 1e-5


def test_binning_bucket_overflow_clusters(oracle_mod, torch_cuda):
    """3,000 random charges in a 30 A box plus two clusters of 300 and 200
    charges in balls of 1.5 A: cells of a0/3 = 3 A hold 3.5 atoms on average,
    their bucket rows 32, so the two cluster cells overflow their rows and
    their atoms are ranked through the overflow list (both clusters' atoms
    on one list).  FP64 against the oracle; the sorted order is that of every
    other case (by original index inside each cell)."""
    print("[parity] clusters3500", flush=True)
    r = np.random.default_rng(21)
    box = np.array([30.0, 30.0, 30.0])
    parts = [r.uniform(0, 1, (3000, 3)) * box]
    for centre, n in (((7.5, 7.5, 7.5), 300), ((22.5, 16.5, 10.5), 200)):
        v = r.normal(size=(n, 3))
        v *= (1.5 * r.uniform(0, 1, n) ** (1 / 3) / np.linalg.norm(v, axis=1))[:, None]
        parts.append(np.asarray(centre) + v)
    pos = np.concatenate(parts)
    perm = r.permutation(len(pos))  # cluster atoms spread over the index range
    pos = pos[perm]
    q = r.uniform(-1, 1, len(pos))
    q -= q.mean()
    d = pos[:, None, :] - pos[None, :, :]
    assert np.sum(d * d, axis=2)[np.triu_indices(len(pos), 1)].min() > 1e-6
    s = inputs.System("clusters3500", pos, q, np.diag(box), None,
                      dict(nu=6, levels=2, top_grid=6, cutoff=9.0))
    m, o = _pair(s, oracle_mod)
    try:
        E, F, phi = m.energy(s.pos, potential=True)
        Eo, Fo, phio, _ = o.energy(s.pos)
        assert rel_err(E, Eo) < TOL, (E, Eo)
        assert force_err(F, Fo, force_scale(s.q, s.edges)) < TOL
        assert np.max(np.abs(phi - phio)) / np.sqrt(np.mean(phio ** 2)) < TOL
        assert m.pair_count == o.pair_count()
    finally:
        m.close()


def _dft_pair(name, layout, oracle_mod):
    """Systems for the DFT build (FF_set_FFT(false)): automatic levels, grid
    and cutoff at a fixed order, and explicit grids that are not 7-smooth
    (13^3 = 2197 points takes the general DFT pass)."""
    from msmpb_amd import MSM
    if name.startswith("auto"):
        n, box, nu = {"auto_L3": (1500, (25.0, 27.0, 29.0), 4),
                      "auto_L4": (20000, (60.0, 60.0, 60.0), 6)}[name]
        r = np.random.default_rng(3)
        pos = r.uniform(0, 1, (n, 3)) * np.asarray(box)
        q = r.uniform(-1, 1, n)
        q -= q.mean()
        edges = np.diag(box)
        kw = dict(nu=nu)
        okw = dict(nu=nu)
    else:
        edges, q, pos, _ = load_dataset("spceN300")
        L, M = {"spce_L1_M13": (1, 13), "spce_L1_M11": (1, 11), "spce_L2_M11": (2, 11)}[name]
        kw = dict(nu=4, levels=L, top_grid=M, cutoff=9.0)
        okw = dict(nu=4, L=L, Mtop=M, a0=9.0)
    m = MSM(pos, q, edges, layout=layout, fft=False, **kw)
    o = oracle_mod.Oracle(pos, q, edges, compat=1 if layout == "asis" else 0, fft=False, **okw)
    return m, o, pos, q, edges


@pytest.mark.parametrize("name", ["auto_L3", "auto_L4", "spce_L1_M13", "spce_L1_M11",
                                  "spce_L2_M11"])
@pytest.mark.parametrize("layout", ["fixed", "asis"])
def test_dft_build(name, layout, oracle_mod, torch_cuda):
    """FF_set_FFT(false): the DFT build's parameter rules and, in the as-is
    layout, its level-1 top-level read (forcefield1.c:483-497), FP64 against
    the oracle to 1e-10."""
    print(f"[parity] dft build {name} {layout}", flush=True)
    m, o, pos, q, edges = _dft_pair(name, layout, oracle_mod)
    try:
        assert (m.levels, m.top_grid, m.nu) == (o.L, o.Mtop, o.nu)
        assert not m.ff.FFT
        E, F, phi = m.energy(pos, potential=True)
        Eo, Fo, phio, to = o.energy(pos)
        assert rel_err(E, Eo) < TOL, (E, Eo)
        assert force_err(F, Fo, force_scale(q, edges)) < TOL
        assert np.max(np.abs(phi - phio)) / np.sqrt(np.mean(phio ** 2)) < TOL
        t = m.terms()
        assert abs(t[2] - to[2]) <= TOL * max(abs(Eo), abs(to[2]))
        assert m.pair_count == o.pair_count()
    finally:
        m.close()


def test_water96k_fp64_and_mixed(oracle_mod, torch_cuda):
    """The benchmark workload (96,000 atoms, 24,389 pair-search cells): FP64 to
    the oracle at 1e-10 and the mixed path's rms relative force error."""
    print("[parity] water96k", flush=True)
    s = inputs.cfg3_water96k(6)
    m, o = _pair(s, oracle_mod)
    E, F = m.energy(s.pos)
    Eo, Fo, _, _ = o.energy(s.pos, want_phi=False)
    assert rel_err(E, Eo) < TOL
    assert force_err(F, Fo) < TOL
    m.set_precision("mixed")
    Em, Fm = m.energy(s.pos)
    assert rms_rel_force(Fm, Fo) < 1e-5
    assert rel_err(Em, Eo) < 1e-6
