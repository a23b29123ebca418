"""Pins for the oracle functions the published results do not reach, and
for the host-side helpers of the drop-in boundary (no GPU needed).

* The reference HEAD's index layout (``compat = 1``, code/forcefield1.c
  HEAD quirks): the padded arrays are rebuilt here in numpy from the
  reference's own index expressions (forcefield1.c:143, 195-219, 338-368,
  406-415, 499-528), filled from the corrected oracle's exported grids, and
  pushed through the top level (numpy FFT), prolongation and level stencil
  (each pinned by brute force in test_oracle.py).  The oracle's compat
  energy and level-1 potential grid must agree; a shifted offset or a wrong
  padded stride is shown to change the result.
* The default-parameter rules of FF_build (forcefield0.c:219-348): the top
  grid is the next 7-smooth size above the equal-spacing estimate (checked
  against trial division), the automatic cutoff and order satisfy their
  defining properties in the a-priori error model (which results/Figure2.txt
  pins), and the product library's host rules agree with the oracle's.
* invert / prod / prodT of the C boundary (forcefield0.c:709-724,
  forcefield1.c:639-649) against numpy.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, load_dataset, rel_err

dp = ctypes.POINTER(ctypes.c_double)


# ------------------------------------------------------------ as-is layout
def _reference_l1_grid_energy(q1, e1, nu, of_shift=0, stride_extra=None):
    """0.5 * sum over the HEAD's L = 1 energy loop (forcefield1.c:195-219):
    q[1] is the (M + nu - 1)^3 array anterpolate fills at offset nu/2 - 1
    (forcefield1.c:338-346, 358-368; zero ghosts, so the fold at :370-397
    changes nothing), read with pad 0 and offset 0, i.e. at the flat
    unpadded index (i*My + j)*Mz + k."""
    M = np.array(q1.shape)
    sd = nu - 1 if stride_extra is None else stride_extra
    of = nu // 2 - 1 + of_shift
    P = np.zeros(tuple(M + sd))
    P[of:of + M[0], of:of + M[1], of:of + M[2]] = q1
    n = int(np.prod(M))
    return 0.5 * np.dot(P.reshape(-1)[:n], e1.reshape(-1))


def _reference_top_input(qL, nu, of_shift=0, stride_extra=None):
    """The HEAD's top-level input for L >= 2: restrict_ writes q^L unpadded
    (sd = of = 0, forcefield1.c:406-415) at the start of the zeroed
    (M + nu - 1)^3 buffer (forcefield1.c:143); FFTg2g reads it at offset
    nu/2 - 1 with the padded stride M + nu - 1 (forcefield1.c:504-515)."""
    M = np.array(qL.shape)
    pad = nu - 1 if stride_extra is None else stride_extra
    of = nu // 2 - 1 + of_shift
    buf = np.zeros(int(np.prod(M + nu - 1)))
    buf[:qL.size] = qL.reshape(-1)
    i, j, k = np.meshgrid(np.arange(M[0]), np.arange(M[1]), np.arange(M[2]), indexing="ij")
    f = ((i + of) * (M[1] + pad) + (j + of)) * (M[2] + pad) + (k + of)
    out = np.zeros(tuple(M))
    ok = f < buf.size
    out[ok] = buf[f[ok]]
    return out


def _asis_chain(fx, nu, L, **mut):
    """Corrected oracle's grids -> HEAD energy and e^1 (numpy index maps,
    numpy FFT at the top, the oracle's brute-force-pinned prolongation and
    level stencils below)."""
    q1 = fx.grid(1, 0)
    if L == 1:
        e1 = fx.grid(1, 1)
        return _reference_l1_grid_energy(q1, e1, nu, **mut), e1
    x = _reference_top_input(fx.grid(L, 0), nu, **mut)
    e = np.real(np.fft.ifftn(np.fft.fftn(x) * fx.stencil(L)))
    for l in range(L - 1, 0, -1):
        e = fx.prolong(l + 1, e) + fx.grid2grid(l, fx.grid(l, 0))
    return 0.5 * np.sum(q1 * e), e


@pytest.mark.parametrize("nu,L,M", [(4, 1, 24), (6, 1, 24), (4, 2, 8), (4, 3, 4)])
def test_asis_layout_from_reference_index_maps(nu, L, M, oracle_mod):
    edges, q, pos, _ = load_dataset("spceN300")
    a0 = 9.0 if L < 3 else 5.0
    fx = oracle_mod.Oracle(pos, q, edges, nu=nu, L=L, Mtop=M, a0=a0)
    ax = oracle_mod.Oracle(pos, q, edges, nu=nu, L=L, Mtop=M, a0=a0, compat=1)
    _, _, _, t = fx.energy(pos)
    Ea, Fa, _, ta = ax.energy(pos)
    grid_e, e1 = _asis_chain(fx, nu, L)
    # the short-range and constant parts do not depend on the layout
    assert ta[0] == t[0] and ta[1] == t[1]
    assert abs(ta[2] - grid_e) <= 1e-12 * abs(Ea), (ta[2], grid_e)
    assert rel_err(Ea, t[0] + t[1] + grid_e) < 1e-13
    assert np.max(np.abs(ax.grid(1, 1) - e1)) <= 1e-12 * np.max(np.abs(e1))
    if L >= 2:
        # the padded read overlaps the stored values: the top term is not zero
        assert np.any(_reference_top_input(fx.grid(L, 0), nu) != 0)
    # a plausible slip in the index map changes the number by far more than
    # the tolerance above (the top term is smallest at L = 3: ~3e-7 of E)
    for mut in (dict(of_shift=1), dict(stride_extra=nu)):
        wrong, _ = _asis_chain(fx, nu, L, **mut)
        assert abs(wrong - grid_e) > 1e-9 * abs(Ea), mut


# ------------------------------------------------------ default parameters
def _next_7smooth(n):
    """Smallest m >= n with no prime factor above 7, by trial division."""
    m = n
    while True:
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


def _equal_spacing_estimate(N, box, L):
    """determineM's target (forcefield0.c:247-269): a finest grid with the
    same spacing h on every axis and eta points per atom (eta = 5.7 for one
    level, 0.83 for more), i.e. h^3 = V / (eta N); the top grid is the
    finest divided by 2^(L-1), rounded up."""
    eta = 5.7 if L == 1 else 0.83
    h = (np.prod(box) / (eta * N)) ** (1.0 / 3.0)
    return [int(np.ceil(a / h / 2 ** (L - 1) - 1e-12)) for a in box]


def _random(n, box, seed):
    r = np.random.default_rng(seed)
    box = np.asarray(box, dtype=float)
    pos = r.uniform(0, 1, (n, 3)) * box
    q = r.normal(0, 1, n)
    q -= q.mean()
    return pos, q, np.diag(box)


# (N, box, L, nu, estErr, raw estimate per axis that is not 7-smooth)
DEFAULT_CASES = [
    (200, (20.0, 20.0, 20.0), None, None, None, 11),
    (300, (20.0, 26.0, 30.0), None, None, None, 13),
    (1700, (26.0, 26.0, 26.0), None, None, None, 22),
    (3000, (30.0, 30.0, 30.0), None, None, None, 26),
    (2000, (30.0, 30.0, 70.0), 2, 6, None, 11),
    (2000, (30.0, 30.0, 70.0), 2, 4, 1e-4, 11),
]


def _product_build(pos, q, edges, L=None, nu=None, a0=None, estErr=None, Mtop=None, fft=True):
    """FF_build of libmsmpb_hip.so through the C ABI.  Without a GPU the
    device part fails, after the host rules have published their choices
    into the public fields (as the reference's FF_build does)."""
    from msmpb_amd import _lib
    lib = _lib.lib()
    q = np.ascontiguousarray(q)
    pos = np.ascontiguousarray(pos)
    edges = np.ascontiguousarray(edges)
    ff = lib.FF_new(len(q), q.ctypes.data_as(dp))
    if L is not None:
        lib.FF_set_maxLevel(ff, L)
    if nu is not None:
        lib.FF_set_orderAcc(ff, nu)
    if a0 is not None:
        lib.FF_set_cutoff(ff, a0)
    if estErr is not None:
        lib.FF_set_estErr(ff, estErr)
    if Mtop is not None:
        lib.FF_set_topGridDim(ff, (ctypes.c_int * 3)(*Mtop))
    lib.FF_set_FFT(ff, fft)
    lib.FF_build(ff, len(q), pos.ctypes.data_as(dp), edges.ctypes.data_as(dp), 0.0)
    return lib, ff, (q, pos, edges)


@pytest.mark.parametrize("case", range(len(DEFAULT_CASES)))
def test_default_parameters(case, oracle_mod):
    N, box, L, nu, est, raw = DEFAULT_CASES[case]
    pos, q, edges = _random(N, box, case)
    o = oracle_mod.Oracle(pos, q, edges, nu=nu or -1, L=L or -1, estErr=est or -1.0)
    Lv = L or 1
    # top grid: the equal-spacing estimate rounded up to the next 7-smooth size
    est_M = _equal_spacing_estimate(N, box, Lv)
    assert raw in est_M and _next_7smooth(raw) != raw
    assert list(o.Mtop) == [_next_7smooth(m) for m in est_M]
    # cutoff: the error model meets the tolerance exactly at the chosen order
    # (forcefield0.c:319-334), and no order does better at that cutoff
    estErr = est or 1e-3
    a0, nu_o = o.a0, o.nu
    assert rel_err(o.errest(nu_o) / a0 ** nu_o, estErr) < 1e-12
    for t in (4, 6, 8, 10):
        if nu is None:  # the order minimises the error at the cutoff (:337-348)
            assert o.errest(t) / a0 ** t >= o.errest(nu_o) / a0 ** nu_o * (1 - 1e-12)
            # and the cutoff is the smallest that meets estErr at any order
            assert (o.errest(t) / estErr) ** (1.0 / t) >= a0 * (1 - 1e-12)
    if nu is not None:
        assert nu_o == nu
    # the product's host rules make the same choices
    lib, ff, keep = _product_build(pos, q, edges, L=L, nu=nu, estErr=est)
    try:
        d = (ctypes.c_int * 3)()
        lib.FF_get_topGridDim(ff, d)
        assert list(d) == list(o.Mtop)
        assert lib.FF_get_orderAcc(ff) == nu_o
        assert lib.FF_get_maxLevel(ff) == Lv
        assert rel_err(lib.FF_get_cutoff(ff), a0) < 1e-14
        assert lib.FF_get_estErr(ff) == estErr
        assert rel_err(lib._FF_get_errEst(ff, nu_o), o.errest(nu_o)) < 1e-13
    finally:
        lib.FF_delete(ff)


def test_example_c_parameters(oracle_mod):
    """example.c:20-39: Box44 (65,850 atoms, 88 A), cutoff 4 and one level;
    the order and the top grid are left to the rules: raw estimate 73 ->
    75, order 10."""
    edges, q, pos, _ = load_dataset("Box44")
    o = oracle_mod.Oracle(pos, q, edges, L=1, a0=4.0)
    assert _equal_spacing_estimate(len(q), np.diag(edges), 1) == [73, 73, 73]
    assert o.Mtop == (75, 75, 75) and o.nu == 10
    lib, ff, keep = _product_build(pos, q, edges, L=1, a0=4.0)
    try:
        d = (ctypes.c_int * 3)()
        lib.FF_get_topGridDim(ff, d)
        assert list(d) == [75, 75, 75] and lib.FF_get_orderAcc(ff) == 10
        assert lib.FF_get_cutoff(ff) == 4.0
    finally:
        lib.FF_delete(ff)


# ------------------------------------------------------------- DFT build
def _equal_spacing_dft(N, box, L):
    """determineM in the DFT build (forcefield0.c:247-269): eta = 1 grid point
    per atom on the finest level, h^3 = V / N, the top grid the finest
    divided by 2^(L-1), rounded up."""
    h = (np.prod(box) / N) ** (1.0 / 3.0)
    return [int(np.ceil(a / h / 2 ** (L - 1) - 1e-12)) for a in box]


# (N, box, nu, maxLevel requested, levels chosen, top grid chosen)
DFT_CASES = [
    (3000, (30.0, 33.0, 37.0), 4, None, 4, (2, 2, 3)),
    (1500, (25.0, 27.0, 29.0), 4, None, 3, (3, 3, 4)),
    (20000, (60.0, 60.0, 60.0), 6, None, 4, (4, 4, 4)),
    (600, (20.0, 26.0, 30.0), 4, 2, 3, (2, 3, 3)),
]


@pytest.mark.parametrize("case", range(len(DFT_CASES)))
def test_dft_build_levels_and_grid(case, oracle_mod, capfd):
    """FF_set_FFT(false) (forcefield0.c:292-307): the level count becomes the
    first whose determineM grid has at most sqrt(N) points (a requested
    maxLevel is replaced, with a note on stderr), and that grid is kept as
    it is: no rounding to 7-smooth sizes (:309-316)."""
    N, box, nu, L, Lw, Mw = DFT_CASES[case]
    pos, q, edges = _random(N, box, case)
    grids = [_equal_spacing_dft(N, box, l) for l in range(1, Lw + 1)]
    # the defining property: the chosen count is the first that fits
    assert [bool(np.prod(g) <= np.sqrt(N)) for g in grids] == [False] * (Lw - 1) + [True]
    assert tuple(grids[-1]) == Mw
    o = oracle_mod.Oracle(pos, q, edges, nu=nu, L=L or -1, fft=False)
    assert o.L == Lw and o.Mtop == Mw
    capfd.readouterr()
    lib, ff, keep = _product_build(pos, q, edges, L=L, nu=nu, fft=False)
    try:
        d = (ctypes.c_int * 3)()
        lib.FF_get_topGridDim(ff, d)
        assert tuple(d) == Mw and lib.FF_get_maxLevel(ff) == Lw
        assert rel_err(lib.FF_get_cutoff(ff), o.a0) < 1e-14
        assert not lib.FF_get_FFT(ff)
    finally:
        lib.FF_delete(ff)
    note = f"L={L or 1} is replaced by L={Lw}"
    assert (note in capfd.readouterr().err) == ((L or 1) != Lw)


def test_dft_build_keeps_grid_size(oracle_mod):
    """An explicit top grid that is not 7-smooth (11) is kept by the DFT
    build and rounded up to 12 by the FFT build."""
    pos, q, edges = _random(300, (20.0, 20.0, 20.0), 3)
    for fft, want in ((False, 11), (True, 12)):
        o = oracle_mod.Oracle(pos, q, edges, nu=4, L=1, Mtop=11, a0=8.0, fft=fft)
        assert o.Mtop == (want,) * 3
        lib, ff, keep = _product_build(pos, q, edges, L=1, nu=4, a0=8.0, Mtop=(11, 11, 11),
                                       fft=fft)
        try:
            d = (ctypes.c_int * 3)()
            lib.FF_get_topGridDim(ff, d)
            assert tuple(d) == (want,) * 3 and bool(lib.FF_get_FFT(ff)) == fft
        finally:
            lib.FF_delete(ff)


def _reference_dft_l1(fx, nu, of_shift=0, stride_extra=None):
    """The HEAD's DFT build at L = 1: DFTg2g (forcefield1.c:171-172, 483-497)
    reads q[1] -- the (M + nu - 1)^3 array anterpolate fills at offset
    nu/2 - 1 (zero ghosts) -- at the flat unpadded index, convolves that with
    the top-level kernel, and the energy loop (:195-219) reads q[1] the same
    way."""
    q1 = fx.grid(1, 0)
    M = np.array(q1.shape)
    sd = nu - 1 if stride_extra is None else stride_extra
    of = nu // 2 - 1 + of_shift
    P = np.zeros(tuple(M + sd))
    P[of:of + M[0], of:of + M[1], of:of + M[2]] = q1
    x = P.reshape(-1)[:q1.size].reshape(q1.shape)
    e = np.real(np.fft.ifftn(np.fft.fftn(x) * fx.stencil(1)))
    return 0.5 * np.sum(x * e), e


@pytest.mark.parametrize("nu,M", [(4, 22), (6, 20)])
def test_asis_dft_build_from_reference_index_maps(nu, M, oracle_mod):
    edges, q, pos, _ = load_dataset("spceN300")
    fx = oracle_mod.Oracle(pos, q, edges, nu=nu, L=1, Mtop=M, a0=9.0, fft=False)
    ax = oracle_mod.Oracle(pos, q, edges, nu=nu, L=1, Mtop=M, a0=9.0, compat=1, fft=False)
    assert fx.Mtop == (M,) * 3
    _, _, _, t = fx.energy(pos)
    Ea, _, _, ta = ax.energy(pos)
    grid_e, e1 = _reference_dft_l1(fx, nu)
    assert ta[0] == t[0] and ta[1] == t[1]
    assert abs(ta[2] - grid_e) <= 1e-12 * abs(Ea), (ta[2], grid_e)
    assert np.max(np.abs(ax.grid(1, 1) - e1)) <= 1e-12 * np.max(np.abs(e1))
    for mut in (dict(of_shift=1), dict(stride_extra=nu)):
        wrong, _ = _reference_dft_l1(fx, nu, **mut)
        assert abs(wrong - grid_e) > 1e-9 * abs(Ea), mut


def test_asis_dft_build_reads_top_grid_directly(oracle_mod):
    """At L >= 2 the DFT build's top level reads q^L, which restrict_ wrote
    unpadded, at the flat index: the HEAD's result equals the corrected one
    in an orthorhombic box, where the layouts differ in nothing else."""
    edges, q, pos, _ = load_dataset("spceN300")
    for L, M in ((2, 11), (3, 5)):
        a0 = 9.0 if L == 2 else 5.0
        fx = oracle_mod.Oracle(pos, q, edges, nu=4, L=L, Mtop=M, a0=a0, fft=False)
        ax = oracle_mod.Oracle(pos, q, edges, nu=4, L=L, Mtop=M, a0=a0, compat=1, fft=False)
        assert fx.Mtop == (M,) * 3 and fx.L == L
        Ef, Ff, _, _ = fx.energy(pos)
        Ea, Fa, _, _ = ax.energy(pos)
        assert Ea == Ef and np.array_equal(Fa, Ff)


# --------------------------------------------------- invert / prod / prodT
class _Vector(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double)]


class _Matrix(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz")]


def _mat(a):
    return _Matrix(*[float(v) for v in np.asarray(a).reshape(-1)])


def test_invert_prod_prodT_values():
    """invert overwrites A with A^-1 and returns |det A| (forcefield0.c:709-724);
    prod(M, v) = M v and prodT(M, v) = M^T v (forcefield1.c:639-649)."""
    from msmpb_amd import _lib
    lib = _lib.lib()
    lib.invert.argtypes = [ctypes.POINTER(_Matrix)]
    lib.invert.restype = ctypes.c_double
    lib.prod.argtypes = [_Matrix, _Vector]
    lib.prod.restype = _Vector
    lib.prodT.argtypes = [_Matrix, _Vector]
    lib.prodT.restype = _Vector
    rng = np.random.default_rng(5)
    for A in (np.diag([14.4, 20.0, 33.0]),
              np.array([[30.0, 5.0, 3.0], [0.0, 32.0, 4.0], [0.0, 0.0, 35.0]]),
              rng.normal(size=(3, 3)),
              -np.eye(3) * 2.0):  # negative determinant: |det| is returned
        m = _mat(A)
        det = lib.invert(ctypes.byref(m))
        Ai = np.array([getattr(m, f) for f, _ in _Matrix._fields_]).reshape(3, 3)
        assert abs(det - abs(np.linalg.det(A))) <= 1e-13 * abs(det)
        assert np.allclose(Ai, np.linalg.inv(A), rtol=0, atol=1e-14 * np.max(np.abs(np.linalg.inv(A))))
        v = rng.normal(size=3)
        r = lib.prod(_mat(A), _Vector(*v))
        assert np.allclose([r.x, r.y, r.z], A @ v, rtol=1e-15, atol=1e-14)
        r = lib.prodT(_mat(A), _Vector(*v))
        assert np.allclose([r.x, r.y, r.z], A.T @ v, rtol=1e-15, atol=1e-14)


# ------------------------------------------------ example.c-style C caller
EXAMPLE_SRC = os.path.join(ROOT, "tests", "c", "example_driver.c")
EXAMPLE_BIN = os.path.join(ROOT, "tests", "c", "example_driver")


def build_example_driver():
    """Compile the from-scratch C caller (example.c's calls, forcefield.h
    only) with gcc and link it against libmsmpb_hip.so."""
    lib_dir = os.path.join(ROOT, "msmpb_amd")
    subprocess.check_call(["gcc", "-O2", "-std=c11", "-Wall", "-Werror", "-I",
                           os.path.join(ROOT, "include"), "-o", EXAMPLE_BIN, EXAMPLE_SRC,
                           "-L", lib_dir, "-lmsmpb_hip", "-Wl,-rpath," + lib_dir, "-lm"])
    return EXAMPLE_BIN


def test_example_driver_compiles_and_links():
    """A caller written against forcefield.h alone compiles without warnings
    and every symbol it uses resolves in libmsmpb_hip.so."""
    exe = build_example_driver()
    out = subprocess.run(["ldd", "-r", exe], capture_output=True, text=True)
    assert "undefined symbol" not in out.stdout + out.stderr
