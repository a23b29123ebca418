"""Input data: the reference's .ini/.acc/.pot formats and seeded synthetic systems.

File formats follow code/msmpb.c of hkayabilisim/msmpb: `.ini` (msmpb.c:93-118)
is one line of nine box numbers (row-major), one line with N, then N lines
`q x y z mass ...`; `.acc` (msmpb.c:254-263) is N then the N x-, y- and
z-accelerations; `.pot` is one energy.  Units: e, Angstrom, e^2/Angstrom.

The synthetic workloads are the five measured configurations: a 300-atom
water box, a 8192-ion two-species crystal, 3000 random charges in an
orthorhombic box, and water boxes of 96,000 and ~1.0M atoms.
"""
import gzip
from dataclasses import dataclass, field

import numpy as np

Q_O, Q_H = -0.8476, 0.4238
R_OH = 1.0
THETA_HOH = np.deg2rad(109.47)
M_O, M_H = 15.9994, 1.00794


@dataclass
class System:
    name: str
    pos: np.ndarray          # (N, 3)
    q: np.ndarray            # (N,)
    edges: np.ndarray        # (3, 3) as passed to FF_build
    mass: np.ndarray = None
    params: dict = field(default_factory=dict)

    @property
    def N(self):
        return int(self.q.shape[0])


def read_ini(path):
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as f:
        edges = np.array([float(v) for v in f.readline().split()[:9]]).reshape(3, 3)
        n = int(f.readline().split()[0])
        rows = [f.readline().split() for _ in range(n)]
    d = np.array([[float(v) for v in r[:5]] for r in rows])
    return edges, d[:, 0].copy(), d[:, 1:4].copy(), d[:, 4].copy()


def write_ini(path, sysm):
    with open(path, "w") as f:
        f.write(" ".join(f"{v:.10g}" for v in sysm.edges.ravel()) + "\n")
        f.write(f"{sysm.N}\n")
        mass = sysm.mass if sysm.mass is not None else np.ones(sysm.N)
        for qi, r, m in zip(sysm.q, sysm.pos, mass):
            f.write(f"{qi:.10f} {r[0]:.12f} {r[1]:.12f} {r[2]:.12f} {m:.5f} 0 0 0\n")


def read_acc(path, n):
    with open(path) as f:
        f.readline()
        v = np.array([float(f.readline().split()[0]) for _ in range(3 * n)])
    return v.reshape(3, n).T.copy()


def read_pot(path):
    with open(path) as f:
        return float(f.readline().split()[0])


def _random_rotations(rng, n):
    """Uniform random rotation matrices from normalised Gaussian quaternions."""
    qt = rng.normal(size=(n, 4))
    qt /= np.linalg.norm(qt, axis=1, keepdims=True)
    w, x, y, z = qt.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_box(nmol, box, lattice, seed, jitter=0.25):
    """Rigid 3-site waters (O-H 1.0 A, H-O-H 109.47 deg, O -0.8476 e, H +0.4238 e)
    with oxygens on a jittered lattice and uniform random orientations."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = lattice
    sp = np.array([box / nx, box / ny, box / nz])
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    sites = (np.stack([ii, jj, kk], -1).reshape(-1, 3) + 0.5) * sp
    if sites.shape[0] > nmol:
        sites = sites[np.sort(rng.permutation(sites.shape[0])[:nmol])]
    sites = sites + rng.uniform(-jitter, jitter, size=sites.shape)
    half = 0.5 * THETA_HOH
    body = np.array([[0.0, 0.0, 0.0],
                     [R_OH * np.sin(half), 0.0, R_OH * np.cos(half)],
                     [-R_OH * np.sin(half), 0.0, R_OH * np.cos(half)]])
    R = _random_rotations(rng, nmol)
    pos = sites[:, None, :] + np.einsum("nij,aj->nai", R, body)
    pos = pos.reshape(-1, 3)
    q = np.tile([Q_O, Q_H, Q_H], nmol)
    mass = np.tile([M_O, M_H, M_H], nmol)
    return pos, q, mass


def cfg0_water300():
    pos, q, mass = water_box(100, 14.4, (5, 5, 4), seed=0)
    return System("water300", pos, q, np.diag([14.4] * 3), mass,
                  dict(nu=4, levels=2, top_grid=4, cutoff=8.0))


def cfg1_crystal8192():
    rng = np.random.default_rng(1)
    a = 2.82
    n = 16
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = np.stack([ii, jj, kk], -1).reshape(-1, 3).astype(float) * a
    pos = np.concatenate([base, base + 0.5 * a])
    q = np.concatenate([np.ones(len(base)), -np.ones(len(base))])
    pos = pos + rng.normal(scale=0.05 * a, size=pos.shape)
    box = n * a
    return System("crystal8192", pos, q, np.diag([box] * 3), np.ones(len(q)),
                  dict(nu=6, levels=3, top_grid=6, cutoff=9.0))


def cfg2_random3000():
    rng = np.random.default_rng(2)
    L = np.array([30.0, 35.0, 40.0])
    pos = rng.uniform(size=(3000, 3)) * L
    q = rng.normal(size=3000)
    q -= q.mean()
    return System("random3000", pos, q, np.diag(L), np.ones(3000),
                  dict(nu=6, levels=2, top_grid=(8, 9, 10), cutoff=9.0))


def cfg3_water96k(nu=6):
    pos, q, mass = water_box(32000, 98.6, (40, 40, 20), seed=3)
    return System(f"water96k_nu{nu}", pos, q, np.diag([98.6] * 3), mass,
                  dict(nu=nu, levels=4, top_grid=6, cutoff=10.0))


def cfg4_water1m():
    pos, q, mass = water_box(333333, 215.0, (70, 70, 70), seed=4)
    return System("water1m", pos, q, np.diag([215.0] * 3), mass,
                  dict(nu=6, levels=5, top_grid=7, cutoff=10.0))


CONFIGS = {
    "water300": cfg0_water300,
    "crystal8192": cfg1_crystal8192,
    "random3000": cfg2_random3000,
    "water96k": cfg3_water96k,
    "water96k_nu4": lambda: cfg3_water96k(4),
    "water1m": cfg4_water1m,
}
