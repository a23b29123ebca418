"""The PME oracle (oracle/pme_oracle.c) pinned before it checks the GPU.

What pins it, function by function (code/pme0.c, code/pme1.c of the
reference):
  - whole evaluation: the total energies, Ewald parameters and pair counts
    the reference's own pme runs published for Box44 (results/Figure4.json,
    tests/golden/published.json "Figure4_pme"), to 1e-10 relative;
  - constants, self term and Fourier normalisation: Madelung energies of
    perfect CsCl and NaCl crystals (closed-form lattice sums);
  - real/reciprocal split: the energy of a charged, disordered system does
    not depend on the splitting parameter beta;
  - forces and potentials: central differences of the energy;
  - B-spline weights: partition of unity, and agreement with the MSM
    oracle's independent piecewise-polynomial evaluation.
"""
import json
import os

import numpy as np
import pytest

from helpers import DATA, load_dataset, rel_err

PUB = json.load(open(os.path.join(os.path.dirname(DATA), "published.json")))["Figure4_pme"]


def _unique_runs():
    seen, out = set(), []
    for t in PUB:
        if (t["M"], t["a0"]) not in seen:
            seen.add((t["M"], t["a0"]))
            out.append(t)
    return out


def test_published_runs_all_used_order_6():
    """The reference's pme ignores --nu (pme.c:49-50): every run reports the
    default order 6 whatever the journal asked for, and runs that differ
    only in the requested order agree to the last digit."""
    assert {t["nu"] for t in PUB} == {6}
    assert {t["nu_arg"] for t in PUB} == {4, 6}
    by = {}
    for t in PUB:
        by.setdefault((t["M"], t["a0"]), set()).add(t["utotal"])
    assert all(len(v) == 1 for v in by.values())
    assert len(_unique_runs()) == 18


@pytest.mark.parametrize("M,a0", [(48, 9), (30, 12)])
def test_oracle_reproduces_published_box44(M, a0):
    from oracle.oracle import PmeOracle
    t = next(t for t in PUB if t["M"] == M and t["a0"] == a0)
    edges, q, pos, _ = load_dataset("Box44")
    o = PmeOracle(pos, q, edges, nu=t["nu"], M=M, a0=a0)
    assert round(o.beta, 6) == t["beta"]
    assert o.pair_count == t["nlist_interaction_count"]
    E, F, phi, terms = o.energy(pos, want_phi=False)
    assert rel_err(E, t["utotal"]) < 1e-10, (E, t["utotal"])


def _cscl(n):
    g = np.arange(n)
    c = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.astype(float)
    pos = np.concatenate([c, c + 0.5])
    q = np.concatenate([np.ones(len(c)), -np.ones(len(c))])
    return pos, q, np.eye(3) * n


def _nacl(n):
    g = np.arange(2 * n)
    c = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    q = np.where(c.sum(axis=1) % 2 == 0, 1.0, -1.0)
    return c.astype(float) * 0.5, q, np.eye(3) * n


# Madelung constants per ion pair, referred to the nearest-neighbour distance
MADELUNG = {"cscl": (1.76267477307099, np.sqrt(3) / 2, _cscl),
            "nacl": (1.74756459463318, 0.5, _nacl)}


@pytest.mark.parametrize("lattice", list(MADELUNG))
def test_oracle_madelung_energy(lattice):
    from oracle.oracle import PmeOracle
    alpha, d, make = MADELUNG[lattice]
    pos, q, edges = make(4)
    o = PmeOracle(pos, q, edges, nu=10, M=48, a0=1.9, tolDir=1e-12)
    E, F, phi, _ = o.energy(pos)
    exact = -(len(q) / 2) * alpha / d
    assert rel_err(E, exact) < 1e-11, (E, exact)
    assert np.max(np.abs(F)) < 1e-10          # every ion is a centre of symmetry
    # every ion sits at potential -q alpha / d
    assert np.max(np.abs(phi + q * alpha / d)) < 1e-10


def _random_charged(n=40, seed=7, box=(11.0, 12.5, 13.0)):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1, (n, 3)) * np.array(box)
    q = rng.normal(size=n)
    q -= q.mean() - 0.3 / n  # net charge +0.3 e: the neutralising term matters
    return pos, q, np.diag(box)


def test_oracle_energy_independent_of_beta():
    from oracle.oracle import PmeOracle
    pos, q, edges = _random_charged()
    Es = []
    for tol in (1e-9, 1e-12):
        o = PmeOracle(pos, q, edges, nu=10, M=(60, 64, 72), a0=5.4, tolDir=tol)
        Es.append(o.energy(pos, want_phi=False)[0])
    assert rel_err(Es[0], Es[1]) < 2e-9, Es


def test_oracle_forces_and_potentials_are_derivatives():
    from oracle.oracle import PmeOracle
    pos, q, edges = _random_charged(n=24, seed=3)
    o = PmeOracle(pos, q, edges, nu=6, M=(20, 21, 24), a0=5.0)
    E, F, phi, _ = o.energy(pos)
    h, hq = 1e-5, 1e-6
    for i in (0, 5, 17):
        for k in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, k] += h
            p2[i, k] -= h
            fd = -(o.energy(p1, want_phi=False)[0] - o.energy(p2, want_phi=False)[0]) / (2 * h)
            assert abs(fd - F[i, k]) < 1e-6 * max(1.0, abs(F[i, k])), (i, k, fd, F[i, k])
        q1, q2 = q.copy(), q.copy()
        q1[i] += hq
        q2[i] -= hq
        fdq = (o.energy(pos, q1, want_phi=False)[0] - o.energy(pos, q2, want_phi=False)[0]) / (2 * hq)
        assert abs(fdq - phi[i]) < 1e-6 * max(1.0, abs(phi[i])), (i, fdq, phi[i])


@pytest.mark.parametrize("nu", [4, 6, 8, 10])
def test_oracle_bspline_weights(nu):
    from oracle.oracle import Oracle, pme_bspline
    pos, q, edges = _random_charged(n=8)
    msm = Oracle(pos, q, edges, nu=nu, L=1, Mtop=16, a0=4.0)
    for t in (0.0, 0.1234, 0.5, 0.97):
        w, dw = pme_bspline(nu, t)
        assert abs(w.sum() - 1.0) < 1e-14 and abs(dw.sum()) < 1e-13
        for i in range(nu):
            v, d = msm.bspline(i, t)  # from the stored pieces (forcefield1.c:609-637)
            assert abs(w[i] - v) < 1e-14 and abs(dw[i] - d) < 1e-13, (i, t)
