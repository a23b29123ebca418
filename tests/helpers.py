"""Shared helpers for the test-suite (data loading, error metrics)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "data")


def load_dataset(name):
    """(edges, q, pos, mass) of a reference data set under tests/golden/data."""
    from msmpb_amd.inputs import read_ini
    path = os.path.join(DATA, name + ".ini")
    if not os.path.exists(path):
        path += ".gz"
    return read_ini(path)


def rel_err(a, b):
    return abs(a - b) / abs(b)


def force_scale(q, edges):
    """Typical Coulomb force of the system, (sum q^2 / N) / (V / N)^(2/3)
    (the reference's Fref, code/msmpb.c:273)."""
    n = len(q)
    return (np.sum(q * q) / n) / (abs(np.linalg.det(edges)) / n) ** (2.0 / 3.0)


def force_err(F, Fref, scale=None):
    """max_i |F_i - Fref_i| / rms(Fref): relative to the typical force, so
    atoms whose reference force is ~0 do not blow the metric up.  For
    near-perfect crystals (all forces ~0) the denominator is floored at
    1e-3 of the system's force scale."""
    rms = np.sqrt(np.mean(np.sum(Fref ** 2, axis=1)))
    if scale is not None:
        rms = max(rms, 1e-3 * scale)
    return np.max(np.linalg.norm(F - Fref, axis=1)) / rms


def rms_rel_force(F, Fref):
    """sqrt(sum |F - Fref|^2 / sum |Fref|^2)."""
    return np.sqrt(np.sum((F - Fref) ** 2) / np.sum(Fref ** 2))
