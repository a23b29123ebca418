"""Python entry points of msmpb's journal scripts, driving this build's CLI.

Mirrors ``code/msmpb_wrapper.py`` of hkayabilisim/msmpb (lines 7-221): the
same function names, arguments, return values and printed table, so the
journal scripts (``code/journal/*.py``) run unchanged against the GPU
binary.  Differences, all deliberate:

* the executable and data directory are module settings (``MSMPB_BINARY``,
  ``DATA_DIR``, overridable by the environment variables of the same names
  or per call); the reference hard-codes ``'../msmpb'`` and
  ``'../../data/'``.  ``DATA_DIR`` keeps the reference's default, the
  binary defaults to the CLI built by ``msmpb_amd.build``;
* ``runMSM`` accepts ``precision`` ("fp64" or "mixed") and ``layout``
  ("fixed" or "asis") keywords, passed to the CLI as ``--precision`` and
  ``--layout``; without them the command line is the reference's;
* ``runPME`` runs the PME CLI built by ``msmpb_amd.build`` (``PME_BINARY``,
  environment variable of the same name, or ``binary=``); the reference
  hard-codes ``'../pme'``.  Like the reference's pme it ignores ``nu`` (the
  order is 6 unless ``order=`` is given) and accepts ``precision=``;
* ``getFactors`` rejects M < 1 with ``ValueError`` (the reference loops
  forever on 0 and negatives).
"""
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MSMPB_BINARY = os.environ.get("MSMPB_BINARY", os.path.join(HERE, "bin", "msmpb"))
PME_BINARY = os.environ.get("PME_BINARY", os.path.join(HERE, "bin", "pme"))
DATA_DIR = os.environ.get("DATA_DIR", "../../data/")


def calculateM(eta, N):
    """Grid size ceil((eta N)^(1/3)) (msmpb_wrapper.py:7-8)."""
    return int(np.ceil(pow(eta * N, 1 / 3)))


def getFactors(M):
    """Distinct prime factors of M in increasing order (msmpb_wrapper.py:10-20)."""
    M = int(M)
    if M < 1:
        raise ValueError(f"getFactors: M must be >= 1, got {M}")
    out, p = [], 2
    while M > 1:
        if M % p == 0:
            out.append(p)
            while M % p == 0:
                M //= p
        p += 1
    return out


def isFFTFriendly(M):
    """True when 2, 3, 5 or 7 divides M -- the reference's test
    (msmpb_wrapper.py:22-27), which is weaker than 7-smoothness: 22 = 2 * 11
    counts as friendly, 11 and 1 do not."""
    return any(p in (2, 3, 5, 7) for p in getFactors(M))


def makeFFTFriendly(M, decrease=True):
    """M itself if friendly, else the nearest friendly size below (decrease)
    or above (msmpb_wrapper.py:29-36)."""
    step = -1 if decrease else 1
    while not isFFTFriendly(M):
        M += step
    return M


def nextFFTFriendly(M):
    """Smallest friendly size strictly above M (msmpb_wrapper.py:38-41)."""
    M += 1
    while not isFFTFriendly(M):
        M += 1
    return M


def prevFFTFriendly(M):
    """Largest friendly size strictly below M (msmpb_wrapper.py:43-46)."""
    M -= 1
    while not isFFTFriendly(M):
        M -= 1
    return M


def _cli_args(binary, data_dir, name, extra):
    return [binary, os.path.join(data_dir, name) if data_dir else name] + extra


def runMSM(benchmark, M, a0, nu, L=1, bincoor=None, potfile=None, accfile=None, *,
           precision=None, layout=None, binary=None, data_dir=None):
    """Run the CLI on ``data_dir + benchmark['name']`` (the path stem of the
    .ini/.pot/.acc files) and return its JSON record (msmpb_wrapper.py:48-66).
    ``benchmark`` needs the key 'name' (the reference also reads 'A' and
    'N' for unused locals)."""
    binary = binary or MSMPB_BINARY
    data_dir = DATA_DIR if data_dir is None else data_dir
    extra = ["-L", str(L), "--a0", str(a0), "--nu", str(nu), "-M", str(M)]
    for flag, val in (("--bincoor", bincoor), ("--potfile", potfile), ("--accfile", accfile)):
        if val is not None:
            extra += [flag, val]
    if precision is not None:
        extra += ["--precision", precision]
    if layout is not None:
        extra += ["--layout", layout]
    out = subprocess.run(_cli_args(binary, data_dir, benchmark["name"], extra),
                         stdout=subprocess.PIPE)
    return json.loads(out.stdout.decode("utf-8"))


def runPME(benchmark, M, a0, nu, bincoor=None, potfile=None, accfile=None, *,
           order=None, precision=None, binary=None, data_dir=None):
    """Run the PME CLI on ``data_dir + benchmark['name']`` and return its JSON
    record (msmpb_wrapper.py:68-81).  ``nu`` is passed as ``--nu``, which
    the PME program ignores as the reference's does; ``order`` sets it."""
    binary = binary or PME_BINARY
    data_dir = DATA_DIR if data_dir is None else data_dir
    extra = ["--a0", str(a0), "--nu", str(nu), "-M", str(M)]
    for flag, val in (("--bincoor", bincoor), ("--potfile", potfile), ("--accfile", accfile)):
        if val is not None:
            extra += [flag, val]
    if order is not None:
        extra += ["--order", str(order)]
    if precision is not None:
        extra += ["--precision", precision]
    out = subprocess.run(_cli_args(binary, data_dir, benchmark["name"], extra),
                         stdout=subprocess.PIPE)
    return json.loads(out.stdout.decode("utf-8"))


def fudgeFactor(results):
    """Per order nu = 4, 6, 8, 10: the geometric-mean ratio of measured to
    estimated relative force error (msmpb_wrapper.py:84-99).  An order with
    no runs yields nan (the reference divides by zero there as well)."""
    true = np.array([r["deltaF/Fref"] for r in results], dtype=float)
    est = np.array([r["deltaFest/Fref"] for r in results], dtype=float)
    nus = np.array([r["nu"] for r in results])
    factors = []
    for order in (4, 6, 8, 10):
        sel = nus == order
        n = int(np.count_nonzero(sel))
        if n == 0:
            factors.append(float("nan"))
            continue
        factors.append(10 ** ((np.sum(np.log10(true[sel])) - np.sum(np.log10(est[sel]))) / n))
    return factors


def displayResult(results, fudgefactors, filename):
    """Writes the estimate-versus-measurement table (msmpb_wrapper.py:101-115)."""
    with open(filename, "w") as fp:
        print("%20s %5s %2s %5s %2s %8s %2s %10s %10s %s" % (
            "data", "N", "A", "nbar", "M", "a0", "nu", "true", "est", "|log(true)-log(est)|"),
            file=fp)
        for r in results:
            M = r["TopLevelMx"]
            A = r["Edge row1"][0]
            nu = r["nu"]
            est = r["deltaFest/Fref"] * fudgefactors[int(nu / 2 - 2)]
            true = r["deltaF/Fref"]
            print("%20s %5d %2d %5.2f %2d %8.3f %2d %10.3e %10.3e %5.3f" % (
                r["data"], r["N"], A, 2 * r["cutoff"] * M / A, M, r["cutoff"], nu, true, est,
                np.abs(np.log(est) - np.log(true))), file=fp)


def _scatter_groups(ax, est, true, keys, label):
    import matplotlib.pyplot as plt
    for k in np.unique(keys):
        sel = keys == k
        ax.scatter(est[sel], true[sel], 500, label=label(k), edgecolor="k")
    plt.yscale("log")
    plt.xscale("log")
    plt.xlabel("estimated")
    plt.ylabel("true")
    plt.xlim(1e-8, 1e-2)
    plt.ylim(1e-8, 1e-2)
    ax.plot([1e-8, 1e-2], [1e-8, 1e-2], label=None)
    ax.grid(linestyle="dotted")
    ax.legend()


def _scaled(results, fudgefactors):
    est = np.array([fudgefactors[int(r["nu"] / 2 - 2)] * r["deltaFest/Fref"] for r in results])
    true = np.array([r["deltaF/Fref"] for r in results])
    return est, true


def plotEstimatedVersusTrue(ax, results, fudgefactors):
    """Estimated (fudge-scaled) versus measured error, grouped by nu
    (msmpb_wrapper.py:117-145); prints each order's fudge factor."""
    for i in range(4):
        print("Fudge factor: %f" % (fudgefactors[i]))
    est, true = _scaled(results, fudgefactors)
    nus = np.array([r["nu"] for r in results])
    _scatter_groups(ax, est, true, nus, lambda k: r"$\nu=" + str(int(k)) + "$")


def plotEstimatedVersusTrueWRTa0(ax, results, fudgefactors):
    """Same, grouped by cutoff (msmpb_wrapper.py:147-173)."""
    est, true = _scaled(results, fudgefactors)
    a0 = np.array([r["cutoff"] for r in results])
    _scatter_groups(ax, est, true, a0, lambda k: r"$a_0=" + str(k) + "$")


def plotEstimatedVersusTrueWRTM(ax, results, fudgefactors):
    """Same, grouped by top-level grid size (msmpb_wrapper.py:175-197)."""
    est, true = _scaled(results, fudgefactors)
    M = np.array([r["TopLevelMx"] for r in results])
    _scatter_groups(ax, est, true, M, lambda k: r"$M=" + str(k) + "$")


def plotEstimatedVersusTrueWRTN(ax, results, fudgefactors):
    """Same, grouped by atom count (msmpb_wrapper.py:199-221)."""
    est, true = _scaled(results, fudgefactors)
    N = np.array([r["N"] for r in results])
    _scatter_groups(ax, est, true, N, lambda k: r"$N=" + str(k) + "$")
