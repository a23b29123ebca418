"""msmpb_amd.msmpb_wrapper: the journal-script entry points of msmpb
(code/msmpb_wrapper.py:7-221) over this build's CLI.

The helpers are pinned by arithmetic (prime factors, the reference's weak
"friendly" test: some factor among 2, 3, 5, 7), runMSM by the exact command
line it hands to the executable (a stub that echoes its argv as JSON), and
by one GPU run of the real CLI against the oracle energy.
"""
import math
import os
import stat
import sys

import pytest

from helpers import ROOT, load_dataset, rel_err
from msmpb_amd import msmpb_wrapper as W


def test_calculate_m():
    assert W.calculateM(5.7, 300) == math.ceil((5.7 * 300) ** (1 / 3)) == 12
    assert W.calculateM(0.83, 96000) == 44


@pytest.mark.parametrize("M,factors", [(1, []), (2, [2]), (12, [2, 3]), (97, [97]),
                                       (360, [2, 3, 5]), (1001, [7, 11, 13])])
def test_get_factors(M, factors):
    assert W.getFactors(M) == factors


def test_get_factors_rejects_nonpositive():
    with pytest.raises(ValueError):
        W.getFactors(0)


def test_friendly_is_the_weak_reference_test():
    # friendly = divisible by 2, 3, 5 or 7 (not 7-smoothness)
    for M in range(1, 200):
        assert W.isFFTFriendly(M) == any(M % p == 0 for p in (2, 3, 5, 7))
    assert W.isFFTFriendly(22) and not W.isFFTFriendly(11) and not W.isFFTFriendly(1)


def test_friendly_walks():
    assert W.makeFFTFriendly(12) == 12
    assert W.makeFFTFriendly(11) == 10
    assert W.makeFFTFriendly(11, decrease=False) == 12
    assert W.makeFFTFriendly(143, decrease=False) == 144
    assert W.nextFFTFriendly(10) == 12 and W.nextFFTFriendly(12) == 14
    assert W.prevFFTFriendly(12) == 10 and W.prevFFTFriendly(14) == 12


def _stub_binary(tmp_path):
    p = tmp_path / "echo_msmpb"
    p.write_text(f"#!{sys.executable}\nimport json, sys\nprint(json.dumps({{'argv': sys.argv[1:]}}))\n")
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    return str(p)


def test_run_msm_command_line(tmp_path):
    stub = _stub_binary(tmp_path)
    bench = {"name": "spceN300", "A": 20.0, "N": 300}
    out = W.runMSM(bench, 24, 9.0, 4, binary=stub, data_dir="/data/")
    assert out["argv"] == ["/data/spceN300", "-L", "1", "--a0", "9.0", "--nu", "4", "-M", "24"]
    out = W.runMSM(bench, 6, 9, 6, L=2, bincoor="c.coor", potfile="p.pot", accfile="a.acc",
                   binary=stub, data_dir="")
    assert out["argv"] == ["spceN300", "-L", "2", "--a0", "9", "--nu", "6", "-M", "6",
                           "--bincoor", "c.coor", "--potfile", "p.pot", "--accfile", "a.acc"]
    out = W.runMSM(bench, 6, 9, 6, precision="mixed", layout="asis", binary=stub, data_dir="")
    assert out["argv"][-4:] == ["--precision", "mixed", "--layout", "asis"]


def test_default_data_dir_matches_reference():
    assert W.DATA_DIR == os.environ.get("DATA_DIR", "../../data/")


def test_run_pme_command_line(tmp_path):
    """runPME passes the reference's arguments (msmpb_wrapper.py:68-81)."""
    stub = _stub_binary(tmp_path)
    bench = {"name": "Box44"}
    out = W.runPME(bench, 48, 9, 4, binary=stub, data_dir="/data/")
    assert out["argv"] == ["/data/Box44", "--a0", "9", "--nu", "4", "-M", "48"]
    out = W.runPME(bench, 30, 12.0, 6, bincoor="c.coor", potfile="p.pot", accfile="a.acc",
                   order=8, precision="mixed", binary=stub, data_dir="")
    assert out["argv"] == ["Box44", "--a0", "12.0", "--nu", "6", "-M", "30", "--bincoor", "c.coor",
                           "--potfile", "p.pot", "--accfile", "a.acc", "--order", "8",
                           "--precision", "mixed"]
    assert W.PME_BINARY == os.environ.get(
        "PME_BINARY", os.path.join(ROOT, "msmpb_amd", "bin", "pme"))


def test_fudge_factor_is_geometric_mean_ratio():
    res = [{"nu": 4, "deltaF/Fref": 2e-4, "deltaFest/Fref": 1e-4},
           {"nu": 4, "deltaF/Fref": 8e-4, "deltaFest/Fref": 1e-4},
           {"nu": 6, "deltaF/Fref": 1e-5, "deltaFest/Fref": 1e-5}]
    f = W.fudgeFactor(res)
    assert f[0] == pytest.approx(4.0) and f[1] == pytest.approx(1.0)
    assert math.isnan(f[2]) and math.isnan(f[3])


def test_display_result_table(tmp_path):
    r = {"TopLevelMx": 24, "Edge row1": [20.0, 0, 0], "nu": 4, "cutoff": 9.0, "data": "spceN300",
         "N": 300, "deltaF/Fref": 2.5e-4, "deltaFest/Fref": 1e-4}
    fn = tmp_path / "t.txt"
    W.displayResult([r], [2.0, 1.0, 1.0, 1.0], str(fn))
    lines = fn.read_text().splitlines()
    assert lines[0].split() == ["data", "N", "A", "nbar", "M", "a0", "nu", "true", "est",
                                "|log(true)-log(est)|"]
    f = lines[1].split()
    assert f[0] == "spceN300" and f[1] == "300" and f[3] == "21.60" and f[4] == "24"
    assert float(f[7]) == pytest.approx(2.5e-4) and float(f[8]) == pytest.approx(2e-4)
    assert float(f[9]) == pytest.approx(abs(math.log(2e-4) - math.log(2.5e-4)), abs=1e-3)


@pytest.mark.gpu
def test_run_msm_real_cli_matches_oracle(oracle_mod, tmp_path, monkeypatch):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    binary = os.path.join(ROOT, "msmpb_amd", "bin", "msmpb")
    data = os.path.join(ROOT, "tests", "golden", "data") + "/"
    monkeypatch.chdir(tmp_path)  # the CLI writes msmpb.acc / msmpb.pot into the working directory
    print("[wrapper] runMSM NaClN512 L=3 nu=6 M=4 a0=3", flush=True)
    out = W.runMSM({"name": "NaClN512"}, 4, 3.0, 6, L=3, binary=binary, data_dir=data)
    edges, q, pos, _ = load_dataset("NaClN512")
    o = oracle_mod.Oracle(pos, q, edges, nu=6, L=3, Mtop=4, a0=3.0)
    E, _, _, _ = o.energy(pos, want_phi=False)
    assert out["nu"] == 6 and out["N"] == 512
    assert rel_err(out["utotal"], E) < 1e-10, (out["utotal"], E)


# ---------------------------------------------------------------- journal sweeps
# The reference's journal scripts (code/journal/Figure1.py, Figure2.py,
# TableII_run.py) drive runMSM over parameter grids and tabulate the CLI's
# JSON; the published tables (results/Figure1.txt, Figure2.txt,
# TableII.json, in tests/golden/published.json) are reproduced here end to
# end through this build's runMSM and GPU CLI (FP64): force errors against
# the Ewald .acc files to the printed 4 digits, the a-priori estimate, and
# the Box44 total energy to full precision.
import json  # noqa: E402

PUB = json.load(open(os.path.join(ROOT, "tests", "golden", "published.json")))
SPCE = {300: "spceN300", 600: "spceN600", 900: "spceN900", 2250: "spceN2250"}


def _cli(monkeypatch, tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    monkeypatch.chdir(tmp_path)
    return os.path.join(ROOT, "msmpb_amd", "bin", "msmpb"), os.path.join(ROOT, "tests", "golden", "data") + "/"


@pytest.mark.gpu
@pytest.mark.parametrize("row", list(range(0, 240, 10)))
def test_journal_figure1_through_cli(row, tmp_path, monkeypatch):
    binary, data = _cli(monkeypatch, tmp_path)
    t = PUB["Figure1"][row]
    print(f"[journal] Figure1 row {row}: N={t['N']} M={t['M']} a0={t['a0']} nu={t['nu']}", flush=True)
    out = W.runMSM({"name": SPCE[t["N"]]}, t["M"], t["a0"], t["nu"], binary=binary, data_dir=data)
    assert rel_err(out["deltaF/Fref"], t["deltaF/Fref"]) < 2e-3, (out["deltaF/Fref"], t)


@pytest.mark.gpu
@pytest.mark.parametrize("row", list(range(0, 96, 8)))
def test_journal_figure2_through_cli(row, tmp_path, monkeypatch):
    binary, data = _cli(monkeypatch, tmp_path)
    t = PUB["Figure2"][row]
    print(f"[journal] Figure2 row {row}: {t['data']} M={t['M']} a0={t['a0']} nu={t['nu']}", flush=True)
    out = W.runMSM({"name": t["data"]}, t["M"], t["a0"], t["nu"], binary=binary, data_dir=data)
    assert rel_err(out["deltaF/Fref"], t["deltaF/Fref"]) < 2e-3, (out["deltaF/Fref"], t)
    assert rel_err(out["deltaFest/Fref"], t["deltaFest/Fref"]) < 2e-3, (out["deltaFest/Fref"], t)


def _box44_plain(data, tmp_path):
    import gzip
    import shutil
    for ext in (".ini", ".pot"):  # the CLIs read plain files; the fixture is compressed
        src = os.path.join(data, "Box44" + ext)
        if os.path.exists(src):
            shutil.copy(src, tmp_path / ("Box44" + ext))
        else:
            with gzip.open(src + ".gz", "rb") as f, open(tmp_path / ("Box44" + ext), "wb") as g:
                shutil.copyfileobj(f, g)


@pytest.mark.gpu
def test_journal_table2_through_cli(tmp_path, monkeypatch):
    binary, data = _cli(monkeypatch, tmp_path)
    _box44_plain(data, tmp_path)
    t = PUB["TableII"][0]
    print(f"[journal] TableII Box44 L={t['L']} nu={t['nu']} M={t['M']} a0={t['a0']}", flush=True)
    out = W.runMSM({"name": "Box44"}, t["M"] // 2 ** (t["L"] - 1), t["a0"], t["nu"], L=t["L"],
                   binary=binary, data_dir=str(tmp_path) + "/")
    assert out["TopLevelMx"] == t["TopLevelMx"]
    assert rel_err(out["utotal"], t["utotal"]) < 1e-10, (out["utotal"], t["utotal"])
    assert out["nlist_interaction_count"] == t["nlist_interaction_count"]
    assert rel_err(out["poterror"], t["poterror"]) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("row", [0, 13])
def test_journal_figure4_pme_through_cli(row, tmp_path, monkeypatch):
    """code/journal/Figure4_run.py calls runPME(Box44, M, a0, nu); the GPU
    PME CLI reproduces the published record (order 6 whatever nu asks)."""
    _, data = _cli(monkeypatch, tmp_path)
    _box44_plain(data, tmp_path)
    t = PUB["Figure4_pme"][row]
    print(f"[journal] Figure4 PME Box44 M={t['M']} a0={t['a0']} nu={t['nu_arg']}", flush=True)
    out = W.runPME({"name": "Box44"}, t["M"], t["a0"], t["nu_arg"],
                   binary=os.path.join(ROOT, "msmpb_amd", "bin", "pme"),
                   data_dir=str(tmp_path) + "/")
    assert out["nu"] == 6 and out["TopLevelMx"] == t["TopLevelMx"]
    assert rel_err(out["utotal"], t["utotal"]) < 1e-10, (out["utotal"], t["utotal"])
    assert out["beta"] == t["beta"] and out["tol_dir"] == t["tol_dir"]
    assert out["nlist_interaction_count"] == t["nlist_interaction_count"]
    assert rel_err(out["poterror"], t["poterror"]) < 1e-6
    assert os.path.exists(tmp_path / "pme.acc") and os.path.exists(tmp_path / "pme.pot")
