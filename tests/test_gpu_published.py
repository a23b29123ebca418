"""GPU FP64 against the reference's own published results.

results/TableII.json, results/Figure4.json and results/Figure6.json of the
reference hold the full-precision total energy (`utotal`) and the
neighbour-list pair count and list capacity of every run on Box44 (65,850 atoms of SPC/E
water, code/journal/TableII_run.py, Figure4_run.py, Figure6_run.py); the
fixture tests/golden/published.json carries them.  The plain FP64 path must
reproduce each energy to the 1e-10 relative bar and each pair count and
capacity exactly, for every level count (1-3), order (4, 6), cutoff (9, 12 A) and
grid the journal swept.  `poterror` is |E_pot - E| / |E_pot| with E_pot the
first line of data/Box44.pot (code/msmpb.c), a function of the energy alone.
"""
import json
import os

import pytest

from helpers import DATA, load_dataset, rel_err

pytestmark = pytest.mark.gpu

PUB = json.load(open(os.path.join(os.path.dirname(DATA), "published.json")))
RUNS = [(fig, k) for fig in ("TableII", "Figure4", "Figure6") for k in range(len(PUB[fig]))]


@pytest.fixture(scope="module")
def box44():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    edges, q, pos, _ = load_dataset("Box44")
    with open(os.path.join(DATA, "Box44.pot")) as f:
        epot = float(f.readline().split()[0])
    return edges, q, pos, epot


@pytest.mark.parametrize("fig,idx", RUNS)
def test_published_box44_energy_gpu(fig, idx, box44):
    from msmpb_amd import MSM
    t = PUB[fig][idx]
    edges, q, pos, epot = box44
    L = t["L"]
    print(f"[published] {fig}[{idx}] L={L} nu={t['nu']} M={t['M']} a0={t['a0']}", flush=True)
    m = MSM(pos, q, edges, nu=t["nu"], levels=L, top_grid=t["M"] // 2 ** (L - 1),
            cutoff=t["a0"], precision="fp64")
    try:
        assert m.top_grid[0] == t["TopLevelMx"]
        E, _ = m.energy(pos)
        assert rel_err(E, t["utotal"]) < 1e-10, (E, t["utotal"])
        assert rel_err(abs(epot - E) / abs(epot), t["poterror"]) < 1e-6
        assert m.pair_count == t["nlist_interaction_count"]
        assert m.nlist_len == t["nlist_len"]  # the pair-list capacity the runs report
    finally:
        m.close()
