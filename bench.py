"""Benchmark: MSM energy+force evaluations per second at 96,000 atoms on one
MI355X, mixed precision against plain FP64, with the rms relative force
error of the mixed path.

One step is one full evaluation (binning, short-range pairs, spreading,
restriction, top-level DFT, level stencils, interpolation; energy and
forces) of the water96k workload: 32,000 rigid 3-site waters in a cubic box
of 98.6 A, order-6 B-splines, cutoff 10 A, 4 levels (grids 48/24/12/6),
positions and charges resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nu 6] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MSM energy+force evals/s at 96k atoms, mixed vs FP64 GPU; rms rel force err"
PEAK = {"hbm": 8000.0, "fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # GB/s, TFLOP/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nu", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    return ap.parse_args()


def algorithmic(stage, m, P, nu):
    """Algorithmic work of one launch-group of a stage: (amount, unit, bound)."""
    N = m.N
    if stage == "partcl2partcl":  # P within-cutoff pairs, 35 + 4(nu-1) FLOP each
        return P * (35 + 4 * (nu - 1)), "flop", "valu"
    if stage == "grid2grid":  # 2 * sum_l prod(M_l) prod(s_l), l < L
        f = 0
        for l in range(1, m.levels):
            f += 2 * np.prod(m.grid_dims(l)) * np.prod(m.stencil_dims(l))
        return f, "flop", "mfma"
    if stage == "anterpolation":  # read atoms (x,y,z,q), write level-1 grid
        return N * 32 + 8 * np.prod(m.grid_dims(1)), "byte", "hbm"
    if stage == "interpolation":  # read atoms + grid, read-modify-write F
        return N * (32 + 4 + 48) + 8 * np.prod(m.grid_dims(1)), "byte", "hbm"
    return None


def run_gpu(args, sysm):
    import torch
    from msmpb_amd import MSM, _lib
    L = _lib.lib()
    dev = torch.device("cuda:0")
    p = dict(sysm.params)
    kw = dict(nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"], cutoff=p["cutoff"])
    m64 = MSM(sysm.pos, sysm.q, sysm.edges, precision="fp64", **kw)
    mmx = MSM(sysm.pos, sysm.q, sysm.edges, precision="mixed", **kw)
    pos = torch.tensor(sysm.pos, dtype=torch.float64, device=dev)
    q = torch.tensor(sysm.q, dtype=torch.float64, device=dev)
    out = torch.empty(4, dtype=torch.float64, device=dev)
    F = torch.empty((sysm.N, 3), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def step(m):
        rc = L.msm_eval_device(m._ff, vp(pos), vp(q), vp(out), vp(F), None, sp)
        if rc != 0:
            raise RuntimeError(_lib.last_error())

    result = {}
    for tag, m in (("fp64", m64), ("mixed", mmx)):
        for _ in range(args.warmup):
            step(m)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(m)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # stage breakdown in a separate, shorter loop: the interval events
        # are kept out of the timed region
        m.set_stage_timing(True)
        for _ in range(max(3, args.steps // 3)):
            step(m)
        torch.cuda.synchronize()
        stages = m.stage_times()
        m.set_stage_timing(False)
        step(m)
        torch.cuda.synchronize()
        result[tag] = dict(ms=(t1 - t0) * 1e3 / args.steps, stages=stages, E=out[0].item(),
                           F=F.cpu().numpy().copy(), m=m)
    # forward + adjoint (autograd) path, mixed precision
    from msmpb_amd import msm_energy
    pg = pos.clone().requires_grad_(True)
    for _ in range(args.warmup):
        e = msm_energy(mmx, pg, q)
        e.backward()
        pg.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e = msm_energy(mmx, pg, q)
        e.backward()
        pg.grad = None
    torch.cuda.synchronize()
    result["mixed_autograd_ms"] = (time.perf_counter() - t0) * 1e3 / args.steps
    result["pairs"] = m64.pair_count
    result["models"] = (m64, mmx)
    return result


def cpu_baseline(sysm):
    """The oracle (single-threaded C restatement of the reference) on one full
    evaluation of the same workload, pair-list build excluded like the
    reference's time_energy."""
    from oracle.oracle import Oracle
    p = sysm.params
    o = Oracle(sysm.pos, sysm.q, sysm.edges, nu=p["nu"], L=p["levels"], Mtop=p["top_grid"],
               a0=p["cutoff"])
    t0 = time.process_time()
    E, F, _, _ = o.energy(sysm.pos, want_phi=False)
    t = time.process_time() - t0
    return dict(value=1.0 / t, unit="evals/s", cores=1, kind="port",
                sample=f"1 full evaluation of the same {sysm.N}-atom workload (energy+forces, "
                       f"pair-list build excluded), {t:.2f} s CPU"), E, F


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is design-only)", file=sys.stderr)
        sys.exit(2)
    from msmpb_amd import inputs
    sysm = inputs.cfg3_water96k(args.nu)
    r = run_gpu(args, sysm)
    F64, Fmx = r["fp64"]["F"], r["mixed"]["F"]
    rms_rel = float(np.sqrt(np.sum((Fmx - F64) ** 2) / np.sum(F64 ** 2)))
    relE = abs(r["mixed"]["E"] - r["fp64"]["E"]) / abs(r["fp64"]["E"])
    ms_mx, ms_64 = r["mixed"]["ms"], r["fp64"]["ms"]
    m64, mmx = r["models"]
    # roofline of the dominant kernel of the mixed path
    stages = r["mixed"]["stages"]
    dom = max(stages, key=stages.get)
    roof = None
    alg = algorithmic(dom, mmx, r["pairs"], mmx.nu)
    if alg is not None and stages[dom] > 0:
        amount, unit, bound = alg
        if unit == "flop":
            ach = amount / stages[dom] / 1e12
            peak = PEAK["bf16"] if bound == "mfma" else PEAK["fp32"]
            roof = dict(bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak,
                        traffic=None, kernel=dom, work_per_launch=float(amount))
        else:
            ach = amount / stages[dom] / 1e9
            roof = dict(bound="hbm", achieved=ach, peak=PEAK["hbm"], unit="GB/s",
                        frac=ach / PEAK["hbm"], traffic=None, kernel=dom,
                        work_per_launch=float(amount))
    line = {
        "metric": METRIC, "value": 1e3 / ms_mx, "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_mx,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pairs + bf16-split mfma stencils, fp64 accumulation", "data": "synthetic",
        "config": {"workload": f"water96k: 32,000 rigid 3-site waters ({sysm.N} atoms), cubic "
                               f"box 98.6 A, nu={mmx.nu}, cutoff 10 A, {mmx.levels} levels, "
                               f"top grid {mmx.top_grid[0]} (finest {mmx.grid_dims(1)[0]})",
                   "precision": "mixed", "global_batch": sysm.N},
        "fp64_evals_per_s": 1e3 / ms_64, "fp64_ms_per_step": ms_64,
        "speedup_mixed_vs_fp64": ms_64 / ms_mx,
        "rms_rel_force_err": rms_rel, "rel_energy_err": relE,
        "mixed_forward_adjoint_ms": r["mixed_autograd_ms"],
        "stage_ms_mixed": {k: v * 1e3 for k, v in stages.items()},
        "stage_ms_fp64": {k: v * 1e3 for k, v in r["fp64"]["stages"].items()},
        "pairs_within_cutoff": r["pairs"],
        "roofline": roof,
    }
    if not args.no_cpu:
        cb, Ec, Fc = cpu_baseline(sysm)
        line["cpu_baseline"] = cb
        line["fp64_vs_cpu_rel_energy"] = abs(r["fp64"]["E"] - Ec) / abs(Ec)
        line["rms_rel_force_err_vs_oracle"] = float(np.sqrt(np.sum((Fmx - Fc) ** 2) / np.sum(Fc ** 2)))
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
