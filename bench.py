"""Benchmark: MSM energy+force evaluations per second at 96,000 atoms on one
MI355X, mixed precision against plain FP64, with the rms relative force
error of the mixed path.

One step is one full evaluation (binning, short-range pairs, spreading,
restriction, level stencils, top-level DFT, prolongation, interpolation;
energy and forces) of the water96k workload: 32,000 rigid 3-site waters in a
cubic box of 98.6 A, order-6 B-splines, cutoff 10 A, 4 levels (grids
48/24/12/6), positions and charges resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nu 6] [--no-cpu]
    python bench.py --workload water1m      # 1M-atom run, its own JSON line
    python bench.py --copy-calibration MIB  # plain copy for counter scaling
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
PME_GRID = 96  # PME comparator grid for water96k: spacing 1.03 A, order 6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "water96k_hbm_traffic.txt")
SQ_FILE = os.path.join(ROOT, "profiles", "water96k_kernel_stats.md")
PME_SQ_FILE = os.path.join(ROOT, "profiles", "water96k_pme_kernel_stats.md")
# VALU issue rate of gfx950: a wave64 VALU instruction occupies one of a
# CU's four SIMDs for 4 cycles; 256 CUs at 2.4 GHz
VALU_WAVE_INSTS_PER_S = 256 * 4 * 2.4e9 / 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nu", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="water96k", choices=["water96k", "water1m", "pme96k"],
                    help="pme96k: only the PME comparator on the water96k system")
    ap.add_argument("--no-pme", action="store_true",
                    help="skip the PME comparator sub-measurement of water96k")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the reference "
                         "kernel for scaling FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """Plain copy kernel of known size: reads and writes mib MiB per launch."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": n * 4, "launches": 5}))


def algorithmic(stage, m, P, nu):
    """Algorithmic work of one launch of a stage's kernel: (amount, unit, kernel)."""
    N = m.N
    mixed = "true" if m.precision == "mixed" else "false"
    if stage == "partcl2partcl":  # P within-cutoff pairs i < j, 35 + 4(nu-1) FLOP each
        # the orthorhombic instance (template arguments: precision, order, triclinic)
        return P * (35 + 4 * (nu - 1)), "flop", f"k_pair_cell<{mixed}, {nu}, false>"
    if stage == "grid2grid":  # 2 * sum_l prod(M_l) prod(s_l), l < L
        f = sum(2 * np.prod(m.grid_dims(l)) * np.prod(m.stencil_dims(l))
                for l in range(1, m.levels))
        return f, "flop", "k_g2g_rows<" if mixed == "true" else "k_g2g_tiles<0"
    if stage == "anterpolation":  # N (2 nu^3 + 6 nu^2) FLOP
        return N * (2 * nu ** 3 + 6 * nu ** 2), "flop", f"k_spread_bricks<{nu},"
    if stage == "interpolation":  # N (8 nu^3 + 6 nu^2) FLOP
        return N * (8 * nu ** 3 + 6 * nu ** 2), "flop", f"k_gather_tile<{nu},"
    return None


def traffic_of(kernel):
    """Calibrated HBM bytes per dispatch of `kernel` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("#") or not line.strip():
            continue
        name, _, _, _, b = line.rsplit(None, 4)
        if name.strip().startswith(kernel):
            return float(b)
    return None


def valu_insts_of(kernel, path=SQ_FILE):
    """SQ_INSTS_VALU per dispatch of `kernel` from a committed rocprofv3
    SQ-counter summary (None if absent)."""
    if not os.path.exists(path):
        return None
    cols = None
    for line in open(path):
        if line.startswith("| kernel | WAVES"):
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            continue
        if cols and line.startswith("| `" + kernel + "`"):
            vals = [c.strip() for c in line.strip().strip("|").split("|")]
            return float(vals[cols.index("INSTS_VALU")])
    return None


def rebuild_ms(sysm, precision="mixed", reps=5):
    """FF_rebuild (box change) time: stencils and MFMA operands regenerated
    on the GPU for a box scaled by 1 + 1e-4 k, median of reps calls."""
    import torch
    from msmpb_amd import MSM, _lib
    p = sysm.params
    m = MSM(sysm.pos, sysm.q, sysm.edges, precision=precision, nu=p["nu"], levels=p["levels"],
            top_grid=p["top_grid"], cutoff=p["cutoff"])
    dp = ctypes.POINTER(ctypes.c_double)
    ts = []
    for k in range(reps + 1):
        e = np.ascontiguousarray(sysm.edges * (1.0 + 1e-4 * (k + 1)))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.lib().FF_rebuild(m._ff, e.ctypes.data_as(dp), None, ctypes.c_double(0.0))
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    m.close()
    return float(np.median(ts[1:]) * 1e3)


def device_free():
    import torch
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info(0)[0]


def run_precisions(args, sysm, precisions=("fp64", "mixed"), steps=None, warmup=None, kw=None):
    """Timed evaluations per precision through msm_eval_device: a plain timed
    region of K steps, then K more steps with per-stage HIP events on the
    streams the kernels run on (kept out of the plain timed region)."""
    import torch
    from msmpb_amd import MSM, _lib
    L = _lib.lib()
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    dev = torch.device("cuda:0")
    p = dict(sysm.params)
    if kw is None:
        kw = dict(nu=p["nu"], levels=p["levels"], top_grid=p["top_grid"], cutoff=p["cutoff"])
    pos = torch.tensor(sysm.pos, dtype=torch.float64, device=dev)
    q = torch.tensor(sysm.q, dtype=torch.float64, device=dev)
    out = torch.empty(4, dtype=torch.float64, device=dev)
    F = torch.empty((sysm.N, 3), dtype=torch.float64, device=dev)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    res = {}
    for tag in precisions:
        free0 = device_free()
        m = MSM(sysm.pos, sysm.q, sysm.edges, precision=tag, **kw)
        footprint = free0 - device_free()

        def step():
            rc = L.msm_eval_device(m._ff, vp(pos), vp(q), vp(out), vp(F), None, sp)
            if rc != 0:
                raise RuntimeError(_lib.last_error())

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        m.set_stage_timing(True)
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        stages = m.stage_times()
        m.set_stage_timing(False)
        step()
        torch.cuda.synchronize()
        res[tag] = dict(ms=(t1 - t0) * 1e3 / steps, stages=stages, E=out[0].item(),
                        F=F.cpu().numpy().copy(), m=m, footprint=footprint)
    res["pos"], res["q"] = pos, q
    return res


def autograd_ms(m, pos, q, steps, warmup, recompute=None):
    """Forward energy + adjoint (dE/dr, dE/dq) through torch.autograd."""
    import torch
    from msmpb_amd import msm_energy
    pg = pos.clone().requires_grad_(True)
    qg = q.clone().requires_grad_(True)
    for _ in range(warmup):
        msm_energy(m, pg, qg, recompute=recompute).backward()
        pg.grad = qg.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        msm_energy(m, pg, qg, recompute=recompute).backward()
        pg.grad = qg.grad = None
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def host_cpu():
    """(CPU model, logical CPUs this process may use) of the host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return model, n


def timed_evals(fn, reps):
    """Median (and min, max) of reps single-thread CPU times of fn(); the
    result of the last call."""
    ts, out = [], None
    for _ in range(reps):
        t0 = time.process_time()
        out = fn()
        ts.append(time.process_time() - t0)
    return float(np.median(ts)), min(ts), max(ts), out


def cpu_baseline(sysm, reps=5):
    """The oracle (single-threaded C restatement of the reference) on full
    evaluations of the same workload, pair-list build excluded like the
    reference's time_energy: the median of reps evaluations."""
    from oracle.oracle import Oracle
    p = sysm.params
    o = Oracle(sysm.pos, sysm.q, sysm.edges, nu=p["nu"], L=p["levels"], Mtop=p["top_grid"],
               a0=p["cutoff"])
    t, tmin, tmax, (E, F, _, _) = timed_evals(lambda: o.energy(sysm.pos, want_phi=False), reps)
    model, ncpu = host_cpu()
    return dict(value=1.0 / t, unit="evals/s", cores=1, kind="port", cpu=model, nproc=ncpu,
                sample=f"median of {reps} full evaluations of the same {sysm.N}-atom workload "
                       f"(energy+forces, pair-list build excluded): {t:.2f} s CPU each "
                       f"(min {tmin:.2f}, max {tmax:.2f}), one thread"), E, F


def rms_rel(Fa, Fb):
    return float(np.sqrt(np.sum((Fa - Fb) ** 2) / np.sum(Fb ** 2)))


def roofline(stages, m, pairs, nu):
    """Dominant kernel of the evaluation -- the pair kernel: the largest GPU
    time per evaluation (profiles/water96k_kernel_stats.md) and the only
    kernel on its stream, so that the HIP events around it time it alone
    (the chain's stage events on the caller's stream also span the waits
    for the other streams) -- its achieved rate from that duration against
    the gfx950 peak, with its calibrated HBM traffic."""
    dom = "partcl2partcl"
    alg = algorithmic(dom, m, pairs, nu)
    if alg is None or stages[dom] <= 0:
        return None
    amount, unit, kernel = alg
    ach = amount / stages[dom] / 1e12
    peak = {"partcl2partcl": PEAK["fp32"] if m.precision == "mixed" else PEAK["fp64"],
            "grid2grid": PEAK["bf16"] if m.precision == "mixed" else PEAK["fp64"]}.get(
                dom, PEAK["fp64"])
    # the pair kernel is bound by VALU issue (FP32 or FP64 vector arithmetic),
    # the level stencils by the matrix cores
    bound = "valu" if dom == "partcl2partcl" else "mfma"
    out = dict(bound=bound, achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak,
               traffic=traffic_of(kernel), kernel=kernel, stage=dom,
               work_per_launch=float(amount), launch_ms=stages[dom] * 1e3)
    vi = valu_insts_of(kernel)
    if vi is not None:  # the pair kernel is bound by VALU issue, not by FLOP
        out["valu"] = {"insts_per_launch": vi,
                       "issue_frac": vi / VALU_WAVE_INSTS_PER_S / stages[dom],
                       "source": "SQ_INSTS_VALU, profiles/water96k_kernel_stats.md"}
    return out


def pme_cpu_baseline(sysm, reps=3):
    """The PME oracle (single-threaded C restatement of pme0.c/pme1.c) on
    evaluations of the same workload, pair count excluded: the median of reps."""
    from oracle.oracle import PmeOracle
    o = PmeOracle(sysm.pos, sysm.q, sysm.edges, nu=6, M=PME_GRID, a0=10.0)
    t, tmin, tmax, (E, F, _, _) = timed_evals(lambda: o.energy(sysm.pos, want_phi=False), reps)
    model, ncpu = host_cpu()
    return dict(value=1.0 / t, unit="evals/s", cores=1, kind="port", cpu=model, nproc=ncpu,
                sample=f"median of {reps} full PME evaluations of the same {sysm.N}-atom "
                       f"workload (energy+forces): {t:.2f} s CPU each (min {tmin:.2f}, "
                       f"max {tmax:.2f}), one thread"), E, F


def pme_comparator(args, sysm, F64=None):
    """Smooth PME (order 6, 96^3 grid, cutoff 10 A) on the same system."""
    kp = dict(nu=6, top_grid=PME_GRID, cutoff=10.0, method="pme")
    rp = run_precisions(args, sysm, kw=kp)
    out = {"config": f"smooth PME, order 6, grid {PME_GRID}^3, cutoff 10 A, tolDir 1e-5",
           "fp64_evals_per_s": 1e3 / rp["fp64"]["ms"], "mixed_evals_per_s": 1e3 / rp["mixed"]["ms"],
           "rms_rel_force_err": rms_rel(rp["mixed"]["F"], rp["fp64"]["F"]),
           "stage_ms_fp64": {k: v * 1e3 for k, v in rp["fp64"]["stages"].items()},
           "stage_ms_mixed": {k: v * 1e3 for k, v in rp["mixed"]["stages"].items()}}
    if F64 is not None:
        out["rms_rel_force_diff_msm_fp64"] = rms_rel(F64, rp["fp64"]["F"])
    # the dominant kernel (the Ewald pairs, mixed): VALU issue against the
    # machine's issue rate; no FLOP roofline (erfc and exp dominate the term)
    st = rp["mixed"]["stages"].get("partcl2partcl", 0.0)
    vi = valu_insts_of("k_pair_cell<true, 0, false>", PME_SQ_FILE)
    if vi is not None and st > 0:
        out["valu"] = {"kernel": "k_pair_cell<true, 0, false>", "insts_per_launch": vi,
                       "issue_frac": vi / VALU_WAVE_INSTS_PER_S / st, "launch_ms": st * 1e3,
                       "source": "SQ_INSTS_VALU, profiles/water96k_pme_kernel_stats.md"}
    if not args.no_cpu:
        cb, Ec, Fc = pme_cpu_baseline(sysm)
        out["cpu_baseline"] = cb
        out["fp64_vs_cpu_rel_energy"] = abs(rp["fp64"]["E"] - Ec) / abs(Ec)
        out["fp64_vs_cpu_rms_rel_force"] = rms_rel(rp["fp64"]["F"], Fc)
    for t in ("fp64", "mixed"):
        rp[t]["m"].close()
    return out


def main_water96k(args):
    from msmpb_amd import inputs
    sysm = inputs.cfg3_water96k(args.nu)
    r = run_precisions(args, sysm)
    F64, Fmx = r["fp64"]["F"], r["mixed"]["F"]
    m64, mmx = r["fp64"]["m"], r["mixed"]["m"]
    pairs = m64.pair_count
    ms_mx, ms_64 = r["mixed"]["ms"], r["fp64"]["ms"]
    ad_mx = autograd_ms(mmx, r["pos"], r["q"], args.steps, args.warmup)
    ad_64 = autograd_ms(m64, r["pos"], r["q"], args.steps, args.warmup)
    stages = r["mixed"]["stages"]
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
        "rms_rel_force_err": rms_rel(Fmx, F64),
        "rel_energy_err": abs(r["mixed"]["E"] - r["fp64"]["E"]) / abs(r["fp64"]["E"]),
        "mixed_forward_adjoint_ms": ad_mx, "fp64_forward_adjoint_ms": ad_64,
        "speedup_adjoint_mixed_vs_fp64": ad_64 / ad_mx,
        "stage_ms_mixed": {k: v * 1e3 for k, v in stages.items()},
        "stage_ms_fp64": {k: v * 1e3 for k, v in r["fp64"]["stages"].items()},
        "hbm_footprint_bytes": {"fp64": r["fp64"]["footprint"], "mixed": r["mixed"]["footprint"]},
        "pairs_within_cutoff": pairs,
        "roofline": roofline(stages, mmx, pairs, mmx.nu),
    }
    line["rebuild_ms"] = {"mixed": rebuild_ms(sysm, "mixed"), "fp64": rebuild_ms(sysm, "fp64")}
    # the same workload at the other B-spline order (nu = 4)
    if args.nu == 6:
        s4 = inputs.cfg3_water96k(4)
        r4 = run_precisions(args, s4)
        line["nu4"] = {"mixed_evals_per_s": 1e3 / r4["mixed"]["ms"],
                       "fp64_evals_per_s": 1e3 / r4["fp64"]["ms"],
                       "rms_rel_force_err": rms_rel(r4["mixed"]["F"], r4["fp64"]["F"]),
                       "rel_energy_err": abs(r4["mixed"]["E"] - r4["fp64"]["E"])
                       / abs(r4["fp64"]["E"])}
        for t in ("fp64", "mixed"):
            r4[t]["m"].close()
    # the reference's PME comparator (pme0.c, pme1.c) on the same workload
    if args.nu == 6 and not args.no_pme:
        line["pme"] = pme_comparator(args, sysm, F64)
    if not args.no_cpu:
        cb, Ec, Fc = cpu_baseline(sysm)
        line["cpu_baseline"] = cb
        line["fp64_vs_cpu_rel_energy"] = abs(r["fp64"]["E"] - Ec) / abs(Ec)
        line["fp64_vs_cpu_rms_rel_force"] = rms_rel(F64, Fc)
        line["rms_rel_force_err_vs_oracle"] = rms_rel(Fmx, Fc)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


def main_water1m(args):
    """1M-atom box: evals/s of both precisions, the adjoint with recomputed
    grid levels, the device-memory footprint of each FF, and the mixed
    path's force error against FP64."""
    from msmpb_amd import inputs
    sysm = inputs.cfg4_water1m()
    steps, warmup = max(2, args.steps // 10), 1
    r = run_precisions(args, sysm, steps=steps, warmup=warmup)
    m64, mmx = r["fp64"]["m"], r["mixed"]["m"]
    ad = autograd_ms(mmx, r["pos"], r["q"], steps, warmup, recompute=True)
    line = {
        "metric": "MSM energy+force evals/s at 1M atoms (water1m)",
        "value": 1e3 / r["mixed"]["ms"], "unit": "evals/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": r["mixed"]["ms"], "higher_is_better": True,
        "data": "synthetic",
        "config": {"workload": f"water1m: 333,333 rigid 3-site waters ({sysm.N} atoms), cubic "
                               f"box 215 A, nu=6, cutoff 10 A, {mmx.levels} levels, top grid "
                               f"{mmx.top_grid[0]} (finest {mmx.grid_dims(1)[0]})"},
        "fp64_evals_per_s": 1e3 / r["fp64"]["ms"],
        "rms_rel_force_err": rms_rel(r["mixed"]["F"], r["fp64"]["F"]),
        "mixed_forward_adjoint_recompute_ms": ad,
        "hbm_footprint_bytes": {"fp64": r["fp64"]["footprint"], "mixed": r["mixed"]["footprint"]},
        "stage_ms_mixed": {k: v * 1e3 for k, v in r["mixed"]["stages"].items()},
        "pairs_within_cutoff": m64.pair_count,
        "rebuild_ms": {"mixed": rebuild_ms(sysm, "mixed"), "fp64": rebuild_ms(sysm, "fp64")},
    }
    print(json.dumps(line))


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "water1m":
        main_water1m(args)
    elif args.workload == "pme96k":
        from msmpb_amd import inputs
        r = pme_comparator(args, inputs.cfg3_water96k(6))
        print(json.dumps({"metric": "PME comparator evals/s at 96k atoms (not the headline)",
                          "value": r["mixed_evals_per_s"], "unit": "evals/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                          "data": "synthetic", "config": {"workload": "water96k " + r["config"]},
                          **r}))
    else:
        main_water96k(args)


if __name__ == "__main__":
    main()
