#!/usr/bin/env python
"""Benchmark of the BSSN right-hand side on one MI355X.

Workload (default): 4096 uniform 13^3 padded blocks (7^3 interior each,
1,404,928 interior points), seeded perturbed-Minkowski data, one full RK4 step
per timed step (4 right-hand side evaluations: derivatives, equations, KO
dissipation, stage update, constraint enforcement per stage), FP64, inputs
resident in HBM. Reports interior Mpoints/s = interior points x 4 / step time.

Prints one JSON line. Also reports:
  roofline      the dominant kernel (the fused kernel on the regular tiles of
                blocks without flagged faces) against HBM: algorithmic bytes
                per launch (192 n^3/(n-6)^3 + 192 + 768 bytes per interior point
                of its tiles: padded state read once, RHS produced once, RK
                stage update) over its average launch time, measured with HIP
                events on the launch stream; `traffic` is the HBM bytes per
                launch measured by rocprofv3 counters (profiles/), when a
                summary for this workload is committed
  cpu_baseline  the oracle's OpenMP C restatement on the host cores, one RK4
                step of a 1024-block sample of the same batch
  rhs_only      right-hand side alone (no RK update), Mpoints/s
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dendro-gr_amd"))

METRIC = "BSSN RHS Mpoints/s on 1 MI355X: plain FP64 vs batched mixed at ≤1e-8 rel err"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_bytes_per_point(n):
    """Algorithmic HBM bytes per interior point per RK stage: padded FP64 state
    read once, RHS written once, plus the stage update (read u0, read and write
    the accumulator, write the next stage)."""
    return 192.0 * n ** 3 / (n - 6) ** 3 + 192.0 + 768.0


def measured_traffic(kind, nblocks):
    """HBM bytes per launch of `kind` from the committed rocprofv3 counter
    summary of this workload (profiles/*traffic*.txt lines
    'traffic <kind> uniform13x<nblocks> <bytes>'), else None."""
    import glob
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.txt"))):
        with open(f) as fh:
            for line in fh:
                p = line.split()
                if len(p) == 4 and p[0] == "traffic" and p[1] == kind and p[2] == f"uniform13x{nblocks}":
                    return float(p[3])
    return None


def cpu_baseline(nblocks, threads, reps=3):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from dgr import workloads as W

    t = W.uniform_13(nblocks)
    st = np.zeros((24, W.total_padded(t)))
    O.init_perturbed_minkowski(st, t, seed=1, amp=1e-3)
    p = O.default_params(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)
    pts = W.interior_points(t)
    times = []
    for r in range(reps + 1):
        s = st.copy()
        t0 = time.perf_counter()
        O.rk4_step(s, t, 0.025, p, enforce_per_stage=True, nthreads=threads)
        dt = time.perf_counter() - t0
        if r > 0:
            times.append(dt)
    med = statistics.median(times)
    return {"value": round(pts * 4 / med / 1e6, 3), "unit": "Mpoints/s", "cores": threads,
            "kind": "port",
            "sample": f"one RK4 step (4 RHS evaluations) of the first {nblocks} blocks of the "
                      f"same uniform 13^3 batch ({pts} interior points), OpenMP over blocks, "
                      f"median of {reps} after 1 warm-up"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nblocks", type=int, default=4096)
    ap.add_argument("--cpu-blocks", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (3 times), to scale "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = (1 << 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path shards "
              f"by blocks with no collective, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch
    import dgr
    from dgr import synth, workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    blocks = W.uniform_13(args.nblocks)
    npts = W.interior_points(blocks)
    n_pad = W.total_padded(blocks)
    log(f"workload: {args.nblocks} x 13^3 blocks, {npts} interior points, state {24 * n_pad * 8 / 1e9:.2f} GB")
    state = synth.perturbed_minkowski(blocks, seed=1, amp=1e-3, device=dev)
    plan = dgr.Plan(blocks, n_pad)
    params = dgr.default_params(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)
    dt = 0.25 * 0.1
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        plan.rk4_step(state, dt, params, enforce_per_stage=True, stream=stream)
    torch.cuda.synchronize()
    if plan.status(stream) != 0:
        raise RuntimeError("constraint violation during warm-up")

    plan.set_timing(True)
    plan.kernel_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.rk4_step(state, dt, params, enforce_per_stage=True, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = plan.kernel_times()
    plan.set_timing(False)
    if plan.status(stream) != 0:
        raise RuntimeError("constraint violation during the timed steps")
    elapsed = t1 - t0
    ms_step = elapsed / args.steps * 1e3
    value = npts * 4 * args.steps / elapsed / 1e6

    # right-hand side alone
    out = torch.empty_like(state)
    for _ in range(2):
        plan.rhs(out, state, params, stream=stream)
    torch.cuda.synchronize()
    r0 = time.perf_counter()
    nrhs = 4 * args.steps
    for _ in range(nrhs):
        plan.rhs(out, state, params, stream=stream)
    torch.cuda.synchronize()
    rhs_only = npts * nrhs / (time.perf_counter() - r0) / 1e6

    stages = 4 * args.steps
    stage_ms = sum(v[0] for v in kt.values()) / stages
    bpp = stage_bytes_per_point(13)
    kind_pts = W.fused_kind_points(blocks)
    per_kernel = {}
    for k, (ms, nl) in kt.items():
        if not nl:
            continue
        entry = {"ms_per_launch": round(ms / nl, 5), "launches_per_stage": nl / stages}
        if k in kind_pts and kind_pts[k]:
            entry["points"] = kind_pts[k]
            entry["achieved_GBs"] = round(bpp * kind_pts[k] / (ms / nl * 1e-3) / 1e9, 1)
        per_kernel[k] = entry
    dom = max(per_kernel, key=lambda k: kt[k][0])
    dom_ms = kt[dom][0] / kt[dom][1]
    dom_bytes = bpp * kind_pts.get(dom, npts)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom, args.nblocks)
    names = {"fused_wide": "k_fused<interior, 256 threads, 7x7x5 tiles>",
             "fused_thin": "k_fused<interior, 128 threads, 7x7x2 tiles>",
             "fused_face": "k_fused<flagged faces, 256 threads>",
             "derivs": "k_derivs", "rhs": "k_rhs", "face": "k_face"}
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": names.get(dom, dom), "launch_ms": round(dom_ms, 5),
                "algorithmic_bytes_per_launch": dom_bytes,
                "algorithmic_bytes_per_point": round(bpp, 1),
                "stage_ms": round(stage_ms, 4),
                "stage_GBs": round(bpp * npts / (stage_ms * 1e-3) / 1e9, 1),
                "per_kernel": per_kernel}
    del out
    plan.close()

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {args.cpu_blocks} blocks on {threads} threads")
        cpu = cpu_baseline(args.cpu_blocks, threads)
        try:
            cpu["cpu_model"] = subprocess.run(["bash", "-c", "lscpu | grep 'Model name' | head -1"],
                                              capture_output=True, text=True).stdout.split(":")[-1].strip()
        except Exception:
            pass

    line = {"metric": METRIC, "value": round(value, 2), "unit": "Mpoints/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"uniform_13^3 x {args.nblocks} blocks, one RK4 step per step",
                       "blocks": args.nblocks, "padded_edge": 13, "interior_points": npts,
                       "precision": "fp64", "rhs_evals_per_step": 4, "ko_sigma": 0.4, "dt": dt},
            "rhs_only_mpoints_s": round(rhs_only, 2),
            "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
