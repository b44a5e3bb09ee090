#!/usr/bin/env python
"""Benchmark of the BSSN right-hand side on one MI355X.

Workloads (interior points are the non-padding points of all blocks):
  uniform13 (default)  4096 uniform 13^3 padded blocks (7^3 interior each,
                       1,404,928 points), seeded perturbed-Minkowski data,
                       KO sigma 0.4, eta 2, dt 0.025
  puncture             the uneven binary-puncture octree: 13^3/19^3/31^3/55^3
                       blocks with counts 60000/8000/1000/150 x --scale
                       (0.42 -> ~30M points), Brill-Lindquist data plus 1e-6
                       noise, KO sigma 0.4, RIT eta, dt 0.25 dx_min
One timed step = one full RK4 step (4 right-hand side evaluations: derivatives,
equations, KO dissipation, stage update, constraint enforcement per stage),
FP64, inputs resident in HBM, replayed from a captured HIP graph.
value = interior points x 4 / step time.

Prints one JSON line. Also reports:
  roofline      the dominant kernel (the fused kernel on the regular tiles of
                blocks without flagged faces) against HBM: algorithmic bytes
                per launch (per block 192 n^3 for reading the padded state once,
                per interior point 192 for the RHS and 768 for the RK stage
                update, split over the launches by the points each covers) over
                its average launch time, measured with HIP events on the launch
                stream; `traffic` is the HBM bytes per launch measured with
                rocprofv3 counters (profiles/), when a summary for this
                workload is committed
  cpu_baseline  the oracle's OpenMP C restatement on the host cores, one RK4
                step of a bounded sample of the same workload
  rhs_only      right-hand side alone (no RK update), Mpoints/s
  mixed_precision  (FP64 runs) the batched mixed-precision path on the same
                workload: its RK4 Mpoints/s, speedup, and errors against FP64
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dendro-gr_amd"))

METRIC = "BSSN RHS Mpoints/s on 1 MI355X: plain FP64 vs batched mixed at ≤1e-8 rel err"
HBM_PEAK_GBS = 8000.0
# MI355X vector-ALU peaks (TFLOP/s) and the algorithmic FLOPs of one RK stage
# per interior point: 6788 for the right-hand side (derivatives 864 + 462 + 396,
# KO 1008 + 96, equations 3952, eta 10) + 96 for the stage update
VALU_PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}
FLOP_PER_POINT_STAGE = 6788 + 96
KERNEL_NAMES = {"fused_wide": "k_fused<interior, 256 threads, 7x7x5 tiles>",
                "fused_thin": "k_fused<interior, 128 threads, 7x7x2 tiles>",
                "fused_face": "k_fused<flagged faces, 256 threads>",
                "fused_full": "k_fused<interior, 384 threads, 7x7x7 tiles, FP32>",
                "derivs": "k_derivs", "rhs": "k_rhs", "face": "k_face"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Workload:
    """Block table, synthetic data and parameters of one benchmark workload."""

    def __init__(self, name, nblocks, scale):
        from dgr import workloads as W
        self.name = name
        if name == "uniform13":
            self.blocks = W.uniform_13(nblocks)
            self.key = f"uniform13x{nblocks}"
            self.desc = f"uniform_13^3 x {nblocks} blocks, one RK4 step per step"
            self.dt = 0.25 * 0.1
            self.params = dict(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)
        elif name == "lattice13":
            # 16^3 lattice of 13^3 blocks sharing boundary nodes: RK4 with the
            # padding refreshed from the neighbours before every stage
            self.blocks = W.lattice(shape=(16, 16, 16), n=13, h=0.05, origin=(-4.8, -4.8, -4.8), order="morton")
            self.key = "lattice13x4096"
            self.desc = ("lattice of 16^3 blocks of 13^3 sharing boundary nodes, stored in Morton order, "
                         "padding refreshed from same-level neighbours before every RK4 stage, "
                         "one RK4 step per step")
            self.dt = 0.25 * 0.05
            self.params = dict(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)
        elif name == "puncture":
            self.blocks = W.puncture_octree(scale)
            self.key = f"puncture{scale:g}"
            sizes = {int(n): int((self.blocks["sz"][:, 0] == n).sum()) for n in (13, 19, 31, 55)}
            self.desc = (f"binary-puncture octree (Brill-Lindquist), blocks 13^3/19^3/31^3/55^3 = "
                         f"{sizes[13]}/{sizes[19]}/{sizes[31]}/{sizes[55]}, one RK4 step per step")
            h = (self.blocks["pmax"][:, 0] - self.blocks["pmin"][:, 0]) / (self.blocks["sz"][:, 0] - 1)
            self.dt = 0.25 * float(h.min())
            self.params = dict(ko_sigma=0.4)
        else:
            raise ValueError(name)
        self.npts = W.interior_points(self.blocks)
        self.n_pad = W.total_padded(self.blocks)

    def device_state(self, dev):
        from dgr import synth
        if self.name in ("uniform13", "lattice13"):
            return synth.perturbed_minkowski(self.blocks, seed=1, amp=1e-3, device=dev)
        return synth.brill_lindquist(self.blocks, seed=1, noise=1e-6, device=dev)

    def cpu_sample(self, nblocks_uniform):
        """A bounded sample of the same workload for the CPU baseline."""
        from dgr import workloads as W
        if self.name == "uniform13":
            return (W.uniform_13(nblocks_uniform),
                    f"the first {nblocks_uniform} blocks of the same uniform 13^3 batch")
        # every k-th block of the octree, keeping the size mix (and enough
        # large blocks that OpenMP over blocks is not bound by one of them)
        k = max(1, len(self.blocks) // 1400)
        sub = self.blocks[::k].copy()
        off = 0
        for b in range(len(sub)):
            sub["offset"][b] = off
            off += int(np.prod(sub["sz"][b].astype(np.int64)))
        return sub, f"every {k}th block of the same octree ({len(sub)} blocks)"


def measured_traffic(kind, key):
    """HBM bytes per launch of `kind` from the committed rocprofv3 counter
    summaries (profiles/*traffic*.txt lines 'traffic <kind> <workload> <bytes>'),
    else None."""
    import glob
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.txt"))):
        with open(f) as fh:
            for line in fh:
                p = line.split()
                if len(p) == 4 and p[0] == "traffic" and p[1] == kind and p[2] == key:
                    return float(p[3])
    return None


def cpu_baseline(wl, nblocks_uniform, threads, reps=3):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from dgr import workloads as W

    t, what = wl.cpu_sample(nblocks_uniform)
    st = np.zeros((24, W.total_padded(t)))
    if wl.name == "uniform13":
        O.init_perturbed_minkowski(st, t, seed=1, amp=1e-3)
    else:
        O.init_brill_lindquist(st, t, seed=1, noise=1e-6)
    p = O.default_params(**wl.params)
    pts = W.interior_points(t)
    times = []
    for r in range(reps + 1):
        s = st.copy()
        t0 = time.perf_counter()
        O.rk4_step(s, t, wl.dt, p, enforce_per_stage=True, nthreads=threads)
        dt = time.perf_counter() - t0
        if r > 0:
            times.append(dt)
    med = statistics.median(times)
    return {"value": round(pts * 4 / med / 1e6, 3), "unit": "Mpoints/s", "cores": threads,
            "kind": "port",
            "sample": f"one RK4 step (4 RHS evaluations) of {what} ({pts} interior points), "
                      f"OpenMP over blocks, median of {reps} after 1 warm-up"}


def mixed_comparison(wl, state0, params, steps, fp64_value, stream, autotune=True):
    """The batched mixed-precision path on the same workload: RK4 Mpoints/s
    (graph replay), and its error against the FP64 path from identical inputs:
    the right-hand side (max over variables of the normwise relative error
    over interior points, max-norm and L2) and the state after one RK4 step
    (max |difference| over the state's scale, max |u|, and per variable, max-
    norm and L2); on the puncture workload also the Hamiltonian-constraint
    norm after 10 RK4 steps (masked r < 5 around each hole) in both precisions."""
    import torch
    import dgr
    from dgr import workloads as W
    mask = torch.from_numpy(interior_mask(wl.blocks, wl.n_pad)).to(state0.device)
    res, hams = {}, {}
    for prec in ("fp64", "mixed"):
        plan = dgr.Plan(wl.blocks, wl.n_pad)
        plan.set_precision(prec)
        out = torch.empty_like(state0)
        if prec == "mixed" and autotune:
            plan.autotune(out, state0, params, stream=stream)
            tiling = {"x".join(str(n) for n in k): v for k, v in plan.tiling(prec).items()}
        plan.rhs(out, state0, params, stream=stream)
        u = state0.clone()
        plan.rk4_step(u, wl.dt, params, stream=stream)
        torch.cuda.synchronize()
        res[prec] = (out[:, mask], u[:, mask])
        if wl.name == "puncture":
            u.copy_(state0)
            plan.rk4_steps(u, wl.dt, params, nsteps=10, stream=stream)
            ham = out[0]  # one padded array as the constraint's output
            plan.ham(ham, u, stream=stream)
            hams[prec] = plan.constraint_norm(ham, stream=stream)
            u.copy_(state0)
        if prec == "mixed":
            plan.rk4_steps(u, wl.dt, params, nsteps=2, stream=stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.rk4_steps(u, wl.dt, params, nsteps=steps, stream=stream)
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - t0
            if plan.status(stream) != 0:
                raise RuntimeError("constraint violation in the mixed-precision steps")
        plan.close()
        del out, u
    value = wl.npts * 4 * steps / elapsed / 1e6
    (r64, u64), (r32, u32) = res["fp64"], res["mixed"]
    e_rhs = ((r32 - r64).abs().amax(dim=1) / r64.abs().amax(dim=1).clamp_min(1e-300)).max().item()
    e_var = ((u32 - u64).abs().amax(dim=1) / u64.abs().amax(dim=1).clamp_min(1e-300)).max().item()
    e_scale = ((u32 - u64).abs().max() / u64.abs().max()).item()
    e_rhs_l2 = ((r32 - r64).norm(dim=1) / r64.norm(dim=1).clamp_min(1e-300)).max().item()
    e_var_l2 = ((u32 - u64).norm(dim=1) / u64.norm(dim=1).clamp_min(1e-300)).max().item()
    err = {"state_after_one_step_vs_state_scale": e_scale,
           "state_after_one_step_max_per_variable": e_var,
           "state_after_one_step_l2_per_variable": e_var_l2,
           "rhs_max_normwise_per_variable": e_rhs,
           "rhs_l2_per_variable": e_rhs_l2}
    if hams:
        err["ham_constraint_norm_after_10_steps"] = {
            "fp64": hams["fp64"], "mixed": hams["mixed"],
            "relative_difference": abs(hams["mixed"] - hams["fp64"]) / abs(hams["fp64"])}
    return {"value": round(value, 2), "unit": "Mpoints/s", "speedup_vs_fp64": round(value / fp64_value, 3),
            "tiling": tiling if autotune else "default",
            "dtype": "f32 arithmetic (alpha, chi staged and differentiated in f64), f64 state and RK",
            "error_vs_fp64": err}


def interior_mask(blocks, npad):
    m = np.zeros(npad, dtype=bool)
    for b in blocks:
        nx, ny, nz = (int(v) for v in b["sz"])
        blk = np.zeros((nz, ny, nx), dtype=bool)
        blk[3:nz - 3, 3:ny - 3, 3:nx - 3] = True
        off = int(b["offset"])
        m[off:off + nx * ny * nz] = blk.ravel()
    return m


def roofline(wl, kt, stages, precision="fp64"):
    from dgr import workloads as W
    kind_pts = W.fused_kind_points(wl.blocks, precision)
    kind_bytes = W.fused_kind_bytes(wl.blocks, precision)
    stage_ms = sum(v[0] for v in kt.values()) / stages
    per_kernel = {}
    for k, (ms, nl) in kt.items():
        if not nl:
            continue
        entry = {"ms_per_launch": round(ms / nl, 5), "launches_per_stage": nl / stages}
        if kind_pts.get(k):
            entry["points"] = kind_pts[k]
            entry["achieved_GBs"] = round(kind_bytes[k] / (ms / nl * 1e-3) / 1e9, 1)
        per_kernel[k] = entry
    dom = max(per_kernel, key=lambda k: kt[k][0])
    dom_ms = kt[dom][0] / kt[dom][1]
    dom_bytes = kind_bytes.get(dom) or W.stage_bytes(wl.blocks)
    precision_tag = "" if precision == "fp64" else "_" + precision
    # algorithmic FLOP rate of the dominant kernel (halo recomputation not counted)
    tflops = kind_pts.get(dom, wl.npts) * FLOP_PER_POINT_STAGE / (dom_ms * 1e-3) / 1e12
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, wl.key + precision_tag),
            "kernel": KERNEL_NAMES.get(dom, dom), "launch_ms": round(dom_ms, 5),
            "algorithmic_bytes_per_launch": dom_bytes,
            "algorithmic_bytes_per_point": round(W.stage_bytes(wl.blocks) / wl.npts, 1),
            "vector_alu": {"flop_per_point": FLOP_PER_POINT_STAGE, "achieved_tflops": round(tflops, 2),
                           "peak_tflops": VALU_PEAK_TFLOPS[precision],
                           "frac": round(tflops / VALU_PEAK_TFLOPS[precision], 4)},
            "stage_ms": round(stage_ms, 4),
            "stage_GBs": round(W.stage_bytes(wl.blocks) / (stage_ms * 1e-3) / 1e9, 1),
            "per_kernel": per_kernel}


def halo_rate(wl, plan, state, stream, calls=20):
    """dgr_halo_fill alone: average launch time from HIP events on the launch
    stream, and its algorithmic HBM rate (every linked padding point read
    from the neighbour's interior and written, 24 variables x 16 B)."""
    import torch
    nbr = plan.link_neighbors()
    n = 13
    ext = {-1: 3, 0: n - 6, 1: 3}
    pts = 0
    for d in range(27):
        di, dj, dk = d % 3 - 1, (d // 3) % 3 - 1, d // 9 - 1
        if di or dj or dk:
            pts += int((nbr[:, d] >= 0).sum()) * ext[di] * ext[dj] * ext[dk]
    byts = pts * 24 * 16
    plan.halo_fill(state, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(calls):
        plan.halo_fill(state, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / calls
    return {"links": int((nbr >= 0).sum()), "padding_points_filled": pts, "bytes_per_launch": byts,
            "ms_per_launch": round(ms, 4), "achieved_GBs": round(byts / ms / 1e6, 1),
            "use": ("FP64 RK4 stages read linked padding inside k_fused (HALO variant); this fill "
                    "runs once per RK4 call, leaving the caller's padding refreshed")}


def constraint_rates(wl, plan, state, stream, calls=5):
    """The constraint diagnostic (dgr_constraints) on the state after the timed
    steps: interior points per second with all six outputs (ham, mom0..2,
    psi4 real/imaginary) and with ham alone, plus the algorithmic HBM rate of
    the all-outputs call (padded state read once, six outputs written once).
    Skipped when HBM cannot hold the six output arrays."""
    import torch
    free, _ = torch.cuda.mem_get_info()
    need = 6 * wl.n_pad * 8
    if free < need * 1.1 + 2e9:
        log("constraint rates skipped: no room for the six output arrays")
        return None
    out = torch.empty((6, wl.n_pad), dtype=torch.float64, device=state.device)
    res = {}
    for name, args in (("all_outputs", (out[0], out[1:4], out[4:6])), ("ham_only", (out[0], None, None))):
        plan.constraints(args[0], args[1], state, psi4=args[2], stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            plan.constraints(args[0], args[1], state, psi4=args[2], stream=stream)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / calls
        res[name + "_mpoints_s"] = round(wl.npts / dt / 1e6, 2)
        if name == "all_outputs":
            res["all_outputs_ms_per_call"] = round(dt * 1e3, 4)
            byts = 24 * 8 * wl.n_pad + 6 * 8 * wl.npts
            res["all_outputs_algorithmic_gb_s"] = round(byts / dt / 1e9, 1)
    res["note"] = ("derivatives staged through the split path's 162-slot workspace, then the "
                   "generated constraint expressions per point; FP64")
    del out
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["uniform13", "puncture", "lattice13"], default="uniform13")
    ap.add_argument("--nblocks", type=int, default=4096, help="uniform13: number of blocks")
    ap.add_argument("--scale", type=float, default=0.42, help="puncture: block-count scale")
    ap.add_argument("--cpu-blocks", type=int, default=1024, help="uniform13: CPU baseline sample")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 derivatives/equations/KO on offset-subtracted fields, FP64 state")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision comparison of the FP64 run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tiling instead of timing the tilings per block size")
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
              f"by blocks, each GPU running its own plan, with no collective on the hot path", file=sys.stderr)
        sys.exit(2)

    import torch
    import dgr

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    wl = Workload(args.workload, args.nblocks, args.scale)
    log(f"workload: {wl.desc}; {wl.npts} interior points, state {24 * wl.n_pad * 8 / 1e9:.2f} GB")
    state = wl.device_state(dev)
    # the mixed comparison and the CPU baseline are for the frozen-padding step
    if wl.name == "lattice13":
        args.no_mixed = args.no_cpu_baseline = True
    state0 = state.clone() if (args.precision == "fp64" and not args.no_mixed) else None
    plan = dgr.Plan(wl.blocks, wl.n_pad)
    plan.set_precision(args.precision)
    if wl.name == "lattice13":
        plan.set_halo_refresh(True)
    params = dgr.default_params(**wl.params)
    stream = torch.cuda.current_stream()
    if not args.no_autotune:
        # tile shapes per block size, timed on this workload's own data
        scratch = torch.empty_like(state)
        plan.autotune(scratch, state, params, stream=stream)
        torch.cuda.synchronize()
        del scratch
    # the plan allocates its RK4 buffers with hipMalloc: hand torch's cached
    # blocks back first (the 120 M-point octree needs ~200 GB of them)
    torch.cuda.empty_cache()
    tiling = {"x".join(str(n) for n in k): v for k, v in plan.tiling(args.precision).items()}
    log(f"tiling per block size ({'autotuned' if not args.no_autotune else 'default'}): {tiling}")

    # the RK4 step is captured once as a HIP graph and replayed (dgr_rk4_steps)
    plan.rk4_steps(state, wl.dt, params, nsteps=max(args.warmup, 1), stream=stream)
    torch.cuda.synchronize()
    if plan.status(stream) != 0:
        raise RuntimeError("constraint violation during warm-up")
    t0 = time.perf_counter()
    plan.rk4_steps(state, wl.dt, params, nsteps=args.steps, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if plan.status(stream) != 0:
        raise RuntimeError("constraint violation during the timed steps")
    elapsed = t1 - t0
    ms_step = elapsed / args.steps * 1e3
    value = wl.npts * 4 * args.steps / elapsed / 1e6

    # per-kernel times: the same steps launched eagerly with HIP events around each launch
    ksteps = min(args.steps, 5)
    plan.set_timing(True)
    plan.kernel_times()
    for _ in range(ksteps):
        plan.rk4_step(state, wl.dt, params, enforce_per_stage=True, stream=stream)
    kt = plan.kernel_times()
    plan.set_timing(False)

    # right-hand side alone (needs room for a padded output next to the state
    # and the plan's RK4 buffers; skipped when HBM cannot hold it)
    rhs_only = None
    free, _ = torch.cuda.mem_get_info()
    if free > state.numel() * state.element_size() * 1.05:
        out = torch.empty_like(state)
        for _ in range(2):
            plan.rhs(out, state, params, stream=stream)
        torch.cuda.synchronize()
        r0 = time.perf_counter()
        nrhs = 4 * args.steps
        for _ in range(nrhs):
            plan.rhs(out, state, params, stream=stream)
        torch.cuda.synchronize()
        rhs_only = round(wl.npts * nrhs / (time.perf_counter() - r0) / 1e6, 2)
        del out
    else:
        log("right-hand-side-only rate skipped: no room for a second padded state")
    roof = roofline(wl, kt, 4 * ksteps, args.precision)
    constraints = constraint_rates(wl, plan, state, stream)
    halo = halo_rate(wl, plan, state, stream) if wl.name == "lattice13" else None
    plan.close()

    mixed = None
    if state0 is not None:
        mixed = mixed_comparison(wl, state0, params, args.steps, value, stream, not args.no_autotune)
    del state

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(wl, args.cpu_blocks, threads)
        try:
            cpu["cpu_model"] = subprocess.run(["bash", "-c", "lscpu | grep 'Model name' | head -1"],
                                              capture_output=True, text=True).stdout.split(":")[-1].strip()
        except Exception:
            pass

    line = {"metric": METRIC, "value": round(value, 2), "unit": "Mpoints/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if args.precision == "fp64" else "f32 arithmetic, f64 state",
            "data": "synthetic",
            "config": {"workload": wl.desc, "blocks": len(wl.blocks), "interior_points": wl.npts,
                       "precision": args.precision, "rhs_evals_per_step": 4, "rk4_step": "hipGraph replay",
                       "tiling": tiling, "autotuned": not args.no_autotune,
                       "dt": wl.dt, **wl.params},
            "rhs_only_mpoints_s": rhs_only, "constraints": constraints, "halo_fill": halo,
            "roofline": roof, "cpu_baseline": cpu, "mixed_precision": mixed}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
