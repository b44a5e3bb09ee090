#!/usr/bin/env python
"""Benchmark of the BSSN right-hand side on one MI355X.

Default line (--workload octree): the uneven binary-puncture octree,
13^3/19^3/31^3/55^3 blocks with counts 60000/8000/1000/150 x 0.42 (29043
blocks, 30.0 M interior points), Brill-Lindquist data plus 1e-6 noise, KO
sigma 0.4, RIT eta, dt 0.25 dx_min. One timed step = one full RK4 step (4
right-hand side evaluations: derivatives, equations, KO dissipation, stage
update, constraint enforcement per stage), FP64, inputs resident in HBM,
replayed from a captured HIP graph; value = interior points x 4 / step time.
Beside it, in the same line:
  mixed_precision  the batched mixed-precision path on the same octree: its
                rate, speedup, per-variable max and L2 errors against FP64
                from identical inputs, and the Hamiltonian-constraint norm
                after 10 RK4 steps in both precisions
  uniform13     4096 uniform 13^3 blocks (1.40 M points, perturbed Minkowski,
                eta 2, dt 0.025): rate, roofline, mixed-precision errors
  large_octree  the same octree with 4x the block counts (120 M points),
                --large-steps RK4 steps in FP64 and in mixed precision
  roofline      the dominant kernel (the fused kernel on the regular tiles of
                blocks without flagged faces) against HBM: algorithmic bytes
                per launch (per block 192 n^3 for reading the padded state once,
                per interior point 192 for the RHS and 768 for the RK stage
                update, split over the launches by the points each covers) over
                its average launch time, measured with HIP events on the launch
                stream; `traffic` is the HBM bytes per launch measured with
                rocprofv3 counters (profiles/), when a summary measured on the
                current kernel sources is committed
  cpu_baseline  the CPU path (the reference's generated equations inside the
                oracle's block loop, OpenMP over blocks on `nproc` threads), one
                RK4 step of a bounded sample of the workload, with the phase
                split; and one 19^3 block on one thread
  rhs_only      right-hand side alone (no RK update), Mpoints/s
--workload uniform13 / puncture / lattice13 runs one workload alone.
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
        elif name == "refined":
            # a 2:1-balanced octree of 13^3 blocks refined toward the punctures:
            # RK4 with the padding of every stage refreshed across level jumps
            self.blocks = W.refined_octree(4, 9, 4.0)
            self.key = "refined4-9"
            h = (self.blocks["pmax"][:, 0] - self.blocks["pmin"][:, 0]) / 12
            self.desc = (f"2:1-balanced octree of {len(self.blocks)} blocks of 13^3 on six levels refined toward "
                         f"two punctures (Brill-Lindquist), domain-boundary faces flagged (radiative boundary), "
                         f"padding of every RK4 stage refreshed across level jumps (copy, injection, order-6 "
                         f"interpolation), one RK4 step per step")
            self.dt = 0.25 * float(h.min())
            self.params = dict(ko_sigma=0.4)
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

    def cpu_sample(self):
        """A bounded sample of the same workload for the CPU baseline (1.4 M
        interior points of uniform blocks, 2.5 M of the octree: one or two
        seconds per RK4 step on 16 host threads)."""
        from dgr import workloads as W
        if self.name == "uniform13":
            n = min(len(self.blocks), 4096)
            return W.uniform_13(n), f"the first {n} blocks of the same uniform 13^3 batch"
        # every k-th block of the octree, keeping the size mix (and enough
        # large blocks that OpenMP over blocks is not bound by one of them)
        k = max(1, len(self.blocks) // 2000)
        sub = self.blocks[::k].copy()
        off = 0
        for b in range(len(sub)):
            sub["offset"][b] = off
            off += int(np.prod(sub["sz"][b].astype(np.int64)))
        return sub, f"every {k}th block of the same octree ({len(sub)} blocks)"


# the sources the kernels are built from: a traffic summary in profiles/ is
# used only when it was measured on a build of these exact files
KERNEL_SOURCES = ("dendro-gr_amd/csrc/dgr_bssn.hip", "dendro-gr_amd/csrc/stencils.h",
                  "dendro-gr_amd/csrc/bssn_rhs_phased.h", "dendro-gr_amd/csrc/bssn_rhs_generated.h",
                  "dendro-gr_amd/csrc/bssn_ham_generated.h", "dendro-gr_amd/csrc/bssn_cons_phased.h",
                  "dendro-gr_amd/csrc/bssn_rhs_phased_gauge.h", "dendro-gr_amd/csrc/bssn_rhs_gauge_generated.h",
                  "dendro-gr_amd/build.py")


def source_sha():
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(kind, key):
    """HBM bytes per launch of `kind` from the committed rocprofv3 counter
    summaries (profiles/*traffic*.txt: a line 'source_sha <sha>' and lines
    'traffic <kind> <workload> <bytes>'), only from a summary whose sha is
    that of the current kernel sources. Returns (bytes or None, note)."""
    import glob
    sha = source_sha()
    stale = None
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.txt"))):
        fsha, val = None, None
        with open(f) as fh:
            for line in fh:
                p = line.split()
                if len(p) == 2 and p[0] == "source_sha":
                    fsha = p[1]
                if len(p) == 4 and p[0] == "traffic" and p[1] == kind and p[2] == key:
                    val = float(p[3])
        if val is None:
            continue
        if fsha == sha:
            return val, f"profiles/{os.path.basename(f)} (source_sha {sha})"
        stale = f"profiles/{os.path.basename(f)} was measured on other kernel sources ({fsha} != {sha})"
    return None, stale or f"no counter summary for {kind} on {key}"


def nproc():
    try:
        return int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except Exception:
        return os.cpu_count() or 1


def host_info():
    info = {"nproc": nproc(), "os_cpu_count": os.cpu_count(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        info["nproc_all"] = int(subprocess.run(["nproc", "--all"], capture_output=True, text=True).stdout.strip())
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except Exception:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return info


def time_cpu_rk4(O, st, blocks, dt, p, threads, reps, algebra):
    """Median seconds of one oracle RK4 step over `reps` runs after a warm-up,
    and the phase split (fractions of the summed thread time) of those runs."""
    import contextlib
    ctx = O.reference_block_algebra() if algebra == "reference" else contextlib.nullcontext()
    times = []
    with ctx:
        for r in range(reps + 1):
            s = st.copy()
            if r == 1:
                O.phase_times(reset=True)
            t0 = time.perf_counter()
            O.rk4_step(s, blocks, dt, p, enforce_per_stage=True, nthreads=threads)
            if r > 0:
                times.append(time.perf_counter() - t0)
    ph = O.phase_times(reset=True)
    tot = sum(ph.values()) or 1.0
    return statistics.median(times), {k: round(v / tot, 4) for k, v in ph.items()}


def cpu_baseline(wl, reps=5):
    """The CPU path on the host cores: an RK4 step (4 right-hand sides) with the
    reference's own generated equations (compiled in place, oracle/_ref) over
    each block, inside the oracle's restatement of rhs.cpp's block loop and
    derivs.cpp's stencils, OpenMP over blocks on `nproc` threads; median of
    `reps` after a warm-up, on a bounded sample of the workload. Also the same
    with the oracle's own tensor-loop algebra ('port'), the single-thread
    19^3 block of the CPU configuration, and the phase split at the
    reference's timer boundaries (deriv, rhs, bc, ko)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from dgr import workloads as W

    info = host_info()
    threads = info["nproc"]
    algebra = "reference" if O.ref_lib() is not None else "port"
    t, what = wl.cpu_sample()
    st = np.zeros((24, W.total_padded(t)))
    if wl.name in ("uniform13", "lattice13"):
        O.init_perturbed_minkowski(st, t, seed=1, amp=1e-3)
    else:
        O.init_brill_lindquist(st, t, seed=1, noise=1e-6)
    p = O.default_params(**wl.params)
    pts = W.interior_points(t)
    log(f"cpu baseline: {what}, {threads} threads, {algebra} algebra")
    med, split = time_cpu_rk4(O, st, t, wl.dt, p, threads, reps, algebra)
    res = {"value": round(pts * 4 / med / 1e6, 3), "unit": "Mpoints/s", "cores": threads, "kind": algebra,
           "sample": (f"one RK4 step (4 right-hand sides, stage updates, constraint enforcement) of {what} "
                      f"({pts} interior points), OpenMP over blocks on {threads} threads, median of {reps} "
                      f"after 1 warm-up; " +
                      ("the reference's generated equations (BSSN_GR/src/bssneqs_eta_const_standard_gauge.cpp, "
                       "compiled in place) over each block, with the oracle's restatement of the derivative "
                       "stencils, boundary pass, KO sum and RK4 around them" if algebra == "reference" else
                       "the oracle's restatement (tensor-loop algebra)")),
           "seconds_per_step": round(med, 4), "phase_split": split, **info}
    if algebra == "reference":
        med2, split2 = time_cpu_rk4(O, st, t, wl.dt, p, threads, reps, "port")
        res["port"] = {"value": round(pts * 4 / med2 / 1e6, 3), "phase_split": split2,
                       "algebra": "the oracle's tensor loops (no common subexpressions)"}
    # the CPU configuration: one 19^3 block, one thread
    t1 = W.single_block_19()
    st1 = np.zeros((24, W.total_padded(t1)))
    O.init_perturbed_minkowski(st1, t1, seed=1, amp=1e-3)
    p1 = O.default_params(ko_sigma=0.4, eta_central=2.0, eta_outer=2.0)
    med1, split1 = time_cpu_rk4(O, st1, t1, 0.025, p1, 1, reps, algebra)
    res["single_block_19_one_thread"] = {"value": round(W.interior_points(t1) * 4 / med1 / 1e6, 4),
                                         "unit": "Mpoints/s", "seconds_per_step": round(med1, 5),
                                         "phase_split": split1, "kind": algebra}
    return res


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
    traffic, traffic_src = measured_traffic(dom, wl.key + precision_tag)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
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


def fill_rate(wl, plan, state, stream, calls=10):
    """dgr_padding_fill (the whole refresh: same-level copies, then the other
    sizes and levels) alone: average time from HIP events on the launch
    stream, the number of padding points each part writes, and the copy-rate
    equivalent (every written point read once and written once, 24 variables
    x 16 B)."""
    import torch
    nbr = plan.link_neighbors()
    ncopy, ninterp = plan.fill_sources()
    sz = wl.blocks["sz"].astype(np.int64)
    linked = 0
    for d in range(27):
        dd = (d % 3 - 1, (d // 3) % 3 - 1, d // 9 - 1)
        if not any(dd):
            continue
        rows = nbr[:, d] >= 0
        ext = [np.full(int(rows.sum()), 3) if dd[a] else sz[rows, a] - 6 for a in range(3)]
        linked += int((ext[0] * ext[1] * ext[2]).sum())
    plan.padding_fill(state, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(calls):
        plan.padding_fill(state, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / calls
    pts = linked + ncopy + ninterp
    return {"linked_points": linked, "copied_points": ncopy, "interpolated_points": ninterp,
            "ms_per_call": round(ms, 4), "copy_equivalent_GBs": round(pts * 24 * 16 / ms / 1e6, 1),
            "use": ("RK4 stages read same-level padding inside k_fused (HALO variant) and get the rest "
                    "written before each stage by k_fill_copy and k_fill_interp; this full fill runs once per RK4 call")}


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
    res["note"] = ("one fused kernel per tile kind (k_cons_fused): the 17 variables the constraints "
                   "differentiate stream through LDS, derivatives in registers, no workspace; all six "
                   "outputs are computed whatever is asked for; FP64")
    del out
    return res


def run_fp64(wl, args, stream, dev, precision="fp64", keep_state0=True):
    """The timed RK4 steps of one workload (graph replay), the per-launch kernel
    times, the right-hand-side-only rate, the roofline of the dominant launch,
    and the constraint diagnostic. Returns (result dict, initial state or None)."""
    import torch
    import dgr
    log(f"workload: {wl.desc}; {wl.npts} interior points, state {24 * wl.n_pad * 8 / 1e9:.2f} GB")
    state = wl.device_state(dev)
    state0 = state.clone() if keep_state0 else None
    plan = dgr.Plan(wl.blocks, wl.n_pad)
    plan.set_precision(precision)
    if wl.name == "lattice13":
        plan.set_halo_refresh(True)
    if wl.name == "refined":
        plan.set_halo_refresh("full")
    params = dgr.default_params(**wl.params)
    if not args.no_autotune:
        # tile shapes per block size, timed on this workload's own data
        scratch = torch.empty_like(state)
        plan.autotune(scratch, state, params, stream=stream)
        torch.cuda.synchronize()
        del scratch
    # the plan allocates its RK4 buffers with hipMalloc: hand torch's cached
    # blocks back first (the 120 M-point octree needs ~200 GB of them)
    torch.cuda.empty_cache()
    tiling = {"x".join(str(n) for n in k): v for k, v in plan.tiling(precision).items()}
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
    res = {"value": round(wl.npts * 4 * args.steps / elapsed / 1e6, 2),
           "ms_per_step": round(elapsed / args.steps * 1e3, 4), "tiling": tiling}
    log(f"{wl.key}: {res['value']} Mpoints/s ({res['ms_per_step']} ms per RK4 step)")

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
    res["rhs_only_mpoints_s"] = None
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
        res["rhs_only_mpoints_s"] = round(wl.npts * nrhs / (time.perf_counter() - r0) / 1e6, 2)
        del out
    else:
        log("right-hand-side-only rate skipped: no room for a second padded state")
    res["roofline"] = roofline(wl, kt, 4 * ksteps, precision)
    res["constraints"] = constraint_rates(wl, plan, state, stream)
    res["halo_fill"] = halo_rate(wl, plan, state, stream) if wl.name == "lattice13" else None
    if wl.name == "refined":
        res["padding_fill"] = fill_rate(wl, plan, state, stream)
        # the same steps with frozen padding, for the cost of the refresh
        plan.set_halo_refresh(False)
        plan.rk4_steps(state, wl.dt, params, nsteps=1, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.rk4_steps(state, wl.dt, params, nsteps=args.steps, stream=stream)
        torch.cuda.synchronize()
        res["frozen_padding_mpoints_s"] = round(wl.npts * 4 * args.steps / (time.perf_counter() - t0) / 1e6, 2)
    plan.close()
    del state
    torch.cuda.empty_cache()
    return res, state0


def refresh_rate(wl, args, stream, dev, precision):
    """RK4 steps (graph replay) of a workload with the full padding refresh in
    one precision: the rate alone."""
    import torch
    import dgr
    state = wl.device_state(dev)
    plan = dgr.Plan(wl.blocks, wl.n_pad)
    plan.set_precision(precision)
    plan.set_halo_refresh("full")
    params = dgr.default_params(**wl.params)
    torch.cuda.empty_cache()
    plan.rk4_steps(state, wl.dt, params, nsteps=max(args.warmup, 1), stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.rk4_steps(state, wl.dt, params, nsteps=args.steps, stream=stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if plan.status(stream) != 0:
        raise RuntimeError(f"constraint violation during the {precision} steps")
    plan.close()
    del state
    res = {"precision": precision, "value": round(wl.npts * 4 * args.steps / elapsed / 1e6, 2),
           "unit": "Mpoints/s", "ms_per_step": round(elapsed / args.steps * 1e3, 4),
           "padding": "refreshed in the stage kernels from same-level neighbours, the rest written before each stage"}
    log(f"{wl.key} {precision}: {res['value']} Mpoints/s")
    return res


def large_octree(args, stream, dev):
    """The 120 M-point octree (the puncture distribution with 4x the block
    counts): `--large-steps` graph-replayed RK4 steps in FP64, then in mixed
    precision from the same initial data; sustained Mpoints/s, the algorithmic
    HBM rate and FP64/FP32 vector-ALU rate of the whole stage, every value
    finite, and the Hamiltonian-constraint norm after the steps in both."""
    import torch
    import dgr
    wl = Workload("puncture", 0, args.large_scale)
    log(f"large octree: {wl.desc}; {wl.npts} interior points, state {24 * wl.n_pad * 8 / 1e9:.1f} GB")
    params = dgr.default_params(**wl.params)
    plan = dgr.Plan(wl.blocks, wl.n_pad)
    out = {"workload": wl.desc, "interior_points": wl.npts, "steps": args.large_steps,
           "state_GB": round(24 * wl.n_pad * 8 / 1e9, 1)}
    sbytes = W_stage_bytes(wl.blocks)
    for prec in ("fp64", "mixed"):
        state = wl.device_state(dev)
        plan.set_precision(prec)
        torch.cuda.empty_cache()
        plan.rk4_steps(state, wl.dt, params, nsteps=1, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.rk4_steps(state, wl.dt, params, nsteps=args.large_steps, stream=stream)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        if plan.status(stream) != 0:
            raise RuntimeError(f"constraint violation in the large octree ({prec})")
        # one variable at a time: a whole-state temporary would not fit beside the plan's buffers
        finite = all(bool(torch.isfinite(state[v]).all().item()) for v in range(24))
        ham = torch.empty(wl.n_pad, dtype=torch.float64, device=dev)
        plan.ham(ham, state, stream=stream)
        norm = plan.constraint_norm(ham, stream=stream)
        del ham, state
        stage_s = el / args.large_steps / 4
        out[prec] = {"value": round(wl.npts * 4 * args.large_steps / el / 1e6, 2), "unit": "Mpoints/s",
                     "ms_per_step": round(el / args.large_steps * 1e3, 3),
                     "stage_algorithmic_GBs": round(sbytes / stage_s / 1e9, 1),
                     "stage_tflops": round(wl.npts * FLOP_PER_POINT_STAGE / stage_s / 1e12, 2),
                     "all_finite": finite,
                     f"ham_constraint_norm_after_{args.large_steps + 1}_steps": norm}
        log(f"large octree {prec}: {out[prec]['value']} Mpoints/s, finite {finite}, ham norm {norm:.6e}")
    n64 = out["fp64"][f"ham_constraint_norm_after_{args.large_steps + 1}_steps"]
    n32 = out["mixed"][f"ham_constraint_norm_after_{args.large_steps + 1}_steps"]
    out["mixed_speedup_vs_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    out["ham_norm_relative_difference"] = abs(n32 - n64) / abs(n64)
    plan.close()
    torch.cuda.empty_cache()
    return out


def W_stage_bytes(blocks):
    from dgr import workloads as W
    return W.stage_bytes(blocks)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["octree", "uniform13", "puncture", "lattice13", "refined"], default="octree",
                    help="octree (default): the puncture octree (scale 0.42) as the headline, with its "
                         "mixed-precision leg, the 120 M-point octree and the uniform 13^3 batch beside it; "
                         "the others run one workload alone")
    ap.add_argument("--nblocks", type=int, default=4096, help="uniform13: number of blocks")
    ap.add_argument("--scale", type=float, default=0.42, help="puncture: block-count scale")
    ap.add_argument("--large-scale", type=float, default=1.68, help="octree: scale of the large octree")
    ap.add_argument("--large-steps", type=int, default=50, help="octree: RK4 steps on the large octree")
    ap.add_argument("--no-large", action="store_true", help="octree: skip the large octree")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 derivatives/equations/KO on offset-subtracted fields, FP64 state")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision comparison of the FP64 run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tiling instead of timing the tilings per block size")
    ap.add_argument("--gauge", default="",
                    help="comma list of etafunc, rochester, ssl: time the right-hand side of the "
                         "reference's other builds instead of the default one (FP64; use --no-mixed)")
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

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    composite = args.workload == "octree"
    wl = Workload("puncture" if composite else args.workload, args.nblocks, args.scale)
    if args.gauge:
        # the reference's other builds of the equations (their own kernels); the
        # traffic profiles are of the default kernels, so none is looked up
        names = args.gauge.split(",")
        wl.params.update(eta_func=int("etafunc" in names), rochester=int("rochester" in names),
                         ssl=int("ssl" in names), xi=(1.0, 1.0, 1.0))
        wl.key += "_" + "_".join(names)
    if wl.name in ("lattice13", "refined"):  # the comparisons are for the frozen-padding step
        args.no_mixed = args.no_cpu_baseline = True
    want_mixed = args.precision == "fp64" and not args.no_mixed
    res, state0 = run_fp64(wl, args, stream, dev, args.precision, keep_state0=want_mixed)

    mixed = None
    if want_mixed:
        import dgr
        mixed = mixed_comparison(wl, state0, dgr.default_params(**wl.params), args.steps, res["value"], stream,
                                 not args.no_autotune)
        log(f"mixed precision: {mixed['value']} Mpoints/s ({mixed['speedup_vs_fp64']}x), errors "
            f"{json.dumps(mixed['error_vs_fp64'])}")
    del state0
    torch.cuda.empty_cache()

    uniform = large = refined = None
    if composite:
        # the uniform 13^3 batch beside the octree: its rate, roofline and mixed errors
        wu = Workload("uniform13", args.nblocks, args.scale)
        ru, su = run_fp64(wu, args, stream, dev, "fp64", keep_state0=not args.no_mixed)
        uniform = {"workload": wu.desc, "interior_points": wu.npts, "value": ru["value"], "unit": "Mpoints/s",
                   "ms_per_step": ru["ms_per_step"], "rhs_only_mpoints_s": ru["rhs_only_mpoints_s"],
                   "roofline": ru["roofline"], "constraints": ru["constraints"], "tiling": ru["tiling"]}
        if su is not None:
            import dgr
            uniform["mixed_precision"] = mixed_comparison(wu, su, dgr.default_params(**wu.params), args.steps,
                                                          ru["value"], stream, not args.no_autotune)
        del su
        torch.cuda.empty_cache()
        # the refined octree: RK4 with the padding refreshed across level jumps
        wr = Workload("refined", 0, 0.0)
        rr, _ = run_fp64(wr, args, stream, dev, "fp64", keep_state0=False)
        refined = {"workload": wr.desc, "blocks": len(wr.blocks), "interior_points": wr.npts,
                   "value": rr["value"], "unit": "Mpoints/s", "ms_per_step": rr["ms_per_step"],
                   "frozen_padding_mpoints_s": rr["frozen_padding_mpoints_s"],
                   "padding_fill": rr["padding_fill"], "roofline": rr["roofline"]}
        torch.cuda.empty_cache()
        if not args.no_mixed:
            refined["mixed_precision"] = refresh_rate(wr, args, stream, dev, "mixed")
            torch.cuda.empty_cache()
        if not args.no_large:
            free, _ = torch.cuda.mem_get_info()
            need = 24 * 8 * (2 * Workload("puncture", 0, args.large_scale).n_pad) * 1.3
            if free > need:
                large = large_octree(args, stream, dev)
            else:
                log(f"large octree skipped: {free / 1e9:.0f} GB free, about {need / 1e9:.0f} GB needed")

    cpu = None if args.no_cpu_baseline else cpu_baseline(wl)

    line = {"metric": METRIC, "value": res["value"], "unit": "Mpoints/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if args.precision == "fp64" else "f32 arithmetic, f64 state",
            "data": "synthetic",
            "config": {"workload": wl.desc, "blocks": len(wl.blocks), "interior_points": wl.npts,
                       "precision": args.precision, "rhs_evals_per_step": 4, "rk4_step": "hipGraph replay",
                       "tiling": res["tiling"], "autotuned": not args.no_autotune,
                       "dt": wl.dt, **wl.params},
            "rhs_only_mpoints_s": res["rhs_only_mpoints_s"], "constraints": res["constraints"],
            "halo_fill": res["halo_fill"], "roofline": res["roofline"], "cpu_baseline": cpu,
            "mixed_precision": mixed, "uniform13": uniform, "refined_octree": refined, "large_octree": large}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
