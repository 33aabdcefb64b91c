"""Placement-step throughput on one MI355X.

One step = for every orientation of the catalog, the overlap counts of the
oriented part over all legal translations of the current occupancy (spectral
correlation), the feasible mask (rounded count == 0), the lowest (z, y, x)
choice with the earliest orientation winning ties, and stamping the chosen part
into the occupancy — all on the device as one graph launch.  The default
workload is the L-bracket in an r=120 mm sphere voxelised at 256^3 with 96
seeded SO(3) orientations.  Kernel rasterisation runs once per run and is
reported separately.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cpp-max-packer_amd"))

METRIC = "placement steps/s at 256³×96 orientations; parts packed vs plain search"
PEAK_HBM_GBS = 8000.0
# MI355X peak engine clock; a CU issues at most one VALU wave-instruction per
# clock (four SIMDs, each taking four clocks per 64-wide instruction)
PEAK_CLOCK_GHZ = 2.4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def step_model_bytes(nx, ny, nz, orientations, precision):
    """Fixed roofline model of one step: B = 48 Nc + O (120 Nc + 4 N) bytes in
    FP32 (every spectrum term doubles in FP64 and the field read is 8 N)."""
    N = nx * ny * nz
    nc = (nx // 2 + 1) * ny * nz
    if precision == 64:
        return 96 * nc + orientations * (240 * nc + 8 * N)
    return 48 * nc + orientations * (120 * nc + 4 * N)


def kernel_bytes(name, n, kernels, ext, elem, fused=False):
    """Algorithmic bytes one launch of each step kernel must move (all
    orientations).  fused: the product pass writes the fused path's
    column-major output, whose kz = 0 and kz = nz/2 rows share one plane
    (nz/2 planes per column instead of nz/2 + 1)."""
    nx, ny, nz = n
    nh = nz // 2 + 1
    planes = nz // 2 if fused else nh
    ex, ey, ez = ext
    fit = [k for k in kernels if all(ext[a] >= k["shape"][a] for a in range(3))]
    if name == "env_r2c_z":
        return ex * ey * ez + nh * ny * nx * elem
    if name in ("env_c2c_y", "env_c2c_x"):
        return 2 * nh * ny * nx * elem
    if name == "kernel_r2c_z":
        return sum(k["shape"][0] * k["shape"][1] * k["shape"][2] + nh * k["shape"][1] * k["shape"][0] * elem for k in fit)
    if name == "kernel_c2c_y":
        return sum(nh * k["shape"][1] * k["shape"][0] * elem + nh * ny * k["shape"][0] * elem for k in fit)
    if name == "mul_x":
        # each orientation's kernel rows (Kx values) in and legal columns (Lx) out;
        # the environment spectrum once per launch (orientations are batched)
        return nh * ny * nx * elem + sum((nh * k["shape"][0] + planes * (ex - k["shape"][0] + 1)) * ny * elem
                                         for k in fit)
    if name == "inv_c2c_y":
        return sum(nh * (ny + ey - k["shape"][1] + 1) * (ex - k["shape"][0] + 1) * elem for k in fit)
    if name == "c2r_z_argmin":
        return sum(nh * (ey - k["shape"][1] + 1) * (ex - k["shape"][0] + 1) * elem for k in fit)
    if name == "yz_argmin":
        # fused inverse y + z: every plane (kz = 0 and nz/2 packed) of every
        # legal x column, read once
        return sum((nz // 2) * ny * (ex - k["shape"][0] + 1) * elem for k in fit)
    return 0


def incremental_timing(sp, ctx, w, env, kernels, prec, args, ms_step=None):
    """Incremental steps (fields updated by each step's new cells) on a
    session from the container occupancy, timed over args.steps steps and
    checked against full steps from the same occupancy (ms_step: the full
    step's time, for the speedup)."""
    import numpy as np
    import torch
    inc = sp.Session(ctx, w.env_shape, w.env_first, len(kernels), precision=prec, capture_graph=not args.no_graph)
    inc.set_environment(env)
    for i, k in enumerate(kernels):
        inc.set_kernel(i, k["cells"], k["shape"], k["first"])
    inc.set_incremental(True)
    ref = sp.Session(ctx, w.env_shape, w.env_first, len(kernels), precision=prec, capture_graph=not args.no_graph)
    ref.set_environment(env)
    for i, k in enumerate(kernels):
        ref.set_kernel(i, k["cells"], k["shape"], k["first"])
    for _ in range(1 + args.warmup):
        inc.step(stamp=True)
        ref.step(stamp=True)
    istream = torch.cuda.ExternalStream(inc.stream())
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    g0.record(istream)
    inc.run(args.steps, stamp=True)
    g1.record(istream)
    g1.synchronize()
    ms_inc = g0.elapsed_time(g1) / args.steps
    ref.run(args.steps, stamp=True)
    same = bool(np.array_equal(inc.environment(), ref.environment())) and \
        inc.last_result()["translation"] == ref.last_result()["translation"]
    out = {"ms_per_step": ms_inc, "steps_per_s": 1000.0 / ms_inc,
           "speedup_vs_full": (ms_step / ms_inc) if ms_step else None,
           "same_occupancy_and_choice_as_full_steps": same,
           "note": "overlap fields updated by the correlation of each step's new cells "
                   "(gap_fill.cpp:426-451,524-584); first step full"}
    log(f"[bench] incremental steps: {ms_inc:.3f} ms/step ({1000.0 / ms_inc:.1f} steps/s), same as full: {same}")
    inc.close()
    ref.close()
    return out


def fp_deviation(sp, ctx, w, env, kernels, bound):
    """One step from the container occupancy in FP32 and in FP64: the largest
    |FP32 - FP64| over every legal overlap value of every orientation."""
    import numpy as np
    raws = {}
    for prec in (32, 64):
        s = sp.Session(ctx, w.env_shape, w.env_first, len(kernels), precision=prec, capture_graph=False)
        s.set_environment(env)
        for i, k in enumerate(kernels):
            s.set_kernel(i, k["cells"], k["shape"], k["first"])
        s.keep_raw(True)
        s.step(stamp=False)
        raws[prec] = [s.raw_values(i) for i in range(len(kernels))]
        s.close()
    dev = 0.0
    for a, b in zip(raws[32], raws[64]):
        if a.size:
            dev = max(dev, float(np.max(np.abs(a - b))))
    out = {"max_abs_dev": dev, "fp32_bound": bound, "orientations": len(kernels),
           "reference": "the device FP64 path on the same occupancy (itself bit-exact to the FP64 oracle on integers)"}
    # against the CPU FP64 oracle on a bounded sample of orientations: the
    # unrounded full-linear correlation (the oracle's real-valued path on the
    # 0/1 occupancy), read at each legal corner u as full-linear index u + K - 1
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    # (grids up to 256^3: one FP64 CPU correlation there takes about 10 s)
    sample = [o for o in range(len(kernels)) if raws[32][o].size][:2] if int(np.prod(w.env_shape)) <= 256 ** 3 else []
    cdev = 0.0
    t0 = time.perf_counter()
    for o in sample:
        k = kernels[o]
        spec = (w.env_shape, k["shape"], w.env_first, k["first"])
        ref = O.correlate(spec, env.astype(np.float64), k["cells"], proximity=True, limits=O.LIMITS_UNBOUNDED)
        full = ref["values"].reshape(int(ref["shape"][2]), int(ref["shape"][1]), int(ref["shape"][0]))
        kx, ky, kz = k["shape"]
        lx, ly, lz = (w.env_shape[0] - kx + 1, w.env_shape[1] - ky + 1, w.env_shape[2] - kz + 1)
        legal = full[kz - 1:kz - 1 + lz, ky - 1:ky - 1 + ly, kx - 1:kx - 1 + lx].ravel()
        cdev = max(cdev, float(np.max(np.abs(raws[32][o] - legal))))
    out["cpu_fp64_oracle"] = {"max_abs_dev": cdev, "orientations": sample, "seconds": time.perf_counter() - t0,
                              "reference": "the CPU FP64 oracle (oracle/correlation.cpp), unrounded, on the same "
                                           "occupancy"}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(workload, env, kernels, threads, sample_orientations):
    """The oracle's FP64 full-linear restatement of the same correlations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    t0 = time.perf_counter()
    done, _ = O.correlate_orientations(env, workload.env_first, kernels, 0, sample_orientations, threads)
    dt = time.perf_counter() - t0
    per_step = dt / max(done, 1) * len(kernels) if threads == 1 else dt * len(kernels) / max(done, 1)
    return dict(value=1.0 / per_step, unit="steps/s", cores=threads, kind="port", cpu_model=cpu_model(),
                nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)),
                sample=f"{done} of {len(kernels)} orientation correlations of one step (FP64 full-linear restatement, "
                       f"P=E+K-1, {threads} thread(s)) in {dt:.1f} s; steps/s = {done}/({dt:.1f} s x {len(kernels)})")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="bracket", choices=["box", "star", "bracket", "knot"])
    ap.add_argument("--precision", type=int, default=0, choices=[0, 32, 64])
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=1)
    ap.add_argument("--profile-steps", type=int, default=4)
    ap.add_argument("--no-deviation", action="store_true", help="skip the FP32-vs-FP64 deviation step")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal packing comparison")
    ap.add_argument("--no-incremental", action="store_true", help="skip the incremental-step timing")
    ap.add_argument("--separate-passes", action="store_true",
                    help="run the inverse y and inverse z passes separately (no fused pass)")
    ap.add_argument("--learned-seeds", default="9000,9001,9002",
                    help="held-out part seeds packed with and without proposals")
    ap.add_argument("--weights", default="", help="proposal weights (.npy) for the learned leg (default: the package's)")
    ap.add_argument("--top-o", type=int, default=12,
                    help="learned leg: orientations searched per lazy step (and per restricted step)")
    ap.add_argument("--top-r", type=int, default=512, help="learned leg: regions of the restricted step")
    ap.add_argument("--prior-weight", type=float, default=2.0,
                    help="learned leg: weight of the orientation prior added to the network's logits")
    ap.add_argument("--restricted-top-o", type=int, default=8,
                    help="learned leg: orientations of the restricted (non-lazy) proposal shown beside it")
    ap.add_argument("--only-learned", action="store_true",
                    help="run only learned-proposal episodes (no plain step timing, no plain episodes); for profiling")
    ap.add_argument("--no-ranked", action="store_true", help="skip the ranked-page (proximity) timing")
    ap.add_argument("--only-ranked", action="store_true", help="run only the ranked-page timing; for profiling")
    ap.add_argument("--only-incremental", action="store_true",
                    help="run only the incremental-step timing (with its check against full steps); for profiling")
    ap.add_argument("--ranked-placements", type=int, default=10, help="greedy placements timed on ranked pages")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 4, "dispatches": 3}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the placement step runs on a single GPU in this build "
              "(multi-GPU orientation sharding is a design note only); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import spectrapack_amd as sp
    from spectrapack_amd import workloads as W

    if args.only_learned:
        ctx = sp.Context(0)
        print(json.dumps({"learned_proposal": learned_comparison(ctx, args, plain=False)}), flush=True)
        return
    log(f"[bench] workload={args.workload} steps={args.steps} warmup={args.warmup}")
    w = W.WORKLOADS[args.workload]()
    ctx = sp.Context(0)
    # container occupancy and oriented kernels, on the device
    t0 = time.perf_counter()
    if w.container is not None:
        env = ctx.mesh_container(w.container.vertices, w.container.triangles, w.env_first, w.env_shape, w.pitch)
    else:
        env = ctx.box_container(w.env_first, w.env_shape, w.pitch, w.box_dims)
    t_container = time.perf_counter() - t0
    t0 = time.perf_counter()
    kernels = ctx.voxelize_objects(w.part.vertices, w.part.triangles, w.catalog, w.pitch)
    t_raster = time.perf_counter() - t0
    log(f"[bench] container {t_container:.3f}s raster {len(kernels)} orientations {t_raster:.3f}s")

    if args.only_ranked:
        print(json.dumps({"ranked_pages": ranked_timing(sp, ctx, w, env, kernels, args.precision or 0, args)}),
              flush=True)
        return
    if args.only_incremental:
        prec = args.precision or (32 if w.precision == "fp32" else 0)
        print(json.dumps({"incremental": incremental_timing(sp, ctx, w, env, kernels, prec, args)}), flush=True)
        return
    s = sp.Session(ctx, w.env_shape, w.env_first, len(kernels), precision=args.precision or
                   (32 if w.precision == "fp32" else 0), capture_graph=not args.no_graph)
    if args.separate_passes:
        s.set_fused_passes(False)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    r = s.step(stamp=True)  # prepares buffers and captures the graph
    prec = r["precision"]
    log(f"[bench] precision fp{prec} bound={r['error_bound']:.3g} device bytes={s.device_bytes() / 2**30:.2f} GiB")
    tiles = {}
    for p, name in enumerate(("mul_x", "inv_c2c_y", "c2r_z_argmin")):
        radix, width, tms = s.tile(p)
        tiles[name] = {"radix": radix, "width": width, "tuned_ms": tms}
    log("[bench] autotuned tiles: " + ", ".join(f"{k} R{v['radix']} W{v['width']} ({v['tuned_ms']:.3f} ms)"
                                                for k, v in tiles.items()))
    for _ in range(args.warmup):
        r = s.step(stamp=True)
    stream = torch.cuda.ExternalStream(s.stream())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    e0.record(stream)
    s.run(args.steps, stamp=True)
    e1.record(stream)
    e1.synchronize()
    t_wall = time.perf_counter() - t_wall
    torch.cuda.synchronize()
    last = s.last_result()
    ms_step = e0.elapsed_time(e1) / args.steps
    log(f"[bench] {args.steps} steps: {ms_step:.3f} ms/step (events), wall {t_wall / args.steps * 1e3:.3f} ms/step; "
        f"last choice {last['orientation']} residual {last['max_residual']:.3g}")

    # the same steps with every orientation's kernel spectrum recomputed in
    # each step (the spectra do not depend on the occupancy; by default they
    # stay resident in HBM)
    s.set_resident_kernels(False)
    s.step(stamp=True)
    f0 = torch.cuda.Event(enable_timing=True)
    f1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    f0.record(stream)
    s.run(args.steps, stamp=True)
    f1.record(stream)
    f1.synchronize()
    ms_step_recompute = f0.elapsed_time(f1) / args.steps
    s.set_resident_kernels(True)
    log(f"[bench] kernel spectra recomputed every step: {ms_step_recompute:.3f} ms/step")

    # incremental steps (fields updated by each step's new cells) on a second
    # session from the same starting occupancy, checked against this one
    incremental = None
    if not args.no_incremental:
        incremental = incremental_timing(sp, ctx, w, env, kernels, prec, args, ms_step)

    ranked = None
    if not args.no_ranked:
        ranked = ranked_timing(sp, ctx, w, env, kernels, prec, args)

    # per-kernel breakdown (no graph, events between kernels), after one
    # untimed profile step
    s.profile_step(stamp=True)
    prof = {}
    for _ in range(args.profile_steps):
        for name, ms in s.profile_step(stamp=True):
            prof.setdefault(name, []).append(ms)
    prof = {k: sum(v) / len(v) for k, v in prof.items()}
    n = tuple(int(sp_len) for sp_len in _fft_shape(w.env_shape))
    elem = 8 if prec == 32 else 16
    dominant = max(prof, key=prof.get)
    dom_bytes = kernel_bytes(dominant, n, kernels, w.env_shape, elem, fused=s.fused_passes())
    # the dominant pass launched back to back (a graph of 20 launches between
    # two events on the session stream), as a rocprof kernel trace times it
    dom_ms = s.time_pass(dominant, 20)
    log(f"[bench] {dominant}: {dom_ms:.3f} ms per launch back to back "
        f"({prof[dominant]:.3f} ms between events in a step)")
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    step_bytes = step_model_bytes(*n, len(kernels), prec)
    step_achieved = step_bytes / (ms_step * 1e-3) / 1e9
    log("[bench] kernels: " + ", ".join(f"{k} {v:.3f} ms" for k, v in prof.items()))

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        if tj.get("precision") == prec:
            traffic = tj.get("per_kernel_hbm_bytes", {}).get(dominant)
    # issue side of the same kernel: VALU wave-instructions per launch measured
    # with rocprofv3 (SQ_INSTS_VALU) on this workload, against the chip's issue
    # rate of one VALU wave-instruction per CU per clock
    issue = None
    ifile = os.path.join(ROOT, "profiles", f"valu_{args.workload}.json")
    if os.path.exists(ifile):
        with open(ifile) as f:
            ij = json.load(f)
        n_valu = ij.get("valu_wave_instructions_per_launch", {}).get(dominant)
        if ij.get("precision") == prec and n_valu:
            cus = torch.cuda.get_device_properties(0).multi_processor_count
            issue_ms = n_valu / (cus * PEAK_CLOCK_GHZ * 1e9) * 1e3
            issue = {"bound": "valu", "kernel": dominant, "valu_wave_instructions": n_valu, "cus": cus,
                     "clock_ghz": PEAK_CLOCK_GHZ, "issue_ms": issue_ms, "avg_launch_ms": dom_ms,
                     "frac": issue_ms / dom_ms, "source": os.path.relpath(ifile, ROOT)}
    out = {
        "metric": METRIC,
        "value": 1000.0 / ms_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f{prec}",
        "data": "synthetic (seeded meshes and catalog, seed 20261203)",
        "config": {"workload": w.name, "grid": list(w.env_shape), "fft_grid": list(n), "orientations": len(kernels),
                   "pitch_mm": w.pitch, "precision": f"fp{prec}", "error_bound": r["error_bound"],
                   "graph": not args.no_graph, "parallelism": "single GPU", "kernel_spectra": "resident",
                   "fused_inverse_yz": s.fused_passes(),
                   "autotuned_tiles": tiles},
        "ms_per_step_kernel_spectra_recomputed": ms_step_recompute,
        "incremental": incremental,
        "ranked_pages": ranked,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms,
                     "avg_launch_timing": "20 launches back to back in one graph, HIP events on the session stream"},
        "issue_roofline": issue,
        "step_model_rate": {"model_bytes_per_step": step_bytes, "model_bytes_per_s": step_achieved,
                            "ratio_to_hbm_peak": step_achieved / PEAK_HBM_GBS, "unit": "GB/s",
                            "note": "the fixed model B_step = 48 Nc + O (120 Nc + 4 N) divided by the step time; "
                                    "the fused passes skip HBM round trips the model counts, so the ratio can "
                                    "exceed 1 and is not a roofline"},
        "kernel_ms": prof,
        "raster_s": t_raster,
        "container_s": t_container,
        "steps_run": 2 + args.warmup + 2 * args.steps,
        "last_step_found_a_placement": last["orientation"] >= 0,
        "max_residual_last_step": last["max_residual"],
    }
    if not args.no_deviation:
        log("[bench] FP32 vs FP64 deviation ...")
        out["fp32_vs_fp64"] = fp_deviation(sp, ctx, w, env, kernels, r["error_bound"])
        log(f"[bench] max |fp32 - fp64| = {out['fp32_vs_fp64']['max_abs_dev']:.3g} (bound {r['error_bound']:.3g})")
    if not args.no_learned:
        out["learned_proposal"] = learned_comparison(ctx, args)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline sample ...")
        out["cpu_baseline"] = cpu_baseline(w, env, kernels, args.cpu_threads, max(1, args.cpu_threads))
        # the same restatement with orientations in parallel over the host's
        # cores (at most 16, this GPU's share of the host)
        # each correlation holds about 40 bytes per padded cell (two FP64
        # complex buffers and the output); keep the parallel sample under 64 GB
        kmax = [max(k["shape"][a] for k in kernels) for a in range(3)]
        padded = 1
        for a in range(3):
            padded *= w.env_shape[a] + kmax[a] - 1
        n_all = max(1, min(16, os.cpu_count() or 1, int(64e9 // (40 * padded))))
        if n_all > 1 and args.cpu_threads == 1:
            log(f"[bench] CPU baseline on {n_all} cores ...")
            out["cpu_baseline_all_cores"] = cpu_baseline(w, env, kernels, n_all, n_all)
    print(json.dumps(out), flush=True)


def ranked_timing(sp, ctx, w, env, kernels, prec, args):
    """The reference's ranked-page path (spectral_pipeline.cpp:1343-1857 +
    spectral.cpp:1485-1590) on the same workload: proximity field, FP64
    proximity correlation, binary overlaps (incremental fields), a 32-entry
    page per orientation, and greedy placements over the pages."""
    import numpy as np
    import torch
    from spectrapack_amd import workloads as W
    s = sp.Session(ctx, w.env_shape, w.env_first, len(kernels), precision=prec, capture_graph=not args.no_graph)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    s.set_incremental(True)
    cmin, cmax = W.container_z_range(w)
    s.set_ranking(0.0, w.pitch, cmin, cmax, W.oriented_max_z(w.part, w.catalog))
    t0 = time.perf_counter()
    s.ranked_fill(max_copies=1)     # autotune, graph capture, first fields
    t_first = time.perf_counter() - t0
    mass0 = int(s.environment().sum())
    stream = torch.cuda.ExternalStream(s.stream())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = max(3, min(args.steps, 10))
    s.rank()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        s.rank()
    e1.record(stream)
    e1.synchronize()
    ms_rank = e0.elapsed_time(e1) / reps
    pages = s.ranked_pages()   # the binding's first call sets up its numpy/ctypes views (tens of ms, once)
    t0 = time.perf_counter()
    for _ in range(reps):
        pages = s.ranked_pages()
    ms_pages = (time.perf_counter() - t0) / reps * 1e3
    t0 = time.perf_counter()
    placed = s.ranked_fill(max_copies=args.ranked_placements)
    ms_fill = (time.perf_counter() - t0) / max(len(placed), 1) * 1e3
    mass1 = int(s.environment().sum())
    masses = [int(k["cells"].sum()) for k in kernels]
    added = sum(masses[o] for o, _, _ in placed)
    heads = sorted((pages[o][0]["score"], o) for o in pages if pages[o])
    out = {"ms_rank": ms_rank, "ms_pages_all_orientations": ms_pages, "ms_per_placement": ms_fill,
           "placements_per_s": 1000.0 / ms_fill if placed else 0.0, "placements": len(placed),
           "first_fill_s": t_first, "nonempty_pages": sum(1 for o in pages if pages[o]),
           "best_head_score": heads[0][0] if heads else None,
           "mass_added_equals_placed_kernels": mass1 - mass0 == added,
           "note": "rank = binary overlaps (incremental fields) + proximity field + FP64 proximity correlation; "
                   "pages = 32 best per orientation + direct recheck; a placement = rank + pages + choice + stamp"}
    log(f"[bench] ranked pages: rank {ms_rank:.3f} ms, pages {ms_pages:.3f} ms, {ms_fill:.3f} ms/placement "
        f"({len(placed)} placed), mass check {out['mass_added_equals_placed_kernels']}")
    s.close()
    return out


def learned_comparison(ctx, args, plain=True):
    """Packs held-out star parts at 256^3 with 96 orientations to exhaustion,
    with the plain search and with the learned proposal in its lazy form
    (each orientation's last exact minimum kept while its corner stays free;
    of the orientations the new occupancy blocks, the top_o by network logit
    plus the orientation prior are searched), and for reference with the
    restricted proposal (the network's top orientations in its top regions,
    full search as fallback); wall-clock steps/s of each loop (one step and
    its result read-back per iteration)."""
    import spectrapack_amd as sp
    from spectrapack_amd import learned as L
    from spectrapack_amd import workloads as W
    wpath = args.weights or L.DEFAULT_WEIGHTS
    if not os.path.exists(wpath):
        return {"skipped": "no trained weights at " + os.path.relpath(wpath, ROOT)}
    prop = sp.Proposal(ctx, max_batch=1)
    prop.set_parameters(L.load_weights(wpath))
    meta_path = os.path.splitext(wpath)[0] + ".json"
    meta = json.load(open(meta_path)) if os.path.exists(meta_path) else {}
    rows = []
    for seed in [int(v) for v in args.learned_seeds.split(",")]:
        sc = L.Scene(ctx, W.star_so3_fine_box(seed))
        learned = L.run_episode_lazy(sc, prop, args.top_o, prior_weight=args.prior_weight)
        if not plain:
            log(f"[bench] learned proposal seed {seed}: {learned['parts']} parts {learned['steps_per_s']:.1f} steps/s")
            rows.append(dict(seed=seed, learned=learned))
            sc.close()
            continue
        restricted = L.run_episode(sc, prop, args.restricted_top_o, args.top_r)
        pl = L.run_episode(sc)
        log(f"[bench] learned proposal seed {seed}: plain {pl['parts']} parts {pl['steps_per_s']:.1f} steps/s, "
            f"lazy {learned['parts']} parts {learned['steps_per_s']:.1f} steps/s "
            f"({learned['searched_per_step']:.1f} orientations searched per step, {learned['full_searches']} full), "
            f"restricted {restricted['parts']} parts {restricted['steps_per_s']:.1f} steps/s")
        rows.append(dict(seed=seed, plain=pl, learned=learned, restricted=restricted))
        sc.close()
    if not plain:
        return {"episodes": rows, "top_orientations": args.top_o, "prior_weight": args.prior_weight}

    def total(key):
        return (sum(r[key]["parts"] for r in rows),
                sum(r[key]["steps"] for r in rows) / sum(r[key]["seconds"] for r in rows))
    pp, ps = total("plain")
    lp, ls = total("learned")
    rp, rsp = total("restricted")
    return {"workload": W.star_so3_fine_box(rows[0]["seed"]).name, "method": "lazy",
            "top_orientations": args.top_o, "prior_weight": args.prior_weight,
            "plain_parts": pp, "learned_parts": lp, "parts_ratio": lp / max(pp, 1),
            "plain_steps_per_s": ps, "learned_steps_per_s": ls, "speedup": ls / ps,
            "searched_per_step": sum(r["learned"]["searched_per_step"] * r["learned"]["steps"] for r in rows) /
            max(sum(r["learned"]["steps"] for r in rows), 1),
            "full_searches": sum(r["learned"]["full_searches"] for r in rows),
            "restricted": {"top_orientations": args.restricted_top_o, "top_regions": args.top_r, "parts": rp,
                           "parts_ratio": rp / max(pp, 1), "steps_per_s": rsp, "speedup": rsp / ps,
                           "fallbacks": sum(r["restricted"]["fallbacks"] for r in rows)},
            "episodes": rows,
            "training": {k: meta.get(k) for k in ("episodes", "samples", "epochs", "scene", "hits_trained",
                                                  "hits_untrained")}}


def _fft_shape(env_shape):
    """The FFT grid the session uses: the smallest 2^a 3^b 5^c length >= each extent."""
    def good(n):
        m = max(n, 1)
        while True:
            r = m
            for p in (2, 3, 5):
                while r % p == 0:
                    r //= p
            if r == 1:
                return m
            m += 1
    return [good(int(e)) for e in env_shape]


if __name__ == "__main__":
    main()
