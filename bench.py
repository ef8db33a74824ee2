#!/usr/bin/env python3
"""RMSD benchmark on one MI355X.

Default workload (the headline): one-vs-all md.rmsd over 100,000 frames x
2,000 atoms, float32, a seeded random 1,500-atom selection, every frame scored
against frame 0.  A step is one full pass: prepare the reference frame
(gather + FP64 sums) and run the per-frame kernel over all 100,000 frames
(fused gather, FP64 accumulation, QCP in FP64).  Coordinates (2.4 GB) are
generated in HBM before timing.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload one-vs-all|all-pairs]

Prints one JSON line.  ``roofline`` is for the per-frame kernel, timed with
HIP events on the stream it is launched on; ``cpu_baseline`` times the
reference's own compiled C++ core (oracle/_ref, OpenMP over frames, with
numpy's single-threaded gather as md.rmsd does) on a bounded sample of the
same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RMSD frame-pairs/s at \u22641e-5 nm vs FP64 oracle; query p50/p99 latency"
HBM_PEAK_GBS = 8000.0
F16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads():
    n = os.environ.get("OMP_NUM_THREADS")
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = os.cpu_count() or 1
    return int(n) if n and n.isdigit() else aff


def traffic_from_profile(name, shape=None):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel of known size), committed under
    profiles/; None when that measurement is absent or was taken on another
    shape than this run's."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    for r in rec.get("records", [rec]):  # one record, or one per shape
        if shape is None or "shape" not in r or list(r["shape"]) == list(shape):
            return r.get("hbm_bytes_per_launch")
    return None


def _probe_rate(entry, x, nbytes, blocks, warmup, steps):
    """GB/s of one streaming-read probe launch configuration (HIP events on
    the stream the probe runs on)."""
    import torch

    from mdtraj_amd import _lib, _device as dev

    sink = torch.zeros(1, device=x.device)

    def run():
        _lib.call(entry, x.data_ptr(), nbytes, sink.data_ptr(), blocks, dev.stream())

    ms = _time_events(run, steps, warmup)
    return nbytes / (ms * 1e-3) / 1e9, ms


def read_ceiling(x, warmup, steps):
    """Best nontemporal streaming-read rate over x's bytes (GB/s); None when
    x is not a 16-byte-aligned buffer of a multiple of 16 bytes."""
    nbytes = x.numel() * x.element_size()
    if not x.is_contiguous() or x.data_ptr() % 16 or nbytes % 16:
        return None
    rates = [_probe_rate("hr_hbm_read_probe_nt", x, nbytes, b, max(warmup, 2), max(steps, 5))[0]
             for b in (4096, 8192, 16384)]
    log("nontemporal read of the trajectory buffer: " + ", ".join(f"{r:.0f}" for r in rates) + " GB/s")
    return max(rates)


def settle_clocks(step, stream, min_launches=40, window=10, tol=0.015, max_s=2.0):
    """Run `step` until its per-launch time has settled, before warmup.

    On MI355X an HBM-bound kernel launched after the GPU has idled runs at
    full rate for a launch or two, then slows by up to ~45 % while the
    device's clocks and power state adjust, and comes back to its steady
    rate over roughly the next 50 launches (~25 ms of work).  This phase
    keeps launching until two consecutive windows of `window` launches have
    medians within `tol` of each other (and at least `min_launches` ran),
    capped at `max_s` seconds.  Returns what it did, for the JSON line."""
    import numpy as np
    import torch

    meds, times = [], []
    t0 = time.perf_counter()
    while True:
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(window)]
        for s, e in ev:
            s.record(stream)
            step()
            e.record(stream)
        torch.cuda.synchronize()
        w = [s.elapsed_time(e) for s, e in ev]
        times.extend(w)
        meds.append(float(np.median(w)))
        done = (len(times) >= min_launches and len(meds) >= 2
                and abs(meds[-1] - meds[-2]) <= tol * meds[-2])
        if done or time.perf_counter() - t0 > max_s:
            break
    out = {"launches": len(times), "seconds": time.perf_counter() - t0,
           "settled": bool(done), "first_ms": float(times[0]), "peak_ms": float(max(times)),
           "last_window_median_ms": meds[-1],
           "rule": f"until two consecutive {window}-launch medians agree within {tol:.1%} "
                   f"(at least {min_launches} launches, at most {max_s:g} s), before warmup"}
    log(f"clock settle: {out['launches']} launches in {out['seconds'] * 1e3:.0f} ms "
        f"(first {out['first_ms']:.3f} ms, slowest {out['peak_ms']:.3f} ms, last window "
        f"median {out['last_window_median_ms']:.3f} ms)")
    return out


# --------------------------------------------------------------- one-vs-all
def cpu_baseline_one_vs_all(X, idx, budget_s=12.0, sample_frames=10_000):
    """Reference core (or the float32 restatement if the reference was not
    built) on a sample of the same trajectory, repeated for ~budget_s."""
    import numpy as np

    from oracle import native

    kind = "reference" if native.reference_available() else "port"
    impl = native.reference() if kind == "reference" else native.port()
    host = X[:sample_frames].cpu().numpy()
    ref_frame = host[0]
    reps, t_total, n_pairs = 0, 0.0, 0
    out = np.zeros(host.shape[0], np.float32)
    first = None
    while t_total < budget_s or reps < 2:
        t0 = time.perf_counter()
        target = np.asarray(host[:, idx, :], order="C", dtype=np.float32)  # _rmsd.pyx:197
        ref = np.asarray(ref_frame[idx, :], order="C", dtype=np.float32)   # _rmsd.pyx:199
        impl.rmsd_inplace(target, ref, out, True)
        t_total += time.perf_counter() - t0
        reps += 1
        n_pairs += host.shape[0]
        if first is None:
            first = out.copy()
    return {
        "value": n_pairs / t_total,
        "unit": "frame-pairs/s",
        "cores": cpu_threads(),
        "kind": kind,
        "sample": f"first {host.shape[0]} frames of the same trajectory, same 1,500-atom "
                  f"selection, {reps} repeats ({t_total:.1f} s); numpy gather + centering + "
                  "QCP per frame, OpenMP over frames",
        "small_case": small_case_cpu_and_gpu(impl),
    }, first


def small_case_cpu_and_gpu(impl, repeats=200):
    """The reference's own CPU-sized case: one reference vs 1,000 frames x 22
    atoms (seeded cube, noise 0.02 nm, random rotation and translation).
    Median frame-pairs/s of the CPU core and of md.rmsd through the GPU path
    (host arrays in and out, so launch and copy costs included)."""
    import numpy as np

    import mdtraj_amd as md
    from mdtraj_amd import gen

    x, _ = gen.small_cube(seed=0, n_frames=1000, n_atoms=22, sigma=0.02)
    out = np.zeros(x.shape[0], np.float32)
    cpu, gpu = [], []
    t = md.Trajectory(x)
    for _ in range(repeats):
        t0 = time.perf_counter()
        target = np.array(x, order="C", dtype=np.float32)
        impl.rmsd_inplace(target, np.array(x[0]), out, True)
        cpu.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        md.rmsd(t, t, 0)
        gpu.append(time.perf_counter() - t0)
    return {"workload": "1 reference vs 1,000 frames x 22 atoms (host arrays)",
            "cpu_frame_pairs_per_s": 1000 / float(np.median(cpu)),
            "gpu_md_rmsd_frame_pairs_per_s": 1000 / float(np.median(gpu)),
            "repeats": repeats}


def bench_one_vs_all(args):
    import numpy as np
    import torch

    from mdtraj_amd import gen
    from mdtraj_amd.plan import OneVsAll

    F, N, NSEL = args.frames, 2000, 1500
    log(f"generating walk trajectory {F} x {N} in HBM")
    X = gen.walk_device(seed=1, n_frames=F, n_atoms=N)
    idx = gen.subset(2, N, NSEL)
    torch.cuda.synchronize()
    plan = OneVsAll(X, idx)
    stream = torch.cuda.current_stream()

    def step():
        plan.set_reference_frame(0)
        plan.run()

    settle = None if args.no_settle else settle_clocks(step, stream)
    for _ in range(args.warmup):
        plan.set_reference_frame(0)
        plan.run()
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps)")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        plan.set_reference_frame(0)
        starts[k].record(stream)
        plan.run()
        ends[k].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per_launch = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    kern_ms = float(np.mean(per_launch))
    log(f"timed {args.steps} steps: {elapsed * 1e3 / args.steps:.3f} ms/step, kernel {kern_ms:.3f} ms "
        f"(min {per_launch.min():.3f}, median {np.median(per_launch):.3f}, first "
        f"{per_launch[0]:.3f}, last {per_launch[-1]:.3f})")

    streamed = 12.0 * F * N   # every line of every frame is read (75% random subset)
    selected = 12.0 * F * NSEL
    achieved = streamed / (kern_ms * 1e-3) / 1e9
    ceiling = read_ceiling(X, args.warmup, args.steps)
    rmsd_gpu = plan.out.cpu().numpy()
    result = {
        "metric": METRIC,
        "value": args.steps * F / elapsed,
        "unit": "frame-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random walk generated in HBM)",
        "config": {
            "workload": f"one-vs-all md.rmsd, {F} frames x {N} atoms float32, seeded random "
                        f"{NSEL}-atom selection, vs frame 0",
            "input_dtype": "f32",
            "accumulate_dtype": "f64",
            "clock_settle": settle,
            "kernel_ms": kern_ms,
            "kernel_ms_min": float(per_launch.min()),
            "kernel_ms_median": float(np.median(per_launch)),
            "kernel_ms_first": float(per_launch[0]),
            "kernel_ms_last": float(per_launch[-1]),
            "selected_GBps": selected / (kern_ms * 1e-3) / 1e9,
            # outside the timed region: nontemporal float4 streaming reads of
            # the same trajectory buffer, best of three grid sizes
            "read_ceiling_GBps": ceiling,
            "frac_of_read_ceiling": achieved / ceiling if ceiling else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profile("one_vs_all_traffic.json"),
        },
    }
    # outside the timed region: the FP64 oracle on a seeded sample of the
    # same frames, same selection, same reference frame
    from oracle import fp64

    n_check = min(2000, F)
    pick = np.sort(np.random.default_rng(7).choice(F, size=n_check, replace=False))
    sample = X[torch.as_tensor(pick, device=X.device)].cpu().numpy()
    ref0 = X[0:1].cpu().numpy()
    want = fp64.rmsd(sample, ref0, 0, idx)
    err = float(np.abs(rmsd_gpu[pick] - want).max())
    result["config"]["max_err_vs_fp64_nm"] = err
    result["config"]["fp64_check"] = (f"FP64 oracle on a seeded {n_check}-frame sample of the "
                                      f"{F} frames, same selection and reference frame")
    log(f"max |gpu - fp64| on {n_check} sampled frames: {err:.3e} nm")
    # outside the timed region: the same query through md.rmsd with the
    # trajectory in host memory (pinned staging chunks, selected atoms only)
    import mdtraj_amd as md

    host = md.Trajectory(X.cpu().numpy())
    md.rmsd(host, host, 0, atom_indices=idx)
    host_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        md.rmsd(host, host, 0, atom_indices=idx)
        host_ms.append((time.perf_counter() - t0) * 1e3)
    host_s = float(np.median(host_ms)) * 1e-3
    result["config"]["host_input_frame_pairs_per_s"] = F / host_s
    result["config"]["host_input_GBps"] = streamed / host_s / 1e9
    log(f"md.rmsd with host arrays (upload included): {host_s * 1e3:.1f} ms "
        f"({streamed / host_s / 1e9:.1f} GB/s of the trajectory's bytes)")
    del host
    if not args.no_cpu_baseline:
        log("cpu baseline (reference core on host cores)")
        cb, ref_out = cpu_baseline_one_vs_all(X, idx)
        n = ref_out.shape[0]
        cb["max_abs_diff_vs_gpu_nm"] = float(np.abs(ref_out - rmsd_gpu[:n]).max())
        result["cpu_baseline"] = cb
    if not args.no_all_pairs:
        log("all-pairs block (20,000 x 500)")
        del X, plan
        torch.cuda.empty_cache()
        result["config"]["all_pairs"] = all_pairs_block()
    if not args.no_batched:
        log("batched block (512 uneven systems, one graph launch)")
        torch.cuda.empty_cache()
        result["config"]["batched"] = batched_block()
    if args.serving_block:
        log("serving block (16 pools, Poisson 2,000 req/s)")
        torch.cuda.empty_cache()
        result["config"]["serving"] = serving_block(args.serve_block_seconds)
    return result


def batched_block():
    """The batched workload (512 uneven systems) at the library's default
    work split, as a block of the default line: its own line is
    `--workload batched`, which also runs the autotuner."""
    import types

    r = bench_batched(types.SimpleNamespace(host_gen=False, batch_shape="64,0,3,3", warmup=20,
                                            steps=50))
    c = r["config"]
    return {"workload": c["workload"], "schedule": c["schedule"],
            "frame_pairs_per_s": r["value"], "graph_ms": c["graph_ms"],
            "GBps": r["roofline"]["achieved"], "frac_of_hbm_peak": r["roofline"]["frac"],
            "traffic_bytes_per_launch": r["roofline"]["traffic"],
            "total_bytes": c["total_bytes"], "steps": 50, "warmup": 20}


def serving_block(seconds):
    """The serving loop (16 pools, Poisson 2,000 req/s) for a shorter
    window, as a block of the default line: its own line is
    `--workload serving`."""
    import types

    r = bench_serving(types.SimpleNamespace(serve_seconds=seconds))
    c = r["config"]
    return {"workload": c["workload"], "frame_pairs_per_s": r["value"],
            "latency_ms_p50": c["latency_ms_p50"], "latency_ms_p99": c["latency_ms_p99"],
            "latency_ms_p99_after_first_second": c["latency_ms_p99_after_first_second"],
            "latency_ms_mean": c["latency_ms_mean"], "requests": c["requests"],
            "achieved_req_per_s": c["achieved_req_per_s"], "batches": c["batches"]}


# ---------------------------------------------------------------- all-pairs
ALL_PAIRS_PRODUCTS = {"fp32": 1, "f16x3": 3, "bf16x3": 3, "f16exact": 4, "bf16exact": 4}


def time_all_pairs(X, mode_name, warmup, steps):
    """Per-launch milliseconds of hr_rmsd_all_pairs over X (F, N, 3) in one
    mode (HIP events around each launch on the stream it runs on), plus the
    condensed output of the last launch."""
    import numpy as np
    import torch

    from mdtraj_amd import _lib, _device as dev
    from mdtraj_amd.allpairs import MODES

    lib = _lib.load()
    F, N = int(X.shape[0]), int(X.shape[1])
    mode = MODES[mode_name]
    out = dev.empty_f32(F * (F - 1) // 2)
    nbytes = lib.hr_rmsd_all_pairs_workspace_bytes(F, N, mode)
    ws = dev.workspace(nbytes)
    stream = torch.cuda.current_stream()

    def step():
        _lib.call("hr_rmsd_all_pairs", X.data_ptr(), F, N, None, N, mode, out.data_ptr(),
                  ws.data_ptr(), nbytes, dev.stream())

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        step()
        e.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per = np.array([s.elapsed_time(e) for s, e in ev])
    del ws
    return per, elapsed, out


def settle_all_pairs(X, mode_name):
    """The clock-settle phase of settle_clocks for the all-pairs kernel
    (windows of 3 launches of ~7 ms)."""
    import torch

    from mdtraj_amd import _lib, _device as dev
    from mdtraj_amd.allpairs import MODES

    lib = _lib.load()
    F, N = int(X.shape[0]), int(X.shape[1])
    mode = MODES[mode_name]
    out = dev.empty_f32(F * (F - 1) // 2)
    nbytes = lib.hr_rmsd_all_pairs_workspace_bytes(F, N, mode)
    ws = dev.workspace(nbytes)

    def step():
        _lib.call("hr_rmsd_all_pairs", X.data_ptr(), F, N, None, N, mode, out.data_ptr(),
                  ws.data_ptr(), nbytes, dev.stream())

    return settle_clocks(step, torch.cuda.current_stream(), min_launches=9, window=3, max_s=2.0)


def all_pairs_figures(F, N, mode_name, ms):
    """Useful and executed TFLOP/s of one all-pairs launch of `ms` ms."""
    useful = 2.0 * (3 * F) ** 2 * N / 2   # upper triangle of the 3F x 3F Gram
    k_pad = (N + 63) // 64 * 64
    executed = ALL_PAIRS_PRODUCTS[mode_name] * useful * k_pad / N
    peak = FP32_PEAK_TFLOPS if mode_name == "fp32" else F16_PEAK_TFLOPS
    return {"useful_tflops": useful / (ms * 1e-3) / 1e12,
            "executed_tflops": executed / (ms * 1e-3) / 1e12,
            "peak_tflops": peak,
            "useful_frac": useful / (ms * 1e-3) / 1e12 / peak,
            "executed_frac": executed / (ms * 1e-3) / 1e12 / peak,
            "mfma_products": ALL_PAIRS_PRODUCTS[mode_name]}


def all_pairs_subset_errors(X, modes, n_sub=2000):
    """Max |GPU - FP64| (nm) of each mode on a seeded n_sub-frame subset of X,
    and the FP64 oracle's time on the host cores."""
    import numpy as np
    import torch

    from mdtraj_amd.allpairs import rmsd_matrix
    from oracle import fp64

    F = int(X.shape[0])
    sub = np.sort(np.random.default_rng(2).choice(F, size=min(n_sub, F), replace=False))
    Xs = X[torch.as_tensor(sub, device=X.device)].contiguous()
    gpu = {m: rmsd_matrix(Xs, mode=m).cpu().numpy().astype(np.float64) for m in modes}
    xs = Xs.cpu().numpy()
    t0 = time.perf_counter()
    want = fp64.condensed(fp64.rmsd_matrix(xs))
    cpu_s = time.perf_counter() - t0
    errs = {m: float(np.abs(v - want).max()) for m, v in gpu.items()}
    log(f"all-pairs subset {len(sub)} frames: FP64 oracle {cpu_s:.1f} s; max err vs FP64 {errs}")
    return errs, want.size, cpu_s, len(sub)


def all_pairs_block(F=20_000, N=500, mode_name="f16exact"):
    """The all-pairs workload (20,000 frames x 500 atoms, seeded random walk)
    measured beside the one-vs-all line: the split-precision kernel and the
    plain FP32 FMA kernel, each timed per launch, and both checked against
    the FP64 oracle on a seeded 2,000-frame subset."""
    import numpy as np
    import torch

    from mdtraj_amd import gen

    X = gen.walk_device(seed=2, n_frames=F, n_atoms=N)
    torch.cuda.synchronize()
    n_pairs = F * (F - 1) // 2
    settle = settle_all_pairs(X, mode_name)
    split_ms, _, out = time_all_pairs(X, mode_name, warmup=3, steps=5)
    del out
    fp32_ms, _, out = time_all_pairs(X, "fp32", warmup=1, steps=2)
    del out
    torch.cuda.empty_cache()
    split_med, fp32_med = float(np.median(split_ms)), float(np.median(fp32_ms))
    errs, n_sub_pairs, cpu_s, n_sub = all_pairs_subset_errors(X, [mode_name, "fp32"])
    fig = all_pairs_figures(F, N, mode_name, split_med)
    log(f"all-pairs {F} x {N}: {mode_name} {split_med:.3f} ms, fp32 {fp32_med:.3f} ms")
    return {
        "workload": f"all-pairs RMSD matrix, {F} frames x {N} atoms (seeded random walk), "
                    "condensed upper triangle",
        "mode": mode_name,
        "frame_pairs_per_s": n_pairs / (split_med * 1e-3),
        "kernel_ms": split_med,
        "kernel_ms_per_launch": [float(v) for v in split_ms],
        "executed_mfma_frac": fig["executed_frac"],
        "useful_mfma_frac": fig["useful_frac"],
        "executed_tflops": fig["executed_tflops"],
        "mfma_products": fig["mfma_products"],
        "fp32_kernel_ms": fp32_med,
        "fp32_frame_pairs_per_s": n_pairs / (fp32_med * 1e-3),
        "speedup_vs_fp32": fp32_med / split_med,
        "max_err_vs_fp64_nm": errs[mode_name],
        "fp32_max_err_vs_fp64_nm": errs["fp32"],
        "fp64_subset": f"seeded {n_sub}-frame subset, {n_sub_pairs} pairs, FP64 oracle {cpu_s:.1f} s",
        "clock_settle": settle,
    }


def bench_all_pairs(args):
    import numpy as np
    import torch

    from mdtraj_amd import gen

    F, N = args.frames, 500
    log(f"generating walk trajectory {F} x {N} in HBM")
    X = gen.walk_device(seed=2, n_frames=F, n_atoms=N)
    torch.cuda.synchronize()
    n_pairs = F * (F - 1) // 2
    settle = None if args.no_settle else settle_all_pairs(X, args.mode)
    per, elapsed, out = time_all_pairs(X, args.mode, args.warmup, args.steps)
    del out
    ms = float(np.mean(per))
    fig = all_pairs_figures(F, N, args.mode, ms)
    log(f"all-pairs {args.mode}: {ms:.3f} ms, useful {fig['useful_tflops']:.1f} TFLOP/s, executed "
        f"{fig['executed_tflops']:.1f} TFLOP/s")
    result = {
        "metric": METRIC,
        "value": args.steps * n_pairs / elapsed,
        "unit": "frame-pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp32": "f32", "f16x3": "f16x3", "bf16x3": "bf16x3", "f16exact": "f16x4-exact-hi",
                  "bf16exact": "bf16x4-exact-hi"}[args.mode],
        "data": "synthetic (seeded random walk generated in HBM)",
        "config": {"workload": f"all-pairs RMSD matrix, {F} frames x {N} atoms, mode {args.mode}",
                   "clock_settle": settle,
                   "kernel_ms": ms, "kernel_ms_per_launch": [float(v) for v in per],
                   "mfma_products": fig["mfma_products"],
                   "executed_tflops": fig["executed_tflops"],
                   "executed_frac": fig["executed_frac"]},
        "roofline": {"bound": "mfma", "achieved": fig["useful_tflops"], "peak": fig["peak_tflops"],
                     "unit": "TFLOP/s", "frac": fig["useful_frac"], "traffic": None},
    }
    if not args.no_cpu_baseline:
        errs, n_sub_pairs, cpu_s, n_sub = all_pairs_subset_errors(X, sorted({args.mode, "fp32"}))
        result["cpu_baseline"] = {
            "value": n_sub_pairs / cpu_s, "unit": "frame-pairs/s", "cores": cpu_threads(),
            "kind": "port",
            "sample": f"FP64 oracle (numpy) on a seeded {n_sub}-frame subset, {n_sub_pairs} pairs",
            "max_err_vs_fp64_nm": errs}
    return result


# ------------------------------------------------------------------ batched
def bench_batched(args):
    """512 independent systems, atoms log-uniform in [50, 5000], frames
    log-uniform in [10, 2000], each scored against its own frame 0, as one
    hipGraph launch whose tile shape the autotuner picks."""
    import torch

    from mdtraj_amd import autotune, gen
    from mdtraj_amd.batch import BatchRMSD

    if args.host_gen:
        # numpy generation and one upload per system: no device RNG kernels
        # in the process (for counter-collection runs)
        log("generating 512 uneven systems on the host, then uploading")
        xs = [torch.from_numpy(x).to("cuda") for x in gen.uneven_systems(seed=3)]
    else:
        log("generating 512 uneven systems in HBM")
        xs = gen.uneven_systems_device(seed=3)
    torch.cuda.synchronize()
    if args.batch_shape:
        best = tuple(int(v) for v in args.batch_shape.split(","))
        if len(best) == 3:
            best = best + (1,)  # a unit shape: the ticket schedule
        log(f"shape given: fpu={best[0]} kib={best[1]} waves={best[2]} schedule={best[3]} "
            "(autotuner skipped)")
    else:
        best, table = autotune.tune_batch(xs, reps=3)
        log("autotuner: " + ", ".join(f"fpu={k[0]} kib={k[1]} waves={k[2]} schedule={k[3]}: "
                                      f"{v:.3f} ms" for k, v in table.items()))
    b = BatchRMSD(xs, frames_per_unit=best[0], unit_kib=best[1], waves=best[2],
                  schedule=best[3])
    # the same clock-settle phase as the one-vs-all line, before the warmup
    settle = settle_clocks(b.run, torch.cuda.current_stream())
    for _ in range(args.warmup):
        b.run()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(args.steps):
        b.run()
    e.record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = s.elapsed_time(e) / args.steps
    gbs = b.total_bytes / (ms * 1e-3) / 1e9
    return {
        "metric": METRIC,
        "value": args.steps * b.total_frames / elapsed, "unit": "frame-pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded systems generated in HBM)",
        "config": {"workload": "512 systems, atoms log-uniform [50, 5000], frames log-uniform "
                               "[10, 2000], sigma 0.03 nm, vs own frame 0",
                   "total_frames": b.total_frames, "total_bytes": b.total_bytes,
                   "units": b.n_units, "frames_per_unit": best[0], "unit_kib": best[1], "waves_per_simd": best[2], "schedule": best[3],
                   "graph_ms": ms, "clock_settle": settle},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile("batched_traffic.json", list(best))},
    }


# ------------------------------------------------------------------ serving
def bench_serving(args):
    """16 device-resident pools (frames log-uniform [1k, 200k], atoms
    log-uniform [100, 3000]); Poisson arrivals at 2,000 req/s, 1-64 query
    frames each (noisy copies of pool frames), one-vs-pool RMSD; HBM budget
    200 GB; continuous batching.  Latency = arrival -> rows on the host."""
    import numpy as np
    import torch

    from mdtraj_amd import gen
    from mdtraj_amd.serve import RMSDServer

    rng = np.random.default_rng(4)
    n_pools, rate, duration = 16, 2000.0, args.serve_seconds
    nf = gen.log_uniform_int(rng, 1000, 200_000, n_pools)
    na = gen.log_uniform_int(rng, 100, 3000, n_pools)
    g = torch.Generator(device=torch.device("cuda", 0))
    g.manual_seed(4)
    # the client consumes rows as they land in pinned host memory (latency
    # endpoint); no per-request copy-out
    srv = RMSDServer(hbm_budget=200e9, keep_results=False)
    samples = []
    for f, n in zip(nf, na):
        x = gen.cube_device(g, int(f), int(n), sigma=0.05, box=3.0)
        srv.add_pool(x)
        pick = torch.randint(0, int(f), (64,), generator=g, device="cuda")
        q = x[pick] + 0.01 * torch.randn((64, int(n), 3), generator=g, device="cuda")
        samples.append(q.cpu().numpy())
        del x
    torch.cuda.synchronize()
    log(f"pools ready: {srv.resident / 1e9:.2f} GB resident")
    n_req = int(rate * duration)
    gaps = rng.exponential(1.0 / rate, n_req)
    pools = rng.integers(0, n_pools, n_req)
    qs = rng.integers(1, 65, n_req)
    srv.start()
    # warm-up, discarded: one request per pool, then one second of arrivals at
    # the same rate from a separate stream of draws (the first second of a
    # cold loop has a longer tail: p99 8.2 ms over a 2 s window, 5.8 ms after
    # its first second)
    for p in range(n_pools):
        srv.submit(p, samples[p][:1]).wait(120)
    wrng = np.random.default_rng(5)
    warm = []
    t_w = time.perf_counter()
    t_next = t_w
    while t_next - t_w < 1.0:
        t_next += wrng.exponential(1.0 / rate)
        while True:
            now = time.perf_counter()
            if now >= t_next:
                break
            time.sleep(min(t_next - now, 0.0005))
        p = int(wrng.integers(0, n_pools))
        warm.append(srv.submit(p, samples[p][: int(wrng.integers(1, 65))], t_arrival=t_next))
    for r in warm:
        r.wait(600)
    srv.completed.clear()
    srv.pairs_done = 0
    batches0 = srv.batches
    t0 = time.perf_counter()
    t_next = t0
    reqs = []
    t_log = t0
    for k in range(n_req):
        t_next += gaps[k]
        while True:
            now = time.perf_counter()
            if now >= t_next:
                break
            time.sleep(min(t_next - now, 0.0005))
        reqs.append(srv.submit(int(pools[k]), samples[pools[k]][: qs[k]], t_arrival=t_next))
        if now - t_log >= 1.0:
            t_log = now
            log(f"serving: {k + 1} requests submitted, {len(srv.completed)} answered, "
                f"{srv.batches} batches")
    for i, r in enumerate(reqs):
        r.wait(600)
        if time.perf_counter() - t_log >= 1.0:
            t_log = time.perf_counter()
            log(f"serving: draining, {i + 1} of {n_req} answered")
    t1 = max(r.t_done for r in reqs)
    srv.stop()
    lat = np.array([r.latency for r in reqs]) * 1e3
    pairs = int(sum(int(q) * int(nf[p]) for q, p in zip(qs, pools)))
    return {
        "metric": METRIC,
        "value": pairs / (t1 - t0), "unit": "frame-pairs/s", "n_gpus": 1,
        "steps": n_req, "warmup": n_pools + len(warm), "ms_per_step": (t1 - t0) * 1e3 / n_req,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f16x4-exact-hi",
        "data": "synthetic (pools generated in HBM, Poisson arrivals)",
        "config": {"workload": f"{n_pools} pools, frames log-uniform [1k, 200k], atoms "
                               f"log-uniform [100, 3000], Poisson {rate:.0f} req/s for "
                               f"{duration:.1f} s, 1-64 query frames per request, HBM budget 200 GB",
                   "requests": n_req, "batches": srv.batches - batches0,
                   "warmup": "one request per pool, then 1 s of arrivals at the same rate",
                   "largest_batch_queries": srv.largest_batch,
                   "resident_GB": srv.resident / 1e9,
                   "latency_ms_p50": float(np.percentile(lat, 50)),
                   "latency_ms_p99": float(np.percentile(lat, 99)),
                   "latency_ms_p99_after_first_second": float(np.percentile(
                       lat[np.array([r.t_arrival for r in reqs]) - t0 >= 1.0], 99))
                   if duration > 1.5 else None,
                   "latency_ms_mean": float(lat.mean()),
                   "offered_req_per_s": rate, "achieved_req_per_s": n_req / (t1 - t0)},
    }


# ------------------------------------------------------------- HBM ceiling
def bench_hbm_ceiling(args):
    """Float4 streaming reads (plain and nontemporal loads) of a 2.4 GB
    buffer (the one-vs-all trajectory's size): the read bandwidth the device
    delivers in practice."""
    import torch

    nbytes = 12 * 100_000 * 2000
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    best = None
    for entry, kind in (("hr_hbm_read_probe", "plain"), ("hr_hbm_read_probe_nt", "nontemporal")):
        for blocks in (1024, 2048, 4096, 8192, 16384):
            gbs, ms = _probe_rate(entry, x, nbytes, blocks, args.warmup, args.steps)
            log(f"stream read ({kind}) {blocks} blocks: {ms:.3f} ms, {gbs:.0f} GB/s")
            if best is None or gbs > best[1]:
                best = (blocks, gbs, ms, kind)
    # a plain device copy of the same 2.4 GB: the known-size kernel that the
    # rocprofv3 FETCH_SIZE / WRITE_SIZE counters are scaled by
    y = torch.empty_like(x)
    y.copy_(x)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(args.steps):
        y.copy_(x)
    e.record()
    torch.cuda.synchronize()
    copy_ms = s.elapsed_time(e) / args.steps
    log(f"device copy of {nbytes} B: {copy_ms:.3f} ms, {2 * nbytes / (copy_ms * 1e-3) / 1e9:.0f} GB/s "
        "(read + write)")
    return {"metric": "HBM streaming-read ceiling", "value": best[1], "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": best[2],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "float4 streaming read of 2.4 GB",
                                             "blocks": best[0], "loads": best[3],
                                             "copy_bytes": nbytes,
                                             "copy_ms": copy_ms},
            "roofline": {"bound": "hbm", "achieved": best[1], "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": best[1] / HBM_PEAK_GBS, "traffic": None}}


def bench_xtc_decode(args):
    """XTC ingestion: decode every frame of a file resident in HBM into
    float32 (F, N, 3).  Input: the 30-frame, 600-atom water fixture
    (tests/golden/xtc_water.xtc, written by the reference's encoder)
    repeated to 90,000 frames; frames are self-contained, so the repeated
    file is a valid trajectory.  Timed: hr_xtc_decode alone (HIP events),
    bytes already in HBM."""
    import tempfile

    import numpy as np
    import torch

    from mdtraj_amd import _lib, _device as dev
    from mdtraj_amd.xtc import XTCTrajectoryFile

    src = open(os.path.join(ROOT, "tests", "golden", "xtc_water.xtc"), "rb").read()
    reps = 3000
    fd, fn = tempfile.mkstemp(suffix=".xtc")
    with os.fdopen(fd, "wb") as fh:
        for _ in range(reps):
            fh.write(src)
    try:
        f = XTCTrajectoryFile(fn)
        F, N = len(f), f.n_atoms
        d = dev.device()
        xyz = torch.empty((F, N, 3), dtype=torch.float32, device=d)
        tm = torch.empty(F, device=d)
        st = torch.empty(F, dtype=torch.int32, device=d)
        bx = torch.empty((F, 9), device=d)
        pr = torch.empty(F, device=d)
        status = torch.empty(F, dtype=torch.int32, device=d)
        data = f._device_bytes()
        offs = torch.from_numpy(f._offsets).to(d)

        def run():
            _lib.call("hr_xtc_decode", data.data_ptr(), data.numel(), offs.data_ptr(), F, N,
                      xyz.data_ptr(), tm.data_ptr(), st.data_ptr(), bx.data_ptr(),
                      pr.data_ptr(), status.data_ptr(), dev.stream())

        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(args.steps):
            run()
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / args.steps
        assert int(status.max().item()) == 0, "decode reported a malformed frame"
        comp = data.numel()
        out_bytes = 12 * F * N
        achieved = (comp + out_bytes) / (ms * 1e-3) / 1e9
        log(f"xtc decode {F} frames x {N} atoms: {ms:.3f} ms, {F / ms * 1e3:.3e} frames/s, "
            f"{out_bytes / (ms * 1e-3) / 1e9:.0f} GB/s decoded")
        res = {"metric": "XTC frames decoded/s (bit-identical to the reference reader)",
               "value": F / (ms * 1e-3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "u32",
               "data": "the reference-encoded 600-atom water fixture repeated",
               "config": {"workload": f"XTC decode, {F} frames x {N} atoms, bytes in HBM",
                          "compressed_bytes": comp, "decoded_bytes": out_bytes,
                          "decoded_GBps": out_bytes / (ms * 1e-3) / 1e9},
               "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                            "traffic": traffic_from_profile("xtc_decode_traffic.json", [F, N])}}
        if not args.no_cpu_baseline:
            res["cpu_baseline"] = cpu_baseline_xtc(f, xyz, fn)
        return res
    finally:
        os.unlink(fn)


def cpu_baseline_xtc(f, xyz_dev, fn, sample_frames=20_000):
    """The reference's own C reader (oracle/_ref/libref_xtc.so, compiled
    from xdrfile.c / xdrfile_xtc.c) on the first frames, one host thread,
    reading the sample from a file as the reference does; its output must
    equal the GPU decode bit for bit."""
    import ctypes
    import tempfile

    import numpy as np

    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_xtc.so"))
    n = min(len(f), sample_frames)
    end = int(f._offsets[n]) if n < len(f) else f._host.size
    fd, sfn = tempfile.mkstemp(suffix=".xtc")
    with os.fdopen(fd, "wb") as fh:
        fh.write(f._host[:end].tobytes())
    N = f.n_atoms
    bufs = [np.zeros((n + 1) * k, np.float32) for k in (3 * N, 1, 1, 9, 1)]
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    t0 = time.perf_counter()
    got = lib.refxtc_read(sfn.encode(), N, n + 1, P(bufs[0]), P(bufs[1]), P(bufs[2]),
                          P(bufs[3]), P(bufs[4]))
    dt = time.perf_counter() - t0
    os.unlink(sfn)
    same = bool(np.array_equal(bufs[0][: 3 * N * got].reshape(got, N, 3),
                               xyz_dev[:got].cpu().numpy()))
    return {"value": got / dt, "unit": "frames/s", "cores": 1, "kind": "reference",
            "sample": f"first {got} frames, read from a file by the reference's read_xtc loop",
            "gpu_equals_reference": same}


def _time_events(fn, steps, warmup):
    """Mean milliseconds of fn() over `steps` calls on torch's current stream
    (HIP events), after `warmup` untimed calls."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def bench_nearest_leader(args):
    """Leader assignment (examples/two-pass-clustering.ipynb, cell 6): the
    nearest of L leaders for each of F frames, the minimum taken inside the
    MFMA cross kernel (hr_rmsd_cross_argmin, f16exact).  Workload: the
    random-walk trajectory, 100,000 frames x 500 atoms, leaders every 400th
    frame (250).  Timed: the fused kernel with both sets prepared in HBM;
    the unfused route (hr_rmsd_cross into an F x L matrix, then torch's
    argmin) is timed beside it."""
    import torch

    from mdtraj_amd import _lib, _device as dev, allpairs, gen

    F, N, step_l = args.frames or 100_000, 500, 400
    X = gen.walk_device(seed=3, n_frames=F, n_atoms=N)
    a = allpairs.SplitSet(X, mode=args.mode)
    b = allpairs.SplitSet(X[::step_l].contiguous(), mode=args.mode)
    L = b.n_frames
    lib = _lib.load()
    mode = allpairs.MODES[args.mode]
    idx = torch.empty(F, dtype=torch.int32, device=X.device)
    val = torch.empty(F, dtype=torch.float32, device=X.device)
    ws = dev.workspace(lib.hr_rmsd_cross_argmin_workspace_bytes(F))
    full = torch.empty((F, L), dtype=torch.float32, device=X.device)

    def fused():
        _lib.call("hr_rmsd_cross_argmin", a.block.data_ptr(), F, b.block.data_ptr(), L, N, mode,
                  idx.data_ptr(), val.data_ptr(), ws.data_ptr(), dev.stream())

    def unfused():
        _lib.call("hr_rmsd_cross", a.block.data_ptr(), F, b.block.data_ptr(), L, N, mode,
                  full.data_ptr(), full.stride(0), dev.stream())
        torch.argmin(full, dim=1)

    K = 10
    tk_idx = torch.empty((F, K), dtype=torch.int32, device=X.device)
    tk_val = torch.empty((F, K), dtype=torch.float32, device=X.device)

    def nearest10():  # each frame's 10 nearest leaders: cross matrix, then hr_topk_rows
        _lib.call("hr_rmsd_cross", a.block.data_ptr(), F, b.block.data_ptr(), L, N, mode,
                  full.data_ptr(), full.stride(0), dev.stream())
        _lib.call("hr_topk_rows", full.data_ptr(), F, L, full.stride(0), K, tk_idx.data_ptr(),
                  tk_val.data_ptr(), dev.stream())

    def topk_only():
        _lib.call("hr_topk_rows", full.data_ptr(), F, L, full.stride(0), K, tk_idx.data_ptr(),
                  tk_val.data_ptr(), dev.stream())

    ms = _time_events(fused, args.steps, args.warmup)
    ms_unfused = _time_events(unfused, args.steps, args.warmup)
    ms_k10 = _time_events(nearest10, args.steps, args.warmup)
    ms_topk = _time_events(topk_only, args.steps, args.warmup)
    unfused()
    assert torch.equal(idx.long(), torch.argmin(full, dim=1)), "fused argmin differs"
    topk_only()
    assert torch.equal(tk_idx[:, 0].long(), idx.long()), "top-10 nearest differs from the argmin"
    useful = 18.0 * N * F * L  # the nine-entry correlation of every pair
    tflops = useful / (ms * 1e-3) / 1e12
    log(f"nearest leader {F} x {L} leaders x {N} atoms: fused {ms:.3f} ms, unfused "
        f"{ms_unfused:.3f} ms, {tflops:.1f} TFLOP/s useful; 10 nearest {ms_k10:.3f} ms "
        f"(top-k pass {ms_topk:.3f} ms)")
    return {"metric": "frames assigned to their nearest leader per second", "value": F / (ms * 1e-3),
            "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": args.mode, "data": "synthetic (seeded random walk generated in HBM)",
            "config": {"workload": f"nearest of {L} leaders for {F} frames x {N} atoms, mode {args.mode}",
                       "unfused_ms": ms_unfused, "speedup_vs_unfused": ms_unfused / ms,
                       "nearest10_ms": ms_k10, "topk10_pass_ms": ms_topk},
            "roofline": {"bound": "mfma", "achieved": tflops, "peak": F16_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": tflops / F16_PEAK_TFLOPS, "traffic": None}}


def bench_centroid(args):
    """Centroid scoring (examples/centroids.ipynb, cells 3-4) on the condensed
    all-pairs matrix of the random-walk trajectory, 20,000 frames x 500
    atoms (computed before the timed region).  Timed: hr_centroid_scores
    (FP64 std, then one pass of row and column sums of exp(-beta d / std))."""
    import torch

    from mdtraj_amd import _lib, _device as dev, allpairs, gen

    F, N = args.frames or 20_000, 500
    X = gen.walk_device(seed=4, n_frames=F, n_atoms=N)
    D = allpairs.rmsd_matrix(X)
    lib = _lib.load()
    scores = torch.empty(F, dtype=torch.float64, device=X.device)
    ws = dev.workspace(lib.hr_centroid_workspace_bytes(F))

    def run():
        _lib.call("hr_centroid_scores", D.data_ptr(), F, 1.0, scores.data_ptr(), None,
                  ws.data_ptr(), dev.stream())

    ms = _time_events(run, args.steps, args.warmup)
    n_pairs = F * (F - 1) // 2
    algo = 2 * 4 * n_pairs  # the stats pass and the tile pass each read the matrix once
    achieved = algo / (ms * 1e-3) / 1e9
    log(f"centroid {F} frames: {ms:.3f} ms, {achieved:.0f} GB/s")
    return {"metric": "all-pairs matrix entries scored per second (centroid)",
            "value": n_pairs / (ms * 1e-3), "unit": "entries/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded random walk generated in HBM)",
            "config": {"workload": f"centroid scores over the {F}-frame condensed matrix",
                       "centroid": int(torch.argmax(scores).item())},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": traffic_from_profile("centroid_traffic.json", [F])}}


def bench_lprmsd(args):
    """LP-RMSD (md.lprmsd) of a solvated system: 100 distinguishable solute
    atoms and 300 water oxygens whose labels may permute, 20,000 frames of
    400 atoms (the oxygens drift and swap places between frames).  Timed:
    hr_lprmsd (one wave per frame: fit on the solute, exact assignment of the
    oxygens, QCP), inputs in HBM."""
    import numpy as np
    import torch

    from mdtraj_amd import _lib, _device as dev

    F, n_sol, n_wat = args.frames or 20_000, 100, 300
    N = n_sol + n_wat
    rng = np.random.default_rng(17)
    base = np.concatenate([rng.normal(0, 0.6, (n_sol, 3)), rng.uniform(-1.5, 1.5, (n_wat, 3))])
    d = dev.device()
    g = torch.Generator(device=d)
    g.manual_seed(17)
    B = torch.from_numpy(base).to(d)
    X = (B[None] + 0.02 * torch.randn((F, N, 3), generator=g, device=d, dtype=torch.float64))
    perm = torch.argsort(torch.rand((F, n_wat), generator=g, device=d), dim=1) + n_sol
    X[:, n_sol:] = torch.gather(X, 1, perm[:, :, None].expand(F, n_wat, 3))
    X = X.float().contiguous()
    ref = X[0].clone()
    lib = _lib.load()
    idx = torch.arange(N, dtype=torch.int32, device=d)
    dis = torch.arange(n_sol, dtype=torch.int32, device=d)
    grp = torch.arange(n_sol, N, dtype=torch.int32, device=d)
    off = torch.tensor([0, n_wat], dtype=torch.int32, device=d)
    out = torch.empty(F, dtype=torch.float32, device=d)
    ws = dev.workspace(lib.hr_lprmsd_workspace_bytes(N, n_sol))

    def run():
        _lib.call("hr_lprmsd", X.data_ptr(), F, N, ref.data_ptr(), idx.data_ptr(), N,
                  dis.data_ptr(), n_sol, grp.data_ptr(), off.data_ptr(), 1, n_wat, 0,
                  out.data_ptr(), ws.data_ptr(), dev.stream())

    ms = _time_events(run, args.steps, args.warmup)
    bytes_in = 12.0 * N * F
    achieved = bytes_in / (ms * 1e-3) / 1e9
    log(f"lprmsd {F} frames x {N} atoms ({n_wat} permutable): {ms:.3f} ms, "
        f"{F / ms * 1e3:.3e} frames/s")
    res = {"metric": "LP-RMSD frames/s (md.lprmsd)", "value": F / (ms * 1e-3), "unit": "frames/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (solute with noise, water oxygens relabelled per frame)",
           "config": {"workload": f"LP-RMSD, {F} frames x {N} atoms, one permute group of {n_wat}"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None}}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_lprmsd(X, ref, out, n_sol, N)
    return res


def cpu_baseline_lprmsd(X, ref, out_dev, n_sol, N, sample=20):
    """The reference's own LP-RMSD loop (oracle/_ref/libref_rmsd.so: its
    Munkres, euclidean_permutation and QCP compiled from source, the serial
    frame loop of _lprmsd.pyx) on the first frames, one host thread."""
    import ctypes

    import numpy as np

    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_rmsd.so"))
    x = np.ascontiguousarray(X[:sample].cpu().numpy())
    r = np.ascontiguousarray(ref.cpu().numpy())
    idx = np.arange(N, dtype=np.int32)
    grp = np.arange(n_sol, N, dtype=np.int32)
    off = np.array([0, N - n_sol], np.int32)
    o = np.zeros(sample, np.float32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    reps, t0 = 0, time.perf_counter()
    while True:
        lib.ref_lprmsd(P(x), sample, N, P(r), P(idx), N, P(grp), P(off), 1, 0, P(o))
        reps += 1
        if time.perf_counter() - t0 > 10.0:
            break
    dt = (time.perf_counter() - t0) / reps
    gpu = out_dev[:sample].cpu().numpy()
    return {"value": sample / dt, "unit": "frames/s", "cores": 1, "kind": "reference",
            "sample": f"first {sample} frames, {reps} passes",
            "max_abs_diff_vs_gpu": float(np.abs(gpu - o).max())}


def bench_superpose(args):
    """Trajectory.superpose of the one-vs-all workload: 100,000 frames x
    2,000 atoms in HBM, fitted on the 1,500-atom subset to frame 0 and every
    atom moved in place (hr_superpose, translate=1: the fit kernel, then the
    rotation kernel, per chunk of 1 GiB).  Each timed call moves the frames
    again (a superposed trajectory superposed once more), so the work per
    call is the same."""
    import ctypes

    import torch

    from mdtraj_amd import _lib, _device as dev, gen

    F, N, NSEL = args.frames or 100_000, 2000, 1500
    X = gen.walk_device(seed=1, n_frames=F, n_atoms=N)
    idx = torch.from_numpy(gen.subset(2, N, NSEL).astype("int32")).to(X.device)
    R = X[0].clone()
    lib = _lib.load()
    ws = dev.workspace(lib.hr_superpose_workspace_bytes(F, NSEL))

    def run():
        _lib.call("hr_superpose", X.data_ptr(), F, N, None, N, idx.data_ptr(), NSEL, R.data_ptr(),
                  idx.data_ptr(), 1, None, None, ws.data_ptr(), dev.stream())

    ms = _time_events(run, args.steps, args.warmup)
    # fit pass: the frames' lines (12 B per atom, every line of the 75 % subset);
    # rotation pass: every atom read and written
    algo = 12.0 * F * N + 2 * 12.0 * F * N
    achieved = algo / (ms * 1e-3) / 1e9
    log(f"superpose {F} x {N} (fit on {NSEL}): {ms:.3f} ms per call, {achieved:.0f} GB/s")
    return {"metric": "frames superposed per second (Trajectory.superpose)",
            "value": F / (ms * 1e-3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded random walk generated in HBM)",
            "config": {"workload": f"superpose {F} frames x {N} atoms on a {NSEL}-atom subset, in place"},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": traffic_from_profile("superpose_traffic.json", [F, N])}}


def bench_dcd_decode(args):
    """DCD ingestion: the axis-major X/Y/Z records of every frame of a file
    resident in HBM written as float32 (F, N, 3).  Input: a CHARMM-format
    file (unit cell in every frame) of the random-walk trajectory of the
    one-vs-all workload, 100,000 frames x 2,000 atoms by default, in the
    layout the reference's writer produces (gen.dcd_image).  Timed:
    hr_dcd_decode alone (HIP events), bytes already in HBM."""
    import numpy as np
    import torch

    from mdtraj_amd import _lib, _device as dev, gen

    F, N = args.frames or 100_000, 2000
    src = gen.walk_device(seed=11, n_frames=F, n_atoms=N)
    img = gen.dcd_image(src.cpu().numpy(), lengths=(60.0, 61.0, 62.0))
    hdr = _lib.DcdHeader()
    import ctypes

    _lib.call("hr_dcd_scan", img.ctypes.data_as(ctypes.c_void_p), img.size, ctypes.byref(hdr),
              None, 0)
    d = dev.device()
    data = torch.from_numpy(img).to(d)
    sample = img[: hdr.first_frame + min(F, 10_000) * hdr.frame_bytes]
    del img
    frames = torch.arange(F, dtype=torch.int64, device=d)
    xyz = torch.empty((F, N, 3), dtype=torch.float32, device=d)
    cell = torch.empty((F, 6), dtype=torch.float32, device=d)
    status = torch.empty(F, dtype=torch.int32, device=d)

    def run():
        _lib.call("hr_dcd_decode", data.data_ptr(), data.numel(), ctypes.byref(hdr),
                  frames.data_ptr(), F, None, N, None, 1.0, xyz.data_ptr(), cell.data_ptr(),
                  status.data_ptr(), dev.stream())

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(args.steps):
        run()
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / args.steps
    assert int(status.max().item()) == 0, "decode reported a malformed frame"
    assert torch.equal(xyz, src), "decoded coordinates differ from the ones written"
    # per frame: 12 N coordinate bytes, 24 marker bytes and the 56-byte cell
    # record read; 12 N coordinate bytes, the 24-byte cell and a status word written
    per_frame = (12 * N + 24 + 56) + (12 * N + 24 + 4)
    achieved = F * per_frame / (ms * 1e-3) / 1e9
    log(f"dcd decode {F} frames x {N} atoms: {ms:.3f} ms, {F / ms * 1e3:.3e} frames/s, "
        f"{achieved:.0f} GB/s")
    res = {"metric": "DCD frames decoded/s (bit-identical to the reference reader)",
           "value": F / (ms * 1e-3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic random walk written in the reference writer's layout",
           "config": {"workload": f"DCD decode, {F} frames x {N} atoms with a unit cell, bytes in HBM",
                      "file_bytes": int(data.numel()), "bytes_per_frame": per_frame},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": traffic_from_profile("dcd_decode_traffic.json", [F, N])}}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_dcd(sample, xyz, N)
    return res


def cpu_baseline_dcd(sample, xyz_dev, N):
    """The reference's own C reader (oracle/_ref/libref_dcd.so, compiled
    from dcdplugin.c) reading the first frames from a file, one host thread,
    as DCDTrajectoryFile.read drives it; its output must equal the GPU
    decode bit for bit."""
    import ctypes
    import tempfile

    import numpy as np

    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_dcd.so"))
    fd, fn = tempfile.mkstemp(suffix=".dcd")
    with os.fdopen(fd, "wb") as fh:
        fh.write(sample.tobytes())
    na, ns = ctypes.c_int(), ctypes.c_int()
    # the reference library reports on stdout; keep the JSON line alone there
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        lib.refdcd_info(fn.encode(), ctypes.byref(na), ctypes.byref(ns))
        n = ns.value
        out = np.zeros((n, N, 3), np.float32)
        cell = np.zeros((n, 6), np.float32)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        reps, t0 = 0, time.perf_counter()
        while True:
            got = lib.refdcd_read(fn.encode(), 0, 1, n, P(out), P(cell))
            reps += 1
            if time.perf_counter() - t0 > 10.0:
                break
        dt = (time.perf_counter() - t0) / reps
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
        os.unlink(fn)
    same = bool(np.array_equal(out[:got], xyz_dev[:got].cpu().numpy()))
    return {"value": got / dt, "unit": "frames/s", "cores": 1, "kind": "reference",
            "sample": f"first {got} frames, read from a file (page cache) by the reference's "
                      f"read_next_timestep loop, {reps} passes",
            "gpu_equals_reference": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults sized per workload below (None): enough warmup that the GPU
    # reaches steady clocks before the timed region
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="one-vs-all",
                    choices=["one-vs-all", "all-pairs", "hbm-ceiling", "batched", "serving",
                             "xtc-decode", "dcd-decode", "nearest-leader", "centroid", "lprmsd", "superpose"])
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--mode", default="f16exact",
                    choices=["fp32", "f16x3", "bf16x3", "f16exact", "bf16exact"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-settle", action="store_true",
                    help="skip the clock-settle phase before warmup (one-vs-all, all-pairs)")
    ap.add_argument("--no-all-pairs", action="store_true",
                    help="one-vs-all line without its all-pairs block")
    ap.add_argument("--no-batched", action="store_true",
                    help="one-vs-all line without its batched block")
    ap.add_argument("--serving-block", action="store_true",
                    help="one-vs-all line with a serving block.  Off by default: run in the "
                         "same process after the other blocks, the loop's latencies are "
                         "higher (p50 1.2 ms, p99 8.2 ms against 0.9 and 6.1 ms in a process "
                         "of its own, `--workload serving`)")
    ap.add_argument("--serve-block-seconds", type=float, default=2.0,
                    help="arrival window of the serving block of the one-vs-all line")
    ap.add_argument("--batch-shape", default=None,
                    help="batched workload: frames_per_unit,unit_kib,waves_per_simd instead of "
                         "autotuning (e.g. 64,192,3), so that a profile holds only that shape")
    ap.add_argument("--host-gen", action="store_true",
                    help="batched workload: generate the systems with numpy on the host")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--frames-waves", type=int, default=0, choices=[0, 3, 4],
                    help="register budget of the per-frame kernel (0 = library default)")
    ap.add_argument("--frames-grid", type=int, default=0, choices=[0, 1, 2],
                    help="grid of the per-frame kernel: 1 equal quad counts, 2 one wave "
                         "per resident slot (0 = library default, 2)")
    ap.add_argument("--pairs-tile", type=int, default=0, choices=[0, 1, 2, 3, 4, 5],
                    help="split-precision Gram kernels: 1 64x64 tiles, 2-5 64x32 half tiles "
                         "(see HR_TUNE_PAIRS_TILE; 0 = library default, 5)")
    ap.add_argument("--frames-row-lanes", type=int, default=0, choices=[0, 4, 8, 16, 32, 64],
                    help="lanes per frame of the per-frame kernel (0 = library default)")
    args = ap.parse_args()
    per_workload = {"one-vs-all": (200, 100), "all-pairs": (5, 3), "batched": (50, 20),
                    "hbm-ceiling": (20, 10), "serving": (0, 0), "xtc-decode": (20, 5),
                    "dcd-decode": (50, 10), "nearest-leader": (20, 5), "centroid": (20, 5),
                    "lprmsd": (5, 2), "superpose": (20, 5)}
    d_steps, d_warm = per_workload[args.workload]
    args.steps = d_steps if args.steps is None else args.steps
    args.warmup = d_warm if args.warmup is None else args.warmup
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is a design (frames shard across GPUs with no exchange in the hot loop), "
              "not implemented.", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    torch.cuda.set_device(0)
    from mdtraj_amd import _lib

    _lib.call("hr_set_tuning", _lib.HR_TUNE_FRAMES_WAVES_PER_SIMD, args.frames_waves)
    _lib.call("hr_set_tuning", _lib.HR_TUNE_FRAMES_GRID, args.frames_grid)
    _lib.call("hr_set_tuning", _lib.HR_TUNE_FRAMES_ROW_LANES, args.frames_row_lanes)
    _lib.call("hr_set_tuning", _lib.HR_TUNE_PAIRS_TILE, args.pairs_tile)
    if args.workload == "hbm-ceiling":
        print(json.dumps(bench_hbm_ceiling(args)), flush=True)
        return
    if args.workload == "serving":
        print(json.dumps(bench_serving(args)), flush=True)
        return
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "xtc-decode":
        print(json.dumps(bench_xtc_decode(args)), flush=True)
        return
    if args.workload == "dcd-decode":
        print(json.dumps(bench_dcd_decode(args)), flush=True)
        return
    if args.workload == "nearest-leader":
        print(json.dumps(bench_nearest_leader(args)), flush=True)
        return
    if args.workload == "centroid":
        print(json.dumps(bench_centroid(args)), flush=True)
        return
    if args.workload == "lprmsd":
        print(json.dumps(bench_lprmsd(args)), flush=True)
        return
    if args.workload == "superpose":
        print(json.dumps(bench_superpose(args)), flush=True)
        return
    if args.workload == "one-vs-all":
        args.frames = args.frames or 100_000
        res = bench_one_vs_all(args)
    else:
        args.frames = args.frames or 20_000
        res = bench_all_pairs(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
