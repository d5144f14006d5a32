"""Benchmark: target velocities/s for the 2-D vortex-blob tree evaluation on one MI355X.

One step = one full pass of the hot path on device-resident blobs: bounding
box, Morton keys, radix sort, level-by-level tree, S2M, M2M and the velocity
at every blob (targets = sources), i.e. make_tree + biot_tree at all N blobs
(e_velo_mpi.f90:32-48).  Workload: N = 4,000,000 blobs from an 8-component
Gaussian mixture (centres U[0,1]^2, std log-uniform in [1e-3, 5e-2], core
radius 5e-4, circulation N(0, 1/N)), seeded.

The headline `value` is the mixed-precision mode with learned per-leaf
order and near-field radius (FP32 pair sums accumulated in FP64, smooth
cells summed through their Taylor coefficients, FP32-MFMA far-field
contraction, FP64 expansions) when it meets rel. L2 <= 1e-6, else
the mixed mode with fixed parameters; the plain FP64 mode is timed on the
same input, and each states its error against the FP64 direct sum on
sampled targets.  The reference-rule treecode is timed beside them.  Inputs
are resident in HBM for `value`; the rate through host arrays (PCIe copies
included) is reported beside it as `host_inclusive`.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--p P]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# complex coefficients of a smooth cell's degree-8 expansion (9 + 36)
SMOOTH_COEFS = 45
sys.path.insert(0, ROOT)

METRIC = "target velocities/s at N=4M, rel.err≤1e-6 vs FP64 direct: mixed+learned vs plain"
PEAK = {"f64": 78.6, "f32": 157.3}  # MI355X TFLOP/s; vector and matrix peaks coincide


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(xy, cir, cor, nterms, num_p, theta, n_targets, seed, threads):
    """Oracle (reference-semantics C restatement of the treecode): the tree of
    all N blobs built single-threaded as the reference does, then a sample of
    targets on `threads` host threads (and a smaller one on 1 thread); the
    rates are extrapolated to N targets."""
    from oracle import oracle as O
    wl = importlib.import_module("2dfmm_amd.workloads")
    n = xy.shape[0]
    t0 = time.perf_counter()
    T = O.Tree(xy, cir, cor, nterms, num_p, theta)
    t_build = time.perf_counter() - t0
    idx = wl.sample_targets(n, n_targets, seed=seed)
    t0 = time.perf_counter()
    T.eval(xy[idx], nthreads=threads)
    t_eval = time.perf_counter() - t0
    total = t_build + t_eval * n / len(idx)
    idx1 = idx[: max(1, len(idx) // 8)]
    t0 = time.perf_counter()
    T.eval(xy[idx1], nthreads=1)
    t_eval1 = time.perf_counter() - t0
    total1 = t_build + t_eval1 * n / len(idx1)
    return {
        "value": n / total,
        "unit": "target velocities/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"reference-rule treecode (oracle restatement, p={nterms}, leaf {num_p}, "
                   f"theta {theta}): tree of all {n} blobs on 1 thread ({t_build:.1f} s) + "
                   f"{len(idx)} sampled targets on {threads} threads ({t_eval:.2f} s), "
                   f"extrapolated to {n} targets"),
        "value_1core": n / total1,
        "sample_1core": f"same tree + {len(idx1)} sampled targets on 1 thread ({t_eval1:.2f} s)",
        "host": {"cpu_model": _cpu_model(), "nproc": os.cpu_count()},
    }


def bench_batch(args, fmm, wl, torch):
    """Batching form: 256 independent systems (N log-uniform in [2e3, 2e5],
    uniform or 2-cluster layout, core 1e-3, p=12) in one forest, evaluated by
    the persistent kernel replayed from a graph; configuration by the
    autotuner.  Reports aggregate velocities/s, and the per-system latency of
    the one-system-at-a-time path (p50/p99) beside it."""
    tune = importlib.import_module("2dfmm_amd.autotune")
    systems = wl.batch_systems(256, seed=4)
    off, xy, cir, cor = fmm.concat_systems(systems)
    n = int(off[-1])
    dev = torch.device("cuda:0")
    d_xy = torch.from_numpy(xy).to(dev)
    d_cir = torch.from_numpy(cir).to(dev)
    d_cor = torch.from_numpy(cor).to(dev)
    d_uv = torch.empty((n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    fmm.set_mode(fmm.MODE_MIXED)
    log(f"[batch] 256 systems, N total {n} (min {np.diff(off).min()}, max {np.diff(off).max()})")
    best, table = tune.tune_forest(fmm, off, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr(),
                                   d_uv.data_ptr(), p=12, sync=torch.cuda.synchronize, log=log)
    log(f"[batch] autotuner picked {best}")

    def step():
        fmm.make_forest_device(off, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr())
        fmm.eval_sources_device(d_uv.data_ptr())

    for _ in range(args.warmup):
        step()
    fmm.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    fmm.synchronize()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    uv = d_uv.cpu().numpy()
    # accuracy: sampled targets of 16 systems against their own FP64 direct sums
    errs = []
    for s in range(0, 256, 16):
        a, b = int(off[s]), int(off[s + 1])
        t = np.arange(a, b, max(1, (b - a) // 300))
        ref = fmm.direct(xy[a:b], cir[a:b], cor[a:b], xy[t], style=1)
        errs.append(float(np.linalg.norm(uv[t] - ref) / np.linalg.norm(ref)))
    # per-system latency through the single-system path (same parameters)
    fmm.set_persistent(0, False)
    lat = []
    for s in range(256):
        a, b = int(off[s]), int(off[s + 1])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        fmm.make_tree_device(b - a, d_xy[a:b].data_ptr(), d_cir[a:b].data_ptr(),
                             d_cor[a:b].data_ptr())
        fmm.eval_sources_device(d_uv[a:b].data_ptr())
        fmm.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = np.array(lat)
    out = {
        "metric": "aggregate target velocities/s, 256 uneven systems in one forest",
        "value": n / (ms * 1e-3), "unit": "target velocities/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (256 seeded systems)",
        "config": {"workload": "batch_systems_256", "n_blobs": n, "mode": "mixed", "p": 12,
                   "core_radius": 1e-3, "autotuned": best},
        "autotune_table": [[c, t] for c, t in table],
        "accuracy": {"vs": "per-system FP64 direct sum", "worst_rel_L2": max(errs),
                     "median_rel_L2": float(np.median(errs))},
        "one_system_at_a_time": {"total_ms": float(lat.sum()), "p50_ms": float(np.percentile(lat, 50)),
                                 "p99_ms": float(np.percentile(lat, 99))},
        "batched_latency_ms": ms,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--p", type=int, default=16, help="expansion terms of the accurate modes")
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--cut", type=float, default=4.0, help="near-field radius in core radii")
    ap.add_argument("--modes", default="learned,mixed,fmm64,treecode")
    ap.add_argument("--judge-targets", type=int, default=10_000)
    ap.add_argument("--cpu-targets", type=int, default=20_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="mixture4m", choices=["mixture4m", "batch"])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device to device (a kernel of known traffic for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--split", type=int, default=8192,
                    help="accurate modes: near work units per chunk of a split target leaf (0: off)")
    ap.add_argument("--persistent", type=int, default=0,
                    help="accurate modes: persistent fused pass with this many workgroups per CU "
                         "(0: one wave per target leaf over a sized grid)")
    ap.add_argument("--smooth", type=float, default=0.3,
                    help="mixed modes: smooth cells of half side <= SMOOTH core radii (0: off)")
    ap.add_argument("--smooth-min", type=int, default=24,
                    help="mixed modes: fewest blobs of a smooth cell")
    ap.add_argument("--traffic",
                    default=os.path.join(ROOT, "profiles", "fmm_eval_gaussian_mixture_4M_traffic.txt"),
                    help="measured HBM bytes per evaluation (rocprofv3 FETCH_SIZE / WRITE_SIZE, "
                         "copy-calibrated), JSON keyed by precision")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: this build runs on one GPU only "
              f"(the path would shard by Morton ranges of targets; not built)", flush=True)
        sys.exit(2)

    import torch

    fmm = importlib.import_module("2dfmm_amd")
    wl = importlib.import_module("2dfmm_amd.workloads")
    if args.workload == "batch":
        bench_batch(args, fmm, wl, torch)
        return

    if args.copy_calibration:
        a = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
        a.fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 256 MiB read + 256 MiB written")
        del a, b

    n = args.n
    log(f"[bench] generating gaussian_mixture N={n} seed={args.seed}")
    xy, cir, cor = wl.gaussian_mixture(n, seed=args.seed)
    dev = torch.device("cuda:0")
    d_xy = torch.from_numpy(xy).to(dev)
    d_cir = torch.from_numpy(cir).to(dev)
    d_cor = torch.from_numpy(cor).to(dev)
    d_uv = torch.empty((n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    fmm.set_timing(True)
    if args.persistent > 0:
        fmm.set_persistent(args.persistent * fmm.device_cus(), False)
    fmm.set_split(args.split)
    fmm.set_smooth(args.smooth, args.smooth_min)

    # FP64 direct sum on sampled targets (the judge)
    idx = wl.sample_targets(n, args.judge_targets, seed=7)
    d_t = torch.from_numpy(np.ascontiguousarray(xy[idx])).to(dev)
    d_ref = torch.empty((len(idx), 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fmm.direct_device(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr(), len(idx),
                      d_t.data_ptr(), d_ref.data_ptr(), 1)
    fmm.synchronize()
    log(f"[bench] FP64 direct judge on {len(idx)} targets: {time.perf_counter() - t0:.2f} s")
    ref = d_ref.cpu().numpy()

    def step():
        fmm.make_tree_device(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr())
        fmm.eval_sources_device(d_uv.data_ptr())

    results = {}
    wpath = os.path.join(ROOT, "2dfmm_amd", "learned_weights.npz")
    for mode in [m for m in args.modes.split(",") if m]:
        fmm.set_learned(None)
        if mode == "learned" and not os.path.exists(wpath):
            log("[bench] no trained weights; skipping the learned mode")
            continue
        if mode == "treecode":
            fmm.set_mode(fmm.MODE_TREECODE)
            fmm.set_params(args.p, args.leaf, args.theta)
            prec, kernel = "f64", "parity_eval_kernel<0>"
        else:
            mixed = mode in ("mixed", "learned")
            fmm.set_mode(fmm.MODE_MIXED if mixed else fmm.MODE_FMM64)
            fmm.set_fmm_params(args.p, args.leaf, args.theta, args.cut)
            prec = "f32" if mixed else "f64"
            # <mixed, one core radius>: the workload has one core radius
            t = "true, true" if mixed else "false, false"
            kernel = f"fmm_leaf_kernel<{t}> + fmm_chunk_kernel<{t}> + fmm_split_sum_kernel"
            if mode == "learned":
                with np.load(wpath) as z:
                    fmm.set_learned(z["weights"], float(z["p_margin"]), float(z["cut_margin"]))
        for w in range(args.warmup):
            step()
            log(f"[bench] {mode} warmup {w + 1}/{args.warmup}: {fmm.get_timing()}")
        phases = []
        torch.cuda.synchronize()
        fmm.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
            phases.append(fmm.get_timing())  # waits for the step (events on the library stream)
        fmm.synchronize()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
        got = d_uv.cpu().numpy()[idx]
        l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        mx = float(np.abs(got - ref).max() / np.abs(ref).max())
        p = args.p
        if mode == "treecode":
            far, pairs = fmm.count_interactions()
            smooth = 0
            flops = far * (8 * (p - 1) + 10) + pairs * 20
        else:
            far, pairs, smooth = fmm.count_work()
            # smooth cell at one target: 8 flops per complex coefficient
            flops = far * 8 * p * p + pairs * 20 + n * 8 * p + smooth * 8 * SMOOTH_COEFS
        ach = flops / (ph["evaluate"] * 1e-3) / 1e12
        results[mode] = {"ms_per_step": ms, "value": n / (ms * 1e-3), "rel_L2": l2, "rel_max": mx,
                         "phases_ms": ph, "tree": fmm.tree_size(), "far": far, "pairs": pairs,
                         "smooth": smooth,
                         "flops": flops, "achieved_tflops": ach, "precision": prec,
                         "kernel": kernel}
        log(f"[bench] {mode}: {ms:.2f} ms/step ({n / (ms * 1e-3):.4g} vel/s), rel_L2={l2:.3e} "
            f"rel_max={mx:.3e}, far={far} pairs={pairs} smooth={smooth}, eval {ph['evaluate']:.2f} ms "
            f"= {ach:.2f} TFLOP/s; phases {ph}")

    # headline: mixed + learned when it meets the 1e-6 bar, else mixed alone
    if "learned" in results and results["learned"]["rel_L2"] <= 1e-6:
        head_mode = "learned"
    elif "mixed" in results:
        head_mode = "mixed"
    else:
        head_mode = next(iter(results))
    head = results[head_mode]
    plain = results.get("fmm64")

    # the same step through host arrays: make_tree copies the blobs in, the
    # velocities come back to host memory (PCIe both ways); never `value`
    host = None
    if head_mode in ("learned", "mixed", "fmm64"):
        fmm.set_mode(fmm.MODE_FMM64 if head_mode == "fmm64" else fmm.MODE_MIXED)
        fmm.set_fmm_params(args.p, args.leaf, args.theta, args.cut)
        if head_mode == "learned":
            with np.load(wpath) as z:
                fmm.set_learned(z["weights"], float(z["p_margin"]), float(z["cut_margin"]))
        else:
            fmm.set_learned(None)
        # one evaluation per step: no velocity prefetch inside make_tree
        fmm.set_prefetch(False)
        fmm.make_tree(xy, cir, cor)
        fmm.eval_sources()
        t0 = time.perf_counter()
        reps = max(1, min(args.steps, 3))
        for _ in range(reps):
            fmm.make_tree(xy, cir, cor)
            fmm.eval_sources()
        hms = (time.perf_counter() - t0) * 1e3 / reps
        fmm.set_prefetch(True)
        host = {"ms_per_step": hms, "value": n / (hms * 1e-3),
                "note": "make_tree from host arrays + velocities copied back (PCIe included)"}
        log(f"[bench] {head_mode} through host arrays: {hms:.2f} ms/step")
    fmm.set_learned(None)

    cpu = None
    if not args.no_cpu:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            threads = min(16, os.cpu_count() or 1)
        log(f"[bench] cpu baseline: oracle on {threads} threads")
        cpu = cpu_baseline(xy, cir, cor, args.p, args.leaf, args.theta, args.cpu_targets, 11,
                           threads)
        log(f"[bench] cpu baseline {cpu['value']:.4g} velocities/s ({cpu['sample']})")

    peak = PEAK[head["precision"]]
    traffic = None
    if args.traffic and os.path.exists(args.traffic) and head_mode != "treecode":
        with open(args.traffic) as f:
            t = json.load(f).get(head["precision"])
        if t:
            traffic = t["bytes"]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "target velocities/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head["precision"] == "f32" else "f64",
        "data": "synthetic (seeded Gaussian mixture)",
        "config": {"workload": "gaussian_mixture_8", "n_blobs": n, "targets": n,
                   "core_radius": 5e-4, "mode": head_mode, "p": args.p, "leaf": args.leaf,
                   "theta": args.theta, "cut_core_radii": args.cut, "parallelism": "single GPU",
                   "persistent_blocks_per_cu": args.persistent, "split_budget": args.split},
        "accuracy": {"vs": "FP64 direct sum (fmm.c:121-142 rule)", "targets": len(idx),
                     "rel_L2": head["rel_L2"], "rel_max": head["rel_max"],
                     "meets_1e-6": head["rel_L2"] <= 1e-6},
        "speedup_vs_plain_fp64": (plain["ms_per_step"] / head["ms_per_step"]) if plain else None,
        "modes": results,
        "roofline": {"kernel": head["kernel"], "bound": "mfma", "achieved": head["achieved_tflops"],
                     "peak": peak, "unit": "TFLOP/s", "frac": head["achieved_tflops"] / peak,
                     "traffic": traffic, "flops_per_launch": head["flops"],
                     "avg_launch_ms": head["phases_ms"]["evaluate"]},
        "cpu_baseline": cpu,
        "host_inclusive": host,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
