"""Benchmark: target velocities/s for the 2-D vortex-blob tree evaluation on one MI355X.

One step = one full pass of the hot path on device-resident blobs: bounding
box, Morton keys, radix sort, level-by-level tree, S2M, M2M and the velocity
at every blob (targets = sources), i.e. make_tree + biot_tree at all N blobs
(e_velo_mpi.f90:32-48).  Default workload: N = 4,000,000 blobs from an
8-component Gaussian mixture (core radius 5e-4, circulation N(0, 1/N)).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--nterms P]

Prints one JSON line.  Besides the throughput it carries the dominant
kernel's roofline (algorithmic FP64 flops / its event-timed duration), the
accuracy against an FP64 direct sum on sampled targets, and a CPU baseline
(the oracle restatement of the reference algorithm on the host cores, on a
bounded sample, extrapolated to all targets).
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
sys.path.insert(0, ROOT)

METRIC = "target velocities/s at N=4M, rel.err≤1e-6 vs FP64 direct: mixed+learned vs plain"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(xy, cir, cor, nterms, num_p, theta, n_targets, seed, threads):
    """Oracle (reference-semantics C restatement): the tree of all N blobs built
    single-threaded as the reference does, then a sample of targets on
    `threads` host threads; the rate is extrapolated to all N targets."""
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
    return {
        "value": n / total,
        "unit": "target velocities/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle tree build of all {n} blobs on 1 thread ({t_build:.1f} s) + "
                   f"{len(idx)} sampled targets on {threads} threads ({t_eval:.2f} s), "
                   f"extrapolated to {n} targets"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--nterms", type=int, default=16)
    ap.add_argument("--num-p", type=int, default=64)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--judge-targets", type=int, default=10_000)
    ap.add_argument("--cpu-targets", type=int, default=20_000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: this build runs on one GPU only "
              f"(the path would shard by Morton ranges of targets; not built)", flush=True)
        sys.exit(2)

    import torch

    fmm = importlib.import_module("2dfmm_amd")
    wl = importlib.import_module("2dfmm_amd.workloads")

    n = args.n
    log(f"[bench] generating gaussian_mixture N={n} seed={args.seed}")
    xy, cir, cor = wl.gaussian_mixture(n, seed=args.seed)
    dev = torch.device("cuda:0")
    d_xy = torch.from_numpy(xy).to(dev)
    d_cir = torch.from_numpy(cir).to(dev)
    d_cor = torch.from_numpy(cor).to(dev)
    d_uv = torch.empty((n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    fmm.set_params(args.nterms, args.num_p, args.theta)
    fmm.set_timing(True)

    def step():
        fmm.make_tree_device(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr())
        fmm.eval_sources_device(d_uv.data_ptr())

    for w in range(args.warmup):
        step()
        fmm.synchronize()
        log(f"[bench] warmup {w + 1}/{args.warmup}: {fmm.get_timing()}")
    phases = []
    torch.cuda.synchronize()
    fmm.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        phases.append(fmm.get_timing())  # waits for the step (events on the library stream)
    fmm.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step; phases(ms) {ph}")

    # algorithmic work of the evaluation kernel (outside the timed region)
    m2s, pairs = fmm.count_interactions()
    p = args.nterms
    flops = m2s * (8 * (p - 1) + 10) + pairs * 20
    eval_ms = ph["evaluate"]
    achieved = flops / (eval_ms * 1e-3) / 1e12
    info = fmm.tree_size()

    # accuracy against the FP64 direct sum on sampled targets
    idx = wl.sample_targets(n, args.judge_targets, seed=7)
    uv = d_uv.cpu().numpy()
    d_t = torch.from_numpy(np.ascontiguousarray(xy[idx])).to(dev)
    d_ref = torch.empty((len(idx), 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    fmm.direct_device(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr(), len(idx),
                      d_t.data_ptr(), d_ref.data_ptr(), 1)
    fmm.synchronize()
    ref = d_ref.cpu().numpy()
    got = uv[idx]
    rel_l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    rel_max = float(np.abs(got - ref).max() / np.abs(ref).max())
    log(f"[bench] accuracy vs FP64 direct on {len(idx)} targets: rel_L2={rel_l2:.3e} "
        f"rel_max={rel_max:.3e}")

    cpu = None
    if not args.no_cpu:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            threads = min(16, os.cpu_count() or 1)
        log(f"[bench] cpu baseline: oracle on {threads} threads")
        cpu = cpu_baseline(xy, cir, cor, args.nterms, args.num_p, args.theta, args.cpu_targets,
                           11, threads)
        log(f"[bench] cpu baseline {cpu['value']:.4g} velocities/s ({cpu['sample']})")

    value = n / (ms_per_step * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "target velocities/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian mixture)",
        "config": {"workload": "gaussian_mixture_8", "n_blobs": n, "targets": n,
                   "core_radius": 5e-4, "mode": "treecode", "nterms": args.nterms,
                   "num_p": args.num_p, "theta": args.theta, "parallelism": "single GPU"},
        "accuracy": {"vs": "FP64 direct sum (direct_evaluate rule)", "targets": len(idx),
                     "rel_L2": rel_l2, "rel_max": rel_max, "meets_1e-6": rel_l2 <= 1e-6},
        "phases_ms": ph,
        "tree": info,
        "roofline": {"kernel": "parity_eval_kernel<0>", "bound": "mfma", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": None, "flops_per_launch": flops, "m2s_per_launch": m2s,
                     "p2p_pairs_per_launch": pairs, "avg_launch_ms": eval_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
