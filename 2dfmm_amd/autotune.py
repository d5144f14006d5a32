"""Autotuner for the batched (forest) evaluation.

Times the whole step (forest build + evaluation) for candidate
configurations on the caller's own device-resident systems and keeps the
fastest.  Knobs, tuned one group after another (each stage keeps the best of
the previous ones):
  1. leaf     leaf capacity: the target tile of the near-field sums (a wave's
              targets: up to `leaf` targets x 64 / leaf source replicas) and,
              through the tree depth, how many far cells each leaf contracts;
     blocks   persistent grid size in units of four waves (0 = one wave per
              target leaf over a sized grid, no work queue);
     graph    replay the persistent evaluation from a captured HIP graph;
  2. split    near-work units per chunk of a heavy target leaf: how the P2P
              tiles of one leaf are dealt to waves (fmm_set_split);
  3. burst    far cells per MFMA burst: the M2L grouping of a leaf's
              contraction (fmm_set_far_burst).
The accuracy of every candidate against the FP64 direct sum is the caller's
to check (see bench.py); the tuner only orders by time.
"""
from __future__ import annotations

import time

DEFAULTS = {"leaf": 64, "blocks": 0, "graph": False, "split": 8192, "burst": 128}


def apply(fmm, cfg, p=12, theta=0.5, cut=4.0) -> None:
    fmm.set_fmm_params(p, cfg["leaf"], theta, cut)
    fmm.set_persistent(cfg["blocks"], cfg["graph"])
    fmm.set_split(cfg["split"])
    fmm.set_far_burst(cfg["burst"])


def reset(fmm) -> None:
    """Back to the library defaults."""
    fmm.set_persistent(0, False)
    fmm.set_split(DEFAULTS["split"])
    fmm.set_far_burst(DEFAULTS["burst"])
    fmm.set_fmm_params()


def tune_forest(fmm, offsets, xy_ptr, cir_ptr, cor_ptr, uv_ptr, p=12, theta=0.5, cut=4.0,
                leaves=(32, 48, 64), blocks_per_cu=(0, 2, 4, 8), splits=(2048, 8192, 32768),
                bursts=(64, 96, 128), reps=2, sync=None, log=None):
    """Returns (best_config, table) with table = [(config, ms_per_step), ...]."""
    cus = max(1, fmm.device_cus())
    table = []

    def timed(cfg):
        apply(fmm, cfg, p, theta, cut)

        def step():
            fmm.make_forest_device(offsets, xy_ptr, cir_ptr, cor_ptr)
            fmm.eval_sources_device(uv_ptr)

        step()  # warm-up (allocation, graph capture)
        fmm.synchronize()
        if sync:
            sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            step()
        fmm.synchronize()
        if sync:
            sync()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        table.append((dict(cfg), ms))
        if log:
            log(f"[autotune] {cfg}: {ms:.2f} ms/step")
        return ms

    best, best_ms = None, float("inf")
    for leaf in leaves:
        for bpc in blocks_per_cu:
            cfg = dict(DEFAULTS, leaf=leaf, blocks=bpc * cus, graph=bpc > 0)
            ms = timed(cfg)
            if ms < best_ms:
                best, best_ms = cfg, ms
    for key, values in (("split", splits), ("burst", bursts)):
        for v in values:
            if v == best[key]:
                continue
            cfg = dict(best, **{key: v})
            ms = timed(cfg)
            if ms < best_ms:
                best, best_ms = cfg, ms
    apply(fmm, best, p, theta, cut)
    return best, table
