"""Autotuner for the batched (forest) evaluation.

Times the whole step (forest build + persistent evaluation) for each candidate
configuration on the caller's own device-resident systems and keeps the
fastest one.  Knobs:
  leaf     leaf capacity: sets the target tile of the near-field sums (the
           wave's targets) and, through the tree depth, how many far cells
           each leaf's MFMA contraction gathers;
  blocks   persistent grid size in workgroups (0 = one workgroup per four
           target leaves, no work queue);
  graph    replay the evaluation from a captured HIP graph.
The accuracy of every candidate against the FP64 direct sum is the caller's
to check (see bench.py); the tuner only orders by time.
"""
from __future__ import annotations

import itertools
import time


def tune_forest(fmm, offsets, xy_ptr, cir_ptr, cor_ptr, uv_ptr, p=12, theta=0.5, cut=4.0,
                leaves=(32, 48, 64), blocks_per_cu=(0, 2, 4, 8), graph=(True,), reps=2,
                sync=None, log=None):
    """Returns (best_config, table) with table = [(config, ms_per_step), ...]."""
    cus = max(1, fmm.device_cus())
    table = []
    for leaf, bpc, gr in itertools.product(leaves, blocks_per_cu, graph):
        cfg = {"leaf": leaf, "blocks": bpc * cus, "graph": bool(gr and bpc > 0)}
        fmm.set_fmm_params(p, leaf, theta, cut)
        fmm.set_persistent(cfg["blocks"], cfg["graph"])

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
        table.append((cfg, ms))
        if log:
            log(f"[autotune] {cfg}: {ms:.2f} ms/step")
    best = min(table, key=lambda t: t[1])[0]
    fmm.set_fmm_params(p, best["leaf"], theta, cut)
    fmm.set_persistent(best["blocks"], best["graph"])
    return best, table
