"""Learned per-leaf expansion order and near-field radius.

Pipeline (all heavy work on the GPU through the library):

1. data   advect the 2M-blob elliptic vortex patch for 50 forward-Euler steps
          (dt = 1e-3) with velocities of the plain FP64 accurate mode, plus
          random-sign Gaussian mixtures of 1M and 4M blobs (seeds other than
          the benchmark's);
2. labels at sampled steps, for sampled target leaves, the smallest near-field
          radius (in core radii) and then the smallest order p whose
          per-target error against the FP64 direct sum stays below
          tol * U_rms (--label-mode: the plain FP64 mode by default, or the
          mixed mode with smooth cells and parent expansions); the error of a
          leaf depends only on its own (p, cut) -- parent expansions take the
          largest order and radius of their children, which only helps -- so
          each candidate is one evaluation with that value set for every leaf.
          Radii below the library's floor (fmm_set_learned_floor, 2.5) are
          labelled too; inference clamps them to the floor;
3. train  the FP32 MLP 8 -> 64 -> 64 -> 2 on (features, labels) with the
          library's hand-written kernels (Adam, minibatches);
4. margins: the smallest on a grid for which held-out patch steps and a
          held-out 4M mixture (seed 5) stay below 0.8 x the tolerance in rel.
          L2 and below 3 x the tolerance in rel. max (with the floor applied);
          the fixed order p = 16, cut = 4 stays the fallback.  The benchmark
          layout (4M, seed 3) is used only for the final report.  With
          --weights the network is not retrained: only steps 1 (the held-out
          patch steps) and 4 run, for the given weights.

    python -m 2dfmm_amd.learned --out 2dfmm_amd/learned_weights.npz
"""
from __future__ import annotations

import argparse
import importlib
import json
import sys
import time

import numpy as np

CUTS = (1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 5.0)
PS = (4, 6, 8, 10, 12, 14, 16)


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def normalise(p, cut):
    return np.stack([(np.asarray(p, np.float32) - 11.0) / 5.0,
                     (np.asarray(cut, np.float32) - 3.0) / 2.0], axis=1)


def leaf_errors(fmm, uv, ref, targets, tleaf, nl, urms):
    """Largest |du| / U_rms per target leaf over the sampled targets."""
    err = np.linalg.norm(uv[targets] - ref, axis=1) / urms
    out = np.zeros(nl)
    np.maximum.at(out, tleaf, err)
    return out


def label_tree(fmm, xy, cir, cor, tol, n_leaves, seed, cut_max=CUTS[-1], p_max=16, mode=None,
               cuts=CUTS):
    """Build the tree (plain FP64 mode unless `mode` says otherwise) and label
    sampled target leaves.

    Returns (features[sampled], (p*, cut*), sampled leaf ids)."""
    fmm.set_mode(fmm.MODE_FMM64 if mode is None else mode)
    fmm.set_fmm_params(p_max, 64, 0.5, cut_max)
    fmm.set_learned(None)
    fmm.make_tree(xy, cir, cor)
    nl = fmm.num_target_leaves()
    feat = fmm.leaf_features()
    bl = fmm.blob_leaf()
    rng = np.random.default_rng(seed)
    leaves = np.sort(rng.choice(nl, min(n_leaves, nl), replace=False))
    mask = np.zeros(nl, bool)
    mask[leaves] = True
    targets = np.nonzero(mask[bl])[0]
    tleaf = bl[targets]
    ref = fmm.direct(xy, cir, cor, xy[targets], style=1)
    urms = np.sqrt((ref ** 2).sum(axis=1).mean())
    ok_cut = np.zeros((len(cuts), nl), bool)
    for k, c in enumerate(cuts):
        fmm.set_leaf_params(np.tile(np.array([p_max, c], np.float32), (nl, 1)))
        e = leaf_errors(fmm, fmm.eval_sources(), ref, targets, tleaf, nl, urms)
        ok_cut[k] = e <= tol
    # smallest cut from which every larger cut also meets the tolerance
    cut_star = np.full(nl, cuts[-1], np.float32)
    good = np.ones(nl, bool)
    for k in range(len(cuts) - 1, -1, -1):
        good &= ok_cut[k]
        cut_star[good] = cuts[k]
    ok_p = np.zeros((len(PS), nl), bool)
    for k, p in enumerate(PS):
        par = np.stack([np.full(nl, p, np.float32), cut_star], axis=1)
        fmm.set_leaf_params(par)
        e = leaf_errors(fmm, fmm.eval_sources(), ref, targets, tleaf, nl, urms)
        ok_p[k] = e <= tol
    p_star = np.full(nl, p_max, np.float32)
    good = np.ones(nl, bool)
    for k in range(len(PS) - 1, -1, -1):
        good &= ok_p[k]
        p_star[good] = PS[k]
    fmm.set_leaf_params(None)
    return feat[leaves], np.stack([p_star[leaves], cut_star[leaves]], 1), leaves


def evaluate(fmm, xy, cir, cor, weights, p_margin, cut_margin, judge=4000, seed=0, mode=None,
             reps=2):
    """(fixed, learned) dicts with rel_L2 vs FP64 direct, ms per evaluation,
    near pairs, mean predicted p and cut."""
    wl = importlib.import_module("2dfmm_amd.workloads")
    n = xy.shape[0]
    idx = wl.sample_targets(n, judge, seed=seed + 17)
    ref = fmm.direct(xy, cir, cor, xy[idx], style=1)
    fmm.set_mode(fmm.MODE_MIXED if mode is None else mode)
    fmm.set_fmm_params(16, 64, 0.5, 4.0)
    out = {}
    for name, w in (("fixed", None), ("learned", weights)):
        fmm.set_learned(w, p_margin, cut_margin)
        fmm.make_tree(xy, cir, cor)
        uv = fmm.eval_sources()
        fmm.set_timing(True)
        t = []
        for _ in range(reps):
            fmm.eval_sources()
            t.append(fmm.get_timing()["evaluate"])
        fmm.set_timing(False)
        far, pairs = fmm.count_interactions()
        par = fmm.leaf_params()
        out[name] = {"rel_L2": float(np.linalg.norm(uv[idx] - ref) / np.linalg.norm(ref)),
                     "rel_max": float(np.abs(uv[idx] - ref).max() / np.abs(ref).max()),
                     "eval_ms": float(np.min(t)), "pairs": int(pairs), "far": int(far),
                     "mean_p": float(par[:, 0].mean()), "mean_cut": float(par[:, 1].mean())}
    fmm.set_learned(None)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="gpurun_out/learned_weights.npz")
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--leaves", type=int, default=3000)
    ap.add_argument("--label-every", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--label-mode", choices=["fmm64", "mixed"], default="fmm64",
                    help="mode whose errors label the leaves")
    ap.add_argument("--cut-min", type=float, default=CUTS[0],
                    help="smallest near radius labelled (core radii)")
    ap.add_argument("--max-rel-max", type=float, default=3.0,
                    help="margins: held-out rel. max at most this many tolerances")
    ap.add_argument("--weights", default=None,
                    help="existing weights: skip labelling and training, choose the margins only")
    ap.add_argument("--mixture-n", type=int, default=4_000_000,
                    help="blobs of the large mixtures: training (--big-seeds), margins "
                         "(--mix-val-seed) and the benchmark layout (seed 3) evaluated last")
    ap.add_argument("--mix-seeds", default="101,102,103,104,105,106",
                    help="Gaussian-mixture layouts (other seeds) added to the training data")
    ap.add_argument("--mix-n", type=int, default=1_000_000)
    ap.add_argument("--big-seeds", default="201,202",
                    help="Gaussian mixtures of --mixture-n blobs added to the training data")
    ap.add_argument("--mix-val-seed", type=int, default=5,
                    help="held-out Gaussian mixture of --mixture-n blobs for the margins")
    args = ap.parse_args()
    fmm = importlib.import_module("2dfmm_amd")
    wl = importlib.import_module("2dfmm_amd.workloads")
    t0 = time.time()
    mode = fmm.MODE_MIXED if args.label_mode == "mixed" else fmm.MODE_FMM64
    cuts = tuple(c for c in CUTS if c >= args.cut_min) or CUTS[-1:]
    xy, cir, cor = wl.elliptic_patch(args.n)
    label_steps = list(range(0, args.steps, args.label_every))
    held_out = [s for s in (args.steps // 2 + 1, args.steps - 1) if s not in label_steps]
    snaps = {}
    X, Y = [], []
    for step in range(args.steps):
        fmm.set_mode(fmm.MODE_FMM64)
        fmm.set_fmm_params(16, 64, 0.5, 4.0)
        fmm.set_learned(None)
        fmm.make_tree(xy, cir, cor)
        uv = fmm.eval_sources()
        if step in label_steps and not args.weights:
            f, lab, _ = label_tree(fmm, xy, cir, cor, args.tol, args.leaves, seed=step, mode=mode,
                                   cuts=cuts)
            X.append(f)
            Y.append(lab)
            _log(f"[learned] step {step}: {len(f)} leaves labelled; mean p* {lab[:, 0].mean():.2f}, "
                 f"mean cut* {lab[:, 1].mean():.2f} ({time.time() - t0:.0f} s)")
        if step in held_out:
            snaps[step] = xy.copy()
        xy = xy + args.dt * uv  # forward Euler
    # random-sign circulation fields in clustered layouts (other seeds than the
    # benchmark layout), so the network sees that regime too
    for sd in [int(s) for s in args.mix_seeds.split(",") if s and not args.weights]:
        mxy, mcir, mcor = wl.gaussian_mixture(args.mix_n, seed=sd)
        f, lab, _ = label_tree(fmm, mxy, mcir, mcor, args.tol, args.leaves, seed=sd, mode=mode,
                               cuts=cuts)
        X.append(f)
        Y.append(lab)
        _log(f"[learned] mixture seed {sd}: {len(f)} leaves labelled; mean p* {lab[:, 0].mean():.2f}, "
             f"mean cut* {lab[:, 1].mean():.2f} ({time.time() - t0:.0f} s)")
    for sd in [int(s) for s in args.big_seeds.split(",") if s and not args.weights]:
        mxy, mcir, mcor = wl.gaussian_mixture(args.mixture_n, seed=sd)
        f, lab, _ = label_tree(fmm, mxy, mcir, mcor, args.tol, args.leaves, seed=sd, mode=mode,
                               cuts=cuts)
        X.append(f)
        Y.append(lab)
        _log(f"[learned] {args.mixture_n}-blob mixture seed {sd}: {len(f)} leaves labelled; "
             f"mean p* {lab[:, 0].mean():.2f}, mean cut* {lab[:, 1].mean():.2f} ({time.time() - t0:.0f} s)")
    if args.weights:
        with np.load(args.weights) as z:
            w = np.ascontiguousarray(z["weights"], np.float32)
            hist = z["loss_hist"] if "loss_hist" in z.files else np.zeros(1, np.float32)
            n_train = int(z["n_train"]) if "n_train" in z.files else 0
        Y = np.zeros((1, 2))
        _log(f"[learned] weights from {args.weights}: margins only")
    else:
        X = np.concatenate(X).astype(np.float32)
        Y = np.concatenate(Y)
        Yn = normalise(Y[:, 0], Y[:, 1])
        rng = np.random.default_rng(0)  # He initialisation
        w0 = np.concatenate([
            rng.standard_normal(64 * 8) * np.sqrt(2 / 8), np.zeros(64),
            rng.standard_normal(64 * 64) * np.sqrt(2 / 64), np.zeros(64),
            rng.standard_normal(2 * 64) * np.sqrt(2 / 64), np.zeros(2)]).astype(np.float32)
        w, hist = fmm.mlp_train(X, Yn, w0, epochs=args.epochs, batch=512, lr=2e-3, seed=1)
        n_train = len(X)
        _log(f"[learned] trained on {n_train} leaves: loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    # margins: smallest that meets the tolerance on the held-out sets (80% of it)
    chosen = None
    results = {}
    vxy, vcir, vcor = wl.gaussian_mixture(args.mixture_n, seed=args.mix_val_seed)
    vals = [(f"step{s}", sxy, cir, cor) for s, sxy in snaps.items()]
    vals.append((f"mixture{args.mixture_n}_seed{args.mix_val_seed}", vxy, vcir, vcor))
    for pm, cm in ((0.0, 0.0), (0.5, 0.125), (1.0, 0.25), (1.5, 0.375), (2.0, 0.5), (3.0, 1.0),
                   (4.0, 1.5)):
        ok = True
        for name, sxy, scir, scor in vals:
            r = evaluate(fmm, sxy, scir, scor, w, pm, cm, seed=len(name))
            results[f"{name}_margin{pm}_{cm}"] = r
            _log(f"[learned] held-out {name} margins ({pm},{cm}): {json.dumps(r)}")
            ok &= (r["learned"]["rel_L2"] <= 0.8 * args.tol
                   and r["learned"]["rel_max"] <= args.max_rel_max * args.tol)
        if ok:
            chosen = (pm, cm)
            break
    if chosen is None:
        chosen = (4.0, 1.5)
    # transfer to the 4M Gaussian mixture layout
    mxy, mcir, mcor = wl.gaussian_mixture(args.mixture_n, seed=3)
    r = evaluate(fmm, mxy, mcir, mcor, w, *chosen, seed=3)
    results["mixture4m"] = r
    _log(f"[learned] mixture 4M with margins {chosen}: {json.dumps(r)}")
    np.savez(args.out, weights=w, p_margin=np.float32(chosen[0]), cut_margin=np.float32(chosen[1]),
             loss_hist=hist, n_train=np.int64(n_train))
    print(json.dumps({"weights": args.out, "margins": chosen, "n_train": int(n_train),
                      "loss": [float(hist[0]), float(hist[-1])], "results": results,
                      "label_mean_p": float(Y[:, 0].mean()), "label_mean_cut": float(Y[:, 1].mean()),
                      "seconds": time.time() - t0}), flush=True)


if __name__ == "__main__":
    main()
