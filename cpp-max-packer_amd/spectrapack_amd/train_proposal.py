"""Train the learned proposal on plain-search episodes and compare packing
with and without it.

    python -m spectrapack_amd.train_proposal --episodes 200 --out DIR
    python -m spectrapack_amd.train_proposal --labels lazy --init W.npy --out DIR

Collects plain-search episodes of the 130^3 star scene with 96 orientations,
trains on all but the last tenth of them, reports ranking hit rates on the
held-out tenth, then packs held-out seeds with the plain search and with
proposals (at 130^3 and at 256^3), writing weights and a JSON report to DIR.
--labels lazy labels the steps for the lazy learned step instead
(learned.lazy_labels: rank the blocked candidates, target the ones reaching
the lowest z) and packs with lazy steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--episodes", type=int, default=200)
    ap.add_argument("--max-steps", type=int, default=100000)
    ap.add_argument("--every", type=int, default=1, help="record every n-th step of an episode")
    ap.add_argument("--threads", type=int, default=1, help="episodes collected concurrently")
    ap.add_argument("--incremental", action="store_true", help="collect with incremental steps (same choices)")
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--heat-weight", type=float, default=1.0)
    ap.add_argument("--top-o", type=int, default=8)
    ap.add_argument("--eval-top-o", default="8,12", help="top orientations tried when packing")
    ap.add_argument("--top-r", type=int, default=512)
    ap.add_argument("--eval-seeds", type=int, default=3)
    ap.add_argument("--eval-fine", type=int, default=2, help="held-out seeds packed at 256^3")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--seed-base", type=int, default=1000, help="episode seeds are seed-base + 0 .. episodes - 1")
    ap.add_argument("--skip-packing", action="store_true", help="train only (no packing comparison)")
    ap.add_argument("--out", default="proposal_run", help="directory for the weights, the sample bank and the report")
    ap.add_argument("--init", default="", help="start from these weights (.npy)")
    ap.add_argument("--labels", choices=("plain", "lazy"), default="plain")
    ap.add_argument("--prior-weight", type=float, default=2.0, help="lazy packing: weight of the orientation prior")
    ap.add_argument("--eval-seed-base", type=int, default=9000, help="first held-out seed packed")
    args = ap.parse_args(argv)
    if args.labels == "lazy":
        return main_lazy(args)

    from spectrapack_amd import Context, Proposal
    from spectrapack_amd import learned as L
    from spectrapack_amd import workloads as W

    def log(msg):
        print(msg, flush=True)

    os.makedirs(args.out, exist_ok=True)
    ctx = Context(0)
    prop = Proposal(ctx, max_batch=args.batch)
    capacity = args.episodes * (400 // args.every + 2)
    prop.reserve(capacity)
    if args.init:
        flat = np.load(args.init, allow_pickle=False)
    else:
        flat = L.initial_parameters(args.seed)
    prop.set_parameters(flat)
    init_flat = prop.parameters()

    seeds = [args.seed_base + i for i in range(args.episodes)]
    t0 = time.perf_counter()
    data = L.collect(ctx, prop, seeds, max_steps=args.max_steps, log=log, every=args.every,
                     incremental=args.incremental, threads=args.threads, log_every=max(1, args.episodes // 40))
    t_collect = time.perf_counter() - t0
    n = len(data["slot"])
    held = max(1, args.episodes // 10)
    test_idx = np.nonzero(data["episode"] >= args.episodes - held)[0]
    train_idx = np.nonzero(data["episode"] < args.episodes - held)[0]
    log(f"[data] samples={n} train={len(train_idx)} held-out={len(test_idx)} collect={t_collect:.1f}s")
    before = L.ranking_hits(prop, data, test_idx, args.top_o, args.top_r)
    log(f"[rank] untrained {before}")
    t0 = time.perf_counter()
    hist = L.train(prop, data, train_idx, epochs=args.epochs, batch=args.batch, lr=args.lr,
                   heat_weight=args.heat_weight, seed=args.seed, log=log)
    t_train = time.perf_counter() - t0
    after = L.ranking_hits(prop, data, test_idx, args.top_o, args.top_r)
    train_hits = L.ranking_hits(prop, data, train_idx[:: max(1, len(train_idx) // 2000)], args.top_o, args.top_r)
    log(f"[rank] trained held-out {after} train {train_hits}")
    flat = prop.parameters()
    meta = dict(episodes=args.episodes, samples=int(n), train_samples=int(len(train_idx)),
                heldout_samples=int(len(test_idx)), epochs=args.epochs, batch=args.batch, lr=args.lr,
                heat_weight=args.heat_weight, collect_seconds=t_collect, train_seconds=t_train,
                history=hist, hits_untrained=before, hits_trained=after, hits_train=train_hits,
                top_o=args.top_o, top_r=args.top_r, scene="star512-in-box128-h1-o96",
                catalog_seed=W.PROPOSAL_CATALOG_SEED, episode_seeds=[seeds[0], seeds[-1]], every=args.every,
                incremental=args.incremental, init=os.path.basename(args.init) if args.init else None)
    L.save_weights(os.path.join(args.out, "proposal_star96.npy"), flat, meta)

    # packing: plain vs learned (and the untrained network) on held-out seeds
    report = dict(meta=meta, packing=[])
    if args.skip_packing:
        with open(os.path.join(args.out, "report.json"), "w") as f:
            json.dump(report, f, indent=1)
        log("[done]")
        return 0
    untrained = Proposal(ctx, max_batch=1)
    untrained.set_parameters(init_flat)
    for fine in (False, True):
        count = args.eval_fine if fine else args.eval_seeds
        for k in range(count):
            seed = 9000 + k
            w = W.star_so3_fine_box(seed) if fine else W.star_so3_box(seed)
            sc = L.Scene(ctx, w)
            plain = L.run_episode(sc)
            rnd = L.run_episode(sc, untrained, args.top_o, args.top_r)
            heights = L.orientation_heights(sc.kernels)
            static = L.run_episode_static(sc, [int(o) for o in np.argsort(-heights, kind="stable")[:args.top_o]])
            row = dict(workload=w.name, seed=seed, plain=plain, untrained=rnd, static_heights=static, learned={})
            log(f"[pack] {w.name} seed={seed} static top-{args.top_o} by height {static['parts']} parts "
                f"{static['steps_per_s']:.1f} steps/s fallbacks {static['fallbacks']}")
            for top_o in [int(v) for v in args.eval_top_o.split(",")]:
                learned = L.run_episode(sc, prop, top_o, args.top_r)
                row["learned"][str(top_o)] = dict(learned, parts_ratio=learned["parts"] / max(plain["parts"], 1),
                                                  speedup=learned["steps_per_s"] / plain["steps_per_s"])
                log(f"[pack] {w.name} seed={seed} plain {plain['parts']} parts {plain['steps_per_s']:.1f} steps/s | "
                    f"learned top-{top_o} {learned['parts']} parts {learned['steps_per_s']:.1f} steps/s fallbacks "
                    f"{learned['fallbacks']} | untrained {rnd['parts']} parts fallbacks {rnd['fallbacks']}")
            report["packing"].append(row)
            sc.close()
            with open(os.path.join(args.out, "report.json"), "w") as f:
                json.dump(report, f, indent=1)
    log("[done]")
    return 0


def main_lazy(args):
    """Training on lazy-step labels (see the module docstring)."""
    from spectrapack_amd import Context, Proposal
    from spectrapack_amd import learned as L
    from spectrapack_amd import workloads as W

    def log(msg):
        print(msg, flush=True)

    os.makedirs(args.out, exist_ok=True)
    ctx = Context(0)
    prop = Proposal(ctx, max_batch=args.batch)
    prop.reserve(args.episodes * 150)
    flat = np.load(args.init, allow_pickle=False) if args.init else L.initial_parameters(args.seed)
    prop.set_parameters(flat)
    init_flat = prop.parameters()
    seeds = [args.seed_base + i for i in range(args.episodes)]
    t0 = time.perf_counter()
    data = L.collect_lazy(ctx, prop, seeds, max_steps=args.max_steps, log=log, threads=args.threads,
                          log_every=max(1, args.episodes // 40))
    t_collect = time.perf_counter() - t0
    n = len(data["slot"])
    held = max(1, args.episodes // 10)
    test_idx = np.nonzero(data["episode"] >= args.episodes - held)[0]
    train_idx = np.nonzero(data["episode"] < args.episodes - held)[0]
    # the prior of each sample's part (its scene's orientation prior)
    priors = {}
    prior = np.zeros((n, 96), np.float32)
    for i in range(n):
        e = int(data["episode"][i])
        if e not in priors:
            sc = L.Scene(ctx, W.star_so3_box(seeds[e]))
            priors[e] = sc.prior
            sc.close()
        prior[i] = priors[e]
    log(f"[data] samples={n} train={len(train_idx)} held-out={len(test_idx)} collect={t_collect:.1f}s")
    hits = {}
    for name, w in (("before", args.prior_weight), ("before_no_prior", 0.0)):
        hits[name] = L.lazy_hits(prop, data, test_idx, args.top_o, prior, w)
    t0 = time.perf_counter()
    hist = L.train(prop, data, train_idx, epochs=args.epochs, batch=args.batch, lr=args.lr,
                   heat_weight=args.heat_weight, seed=args.seed, log=log)
    t_train = time.perf_counter() - t0
    hits["after"] = L.lazy_hits(prop, data, test_idx, args.top_o, prior, args.prior_weight)
    hits["after_no_prior"] = L.lazy_hits(prop, data, test_idx, args.top_o, prior, 0.0)
    hits["prior_only"] = L.lazy_hits(prop, data, test_idx, args.top_o, prior * 1e6, 1.0)
    log(f"[rank] top-{args.top_o} held-out hits {hits}")
    meta = dict(labels="lazy", episodes=args.episodes, samples=int(n), train_samples=int(len(train_idx)),
                heldout_samples=int(len(test_idx)), epochs=args.epochs, batch=args.batch, lr=args.lr,
                heat_weight=args.heat_weight, collect_seconds=t_collect, train_seconds=t_train, history=hist,
                lazy_hits=hits, top_o=args.top_o, prior_weight=args.prior_weight,
                scene="star512-in-box128-h1-o96", catalog_seed=W.PROPOSAL_CATALOG_SEED,
                episode_seeds=[seeds[0], seeds[-1]], init=os.path.basename(args.init) if args.init else None)
    L.save_weights(os.path.join(args.out, "proposal_star96.npy"), prop.parameters(), meta)
    report = dict(meta=meta, packing=[])
    if not args.skip_packing:
        before = Proposal(ctx, max_batch=1)
        before.set_parameters(init_flat)
        for k in range(args.eval_fine):
            seed = args.eval_seed_base + k
            sc = L.Scene(ctx, W.star_so3_fine_box(seed))
            plain = L.run_episode(sc)
            lazy = L.run_episode_lazy(sc, prop, args.top_o, args.prior_weight)
            lazy0 = L.run_episode_lazy(sc, before, args.top_o, args.prior_weight)
            log(f"[pack] seed={seed} plain {plain['parts']} | lazy trained {lazy['parts']} "
                f"{lazy['steps_per_s'] / plain['steps_per_s']:.2f}x | lazy before {lazy0['parts']}")
            report["packing"].append(dict(seed=seed, plain=plain, lazy=lazy, lazy_before=lazy0))
            sc.close()
    with open(os.path.join(args.out, "report.json"), "w") as f:
        json.dump(report, f, indent=1)
    log("[done]")
    return 0


if __name__ == "__main__":
    sys.exit(main())
