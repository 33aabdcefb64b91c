"""Learned proposal pipeline: plain-search episodes -> training samples ->
training on the GPU -> packing with proposals, compared with plain search.

Everything that computes runs on the device through the C ABI
(sp_hip_session_*, sp_hip_proposal_*); this module holds the bookkeeping:
the part descriptor, the labels taken from plain-search traces, the training
schedule and the episode loops.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import Context, Proposal, Session
from . import workloads as W

REGIONS = 16
DESCRIPTOR = 32
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "proposal_star96.npy")


SH_DEGREE = 4  # (SH_DEGREE + 1)^2 = 25 coefficients


def real_spherical_harmonics(dirs, degree=SH_DEGREE):
    """Orthonormal real spherical harmonics Y_lm (l <= degree) at unit
    vectors dirs [n, 3], columns ordered by l then m = -l..l."""
    from scipy.special import sph_harm_y
    d = np.asarray(dirs, np.float64)
    theta = np.arccos(np.clip(d[:, 2], -1.0, 1.0))
    phi = np.arctan2(d[:, 1], d[:, 0])
    cols = []
    for l in range(degree + 1):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            cols.append(np.sqrt(2.0) * (y.imag if m < 0 else y.real) if m != 0 else y.real)
    return np.stack(cols, axis=1)


def down_directions(catalog):
    """Per orientation, the part-frame direction that the rotation maps to
    world +z (the third row of its rotation matrix)."""
    return np.array([W.rotation_matrix(q)[2] for q in catalog])


def orientation_heights(kernels):
    """Per oriented kernel: mean height of its occupied cells above its lowest
    occupied layer, in cells (high for orientations resting on a point)."""
    out = []
    for k in kernels:
        z = np.nonzero(np.asarray(k["cells"]) != 0)[0]
        out.append(float((z - z.min()).mean()) if z.size else 0.0)
    return np.array(out)


def orientation_prior(kernels, valid):
    """The orientation prior of a learned step (96 entries): per fitting
    orientation the negated standardised mean resting height of its kernel
    (orientation_heights), so that orientations resting low rank high; 0 for
    the others."""
    h = orientation_heights(kernels)
    v = np.asarray(valid[:len(kernels)], bool)
    out = np.zeros(96, np.float32)
    if v.any():
        sd = float(h[v].std())
        vals = np.zeros(len(kernels), np.float32)
        vals[v] = -(h[v] - h[v].mean()) / (sd if sd > 0 else 1.0)
        out[:len(kernels)] = vals
    return out


def descriptor(kernels, catalog, env_shape):
    """32 floats describing a part through its oriented kernels (cells
    [z, y, x], shape (x, y, z)) and the catalog that produced them:
    [0] rms radius of the first kernel's occupied cells about their centroid,
        in units of the 64^3 network input, / 8;
    [1:26] least-squares coefficients, over the catalog's down directions, of
        the real spherical harmonics of degree <= 4 fitted to the standardised
        per-orientation mean heights (orientation_heights);
    [26:29] eigenvalues of the first kernel's cell covariance / rms^2 (sorted);
    [29] mean height / rms; [30] height spread / rms; [31] fill fraction of
    the first kernel's box.  Each orientation's fitted height is linear in
    [1:26]."""
    d = np.zeros(DESCRIPTOR, np.float64)
    k0 = kernels[0]
    c = np.asarray(k0["cells"]) != 0
    z, y, x = np.nonzero(c)
    if len(x) == 0:
        return d.astype(np.float32)
    scale = 64.0 / np.asarray(env_shape, np.float64)
    p = np.stack([x, y, z], axis=1).astype(np.float64) + 0.5
    q = p - p.mean(axis=0)
    rms_cells = float(np.sqrt((q * q).sum(axis=1).mean()))
    d[0] = float(np.sqrt(((q * scale) ** 2).sum(axis=1).mean())) / 8.0
    h = orientation_heights(kernels)
    mu, sd = float(h.mean()), float(h.std())
    hs = (h - mu) / sd if sd > 0 else np.zeros_like(h)
    B = real_spherical_harmonics(down_directions(catalog[:len(kernels)]))
    coef, *_ = np.linalg.lstsq(B, hs, rcond=None)
    d[1:26] = coef
    if rms_cells > 0:
        ev = np.sort(np.linalg.eigvalsh(q.T @ q / len(q)))
        d[26:29] = ev / rms_cells ** 2
        d[29] = mu / rms_cells
        d[30] = sd / rms_cells
    d[31] = len(x) / float(np.prod(k0["shape"]))
    return d.astype(np.float32)


# Flat parameter layout of the network (sp_hip_proposal_set_parameters).
LAYOUT = [("w1", (32, 27)), ("b1", (32,)), ("w2", (32, 27, 32)), ("b2", (32,)), ("w3", (32, 27, 32)), ("b3", (32,)),
          ("w4", (32, 27, 32)), ("b4", (32,)), ("wf", (32, 32)), ("wh", (32,)), ("bh", (1,)), ("wa", (64, 96)),
          ("ba", (64,)), ("wo", (96, 64)), ("bo", (96,))]
FAN_IN = {"w1": 27, "w2": 864, "w3": 864, "w4": 864, "wf": 32, "wh": 32, "wa": 96, "wo": 64}


def initial_parameters(seed):
    """He-normal weights (descriptor weights scaled by 0.1), zero biases."""
    rs = np.random.default_rng(seed)
    parts = []
    for name, shape in LAYOUT:
        if name.startswith("b"):
            parts.append(np.zeros(int(np.prod(shape))))
            continue
        w = rs.normal(0.0, np.sqrt(2.0 / FAN_IN[name]), int(np.prod(shape)))
        parts.append(w * (0.1 if name == "wf" else 1.0))
    return np.concatenate(parts).astype(np.float32)


def region_index(u_xyz, env_shape):
    """Region (16 per axis, X-fastest) of kernel corner u."""
    r = [int(u_xyz[a]) * REGIONS // int(env_shape[a]) for a in range(3)]
    return r[0] + REGIONS * (r[1] + REGIONS * r[2])


def orientation_target(minima, chosen):
    """Target over orientations from one plain step: uniform over the
    orientations whose lowest feasible corner reaches the step's lowest z
    (the chosen orientation is one of them).  A proposal that contains any of
    them places the part at the same height as the plain search."""
    t = np.zeros(96, np.float32)
    zs = [(m[0], o) for o, m in enumerate(minima) if m is not None]
    if not zs or chosen < 0:
        return t
    zmin = min(z for z, _ in zs)
    tied = [o for z, o in zs if z == zmin]
    assert chosen in tied
    t[tied] = 1.0 / len(tied)
    return t


class Scene:
    """A workload's container and kernels on the device, and a session over them."""

    def __init__(self, ctx: Context, w: W.Workload, capture_graph=True):
        self.w = w
        if w.container is not None:
            self.env0 = ctx.mesh_container(w.container.vertices, w.container.triangles, w.env_first, w.env_shape,
                                           w.pitch)
        else:
            self.env0 = ctx.box_container(w.env_first, w.env_shape, w.pitch, w.box_dims)
        self.kernels = ctx.voxelize_objects(w.part.vertices, w.part.triangles, w.catalog, w.pitch)
        prec = 32 if w.precision == "fp32" else 0
        self.session = Session(ctx, w.env_shape, w.env_first, len(self.kernels), precision=prec,
                               capture_graph=capture_graph)
        for i, k in enumerate(self.kernels):
            self.session.set_kernel(i, k["cells"], k["shape"], k["first"])
        self.desc = descriptor(self.kernels, w.catalog, w.env_shape)
        es = w.env_shape
        self.fits = np.array([all(k["shape"][a] <= es[a] for a in range(3)) for k in self.kernels] +
                             [False] * (96 - len(self.kernels)), np.uint8)
        self.prior = orientation_prior(self.kernels, self.fits)
        self.reset()

    def reset(self):
        self.session.set_active()
        self.session.set_environment(self.env0)

    def close(self):
        self.session.close()


def collect(ctx, proposal: Proposal, seeds, max_steps=100000, log=print, first_slot=0, every=1, incremental=False,
            threads=1, log_every=1):
    """Plain-search episodes (stamping every choice) recorded as samples in
    the proposal's bank: every `every`-th step is recorded (its input before
    the step, its choice and every orientation's lowest feasible corner); the
    steps between run without reading results back.  incremental: the
    episodes run incremental steps (identical choices).  threads: episodes
    run concurrently on that many sessions (one stream each).  Returns the
    labels and per-episode statistics."""
    import threading
    lock = threading.Lock()
    state = dict(slot=first_slot, next=0, done=0, t0=time.perf_counter())
    rec = dict(slot=[], desc=[], target=[], mask=[], region=[], orientation=[], episode=[], tied=[])
    episodes = [None] * len(seeds)
    errors = []

    def take_slot():
        with lock:
            if state["slot"] >= proposal.samples:
                raise RuntimeError("sample bank full")
            state["slot"] += 1
            return state["slot"] - 1

    local = threading.local()

    def episode(e, seed):
        t0 = time.perf_counter()
        # one context per thread: a context serialises its own calls
        if threads > 1 and not hasattr(local, "ctx"):
            local.ctx = Context(ctx.device)
        sc = Scene(local.ctx if threads > 1 else ctx, W.star_so3_box(int(seed)))
        s = sc.session
        s.set_autotune(False)  # every tile gives identical results
        if incremental:
            s.set_incremental(True)
        steps = 0
        rows = []
        while steps < max_steps:
            gap = min(every - 1, max_steps - steps - 1)
            if gap > 0:
                s.run(gap, stamp=True)
                steps += gap
            slot = take_slot()
            proposal.record(s, slot)
            r = s.step(stamp=True)
            steps += 1
            if r["orientation"] < 0:
                break
            minima = s.orientation_minima()
            u = tuple(r["key"][a] - sc.w.env_first[a] for a in range(3))
            zmin = min(m[0] for m in minima if m is not None)
            rows.append((slot, orientation_target(minima, r["orientation"]), region_index(u, sc.w.env_shape),
                         r["orientation"], np.array([m is not None and m[0] == zmin for m in minima] +
                                                    [False] * (96 - len(minima)), bool)))
        parts = sum(1 for c in s.copies())
        dt = time.perf_counter() - t0
        with lock:
            for slot, target, region, o, tied in rows:
                rec["slot"].append(slot)
                rec["desc"].append(sc.desc)
                rec["target"].append(target)
                rec["mask"].append(sc.fits)
                rec["region"].append(region)
                rec["orientation"].append(o)
                rec["episode"].append(e)
                rec["tied"].append(tied)
            episodes[e] = dict(seed=int(seed), parts=parts, seconds=dt, samples=len(rows))
            state["done"] += 1
            if state["done"] % log_every == 0 or state["done"] == len(seeds):
                log(f"[collect] {state['done']}/{len(seeds)} episodes, seed={seed} parts={parts} "
                    f"samples={state['slot'] - first_slot} {time.perf_counter() - state['t0']:.1f}s")
        sc.close()

    def worker():
        while True:
            with lock:
                e = state["next"]
                state["next"] += 1
            if e >= len(seeds) or errors:
                return
            try:
                episode(e, seeds[e])
            except Exception as ex:  # reported after the join
                errors.append(ex)
                return

    pool = [threading.Thread(target=worker) for _ in range(max(1, threads))]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    if errors:
        raise errors[0]
    order = np.lexsort((np.asarray(rec["slot"]), np.asarray(rec["episode"])))
    out = {k: np.asarray(v)[order] for k, v in rec.items()}
    out["episodes"] = episodes
    return out


def lazy_labels(prev, cur):
    """Labels of one step for the lazy learned step, from the plain search's
    per-orientation minima before (prev, the last step) and after (cur) the
    last placement (None: no feasible corner).  An orientation's minimum
    changes only when the placement blocked it, and then the old minimum
    bounds the new one from below; the orientations whose minimum is
    unchanged are current.  Candidates: blocked orientations whose bound
    lies below the lowest current minimum (a lazy step ranks and searches
    them).  Returns (candidate mask, target), the target uniform over the
    candidates whose new minimum reaches the step's lowest z, or None when
    the ranking cannot change the step's height (a current minimum lies at
    the lowest z, or no candidate reaches it)."""
    def key(m, o):
        return (m[0], m[1], m[2], o)
    n = len(cur)
    best = None
    for o in range(n):
        if prev[o] is not None and cur[o] == prev[o]:
            k = key(cur[o], o)
            best = k if best is None or k < best else best
    cand = np.zeros(96, bool)
    for o in range(n):
        if prev[o] is None or cur[o] == prev[o]:
            continue  # infeasible stays infeasible; unchanged minima are current
        if best is None or key(prev[o], o) < best:
            cand[o] = True
    zs = [m[0] for m in cur if m is not None]
    if not zs or not cand.any():
        return cand, None
    zmin = min(zs)
    if best is not None and best[0] <= zmin:
        return cand, None  # a current minimum already lies at the lowest z
    reach = np.array([bool(cand[o]) and cur[o] is not None and cur[o][0] == zmin for o in range(n)] +
                     [False] * (96 - n))
    if not reach.any():
        return cand, None
    t = np.zeros(96, np.float32)
    t[reach] = 1.0 / reach.sum()
    return cand, t


def collect_lazy(ctx, proposal: Proposal, seeds, max_steps=100000, log=print, first_slot=0, threads=1, log_every=1,
                 min_candidates=2):
    """Plain-search episodes (incremental steps, identical choices) labelled
    for the lazy learned step (lazy_labels): the occupancy before each step
    is recorded, and kept as a sample when at least min_candidates
    orientations are candidates and one of them reaches the step's lowest z;
    the sample's orientation mask is the candidate set.  Returns the labels
    and per-episode statistics, as collect does."""
    import threading
    lock = threading.Lock()
    state = dict(slot=first_slot, next=0, done=0, t0=time.perf_counter())
    rec = dict(slot=[], desc=[], target=[], mask=[], region=[], orientation=[], episode=[], tied=[])
    episodes = [None] * len(seeds)
    errors = []
    local = threading.local()

    def take_slot():
        with lock:
            if state["slot"] >= proposal.samples:
                raise RuntimeError("sample bank full")
            state["slot"] += 1
            return state["slot"] - 1

    def episode(e, seed):
        t0 = time.perf_counter()
        if threads > 1 and not hasattr(local, "ctx"):
            local.ctx = Context(ctx.device)
        sc = Scene(local.ctx if threads > 1 else ctx, W.star_so3_box(int(seed)))
        s = sc.session
        s.set_autotune(False)
        s.set_incremental(True)
        rows = []
        prev = None
        slot = None
        steps = 0
        while steps < max_steps:
            if slot is None:
                slot = take_slot()
            proposal.record(s, slot)
            r = s.step(stamp=True)
            steps += 1
            if r["orientation"] < 0:
                break
            cur = s.orientation_minima()
            if prev is not None:
                cand, target = lazy_labels(prev, cur)
                if target is not None and cand.sum() >= min_candidates:
                    u = tuple(r["key"][a] - sc.w.env_first[a] for a in range(3))
                    rows.append((slot, target, cand.astype(np.uint8), region_index(u, sc.w.env_shape),
                                 r["orientation"], target > 0))
                    slot = None  # kept; the next step takes a new slot
            prev = cur
        parts = sum(1 for c in s.copies())
        dt = time.perf_counter() - t0
        with lock:
            for slot_, target, mask, region, o, tied in rows:
                rec["slot"].append(slot_)
                rec["desc"].append(sc.desc)
                rec["target"].append(target)
                rec["mask"].append(mask)
                rec["region"].append(region)
                rec["orientation"].append(o)
                rec["episode"].append(e)
                rec["tied"].append(tied)
            episodes[e] = dict(seed=int(seed), parts=parts, seconds=dt, samples=len(rows))
            state["done"] += 1
            if state["done"] % log_every == 0 or state["done"] == len(seeds):
                log(f"[collect-lazy] {state['done']}/{len(seeds)} episodes, seed={seed} parts={parts} "
                    f"samples={len(rec['slot'])} {time.perf_counter() - state['t0']:.1f}s")
        sc.close()

    def worker():
        while True:
            with lock:
                e = state["next"]
                state["next"] += 1
            if e >= len(seeds) or errors:
                return
            try:
                episode(e, seeds[e])
            except Exception as ex:  # reported after the join
                errors.append(ex)
                return

    pool = [threading.Thread(target=worker) for _ in range(max(1, threads))]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    if errors:
        raise errors[0]
    order = np.lexsort((np.asarray(rec["slot"]), np.asarray(rec["episode"])))
    out = {k: np.asarray(v)[order] for k, v in rec.items()}
    out["episodes"] = episodes
    return out


def lazy_hits(proposal: Proposal, data, idx, top_o=12, prior=None, prior_weight=0.0, batch=32):
    """Fraction of lazy-labelled samples whose top_o candidates by score
    (logit + prior_weight x prior; prior: [n][96] or None) contain a
    candidate that reaches the step's lowest z."""
    hits = 0
    for i in range(0, len(idx), batch):
        j = idx[i:i + batch]
        logits, _ = proposal.forward(data["slot"][j], data["desc"][j])
        for k, s in enumerate(j):
            sc = logits[k] + (prior_weight * prior[s] if prior is not None else 0.0)
            sc = np.where(data["mask"][s] != 0, sc, -np.inf)
            top = np.argsort(-sc, kind="stable")[:top_o]
            hits += bool(data["tied"][s][top].any())
    return hits / max(len(idx), 1)


def ranking_hits(proposal: Proposal, data, idx, top_o=8, top_r=512, batch=32):
    """Fractions of samples whose plain choice lies in the proposal: chosen
    orientation among the top_o valid logits, chosen region among the top_r
    heat logits, both, and some proposed orientation reaching the step's
    lowest z."""
    ho = hr = hb = hz = 0
    for i in range(0, len(idx), batch):
        j = idx[i:i + batch]
        logits, heat = proposal.forward(data["slot"][j], data["desc"][j])
        for k, s in enumerate(j):
            lg = np.where(data["mask"][s] != 0, logits[k], -np.inf)
            to = np.argsort(-lg, kind="stable")[:top_o]
            tr = np.argsort(-heat[k], kind="stable")[:top_r]
            a = data["orientation"][s] in to
            b = data["region"][s] in tr
            ho += a
            hr += b
            hb += a and b
            hz += bool(data["tied"][s][to].any())
    n = max(len(idx), 1)
    return dict(orientation=ho / n, region=hr / n, both=hb / n, lowest_z=hz / n)


def train(proposal: Proposal, data, train_idx, epochs=6, batch=32, lr=1e-3, heat_weight=1.0, seed=0, log=print):
    rs = np.random.default_rng(seed)
    n = len(train_idx)
    iters = epochs * ((n + batch - 1) // batch)
    it = 0
    history = []
    for ep in range(epochs):
        order = rs.permutation(train_idx)
        lo_sum = lh_sum = 0.0
        nb = 0
        t0 = time.perf_counter()
        for i in range(0, n, batch):
            j = order[i:i + batch]
            cur = lr * 0.5 * (1.0 + np.cos(np.pi * it / iters))
            (lo, lh), _ = proposal.train(data["slot"][j], data["desc"][j], data["target"][j], data["mask"][j],
                                         data["region"][j], heat_weight=heat_weight, lr=cur, update=True)
            lo_sum += lo
            lh_sum += lh
            nb += 1
            it += 1
        history.append(dict(epoch=ep, loss_orientation=lo_sum / nb, loss_region=lh_sum / nb,
                            seconds=time.perf_counter() - t0))
        log(f"[train] epoch {ep + 1}/{epochs} loss_o={lo_sum / nb:.3f} loss_r={lh_sum / nb:.3f} "
            f"{time.perf_counter() - t0:.1f}s")
    return history


def run_episode(scene: Scene, proposal: Proposal | None = None, top_o=8, top_r=512, max_steps=100000,
                verify_z=False, prior_weight=0.0):
    """Packs until nothing fits: the plain search, or proposals with the full
    search as fallback.  Every choice is stamped.  Returns parts, steps,
    seconds and fallbacks; steps include the final one that finds nothing.

    prior_weight: proposals rank orientations by logit + prior_weight x the
    scene's orientation prior (orientation_prior).

    verify_z: a proposal's choice is kept only when it lies at the lowest z
    any step can still reach, else the step is the full search.  The lowest
    key over every orientation never decreases (occupancy only grows), so the
    z of the last full search's choice bounds every later choice from below;
    a proposal at that z is at the plain rule's z.  The full search's choice
    raises the bound."""
    s = scene.session
    ez = scene.w.env_first[2]
    if proposal is not None:
        proposal.set_prior(scene.prior, prior_weight)
    # warm-up outside the timed loop: tuning and graph capture of both modes
    scene.reset()
    if proposal is not None:
        proposal.step(s, scene.desc, top_orientations=top_o, top_regions=top_r, stamp=not verify_z,
                      fallback=False)
    s.step(stamp=True)
    scene.reset()
    parts = steps = fallbacks = 0
    lb_z = -1
    t0 = time.perf_counter()
    while steps < max_steps:
        if proposal is None:
            r = s.step(stamp=True)
        elif not verify_z:
            r = proposal.step(s, scene.desc, top_orientations=top_o, top_regions=top_r, stamp=True, fallback=True)
            fallbacks += r["fallback"]
        else:
            r = proposal.step(s, scene.desc, top_orientations=top_o, top_regions=top_r, stamp=False, fallback=False)
            if r["orientation"] >= 0 and r["key"][2] - ez <= lb_z:
                s.stamp(r["orientation"], r["translation"])
            else:
                r = s.step(stamp=True)
                fallbacks += 1
                if r["orientation"] >= 0:
                    lb_z = r["key"][2] - ez
        steps += 1
        if r["orientation"] < 0:
            break
        parts += 1
    dt = time.perf_counter() - t0
    return dict(parts=parts, steps=steps, seconds=dt, steps_per_s=steps / dt, fallbacks=fallbacks)


def run_episode_lazy(scene: Scene, proposal: Proposal, top_o=16, prior_weight=0.0, max_steps=100000):
    """Packs until nothing fits with lazy learned steps (Proposal.lazy_step):
    every orientation's last exact minimum is kept while its corner stays
    free, and of the orientations the new occupancy blocked, the top_o by
    score are searched.  Returns parts, steps, seconds, the full searches and
    the mean number of orientations searched per step."""
    s = scene.session
    proposal.set_prior(scene.prior, prior_weight)
    # warm-up outside the timed loop: tuning and graph capture of both modes
    scene.reset()
    proposal.lazy_reset()
    for _ in range(2):
        proposal.lazy_step(s, scene.desc, top_o, stamp=True)
    scene.reset()
    proposal.lazy_reset()
    parts = steps = full = searched = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        r = proposal.lazy_step(s, scene.desc, top_o, stamp=True)
        full += r["used_full"]
        searched += r["searched"]
        steps += 1
        if r["orientation"] < 0:
            break
        parts += 1
    dt = time.perf_counter() - t0
    proposal.lazy_reset()
    return dict(parts=parts, steps=steps, seconds=dt, steps_per_s=steps / dt, full_searches=full,
                searched_per_step=searched / max(steps, 1))


def run_episode_static(scene: Scene, orientations, max_steps=100000):
    """Packs with a fixed orientation subset (full search as fallback): the
    proposal without a network."""
    s = scene.session
    scene.reset()
    s.set_active(orientations)
    s.step(stamp=True)
    s.set_active()
    s.step(stamp=True)
    scene.reset()
    parts = steps = fallbacks = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        s.set_active(orientations)
        r = s.step(stamp=True)
        if r["orientation"] < 0:
            s.set_active()
            r = s.step(stamp=True)
            fallbacks += 1
        steps += 1
        if r["orientation"] < 0:
            break
        parts += 1
    s.set_active()
    dt = time.perf_counter() - t0
    return dict(parts=parts, steps=steps, seconds=dt, steps_per_s=steps / dt, fallbacks=fallbacks)


def save_weights(path, flat, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(flat, np.float32), allow_pickle=False)
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)


def load_weights(path=DEFAULT_WEIGHTS):
    return np.load(path, allow_pickle=False)
