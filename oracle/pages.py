"""Ranked candidate pages restated in numpy -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, as the checker; the product path never does.

  page_scores         spectral_pipeline.cpp:1795-1801 (score of a legal translation)
  ranked_page         spectral_pipeline.cpp:1740-1824 (order, filter, cursor, page of 32)
                      + 1825-1855 (direct recheck of ordinary pages)
  proximity_direct    correlate_proximity_cpu (correlation.cpp:650-655) as a direct
                      FP64 sum over the kernel's cells at the legal translations
  ranked_fill         spectral.cpp:1485-1590, the ordinary part of the greedy
                      trial loop: pages per orientation, best head by rank_less,
                      accept direct-zero heads, refill consumed pages by cursor

Pins (tests/test_oracle_pages.py): ranked_page against a literal sort with the
reference's comparator (spectral_pipeline.cpp:1749-1766) on inputs with forced
ties, cursor pages against slices of the full order, page_scores against a
hand-computed known answer, proximity_direct against the FFT restatement of
correlate_proximity_cpu.
"""
from __future__ import annotations

import numpy as np

PAGE_SIZE = 32


def page_scores(prox, mass, origin_z, pitch, container_min_z, container_max_z, object_max_z, first_z):
    """Score of every legal translation of one orientation.

    prox: proximity correlation at the legal corners, [lz, ly, lx]; first_z:
    the translation z of corner 0 (environment first z - kernel first z).
    Evaluated in the pipeline's order: world_z = origin + pitch * t_z;
    top = (world_z + object_max_z - container_min_z) / height; mean =
    prox / mass; score = .65 * top - .35 * mean."""
    prox = np.asarray(prox, dtype=np.float64)
    lz = prox.shape[0]
    tz = (np.int64(first_z) + np.arange(lz, dtype=np.int64)).astype(np.float64)
    world_z = np.float64(origin_z) + np.float64(pitch) * tz
    height = np.float64(container_max_z) - np.float64(container_min_z)
    top = (world_z + np.float64(object_max_z) - np.float64(container_min_z)) / height
    mean = prox / np.float64(mass)
    return np.float64(.65) * top[:, None, None] - np.float64(.35) * mean


def _direct_zero(env, kernel_cells, u):
    kz, ky, kx = kernel_cells.shape
    win = env[u[2]:u[2] + kz, u[1]:u[1] + ky, u[0]:u[0] + kx]
    return not bool(np.any(win[kernel_cells != 0]))


def ranked_page(overlap, scores, translation_first, mode=0, cursor=None, excluded=(), env=None, kernel_cells=None,
                size=PAGE_SIZE):
    """One orientation's page.

    overlap: exact binary overlap per legal corner [lz, ly, lx]; scores: the
    same shape; translation_first (x, y, z) of corner 0.  mode 0 (ordinary)
    keeps zero overlaps and orders by (score, z, y, x); mode 1 (low overlap)
    orders by (overlap, score, z, y, x).  cursor: a candidate dict -- only
    candidates strictly after it.  excluded: translations never offered.
    With env and kernel_cells, ordinary entries get the direct recheck."""
    ov = np.asarray(overlap).astype(np.int64).ravel()
    sc = np.asarray(scores, dtype=np.float64).ravel()
    lz, ly, lx = np.asarray(overlap).shape
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    z, y, x = z.ravel(), y.ravel(), x.ravel()
    f = [int(v) for v in translation_first]
    tx, ty, tz = x + f[0], y + f[1], z + f[2]
    keep = ov == 0 if mode == 0 else np.ones(ov.size, bool)
    for t in excluded:
        keep &= ~((tx == t[0]) & (ty == t[1]) & (tz == t[2]))
    okey = ov if mode == 1 else np.zeros_like(ov)
    if cursor is not None:
        co = int(cursor["binary_overlap"]) if mode == 1 else 0
        cs = float(cursor["score"])
        ct = tuple(int(v) for v in cursor["translation"])
        later_pos = (tz > ct[2]) | ((tz == ct[2]) & ((ty > ct[1]) | ((ty == ct[1]) & (tx > ct[0]))))
        keep &= (okey > co) | ((okey == co) & ((sc > cs) | ((sc == cs) & later_pos)))
    idx = np.nonzero(keep)[0]
    order = idx[np.lexsort((tx[idx], ty[idx], tz[idx], sc[idx], okey[idx]))][:size]
    out = []
    for i in order:
        c = dict(translation=(int(tx[i]), int(ty[i]), int(tz[i])), score=float(sc[i]),
                 binary_overlap=float(ov[i]), direct_zero_overlap=False)
        if mode == 0 and env is not None:
            c["direct_zero_overlap"] = _direct_zero(env, np.asarray(kernel_cells), (int(x[i]), int(y[i]), int(z[i])))
        out.append(c)
    return out


def proximity_direct(field, kernel_cells):
    """Sum of the proximity field under the kernel's cells at every legal
    corner ([lz, ly, lx]; None when the kernel does not fit)."""
    field = np.asarray(field, dtype=np.float64)
    k = np.asarray(kernel_cells)
    legal = [field.shape[a] - k.shape[a] + 1 for a in range(3)]
    if min(legal) <= 0:
        return None
    acc = np.zeros(legal, np.float64)
    for cz, cy, cx in np.argwhere(k != 0):
        acc += field[cz:cz + legal[0], cy:cy + legal[1], cx:cx + legal[2]]
    return acc


def rank_less(a, b):
    """spectral_pipeline.cpp:1749-1759 (score, orientation, translation z, y, x)."""
    if a["score"] != b["score"]:
        return a["score"] < b["score"]
    if a["orientation"] != b["orientation"]:
        return a["orientation"] < b["orientation"]
    return tuple(a["translation"][::-1]) < tuple(b["translation"][::-1])


def ranked_fill(pages_of, stamp, orientations, max_copies=None, validate=None):
    """The ordinary greedy loop of spectral.cpp:1485-1590 over a page source.

    pages_of(list_of_orientations, cursors) -> {o: page}; stamp(o, t) places a
    copy (the occupancy changes, so every page is fetched again); validate(o,
    t) -> bool stands for the exact validation (all accepted when None).
    Returns the placements [(o, translation, score)]."""
    placed = []
    while max_copies is None or len(placed) < max_copies:
        pages = pages_of(list(orientations), {})
        index = {o: 0 for o in orientations}
        cursors = {}
        exhausted = {o: len(pages[o]) == 0 for o in orientations}
        admitted = False
        while True:
            refill = [o for o in orientations if not exhausted[o] and index[o] == len(pages[o])]
            short = [o for o in refill if len(pages[o]) < PAGE_SIZE]
            for o in short:
                exhausted[o] = True
            refill = [o for o in refill if len(pages[o]) == PAGE_SIZE]
            if refill:
                more = pages_of(refill, {o: cursors[o] for o in refill})
                for o in refill:
                    pages[o] = more[o]
                    index[o] = 0
                    exhausted[o] = len(more[o]) == 0
            best = None
            for o in orientations:
                if index[o] == len(pages[o]):
                    continue
                head = dict(pages[o][index[o]], orientation=o)
                if best is None or rank_less(head, best):
                    best = head
            if best is None:
                break
            o = best["orientation"]
            index[o] += 1
            cursors[o] = best
            if not best["direct_zero_overlap"]:
                continue
            if validate is not None and not validate(o, best["translation"]):
                continue
            stamp(o, best["translation"])
            placed.append((o, tuple(best["translation"]), best["score"]))
            admitted = True
            break
        if not admitted:
            break
    return placed
