"""Writes the step fixtures of the box and star workloads from the CPU oracle.

Run in a container that has the built oracle (oracle/Makefile):

    python tests/golden/make_fixtures.py box     # 10x6x4 box in the 64 mm box, to exhaustion
    python tests/golden/make_fixtures.py star    # 512-face star in the 128 mm box, 12 steps
    python tests/golden/make_fixtures.py bracket # L-bracket in the r=120 mm sphere, 256^3, 96 orientations, 3 steps

Each step of the plain greedy loop (engine/pack_solver/src/gap_fill.cpp:380-509:
every orientation's legal overlap field, the lowest (z, y, x) feasible corner,
earliest orientation on ties, then the chosen copy stamped) is computed by the
oracle's checked FP64 FFT restatement (star, bracket) or its direct integer path (box),
orientations spread over worker processes.  A fixture holds, per step, the
choice (orientation, translation) and the SHA-256 of every orientation's
int32 legal field; the GPU tests recompute both and compare.
"""
import hashlib
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "cpp-max-packer_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402
from spectrapack_amd import workloads as W  # noqa: E402

from fixture_hash import cells_digest, fields_digest  # noqa: E402

CONFIGS = {
    # name: (workload, oracle mode, steps (None: to exhaustion))
    "box": (W.box_in_box, True, None),
    "star": (W.star_in_box, False, 12),
    "bracket": (W.bracket_in_sphere, False, 3),
}

_STATE = {}


def _init(env_first, kernels, direct):
    _STATE.update(env_first=env_first, kernels=kernels, direct=direct)


def _fields(args):
    env, lo, hi = args
    _, _, f = O.place_step(env, _STATE["env_first"], _STATE["kernels"][lo:hi], direct=_STATE["direct"], fields=True)
    return lo, f


def choice(fields, env_shape):
    """The lowest (z, y, x) zero of every field, earliest orientation on ties."""
    best = None
    for o, f in enumerate(fields):
        if f is None:
            continue
        zs, ys, xs = np.nonzero(f == 0)
        if zs.size == 0:
            continue
        i = np.lexsort((xs, ys, zs))[0]
        cand = ((int(zs[i]), int(ys[i]), int(xs[i])), o)
        if best is None or cand < best:
            best = cand
    return best


def run(name, workers):
    make, direct, steps = CONFIGS[name]
    w = make()
    if w.container is not None:
        env = O.mesh_container(w.container.vertices, w.container.triangles, w.env_first, w.env_shape, w.pitch)
    else:
        env = O.box_container(w.env_first, w.env_shape, w.pitch, w.box_dims)
    kernels = []
    for q in w.catalog:
        f, s, c = O.voxelize_object(w.part.vertices, w.part.triangles, q, w.pitch)
        kernels.append(dict(first=tuple(int(v) for v in f), shape=tuple(int(v) for v in s), cells=c))
    out = dict(workload=w.name, pitch=w.pitch, env_first=list(w.env_first), env_shape=list(w.env_shape),
               box_dims=list(w.box_dims) if w.box_dims else None, orientations=len(kernels),
               environment_sha256=cells_digest(env),
               kernels=[dict(first=list(k["first"]), shape=list(k["shape"]), sha256=cells_digest(k["cells"]))
                        for k in kernels],
               oracle="direct integer" if direct else "checked FP64 FFT", steps=[])
    n = len(kernels)
    chunks = [(i * n // workers, (i + 1) * n // workers) for i in range(workers)]
    t0 = time.time()
    with mp.get_context("fork").Pool(workers, _init, (w.env_first, kernels, direct)) as pool:
        step = 0
        while steps is None or step < steps:
            parts = pool.map(_fields, [(env, lo, hi) for lo, hi in chunks if hi > lo])
            fields = [None] * n
            for lo, f in parts:
                fields[lo:lo + len(f)] = f
            best = choice(fields, w.env_shape)
            digest = fields_digest(fields)
            if best is None:
                out["steps"].append([-1, 0, 0, 0, digest])
                break
            (uz, uy, ux), o = best
            u = (ux, uy, uz)
            t = [int(w.env_first[a] + u[a] - kernels[o]["first"][a]) for a in range(3)]
            out["steps"].append([o] + t + [digest])
            env = O.stamp(env, w.env_first, kernels[o], t)
            step += 1
            if step % 25 == 0 or steps is not None:
                print(f"[{name}] step {step} ({time.time() - t0:.0f} s)", flush=True)
    out["final_environment_sha256"] = cells_digest(env)
    path = os.path.join(HERE, f"steps_{name}.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")
    print(f"[{name}] {len(out['steps'])} steps in {time.time() - t0:.0f} s -> {path}", flush=True)


if __name__ == "__main__":
    if not os.path.exists(O.LIB_PATH):
        O.build()
    for name in sys.argv[1:] or list(CONFIGS):
        run(name, int(os.environ.get("FIXTURE_WORKERS", "8")))
