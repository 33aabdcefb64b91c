"""The C++ host mirror (spectrapack::compute drop-in and the placement loop
with accept / reject / stop verdicts) against the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cpp-max-packer_amd", "spectrapack_amd", "spectrapack_host_test")


def test_cpp_correlation_dropin():
    print("[host] correlate checks", flush=True)
    r = subprocess.run([DRIVER, "correlate"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def write_scene(path, env, first, pitch, kernels, capacity, reject_every, stop_after, tail=None):
    with open(path, "w") as f:
        f.write(" ".join(str(v) for v in env.shape[::-1]) + "\n")
        f.write(" ".join(str(int(v)) for v in first) + f" {pitch!r}\n")
        f.write(" ".join(str(int(v)) for v in env.ravel()) + "\n")
        f.write(f"{len(kernels)}\n")
        for k in kernels:
            f.write(" ".join(repr(float(v)) for v in k["q"]) + " ")
            f.write(" ".join(str(int(v)) for v in k["first"]) + " " + " ".join(str(int(v)) for v in k["shape"]) + "\n")
            f.write(" ".join(str(int(v)) for v in k["cells"].ravel()) + "\n")
        f.write(tail if tail is not None else f"{capacity} {reject_every} {stop_after}\n")


@pytest.mark.parametrize("dims,reject_every,stop_after,capacity", [
    ((26.0, 14.0, 12.0), 0, 0, 1000), ((26.0, 14.0, 12.0), 3, 0, 1000), ((26.0, 14.0, 12.0), 0, 7, 1000),
    ((26.0, 14.0, 12.0), 2, 0, 5),
    # large enough for incremental steps (update grid 27^3 against a 45^3 grid)
    ((40.0, 40.0, 40.0), 3, 0, 24)])
def test_cpp_placement_loop(tmp_path, oracle, dims, reject_every, stop_after, capacity):
    from spectrapack_amd import workloads as W
    first, shape = W.box_window(dims, 1.0)
    env = oracle.box_container(first, shape, 1.0, dims)
    box = W.cuboid(10.0, 6.0, 4.0)
    kernels = []
    for q in oracle.cube_catalog()[:8]:
        f, s, c = oracle.voxelize_object(box.vertices, box.triangles, q, 1.0)
        kernels.append(dict(q=q, first=f, shape=s, cells=c))
    path = tmp_path / "scene.txt"
    write_scene(path, env, first, 1.0, kernels, capacity, reject_every, stop_after)
    print(f"[host] place dims={dims} reject_every={reject_every} stop_after={stop_after} capacity={capacity}",
          flush=True)
    r = subprocess.run([DRIVER, "place", str(path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    props = [tuple(int(v) for v in l.split()[1:]) for l in lines if l.startswith("P ")]
    summary = [l for l in lines if l.startswith("S ")][0].split()
    # the same loop on the oracle
    cur = env.copy()
    excluded = []
    count = accepted = 0
    outcome = None
    expected = []
    while True:
        if accepted >= capacity:
            outcome = "GAP_FILL_CAPACITY"
            break
        o, t, _ = oracle.place_step(cur, first, kernels, direct=True, excluded=excluded)
        if o < 0:
            outcome = "GAP_FILL_EXHAUSTED"
            break
        count += 1
        verdict = 0
        if stop_after and count >= stop_after:
            verdict = 2
        elif reject_every and count % reject_every == 0:
            verdict = 1
        expected.append((o, *[int(v) for v in t], verdict))
        if verdict == 2:
            outcome = "GAP_FILL_STOPPED"
            break
        if verdict == 1:
            excluded.append((o, tuple(int(v) for v in t)))
        else:
            cur = oracle.stamp(cur, first, kernels[o], t)
            accepted += 1
    assert props == expected
    assert summary[4] == outcome
    assert int(summary[3]) == accepted
    assert int(summary[5]) == int(cur.sum())


@pytest.mark.parametrize("reject_every,capacity", [(0, 1000), (3, 1000), (2, 6)])
def test_cpp_ranked_loop(tmp_path, ctx, reject_every, capacity):
    """rank_parts (the C++ mirror of the spectral pipeline's ranked loop) in a
    sphere-mesh container makes the same placements, with the same score bits,
    as the session's native ranked loop under the same validation verdicts."""
    import spectrapack_amd as sp
    from spectrapack_amd import workloads as W
    sph = W.sphere(14.0, n_seg=24, n_lat=11)
    first, shape = W.mesh_window(sph, 1.0)
    env = ctx.mesh_container(sph.vertices, sph.triangles, first, shape, 1.0)
    part = W.star(7, r_min=1.5, r_max=3.5)
    cat = W.cube_catalog()[:6]
    raw = ctx.voxelize_objects(part.vertices, part.triangles, cat, 1.0)
    kernels = [dict(q=q, first=k["first"], shape=k["shape"], cells=k["cells"]) for q, k in zip(cat, raw)]
    lo, hi = sph.bounds()
    max_z = W.oriented_max_z(part, cat)
    tail = f"{0.0!r} {float(lo[2])!r} {float(hi[2])!r}\n" + " ".join(repr(float(v)) for v in max_z) + \
        f"\n{capacity} {reject_every}\n"
    path = tmp_path / "scene.txt"
    write_scene(path, env, first, 1.0, kernels, 0, 0, 0, tail=tail)
    print(f"[host] rank reject_every={reject_every} capacity={capacity}", flush=True)
    r = subprocess.run([DRIVER, "rank", str(path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    got = [(int(l.split()[1]), tuple(int(v) for v in l.split()[2:5]), l.split()[5]) for l in lines
           if l.startswith("R ")]
    summary = [l for l in lines if l.startswith("S ")][0].split()
    s = sp.Session(ctx, shape, first, len(kernels))
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    s.set_incremental(True)
    s.set_ranking(0.0, 1.0, float(lo[2]), float(hi[2]), max_z)
    calls = [0]

    def validate(o, t):
        calls[0] += 1
        return not (reject_every and calls[0] % reject_every == 0)

    want = s.ranked_fill(max_copies=capacity, validate=validate)
    assert len(want) >= min(capacity, 8)
    assert got == [(o, tuple(t), f"{np.float64(v).view(np.uint64):016x}") for o, t, v in want]
    assert int(summary[1]) == len(want) and int(summary[2]) == calls[0]
    assert summary[3] == ("SPECTRAL_COPY_LIMIT" if len(want) == capacity else "SPECTRAL_PIPELINE_COMPLETE")
    assert int(summary[4]) == int(s.environment().sum())


def test_cpp_placement_with_exact_validation(tmp_path, oracle):
    """place_parts with the exact host validation as its accept verdict
    (gap_fill.cpp:512-523 with validation.cpp:517-634): an L-bracket at 12
    seeded rotations in a 40 mm box; every spectral proposal is certified
    valid, so the loop equals the plain one, and the accepted copies validate
    together."""
    from spectrapack_amd import workloads as W
    dims = (40.0, 40.0, 40.0)
    first, shape = W.box_window(dims, 1.0)
    env = oracle.box_container(first, shape, 1.0, dims)
    part = W.l_bracket(11, 4.0, 12.0)
    cat = W.so3_catalog(12, seed=5)
    kernels = []
    for q in cat:
        f, s, c = oracle.voxelize_object(part.vertices, part.triangles, q, 1.0)
        kernels.append(dict(q=q, first=f, shape=s, cells=c))
    tail = f"40 {len(part.vertices)} " + " ".join(repr(float(v)) for v in part.vertices.ravel()) + \
        f" {len(part.triangles)} " + " ".join(str(int(v)) for v in part.triangles.ravel()) + \
        " " + " ".join(repr(float(v)) for v in dims) + "\n"
    path = tmp_path / "scene.txt"
    write_scene(path, env, first, 1.0, kernels, 0, 0, 0, tail=tail)
    print("[host] place-exact bracket in 40 mm box", flush=True)
    r = subprocess.run([DRIVER, "place-exact", str(path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    props = [l.split() for l in lines if l.startswith("P ")]
    assert props and all(p[5] == "0" and p[6] == "VALID" for p in props), props[:5]
    summary = [l for l in lines if l.startswith("S ")][0].split()
    assert int(summary[3]) == len(props)
    assert [l for l in lines if l.startswith("F ")][0].split()[1] == "VALID"
    # the same proposals as the plain oracle loop
    cur = env.copy()
    for p in props:
        o, t, _ = oracle.place_step(cur, first, kernels, direct=True)
        assert (o, tuple(int(v) for v in t)) == (int(p[1]), (int(p[2]), int(p[3]), int(p[4])))
        cur = oracle.stamp(cur, first, kernels[o], t)
