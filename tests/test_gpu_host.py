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


def write_scene(path, env, first, pitch, kernels, capacity, reject_every, stop_after):
    with open(path, "w") as f:
        f.write(" ".join(str(v) for v in env.shape[::-1]) + "\n")
        f.write(" ".join(str(int(v)) for v in first) + f" {pitch!r}\n")
        f.write(" ".join(str(int(v)) for v in env.ravel()) + "\n")
        f.write(f"{len(kernels)}\n")
        for k in kernels:
            f.write(" ".join(repr(float(v)) for v in k["q"]) + " ")
            f.write(" ".join(str(int(v)) for v in k["first"]) + " " + " ".join(str(int(v)) for v in k["shape"]) + "\n")
            f.write(" ".join(str(int(v)) for v in k["cells"].ravel()) + "\n")
        f.write(f"{capacity} {reject_every} {stop_after}\n")


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
