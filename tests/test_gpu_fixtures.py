"""The box, star and bracket workloads step by step against fixtures the CPU
oracle wrote (tests/golden/make_fixtures.py): every step's choice
(orientation, translation) and the SHA-256 of every orientation's int32 legal
overlap field, bit-exact; the box to exhaustion (gap_fill.cpp:380-509 runs
until no feasible placement remains).  The star also reports max |FP32 - FP64|
of the unrounded values per step against the a-priori bound.  The bracket
(256^3 sphere container, 96 orientations) holds all 96 fields of its first
steps, in FP32 and FP64, and the fused default path's choices."""
import json
import os

import numpy as np
import pytest

from fixture_hash import cells_digest, fields_digest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with open(os.path.join(HERE, "golden", f"steps_{name}.json")) as f:
        return json.load(f)


def setup(ctx, fx, workload, precision=0):
    import spectrapack_amd as sp
    w = workload()
    assert w.name == fx["workload"]
    if w.container is not None:
        env = ctx.mesh_container(w.container.vertices, w.container.triangles, w.env_first, w.env_shape, w.pitch)
    else:
        env = ctx.box_container(w.env_first, w.env_shape, w.pitch, w.box_dims)
    assert cells_digest(env) == fx["environment_sha256"]
    kernels = ctx.voxelize_objects(w.part.vertices, w.part.triangles, w.catalog, w.pitch)
    for k, want in zip(kernels, fx["kernels"]):
        assert list(k["first"]) == want["first"] and list(k["shape"]) == want["shape"]
        assert cells_digest(k["cells"]) == want["sha256"]
    s = sp.Session(ctx, w.env_shape, w.env_first, len(kernels), precision=precision)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    return s, kernels


def step_digest(s, n):
    return fields_digest(s.overlap_field(o) for o in range(n))


@pytest.mark.parametrize("incremental", [False, True])
def test_box_to_exhaustion_matches_oracle_fixture(ctx, incremental):
    from spectrapack_amd import workloads as W
    fx = load("box")
    print(f"[fixture] box to exhaustion, {len(fx['steps'])} steps, incremental={incremental}", flush=True)
    s, kernels = setup(ctx, fx, W.box_in_box)
    s.keep_fields(True)
    if incremental:
        s.set_incremental(True)
    for i, (o, tx, ty, tz, digest) in enumerate(fx["steps"]):
        r = s.step(stamp=True)
        assert r["orientation"] == o, (i, r, o)
        if o >= 0:
            assert tuple(r["translation"]) == (tx, ty, tz), (i, r)
        assert step_digest(s, len(kernels)) == digest, i
    assert fx["steps"][-1][0] == -1
    assert cells_digest(s.environment()) == fx["final_environment_sha256"]
    s.close()


def test_star_steps_match_oracle_fixture_with_fp64_deviation(ctx):
    from spectrapack_amd import workloads as W
    fx = load("star")
    print(f"[fixture] star, {len(fx['steps'])} steps", flush=True)
    s32, kernels = setup(ctx, fx, W.star_in_box, precision=32)
    s64, _ = setup(ctx, fx, W.star_in_box, precision=64)
    for s in (s32, s64):
        s.keep_fields(True)
        s.keep_raw(True)
    devs = []
    for i, (o, tx, ty, tz, digest) in enumerate(fx["steps"]):
        r32 = s32.step(stamp=True)
        r64 = s64.step(stamp=True)
        for r in (r32, r64):
            assert r["orientation"] == o and tuple(r["translation"]) == (tx, ty, tz), (i, r)
        assert step_digest(s32, len(kernels)) == digest, i
        assert step_digest(s64, len(kernels)) == digest, i
        dev = max(float(np.abs(s32.raw_values(k) - s64.raw_values(k)).max()) for k in range(len(kernels)))
        devs.append(dev)
        print(f"[fixture] star step {i}: orientation {o}, max |FP32 - FP64| {dev:.3e}, "
              f"bound {r32['error_bound']:.4f}, FP32 residual {r32['max_residual']:.3e}", flush=True)
        assert dev < r32["error_bound"] and r32["precision"] == 32 and r64["precision"] == 64
    assert cells_digest(s32.environment()) == fx["final_environment_sha256"]
    s32.close()
    s64.close()


def test_bracket_256_steps_match_oracle_fixture(ctx):
    """The headline workload at full size: every one of the 96 orientations'
    legal fields (about 9 M corners each) bit-exact to the oracle's checked
    FP64 FFT, with FP32 (the bound's choice) and FP64 sessions, and the fused
    inverse y + z path (no field output) making the same choices."""
    from spectrapack_amd import workloads as W
    fx = load("bracket")
    print(f"[fixture] bracket 256^3, {len(fx['steps'])} steps", flush=True)
    sessions = {}
    for name, precision, keep in (("fp32", 0, True), ("fp64", 64, True), ("fused", 0, False)):
        s, kernels = setup(ctx, fx, W.bracket_in_sphere, precision=precision)
        s.keep_fields(keep)
        sessions[name] = s
    assert sessions["fused"].fused_passes()
    for i, (o, tx, ty, tz, digest) in enumerate(fx["steps"]):
        for name, s in sessions.items():
            r = s.step(stamp=True)
            print(f"[fixture] bracket step {i} {name}: orientation {r['orientation']}", flush=True)
            assert r["orientation"] == o and tuple(r["translation"]) == (tx, ty, tz), (i, name, r)
            assert r["precision"] == (64 if name == "fp64" else 32) and r["numeric_flags"] == 0
            if name != "fused":
                assert step_digest(s, len(kernels)) == digest, (i, name)
    for s in sessions.values():
        assert cells_digest(s.environment()) == fx["final_environment_sha256"]
        s.close()
