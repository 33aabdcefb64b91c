"""Exact host validation of the GPU's final placements on the 256^3
L-bracket-in-sphere workload: every copy inside the 4k-face sphere mesh
container and no two copies overlapping (validation.cpp:517-634)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_bracket_in_sphere_placements_validate(ctx):
    import spectrapack_amd as sp
    from spectrapack_amd import validation as V
    from spectrapack_amd import workloads as W
    w = W.bracket_in_sphere()
    env = ctx.mesh_container(w.container.vertices, w.container.triangles, w.env_first, w.env_shape, w.pitch)
    kernels = ctx.voxelize_objects(w.part.vertices, w.part.triangles, w.catalog, w.pitch)
    s = sp.Session(ctx, w.env_shape, w.env_first, len(kernels))
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    poses = []
    for step in range(30):
        r = s.step(stamp=True)
        assert r["orientation"] >= 0
        poses.append((tuple(w.pitch * np.array(r["translation"], dtype=np.float64)), tuple(w.catalog[r["orientation"]])))
    print(f"[validation] {len(poses)} bracket copies", flush=True)
    part = (w.part.vertices, w.part.triangles)
    cont = (w.container.vertices, w.container.triangles)
    rep = V.validate_placements(part, poses, container=cont)
    print(f"[validation] {rep}", flush=True)
    assert rep["code"] == "VALID", rep
    # the last copy moved half-way onto the one before it is caught
    a = np.array(poses[-2][0])
    b = np.array(poses[-1][0])
    bad = poses[:-1] + [(tuple(a + 0.5 * (b - a)), poses[-1][1])]
    assert V.validate_placements(part, bad, container=cont)["code"] != "VALID"
    s.close()
