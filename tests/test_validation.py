"""Exact host validation of final placements (validation.cpp:517-634 role),
on the CPU: known answers for disjoint, touching and overlapping pairs, box
and mesh containers, and the placements of the box and star fixtures."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def V():
    from spectrapack_amd import validation
    return validation


def cube(side=2.0):
    from spectrapack_amd import workloads as W
    m = W.cuboid(side, side, side)
    return m.vertices, m.triangles


ID = (0.0, 0.0, 0.0, 1.0)
RZ45 = (0.0, 0.0, np.sin(np.pi / 8), np.cos(np.pi / 8))


@pytest.mark.parametrize("second,expect", [
    ((2.0, 0.0, 0.0), ("valid", "VALID")),                    # faces in contact: the bounds certify it
    ((2.5, 0.0, 0.0), ("valid", "VALID")),
    ((1.5, 0.5, 0.3), ("invalid", "VALIDATION_PAIR")),        # edges cross faces strictly
    ((1.5, 0.0, 0.0), ("invalid", "VALIDATION_PAIR")),        # degenerate contacts; a face centroid inside
    ((0.0, 0.0, 0.0), ("indeterminate", "VALIDATION_PAIR_UNRESOLVED")),  # coincident surfaces
])
def test_cube_pairs(second, expect):
    base = np.array([10.0, 10.0, 10.0])
    r = V().validate_placements(cube(), [(tuple(base), ID), (tuple(base + second), ID)], box_dims=(100, 100, 100))
    assert (r["validity"], r["code"]) == expect, r


def test_rotated_cubes_disjoint_and_overlapping():
    """Two cubes turned 45 degrees about z, side by side along (1, 1): their
    bounds overlap for offsets d < 2 sqrt 2 while the facing faces are apart
    for d > sqrt 2, so the surfaces decide."""
    for d, expect in ((1.1 * np.sqrt(2.0), "VALID"), (0.9 * np.sqrt(2.0), "VALIDATION_PAIR")):
        r = V().validate_placements(cube(), [((10, 10, 10), RZ45), ((10 + d, 10 + d, 10.2), RZ45)],
                                    box_dims=(100, 100, 100))
        assert r["code"] == expect, (d, r)
        assert r["pair_tests"] == 1 and r["triangle_tests"] > 0


def test_planted_pose_error_is_caught():
    # the disjoint diagonal pair above with the second quaternion's z sign
    # flipped (a turn of -45 degrees, the same square) stays valid, while a
    # turn of 22.5 degrees brings a corner through the facing face
    d = 1.1 * np.sqrt(2.0)
    flipped = (RZ45[0], RZ45[1], -RZ45[2], RZ45[3])
    r = V().validate_placements(cube(), [((10, 10, 10), RZ45), ((10 + d, 10 + d, 10.2), flipped)],
                                box_dims=(100, 100, 100))
    assert r["code"] == "VALID"
    half = (0.0, 0.0, np.sin(np.pi / 16), np.cos(np.pi / 16))
    r = V().validate_placements(cube(), [((10, 10, 10), RZ45), ((10 + d, 10 + d, 10.2), half)],
                                box_dims=(100, 100, 100))
    assert r["code"] == "VALIDATION_PAIR"


def test_box_container_and_wall_clearance():
    inside = [((5.0, 5.0, 5.0), ID)]
    assert V().validate_placements(cube(), inside, box_dims=(10, 10, 10))["code"] == "VALID"
    on_wall = [((1.0, 5.0, 5.0), ID)]  # x from 0 to 2: touches the wall
    assert V().validate_placements(cube(), on_wall, box_dims=(10, 10, 10))["code"] == "VALID"
    r = V().validate_placements(cube(), on_wall, box_dims=(10, 10, 10), wall_clearance=0.5)
    assert r["code"] == "VALIDATION_CONTAINER" and r["first"] == 0
    out = [((9.5, 5.0, 5.0), ID)]
    assert V().validate_placements(cube(), out, box_dims=(10, 10, 10))["code"] == "VALIDATION_CONTAINER"
    # a turned cube whose corner passes the wall by 0.01 mm
    x = 10.0 - np.sqrt(2.0) + 0.01
    assert V().validate_placements(cube(), [((x, 5, 5), RZ45)], box_dims=(10, 10, 10))["code"] == "VALIDATION_CONTAINER"
    x = 10.0 - np.sqrt(2.0) - 0.01
    assert V().validate_placements(cube(), [((x, 5, 5), RZ45)], box_dims=(10, 10, 10))["code"] == "VALID"


def test_sphere_mesh_container():
    from spectrapack_amd import workloads as W
    s = W.sphere(20.0, n_seg=32, n_lat=15)
    lo, hi = s.bounds()
    c = (lo + hi) / 2
    sph = (s.vertices, s.triangles)
    assert V().validate_placements(cube(), [(tuple(c), ID)], container=sph)["code"] == "VALID"
    crossing = (c[0] + 19.5, c[1], c[2])
    assert V().validate_placements(cube(), [(crossing, ID)], container=sph)["code"] == "VALIDATION_CONTAINER"
    outside = (c[0] + 40.0, c[1], c[2])
    r = V().validate_placements(cube(), [(outside, ID)], container=sph)
    assert r["code"] == "VALIDATION_CONTAINER", r


def test_star_pairs_against_sampling():
    """Seeded star pairs: VALID only where no sampled point of one copy lies in
    the other; deep overlaps are VALIDATION_PAIR."""
    from spectrapack_amd import workloads as W
    star = W.star(7)
    part = (star.vertices, star.triangles)
    rs = np.random.default_rng(4)
    verdicts = {}
    for trial in range(12):
        q = W.canonical(rs.normal(size=4))
        d = rs.normal(size=3)
        d /= np.linalg.norm(d)
        dist = [2.0, 14.0, 19.0, 30.0][trial % 4]
        r = V().validate_placements(part, [((100, 100, 100), ID), (tuple(100 + dist * d), tuple(q))],
                                    box_dims=(1e3, 1e3, 1e3))
        verdicts.setdefault(dist, []).append(r["code"])
        if r["code"] == "VALID":
            assert not sampled_overlap(star, np.eye(3), np.zeros(3), W.rotation_matrix(q), dist * d, rs)
    assert set(verdicts[2.0]) == {"VALIDATION_PAIR"}
    assert set(verdicts[30.0]) == {"VALID"}  # radii at most 12 mm: bounds apart


def inside_star(star, p):
    """Ray parity in floating point (test helper for random points)."""
    v, t = star.vertices, star.triangles
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    d = np.array([0.3718420711, 0.1913264537, 1.0])
    e1, e2 = b - a, c - a
    h = np.cross(d, e2)
    det = (e1 * h).sum(1)
    f = 1.0 / det
    s = p - a
    u = f * (s * h).sum(1)
    qv = np.cross(s, e1)
    w = f * (qv @ d)
    tt = f * (e2 * qv).sum(1)
    hit = (u > 0) & (w > 0) & (u + w < 1) & (tt > 0)
    return hit.sum() % 2 == 1


def sampled_overlap(star, Ra, ta, Rb, tb, rs, n=400):
    lo, hi = star.vertices.min(0), star.vertices.max(0)
    for _ in range(n):
        p = rs.uniform(lo, hi)
        if not inside_star(star, p):
            continue
        w = Ra @ p + ta
        local_b = Rb.T @ (w - tb)
        if inside_star(star, local_b):
            return True
    return False


def test_incremental_validator_keeps_only_valid_copies():
    val = V().Validator(cube(), box_dims=(10, 10, 10))
    for x in (1.0, 3.0, 5.0):
        assert val.try_add((x, 1.0, 1.0), ID)["code"] == "VALID"
    r = val.try_add((4.5, 1.5, 1.2), ID)
    assert r["code"] == "VALIDATION_PAIR" and r["second"] == 3
    assert val.try_add((7.0, 1.0, 1.0), ID)["code"] == "VALID"
    assert val.try_add((11.0, 1.0, 1.0), ID)["code"] == "VALIDATION_CONTAINER"
    val.close()


@pytest.mark.parametrize("name", ["box", "star"])
def test_fixture_placements_validate_exactly(name):
    """The oracle's placements of the box (to exhaustion) and star workloads,
    which the GPU reproduces step for step, pass exact validation: the
    rotation of each chosen orientation and translation pitch * t
    (gap_fill.cpp:507-509) against the box container."""
    from spectrapack_amd import workloads as W
    with open(os.path.join(HERE, "golden", f"steps_{name}.json")) as f:
        fx = json.load(f)
    w = {"box": W.box_in_box, "star": W.star_in_box}[name]()
    poses = [(tuple(w.pitch * np.array(s[1:4], dtype=np.float64)), tuple(w.catalog[s[0]]))
             for s in fx["steps"] if s[0] >= 0]
    r = V().validate_placements((w.part.vertices, w.part.triangles), poses, box_dims=w.box_dims)
    print(name, len(poses), r)
    assert r["code"] == "VALID", r
    # a copy moved by one cell into its neighbour is caught
    o, t = fx["steps"][1][0], np.array(fx["steps"][1][1:4], dtype=np.float64)
    t0 = np.array(fx["steps"][0][1:4], dtype=np.float64)
    bad = poses[:1] + [(tuple(w.pitch * (t0 + 0.5 * (t - t0))), tuple(w.catalog[o]))]
    assert V().validate_placements((w.part.vertices, w.part.triangles), bad, box_dims=w.box_dims)["code"] != "VALID"
