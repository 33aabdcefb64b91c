"""Seeded synthetic parts, containers and orientation catalogs for the
placement-step workloads (box part in a box, star part in a box, L-bracket in a
sphere, torus-knot tube in a fine box).

Meshes are closed triangle meshes in a local frame centred on their bounding
box (p_local = p - bbox_center, the part frame of doc/spectrapack-spec.md:166-172);
containers keep their bounding-box minimum at the origin.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

DEFAULT_SEED = 20261203


@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 3) float64
    triangles: np.ndarray  # (T, 3) uint32

    def bounds(self):
        return self.vertices.min(axis=0), self.vertices.max(axis=0)


def centred(m: Mesh) -> Mesh:
    lo, hi = m.bounds()
    c = (lo + hi) / 2
    return Mesh(m.vertices - c, m.triangles)


def cuboid(w: float, d: float, h: float) -> Mesh:
    """8-vertex, 12-triangle axis-aligned box centred on the origin (outward winding)."""
    x, y, z = w / 2, d / 2, h / 2
    v = np.array([[-x, -y, -z], [x, -y, -z], [x, y, -z], [-x, y, -z],
                  [-x, -y, z], [x, -y, z], [x, y, z], [-x, y, z]], dtype=np.float64)
    t = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [1, 2, 6], [1, 6, 5], [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], dtype=np.uint32)
    return Mesh(v, t)


def _uv_sphere_topology(n_seg: int, n_lat: int):
    """Unit directions and triangles of a UV sphere with poles and n_lat rings."""
    dirs = [[0.0, 0.0, 1.0]]
    for i in range(1, n_lat + 1):
        th = math.pi * i / (n_lat + 1)
        for j in range(n_seg):
            ph = 2 * math.pi * j / n_seg
            dirs.append([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    dirs.append([0.0, 0.0, -1.0])
    dirs = np.array(dirs, dtype=np.float64)
    south = len(dirs) - 1
    tris = []
    ring = lambda i, j: 1 + (i - 1) * n_seg + (j % n_seg)
    for j in range(n_seg):
        tris.append([0, ring(1, j), ring(1, j + 1)])
    for i in range(1, n_lat):
        for j in range(n_seg):
            a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j), ring(i + 1, j + 1)
            tris.append([a, c, b])
            tris.append([b, c, d])
    for j in range(n_seg):
        tris.append([south, ring(n_lat, j + 1), ring(n_lat, j)])
    return dirs, np.array(tris, dtype=np.uint32)


def star(seed: int = DEFAULT_SEED, faces: int = 512, r_min: float = 4.0, r_max: float = 12.0) -> Mesh:
    """Star-shaped closed mesh: a 16-segment, 16-ring UV sphere (512 faces) with
    per-vertex radii uniform in [r_min, r_max] about the origin."""
    assert faces == 512
    dirs, tris = _uv_sphere_topology(16, 16)
    rng = np.random.default_rng(seed)
    r = rng.uniform(r_min, r_max, size=len(dirs))
    return centred(Mesh(dirs * r[:, None], tris))


def sphere(radius: float = 120.0, n_seg: int = 64, n_lat: int = 31, seed: int = DEFAULT_SEED) -> Mesh:
    """Closed UV sphere (64 x 31 rings: 3968 faces), given a seeded rotation,
    translated so its bounding-box minimum is the origin."""
    dirs, tris = _uv_sphere_topology(n_seg, n_lat)
    rng = np.random.default_rng(seed)
    R = rotation_matrix(random_quaternions(1, rng)[0])
    v = (dirs * radius) @ R.T
    v = v - v.min(axis=0)
    return Mesh(v, tris)


def union_of_boxes(boxes) -> Mesh:
    """Watertight boundary of a union of axis-aligned boxes [(lo, hi), ...]:
    quads of the arrangement grid between inside and outside cells."""
    xs = [sorted({b[0][a] for b in boxes} | {b[1][a] for b in boxes}) for a in range(3)]
    nx, ny, nz = (len(x) - 1 for x in xs)
    inside = np.zeros((nx, ny, nz), dtype=bool)
    for i in range(nx):
        cx = (xs[0][i] + xs[0][i + 1]) / 2
        for j in range(ny):
            cy = (xs[1][j] + xs[1][j + 1]) / 2
            for k in range(nz):
                cz = (xs[2][k] + xs[2][k + 1]) / 2
                inside[i, j, k] = any(lo[0] < cx < hi[0] and lo[1] < cy < hi[1] and lo[2] < cz < hi[2]
                                      for lo, hi in boxes)
    verts = {}
    tris = []

    def vid(p):
        key = tuple(p)
        if key not in verts:
            verts[key] = len(verts)
        return verts[key]

    def occ(i, j, k):
        return 0 <= i < nx and 0 <= j < ny and 0 <= k < nz and inside[i, j, k]

    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if not inside[i, j, k]:
                    continue
                for axis in range(3):
                    for side in (0, 1):
                        n = [i, j, k]
                        n[axis] += 1 if side else -1
                        if occ(*n):
                            continue
                        idx = [i, j, k]
                        plane = xs[axis][idx[axis] + side]
                        b, c = (axis + 1) % 3, (axis + 2) % 3
                        corners = []
                        for (db, dc) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                            p = [0.0, 0.0, 0.0]
                            p[axis] = plane
                            p[b] = xs[b][idx[b] + db]
                            p[c] = xs[c][idx[c] + dc]
                            corners.append(vid(p))
                        if side == 0:
                            corners = corners[::-1]
                        tris.append([corners[0], corners[1], corners[2]])
                        tris.append([corners[0], corners[2], corners[3]])
    v = np.zeros((len(verts), 3), dtype=np.float64)
    for p, i in verts.items():
        v[i] = p
    return Mesh(v, np.array(tris, dtype=np.uint32))


def l_bracket(seed: int = DEFAULT_SEED, lo: float = 5.0, hi: float = 30.0) -> Mesh:
    """Union of three corner-anchored boxes with extents uniform in [lo, hi]:
    a base plate, an upright plate and a web, all meeting at one corner."""
    rng = np.random.default_rng(seed)
    e = rng.uniform(lo, hi, size=(3, 3))
    # thin axis per plate: base thin in z, upright thin in y, web thin in x
    base = ((0.0, 0.0, 0.0), (e[0, 0], e[0, 1], min(e[0, 2], 0.35 * e[0, 2] + 0.65 * lo)))
    upright = ((0.0, 0.0, 0.0), (e[1, 0], min(e[1, 1], 0.35 * e[1, 1] + 0.65 * lo), e[1, 2]))
    web = ((0.0, 0.0, 0.0), (min(e[2, 0], 0.35 * e[2, 0] + 0.65 * lo), e[2, 1], e[2, 2]))
    return centred(union_of_boxes([base, upright, web]))


def torus_knot(p: int = 2, q: int = 3, major: float = 40.0, minor: float = 15.0, tube: float = 3.0,
               n_along: int = 256, n_around: int = 16) -> Mesh:
    """(p, q) torus-knot tube: n_along x n_around x 2 faces (8192 by default)."""
    t = np.linspace(0, 2 * np.pi, n_along, endpoint=False)

    def curve(tt):
        r = major + minor * np.cos(q * tt)
        return np.stack([r * np.cos(p * tt), r * np.sin(p * tt), minor * np.sin(q * tt)], axis=-1)

    c = curve(t)
    d = curve(t + 1e-4) - curve(t - 1e-4)
    T = d / np.linalg.norm(d, axis=1, keepdims=True)
    ref = np.array([0.0, 0.0, 1.0])
    N = np.cross(T, ref)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    B = np.cross(T, N)
    a = np.linspace(0, 2 * np.pi, n_around, endpoint=False)
    v = (c[:, None, :] + tube * (np.cos(a)[None, :, None] * N[:, None, :] + np.sin(a)[None, :, None] * B[:, None, :]))
    v = v.reshape(-1, 3)
    tris = []
    for i in range(n_along):
        for j in range(n_around):
            a0 = i * n_around + j
            a1 = i * n_around + (j + 1) % n_around
            b0 = ((i + 1) % n_along) * n_around + j
            b1 = ((i + 1) % n_along) * n_around + (j + 1) % n_around
            tris.append([a0, b0, a1])
            tris.append([a1, b0, b1])
    return centred(Mesh(v, np.array(tris, dtype=np.uint32)))


def random_quaternions(n: int, rng) -> np.ndarray:
    """Uniform SO(3) sample (Shoemake 1992, "Uniform random rotations"), xyzw."""
    u = rng.uniform(0.0, 1.0, size=(n, 3))
    a = np.sqrt(1 - u[:, 0])
    b = np.sqrt(u[:, 0])
    return np.stack([a * np.sin(2 * np.pi * u[:, 1]), a * np.cos(2 * np.pi * u[:, 1]),
                     b * np.sin(2 * np.pi * u[:, 2]), b * np.cos(2 * np.pi * u[:, 2])], axis=1)


def rotation_matrix(q) -> np.ndarray:
    x, y, z, w = q
    return np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), w * w - x * x + y * y - z * z, 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), w * w - x * x - y * y + z * z]])


def oriented_max_z(mesh: Mesh, quaternions) -> np.ndarray:
    """Per orientation, the rotated part's highest vertex z in mm (the max z of
    its oriented bounds, geometry::oriented_bounds(...).bounds_mm.max[2])."""
    return np.array([float((mesh.vertices @ rotation_matrix(q).T)[:, 2].max()) for q in quaternions])


def container_z_range(w) -> tuple:
    """Container bounds in z (mm): the mesh's bounds, or [0, height] of a box."""
    if w.container is not None:
        lo, hi = w.container.bounds()
        return float(lo[2]), float(hi[2])
    return 0.0, float(w.box_dims[2])


def canonical(q) -> np.ndarray:
    """orientations.cpp:44-73: unit norm, w > 0 (else first nonzero > 0), no -0."""
    q = np.asarray(q, dtype=np.float64).copy()
    n = math.hypot(math.hypot(q[0], q[1]), math.hypot(q[2], q[3]))
    q = q / n
    q[q == 0] = 0.0
    neg = q[3] < 0
    if q[3] == 0:
        for i in range(3):
            if q[i] != 0:
                neg = q[i] < 0
                break
    if neg:
        q = -q
    q[q == 0] = 0.0
    return q


def so3_catalog(n: int, seed: int = DEFAULT_SEED) -> np.ndarray:
    """Seeded uniform SO(3) catalog, canonicalised and deduplicated in order."""
    rng = np.random.default_rng(seed + 96)
    out = []
    for q in random_quaternions(n, rng):
        c = canonical(q)
        if not any(np.array_equal(c, o) for o in out):
            out.append(c)
    return np.array(out)


def cube_catalog() -> np.ndarray:
    """orientation_cube.hpp:10-80: the 24 proper signed permutations sorted
    lexicographically as row-major matrices, as canonical quaternions."""
    import itertools
    mats = []
    for perm in itertools.permutations(range(3)):
        for s in itertools.product((-1, 1), repeat=3):
            m = np.zeros((3, 3), dtype=int)
            for r in range(3):
                m[r, perm[r]] = s[r]
            if round(np.linalg.det(m)) == 1:
                mats.append(m)
    mats.sort(key=lambda m: tuple(m.ravel()))
    out = []
    for m in mats:
        m = m.ravel()
        tr = m[0] + m[4] + m[8]
        if tr > 0:
            n = [m[7] - m[5], m[2] - m[6], m[3] - m[1], tr + 1]
        elif m[0] > m[4] and m[0] > m[8]:
            n = [1 + m[0] - m[4] - m[8], m[1] + m[3], m[2] + m[6], m[7] - m[5]]
        elif m[4] > m[8]:
            n = [m[1] + m[3], 1 + m[4] - m[0] - m[8], m[5] + m[7], m[2] - m[6]]
        else:
            n = [m[2] + m[6], m[5] + m[7], 1 + m[8] - m[0] - m[4], m[3] - m[1]]
        nz = sum(1 for v in n if v != 0)
        mag = 1.0 if nz == 1 else 0.7071067811865475244 if nz == 2 else 0.5
        q = np.array([(-mag if v < 0 else mag if v else 0.0) for v in n])
        neg = q[3] < 0
        if q[3] == 0:
            for i in range(3):
                if q[i] != 0:
                    neg = q[i] < 0
                    break
        if neg:
            q = -q
        q[q == 0] = 0.0
        out.append(q)
    return np.array(out)


@dataclass
class Workload:
    """A placement workload: part mesh, container, lattice and catalog."""
    name: str
    part: Mesh
    pitch: float
    env_first: tuple
    env_shape: tuple  # (x, y, z)
    catalog: np.ndarray
    box_dims: tuple | None = None  # analytic box container (mm)
    container: Mesh | None = None  # mesh container (interior is free)
    precision: str = "auto"
    notes: dict = field(default_factory=dict)


def box_window(dims, pitch):
    """gap_fill.cpp:39-53: first = -1, shape = ceil(L/h) + 2 per axis."""
    return (-1, -1, -1), tuple(int(math.ceil(d / pitch)) + 2 for d in dims)


def mesh_window(mesh: Mesh, pitch: float):
    """spectral_pipeline.cpp:107-140: first = floor(min/h) - 1, last = ceil(max/h)."""
    lo, hi = mesh.bounds()
    first = tuple(int(math.floor(l / pitch)) - 1 for l in lo)
    last = tuple(int(math.ceil(h / pitch)) for h in hi)
    return first, tuple(b - a + 1 for a, b in zip(first, last))


def box_in_box(seed: int = DEFAULT_SEED) -> Workload:
    """12-triangle 10x6x4 mm box part in a 64 mm box container at 1 mm (66^3 window), 24 cube orientations."""
    first, shape = box_window((64.0, 64.0, 64.0), 1.0)
    return Workload("box10x6x4-in-box64-h1", cuboid(10.0, 6.0, 4.0), 1.0, first, shape, cube_catalog(),
                    box_dims=(64.0, 64.0, 64.0), precision="fp32")


def star_in_box(seed: int = DEFAULT_SEED) -> Workload:
    """512-face star part (radii U[4,12] mm) in a 128 mm box at 1 mm (130^3 window), 24 cube orientations."""
    first, shape = box_window((128.0, 128.0, 128.0), 1.0)
    return Workload("star512-in-box128-h1", star(seed), 1.0, first, shape, cube_catalog(),
                    box_dims=(128.0, 128.0, 128.0), precision="fp32")


def bracket_in_sphere(seed: int = DEFAULT_SEED, orientations: int = 96) -> Workload:
    """L-bracket part in the interior of an r=120 mm sphere mesh voxelised at
    256^3, with a seeded uniform SO(3) catalog of 96 orientations."""
    sph = sphere(120.0, seed=seed)
    lo, hi = sph.bounds()
    pitch = float(np.max(hi - lo)) / 253.5
    first, shape = mesh_window(sph, pitch)
    return Workload("lbracket-in-sphere120-256cubed-o96", l_bracket(seed), pitch, first, shape,
                    so3_catalog(orientations, seed), container=sph)


def knot_in_fine_box(seed: int = DEFAULT_SEED) -> Workload:
    """8192-face torus-knot tube (r=3 mm) in a 255 mm box at 0.5 mm (512^3 window), 24 cube orientations."""
    first, shape = box_window((255.0, 255.0, 255.0), 0.5)
    return Workload("torusknot8k-in-box255-h0.5", torus_knot(), 0.5, first, shape, cube_catalog(),
                    box_dims=(255.0, 255.0, 255.0), precision="auto")


# Fixed orientation catalog of the learned proposal (its 96 logits index it).
PROPOSAL_CATALOG_SEED = 20261203


def star_so3_box(seed: int = DEFAULT_SEED) -> Workload:
    """Learned-proposal episodes: a 512-face star part (radii U[4,12] mm, seeded)
    in a 128 mm box at 1 mm (130^3 window) with the fixed 96-orientation SO(3)
    catalog."""
    first, shape = box_window((128.0, 128.0, 128.0), 1.0)
    return Workload("star512-in-box128-h1-o96", star(seed), 1.0, first, shape, so3_catalog(96, PROPOSAL_CATALOG_SEED),
                    box_dims=(128.0, 128.0, 128.0), precision="fp32")


def star_so3_fine_box(seed: int = DEFAULT_SEED) -> Workload:
    """The same scene at 0.5 mm: a 127 mm box gives a 256^3 window; 96 orientations."""
    first, shape = box_window((127.0, 127.0, 127.0), 0.5)
    return Workload("star512-in-box127-h0.5-256cubed-o96", star(seed), 0.5, first, shape,
                    so3_catalog(96, PROPOSAL_CATALOG_SEED), box_dims=(127.0, 127.0, 127.0), precision="fp32")


WORKLOADS = {
    "box": box_in_box,
    "star": star_in_box,
    "bracket": bracket_in_sphere,
    "knot": knot_in_fine_box,
    "star96": star_so3_box,
    "star96_fine": star_so3_fine_box,
}
