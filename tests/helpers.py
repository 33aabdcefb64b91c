"""Shared test data: the reference's own test patterns (restated) and seeded fields."""
import numpy as np


def binary_pattern(shape, salt):
    """engine/pack_compute/tests/correlation_tests.cpp:184-198 (X-fastest; returns [z, y, x])."""
    sx, sy, sz = shape
    z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    h = 17 * x + 11 * y + 5 * z + 3 * x * y + 7 * y * z + 13 * x * z + salt
    cells = ((h % 19) < 7).astype(np.uint8)
    cells.flat[0] = 1
    cells.flat[-1] = 1
    return cells


def threading_pattern(shape, kind):
    """engine/pack_compute/tests/fft_threading_tests.cpp:68-70 patterns (i*7%11)<5 and i%3!=1."""
    n = int(np.prod(shape))
    i = np.arange(n)
    v = ((i * 7 % 11) < 5) if kind == 0 else (i % 3 != 1)
    return v.astype(np.uint8).reshape(shape[2], shape[1], shape[0])


def proper_cube_rotations():
    """All 24 proper signed permutations as (source_axis, sign) per output axis."""
    import itertools
    out = []
    for perm in itertools.permutations(range(3)):
        for s in itertools.product((-1, 1), repeat=3):
            m = np.zeros((3, 3), dtype=int)
            for r in range(3):
                m[r, perm[r]] = s[r]
            if round(np.linalg.det(m)) == 1:
                out.append((list(perm), list(s)))
    return out


def rotate_binary_field(shape, first, cells_zyx, rotation):
    """Rotate a kernel window by a signed permutation (correlation_tests.cpp rotate_binary_field):
    output axis a takes source axis src[a] with sign s[a]."""
    src, sgn = rotation
    a = cells_zyx.transpose(2, 1, 0)  # -> [x, y, z]
    out = np.transpose(a, src)
    for ax in range(3):
        if sgn[ax] < 0:
            out = np.flip(out, axis=ax)
    new_shape = tuple(shape[src[a_]] for a_ in range(3))
    new_first = tuple(first[src[a_]] if sgn[a_] > 0 else -(first[src[a_]] + shape[src[a_]] - 1) for a_ in range(3))
    return new_shape, new_first, np.ascontiguousarray(out.transpose(2, 1, 0))


def direct_expected(oracle, spec, env_zyx, ker_zyx):
    """Expected full-linear values from the independent integer oracle, in the layout of the result."""
    es, ks, ef, kf = spec
    origin = tuple(int(kf[a] - ef[a]) for a in range(3))
    rng, ext, vals = oracle.direct_correlation(env_zyx.ravel().astype(np.int32), es, ker_zyx.ravel().astype(np.int32),
                                               ks, origin)
    # the oracle's translations are world translations (correlation_tests.cpp:62-66)
    first = (rng[0], rng[2], rng[4])
    return np.array(first), vals


def restricted_choice(fields, env_shape, orientations=(), region_mask=None):
    """Expected choice of a step restricted to `orientations` (empty: all) and
    to kernel corners u whose region floor(u * regions / env_shape) has a
    nonzero byte in region_mask [rz, ry, rx].  fields: the oracle's legal
    overlap fields, one [lz, ly, lx] array (or None) per orientation; u is the
    index into that field.  Returns (orientation, (uz, uy, ux)) or (-1, None),
    and the per-orientation minima."""
    ex, ey, ez = env_shape
    allowed = list(orientations) if len(orientations) else list(range(len(fields)))
    minima = [None] * len(fields)
    best = None
    for o in allowed:
        f = fields[o]
        if f is None:
            continue
        zs, ys, xs = np.nonzero(f == 0)
        if region_mask is not None:
            rz, ry, rx = region_mask.shape
            keep = region_mask[zs * rz // ez, ys * ry // ey, xs * rx // ex] != 0
            zs, ys, xs = zs[keep], ys[keep], xs[keep]
        if zs.size == 0:
            continue
        order = np.lexsort((xs, ys, zs))
        i = order[0]
        minima[o] = (int(zs[i]), int(ys[i]), int(xs[i]))
        cand = (minima[o], o)
        if best is None or cand < best:
            best = cand
    if best is None:
        return -1, None, minima
    return best[1], best[0], minima
