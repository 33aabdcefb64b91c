"""Drop-in full-linear correlation on the GPU against the oracle (bit-exact
integers) — the reference's own correlation test cases, restated."""
import numpy as np
import pytest

from helpers import binary_pattern, direct_expected, proper_cube_rotations, rotate_binary_field, threading_pattern

pytestmark = pytest.mark.gpu


def check(ctx, oracle, spec, env, ker, precision=0):
    print(f"[correlate] spec={spec} precision={precision}", flush=True)
    g = ctx.correlate(spec, env, ker, precision=precision)
    f, vals = direct_expected(oracle, spec, env, ker)
    assert np.array_equal(g["translation_first"], f)
    assert np.array_equal(g["values"], vals.astype(np.float64))
    o = oracle.correlate(spec, env, ker, limits=oracle.LIMITS_UNBOUNDED)
    assert np.array_equal(g["values"], o["values"])
    assert np.array_equal(g["shape"], o["shape"])
    return g


@pytest.mark.parametrize("precision", [0, 32, 64])
def test_golden_2d(ctx, oracle, precision):
    spec = ((3, 2, 1), (2, 2, 1), (4, -3, 0), (-1, 1, 0))
    g = check(ctx, oracle, spec, np.array([1, 0, 1, 0, 1, 1], np.uint8).reshape(1, 2, 3),
              np.array([1, 0, 1, 1], np.uint8).reshape(1, 2, 2), precision)
    assert list(g["translation_first"]) == [4, -5, 0] and list(g["shape"]) == [4, 3, 1]


def test_true_3d(ctx, oracle):
    spec = ((2, 3, 2), (1, 2, 2), (7, -4, 9), (-3, 5, -2))
    g = check(ctx, oracle, spec, np.array([1, 0, 1, 0, 1, 1, 0, 1, 0, 1, 1, 0], np.uint8).reshape(2, 3, 2),
              np.array([1, 0, 1, 1], np.uint8).reshape(2, 2, 1))
    assert list(g["translation_first"]) == [10, -10, 10]


@pytest.mark.parametrize("precision", [32, 64])
def test_all_24_rotations(ctx, oracle, precision):
    es, ef = (4, 5, 6), (8, -5, 11)
    env = binary_pattern(es, 2)
    ks, kf = (2, 3, 5), (-4, 2, -7)
    ker = binary_pattern(ks, 7)
    for rot in proper_cube_rotations():
        shape, first, cells = rotate_binary_field(ks, kf, ker, rot)
        check(ctx, oracle, (es, shape, ef, first), env, cells, precision)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("ext", [(8, 6), (64, 64)])
def test_lengths_13_and_127(ctx, oracle, axis, ext):
    e, k = ext
    es, ks = [1, 1, 1], [1, 1, 1]
    es[axis], ks[axis] = e, k
    check(ctx, oracle, (tuple(es), tuple(ks), (0, 0, 0), (0, 0, 0)), threading_pattern(es, 0),
          threading_pattern(ks, 1))


@pytest.mark.parametrize("seed", range(6))
def test_random_fields(ctx, oracle, seed):
    rs = np.random.default_rng(100 + seed)
    es = tuple(int(v) for v in rs.integers(5, 40, 3))
    ks = tuple(int(v) for v in rs.integers(1, 12, 3))
    env = (rs.uniform(0, 1, es[::-1]) < rs.uniform(0.1, 0.9)).astype(np.uint8)
    ker = (rs.uniform(0, 1, ks[::-1]) < 0.6).astype(np.uint8)
    ef = tuple(int(v) for v in rs.integers(-5, 5, 3))
    kf = tuple(int(v) for v in rs.integers(-5, 5, 3))
    check(ctx, oracle, (es, ks, ef, kf), env, ker, precision=[0, 32, 64][seed % 3])


def test_empty_fields(ctx, oracle):
    es, ks = (6, 5, 4), (3, 2, 2)
    check(ctx, oracle, (es, ks, (0, 0, 0), (0, 0, 0)), np.zeros(es[::-1], np.uint8), np.zeros(ks[::-1], np.uint8))


def test_proximity_within_1e10(ctx, oracle):
    rs = np.random.default_rng(11)
    es, ks = (9, 7, 5), (3, 4, 2)
    env = rs.uniform(0, 1, es[::-1])
    ker = binary_pattern(ks, 3)
    spec = (es, ks, (1, -2, 3), (0, 4, -1))
    g = ctx.correlate(spec, env, ker, proximity=True)
    o = oracle.correlate(spec, env, ker, proximity=True, limits=oracle.LIMITS_UNBOUNDED)
    assert np.array_equal(g["translation_first"], o["translation_first"])
    scale = np.maximum(1.0, np.abs(o["values"]))
    assert (np.abs(g["values"] - o["values"]) <= 1e-10 * scale).all()


def test_failure_codes(ctx):
    import spectrapack_amd as sp
    spec = ((3, 2, 1), (2, 2, 1), (0, 0, 0), (0, 0, 0))
    env = np.array([1, 0, 1, 0, 1, 1], np.uint8)
    ker = np.array([1, 0, 1, 1], np.uint8)
    cases = [
        (dict(environment=np.array([1, 0, 2, 0, 1, 1], np.uint8)), "CORRELATION_INVALID_ENVIRONMENT"),
        (dict(kernel=np.array([1, 0, 2, 1], np.uint8)), "CORRELATION_INVALID_KERNEL"),
        (dict(environment=env[:5]), "CORRELATION_INVALID_FIELD"),
        (dict(spec=((0, 2, 1), (2, 2, 1), (0, 0, 0), (0, 0, 0))), "CORRELATION_SHAPE_OVERFLOW"),
        (dict(spec=((3, 2, 1), (2, 2, 1), (-(2 ** 63), 0, 0), (0, 0, 0))), "CORRELATION_INDEX_OVERFLOW"),
        (dict(limits=(512 << 20, 0, 11, 200000000)), "CORRELATION_PADDED_CELL_LIMIT"),
        (dict(limits=(512 << 20, 0, 16777216, 3)), "CORRELATION_DIRECT_TERM_LIMIT"),
        (dict(limits=(100, 0, 16777216, 200000000)), "CORRELATION_MEMORY_LIMIT"),
    ]
    for over, code in cases:
        args = dict(spec=spec, environment=env, kernel=ker)
        args.update(over)
        with pytest.raises(sp.Failure) as e:
            ctx.correlate(**args)
        assert e.value.code == code, (over, e.value.code)
