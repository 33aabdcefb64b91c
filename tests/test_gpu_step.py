"""Batched placement step on the GPU against the oracle: overlap fields, the
chosen (orientation, translation) and the stamped occupancy, bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def kernels_for(oracle, mesh, catalog, pitch):
    out = []
    for q in catalog:
        f, s, c = oracle.voxelize_object(mesh.vertices, mesh.triangles, q, pitch)
        out.append(dict(first=f, shape=s, cells=c))
    return out


def make_session(ctx, env, first, kernels, precision=0, graph=True):
    import spectrapack_amd as sp
    s = sp.Session(ctx, env.shape[::-1], first, len(kernels), precision=precision, capture_graph=graph)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    return s


@pytest.mark.parametrize("precision,graph", [(32, True), (64, False), (0, True)])
def test_box_in_small_box_steps(ctx, oracle, precision, graph):
    from spectrapack_amd import workloads as W
    dims = (24.0, 18.0, 14.0)
    first, shape = W.box_window(dims, 1.0)
    env = oracle.box_container(first, shape, 1.0, dims)
    kernels = kernels_for(oracle, W.cuboid(10.0, 6.0, 4.0), oracle.cube_catalog(), 1.0)
    s = make_session(ctx, env, first, kernels, precision, graph)
    s.keep_fields(True)
    for step in range(6):
        print(f"[step] box-in-box precision={precision} graph={graph} step={step}", flush=True)
        o, t, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
        r = s.step(stamp=True)
        for i in range(len(kernels)):
            g = s.overlap_field(i)
            assert (g is None and fields[i] is None) or np.array_equal(g, fields[i]), (step, i)
        assert r["orientation"] == o
        if o < 0:
            break
        assert tuple(r["translation"]) == tuple(t)
        env = oracle.stamp(env, first, kernels[o], t)
        assert np.array_equal(s.environment(), env)


@pytest.mark.parametrize("seed", range(3))
def test_random_scene_steps(ctx, oracle, seed):
    rs = np.random.default_rng(seed)
    shape = tuple(int(v) for v in rs.integers(20, 40, 3))
    first = tuple(int(v) for v in rs.integers(-3, 3, 3))
    env = (rs.uniform(0, 1, shape[::-1]) < 0.05).astype(np.uint8)
    env[0], env[-1] = 1, 1
    kernels = []
    for _ in range(int(rs.integers(2, 7))):
        s = tuple(int(v) for v in rs.integers(2, 9, 3))
        c = (rs.uniform(0, 1, s[::-1]) < 0.8).astype(np.uint8)
        f = tuple(int(v) for v in rs.integers(-5, 5, 3))
        kernels.append(dict(first=f, shape=s, cells=c))
    sess = make_session(ctx, env, first, kernels, precision=[32, 64, 0][seed])
    sess.keep_fields(True)
    for step in range(4):
        print(f"[step] random seed={seed} step={step}", flush=True)
        o, t, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
        r = sess.step(stamp=True)
        for i in range(len(kernels)):
            g = sess.overlap_field(i)
            assert (g is None and fields[i] is None) or np.array_equal(g, fields[i])
        assert r["orientation"] == o
        if o < 0:
            break
        assert tuple(r["translation"]) == tuple(t)
        env = oracle.stamp(env, first, kernels[o], t)


def test_kernel_larger_than_window_is_skipped(ctx, oracle):
    env = np.zeros((6, 5, 7), np.uint8)
    big = dict(first=(0, 0, 0), shape=(8, 2, 2), cells=np.ones((2, 2, 8), np.uint8))
    small = dict(first=(0, 0, 0), shape=(2, 2, 2), cells=np.ones((2, 2, 2), np.uint8))
    s = make_session(ctx, env, (0, 0, 0), [big, small])
    r = s.step(stamp=False)
    o, t, _ = oracle.place_step(env, (0, 0, 0), [big, small], direct=True)
    assert r["orientation"] == o == 1 and tuple(r["translation"]) == tuple(t)


def test_exhausted_returns_minus_one(ctx, oracle):
    env = np.ones((6, 6, 6), np.uint8)
    k = dict(first=(0, 0, 0), shape=(2, 2, 2), cells=np.ones((2, 2, 2), np.uint8))
    s = make_session(ctx, env, (0, 0, 0), [k])
    assert s.step()["orientation"] == -1


@pytest.mark.parametrize("resident", [False, True])
def test_kernel_spectra_resident_or_recomputed(ctx, oracle, resident):
    rs = np.random.default_rng(11)
    shape = (33, 27, 30)
    first = (1, -2, 0)
    env = (rs.uniform(0, 1, shape[::-1]) < 0.06).astype(np.uint8)
    kernels = []
    for _ in range(5):
        s = tuple(int(v) for v in rs.integers(2, 9, 3))
        kernels.append(dict(first=tuple(int(v) for v in rs.integers(-4, 4, 3)), shape=s,
                            cells=(rs.uniform(0, 1, s[::-1]) < 0.8).astype(np.uint8)))
    sess = make_session(ctx, env, first, kernels, precision=32)
    sess.set_resident_kernels(resident)
    sess.keep_fields(True)
    for step in range(4):
        print(f"[step] resident={resident} step={step}", flush=True)
        o, t, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
        r = sess.step(stamp=True)
        for i in range(len(kernels)):
            g = sess.overlap_field(i)
            assert (g is None and fields[i] is None) or np.array_equal(g, fields[i])
        assert r["orientation"] == o
        if o < 0:
            break
        assert tuple(r["translation"]) == tuple(t)
        env = oracle.stamp(env, first, kernels[o], t)
