"""Restricted search (orientation subsets and region masks) against the
oracle's overlap fields: the chosen (orientation, corner) and the
per-orientation minima, bit-exact."""
import numpy as np
import pytest

from helpers import restricted_choice

pytestmark = pytest.mark.gpu


def scene(seed):
    rs = np.random.default_rng(100 + seed)
    shape = tuple(int(v) for v in rs.integers(24, 40, 3))
    first = tuple(int(v) for v in rs.integers(-3, 3, 3))
    env = (rs.uniform(0, 1, shape[::-1]) < 0.08).astype(np.uint8)
    env[:4] = 1
    kernels = []
    for _ in range(10):
        s = tuple(int(v) for v in rs.integers(2, 10, 3))
        c = (rs.uniform(0, 1, s[::-1]) < 0.8).astype(np.uint8)
        f = tuple(int(v) for v in rs.integers(-5, 5, 3))
        kernels.append(dict(first=f, shape=s, cells=c))
    kernels.append(dict(first=(0, 0, 0), shape=(shape[0] + 1, 2, 2), cells=np.ones((2, 2, shape[0] + 1), np.uint8)))
    return rs, first, env, kernels


def session(ctx, env, first, kernels, precision, graph):
    import spectrapack_amd as sp
    s = sp.Session(ctx, env.shape[::-1], first, len(kernels), precision=precision, capture_graph=graph)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    return s


@pytest.mark.parametrize("seed,precision,graph", [(0, 32, True), (1, 64, True), (2, 32, False)])
def test_subsets_and_regions(ctx, oracle, seed, precision, graph):
    rs, first, env, kernels = scene(seed)
    s = session(ctx, env, first, kernels, precision, graph)
    es = env.shape[::-1]
    for trial in range(6):
        print(f"[subset] seed={seed} precision={precision} graph={graph} trial={trial}", flush=True)
        _, _, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
        k = int(rs.integers(1, len(kernels) + 1))
        orients = [int(v) for v in rs.permutation(len(kernels))[:k]]
        if trial % 3 == 0:
            mask = None
        else:
            regions = tuple(int(v) for v in rs.integers(1, 9, 3))
            mask = (rs.uniform(0, 1, regions[::-1]) < 0.4).astype(np.uint8)
        if trial == 5:
            orients = []
        s.set_active(orients, mask)
        o_exp, u_exp, minima = restricted_choice(fields, es, orients, mask)
        r = s.step(stamp=False)
        assert r["orientation"] == o_exp
        got = s.orientation_minima()
        searched = set(orients) if orients else set(range(len(kernels)))
        for o in range(len(kernels)):
            assert got[o] == (minima[o] if o in searched else None), (trial, o)
        if o_exp >= 0:
            u = tuple(r["key"][a] - first[a] for a in range(3))
            assert u[::-1] == u_exp
            t = tuple(int(v) for v in r["translation"])
            s.stamp(o_exp, t)
            env = oracle.stamp(env, first, kernels[o_exp], t)
        # the plain search again
        s.set_active()
        o_full, t_full, _ = oracle.place_step(env, first, kernels, direct=True, fields=False)
        r = s.step(stamp=False)
        assert r["orientation"] == o_full
        if o_full >= 0:
            assert tuple(r["translation"]) == tuple(t_full)


def test_set_active_rejects_bad_lists(ctx, oracle):
    import spectrapack_amd as sp
    _, first, env, kernels = scene(3)
    s = session(ctx, env, first, kernels, 32, True)
    with pytest.raises(sp.Failure):
        s.set_active([0, 0])
    with pytest.raises(sp.Failure):
        s.set_active([len(kernels)])
    with pytest.raises(sp.Failure):
        s.set_active([], np.ones((65, 1, 1), np.uint8))
    # a list of only non-fitting orientations finds nothing
    s.set_active([len(kernels) - 1])
    assert s.step(stamp=False)["orientation"] == -1
