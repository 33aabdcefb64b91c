"""Pair clearance between placed copies: every stamped copy also blocks the
cells within the clearance stencil of its cells.  Choices, overlap fields and
occupancy against the oracle (dilated stamps), in full and incremental steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def scene(seed):
    rs = np.random.default_rng(seed)
    shape = (52, 48, 44)
    first = (-1, -1, -1)
    env = np.zeros(shape[::-1], np.uint8)
    env[0], env[-1] = 1, 1
    env[:, 0], env[:, -1] = 1, 1
    env[:, :, 0], env[:, :, -1] = 1, 1
    kernels = []
    for _ in range(5):
        s = tuple(int(v) for v in rs.integers(4, 8, 3))
        kernels.append(dict(first=tuple(int(v) for v in rs.integers(-3, 3, 3)), shape=s,
                            cells=(rs.uniform(0, 1, s[::-1]) < 0.8).astype(np.uint8)))
    return first, env, kernels


@pytest.mark.parametrize("clearance,pitch,incremental", [(1.5, 1.0, False), (1.5, 1.0, True), (0.8, 0.5, True),
                                                         (0.0, 1.0, True)])
def test_pair_clearance_against_oracle(ctx, oracle, clearance, pitch, incremental):
    import spectrapack_amd as sp
    first, env, kernels = scene(int(clearance * 10) + incremental)
    stencil = oracle.clearance_stencil(pitch, clearance)
    s = sp.Session(ctx, env.shape[::-1], first, len(kernels), precision=32)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    s.set_pair_clearance(clearance, pitch)
    s.set_incremental(incremental)
    for step in range(12):
        print(f"[clearance] c={clearance} h={pitch} incremental={incremental} step={step}", flush=True)
        o, t, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
        r = s.step(stamp=True)
        assert r["orientation"] == o
        if incremental:
            for i in range(len(kernels)):
                g = s.overlap_field(i)
                assert (g is None and fields[i] is None) or np.array_equal(g, fields[i]), (step, i)
        if o < 0:
            break
        assert tuple(r["translation"]) == tuple(t)
        env = oracle.stamp(env, first, kernels[o], t, stencil=stencil)
        assert np.array_equal(s.environment(), env), step


def test_pair_clearance_rejects_bad_values(ctx):
    import spectrapack_amd as sp
    first, env, kernels = scene(9)
    s = sp.Session(ctx, env.shape[::-1], first, len(kernels))
    with pytest.raises(sp.Failure):
        s.set_pair_clearance(-1.0, 1.0)
    with pytest.raises(sp.Failure):
        s.set_pair_clearance(1.0, 0.0)
