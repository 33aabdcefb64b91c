"""The register-FFT passes (128-, 256- and 512-point axes) against the oracle, and
the full bench workload checked at scale by size-independent properties."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_kernels(rs, n, max_side=4):
    out = []
    for _ in range(n):
        s = tuple(int(v) for v in rs.integers(1, max_side + 1, 3))
        c = (rs.uniform(0, 1, s[::-1]) < 0.7).astype(np.uint8)
        c.flat[0] = 1
        f = tuple(int(v) for v in rs.integers(-4, 4, 3))
        out.append(dict(first=f, shape=s, cells=c))
    return out


def run_against_oracle(ctx, oracle, shape, seed, precision, steps=2, density=0.02):
    import spectrapack_amd as sp
    rs = np.random.default_rng(seed)
    env = (rs.uniform(0, 1, shape[::-1]) < density).astype(np.uint8)
    env[: 2] = 1  # the lowest layers are blocked, so choices are not all at the origin
    first = (-1, -1, -1)
    kernels = small_kernels(rs, 3)
    s = sp.Session(ctx, shape, first, len(kernels), precision=precision)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    s.keep_fields(True)
    for step in range(steps):
        print(f"[fast] shape={shape} precision={precision} step={step}", flush=True)
        o, t, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
        r = s.step(stamp=True)
        for i in range(len(kernels)):
            g = s.overlap_field(i)
            assert np.array_equal(g, fields[i]), (step, i)
        assert r["orientation"] == o and tuple(r["translation"]) == tuple(t)
        env = oracle.stamp(env, first, kernels[o], t)
    assert np.array_equal(s.environment(), env)


@pytest.mark.parametrize("precision", [32, 64])
def test_256_cubed(ctx, oracle, precision):
    run_against_oracle(ctx, oracle, (256, 256, 256), 1, precision, steps=2)


@pytest.mark.parametrize("shape,precision", [((128, 128, 128), 32), ((128, 20, 128), 64), ((20, 128, 24), 32)])
def test_128_axes(ctx, oracle, shape, precision):
    """128-point axes: the mixed-radix (16 x 8) register transforms."""
    run_against_oracle(ctx, oracle, shape, 3, precision, steps=2)


@pytest.mark.parametrize("shape", [(512, 24, 512), (24, 512, 20)])
def test_512_axes(ctx, oracle, shape):
    run_against_oracle(ctx, oracle, shape, 2, 32, steps=2)


def test_bracket_workload_fast_equals_generic_and_properties(ctx, oracle):
    """The bench workload (256^3 sphere container, L-bracket kernels), 8 of its
    orientations: the register-FFT step and the generic LDS step agree on every
    legal overlap count; sampled counts equal direct sums; the chosen
    translation is free and lexicographically lowest."""
    import spectrapack_amd as sp
    from spectrapack_amd import workloads as W
    w = W.bracket_in_sphere()
    env = ctx.mesh_container(w.container.vertices, w.container.triangles, w.env_first, w.env_shape, w.pitch)
    kernels = ctx.voxelize_objects(w.part.vertices, w.part.triangles, w.catalog[:8], w.pitch)
    sess = {}
    for fast in (True, False):
        s = sp.Session(ctx, w.env_shape, w.env_first, len(kernels), precision=32)
        s.set_fast_path(fast)
        s.set_environment(env)
        for i, k in enumerate(kernels):
            s.set_kernel(i, k["cells"], k["shape"], k["first"])
        s.keep_fields(True)
        sess[fast] = s
    rs = np.random.default_rng(4)
    cur = env.copy()
    for step in range(2):
        print(f"[fast] bracket step={step}", flush=True)
        ra = sess[True].step(stamp=True)
        rb = sess[False].step(stamp=True)
        assert ra["orientation"] == rb["orientation"] and ra["translation"] == rb["translation"]
        best = None
        for i, k in enumerate(kernels):
            fa = sess[True].overlap_field(i)
            fb = sess[False].overlap_field(i)
            assert np.array_equal(fa, fb), i
            kz, ky, kx = k["cells"].shape
            for _ in range(40):
                uz, uy, ux = (int(rs.integers(0, fa.shape[a])) for a in range(3))
                direct = int((cur[uz:uz + kz, uy:uy + ky, ux:ux + kx] & k["cells"]).sum())
                assert fa[uz, uy, ux] == direct
            zeros = np.argwhere(fa == 0)
            if len(zeros):
                key = tuple(zeros[0])  # argwhere is lexicographic in (z, y, x)
                if best is None or key < best[0]:
                    best = (key, i)
        assert ra["orientation"] == best[1]
        o = best[1]
        uz, uy, ux = best[0]
        assert tuple(ra["translation"]) == (w.env_first[0] + ux - kernels[o]["first"][0],
                                            w.env_first[1] + uy - kernels[o]["first"][1],
                                            w.env_first[2] + uz - kernels[o]["first"][2])
        k = kernels[o]
        kz, ky, kx = k["cells"].shape
        assert not (cur[uz:uz + kz, uy:uy + ky, ux:ux + kx] & k["cells"]).any()
        cur[uz:uz + kz, uy:uy + ky, ux:ux + kx] |= k["cells"]
        assert np.array_equal(sess[True].environment(), cur)


def test_knot_512_fp64_sampled_steps(ctx):
    """The 512^3 torus-knot workload, 24 cube orientations, FP64 chosen by the
    bound: on two stamped steps, sampled overlap counts of three orientations
    (the chosen one included) equal direct sums, each sampled field's first zero
    in (z, y, x) order is that orientation's reported minimum, the chosen
    corner is free and is the lowest minimum over all orientations."""
    import spectrapack_amd as sp
    from spectrapack_amd import workloads as W
    w = W.knot_in_fine_box()
    env = ctx.box_container(w.env_first, w.env_shape, w.pitch, w.box_dims)
    kernels = ctx.voxelize_objects(w.part.vertices, w.part.triangles, w.catalog, w.pitch)
    s = sp.Session(ctx, w.env_shape, w.env_first, len(kernels))
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    s.keep_fields(True)
    rs = np.random.default_rng(12)
    cur = env.copy()
    for step in range(2):
        print(f"[fast] knot 512^3 step={step}", flush=True)
        r = s.step(stamp=True)
        assert r["precision"] == 64 and r["error_bound"] > 0.4 and r["numeric_flags"] == 0
        o = r["orientation"]
        minima = s.orientation_minima()
        keys = [(m, i) for i, m in enumerate(minima) if m is not None]
        assert min(keys)[1] == o
        for i in sorted({o, int(rs.integers(0, 24)), int(rs.integers(0, 24))}):
            f = s.overlap_field(i)
            k = kernels[i]
            kz, ky, kx = k["cells"].shape
            for _ in range(30):
                uz, uy, ux = (int(rs.integers(0, f.shape[a])) for a in range(3))
                assert f[uz, uy, ux] == int((cur[uz:uz + kz, uy:uy + ky, ux:ux + kx] & k["cells"]).sum())
            zeros = np.argwhere(f == 0)
            assert (tuple(zeros[0]) if len(zeros) else None) == minima[i]
        uz, uy, ux = minima[o]
        k = kernels[o]
        kz, ky, kx = k["cells"].shape
        assert not (cur[uz:uz + kz, uy:uy + ky, ux:ux + kx] & k["cells"]).any()
        cur[uz:uz + kz, uy:uy + ky, ux:ux + kx] |= k["cells"]
        del f
    assert np.array_equal(s.environment(), cur)
    s.close()
