"""The fused inverse y + inverse z pass (FP32, 256^3 FFT grid, steps without
per-value output) against the oracle, and against the separate passes on the
bench workload with rejections and region masks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def kernels_with_long_axes(rs, small=1):
    """Small random kernels plus kernels long in y, x and z: the long y kernel
    leaves 117 legal rows (the fused pass's single-chunk case), the long x
    kernel 56 legal columns and the long z kernel 36 legal layers."""
    out = []
    for _ in range(small):
        s = tuple(int(v) for v in rs.integers(1, 5, 3))
        c = (rs.uniform(0, 1, s[::-1]) < 0.7).astype(np.uint8)
        c.flat[0] = 1
        out.append(dict(first=tuple(int(v) for v in rs.integers(-4, 4, 3)), shape=s, cells=c))
    for s in ((1, 140, 1), (201, 1, 1), (1, 1, 221)):
        c = np.ones(s[::-1], np.uint8)
        c[0, 0, 0] = 0
        out.append(dict(first=(0, 0, 0), shape=s, cells=c))
    return out


def make_session(ctx, env, first, kernels, fused, precision=32):
    import spectrapack_amd as sp
    s = sp.Session(ctx, env.shape[::-1], first, len(kernels), precision=precision)
    s.set_fused_passes(fused)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    return s


def test_fused_steps_against_oracle(ctx, oracle):
    rs = np.random.default_rng(31)
    env = (rs.uniform(0, 1, (256, 256, 256)) < 0.02).astype(np.uint8)
    env[:2] = 1
    first = (-2, 0, -1)
    kernels = kernels_with_long_axes(rs)
    s = make_session(ctx, env, first, kernels, True)
    assert s.fused_passes()
    excluded = []
    for step in range(3):
        print(f"[fused] oracle step={step}", flush=True)
        o, t, _ = oracle.place_step(env, first, kernels, direct=True, excluded=excluded)
        if step == 1:
            # reject the oracle's choice: both must take the next one
            s.exclude(o, t)
            excluded.append((o, tuple(t)))
            o, t, _ = oracle.place_step(env, first, kernels, direct=True, excluded=excluded)
        r = s.step(stamp=True)
        assert r["orientation"] == o and tuple(r["translation"]) == tuple(t), (step, r, o, t)
        assert r["max_residual"] < 0.05
        env = oracle.stamp(env, first, kernels[o], t)
    assert np.array_equal(s.environment(), env)


def test_fused_off_and_fields_use_separate_passes(ctx):
    rs = np.random.default_rng(5)
    env = (rs.uniform(0, 1, (256, 256, 256)) < 0.01).astype(np.uint8)
    kernels = kernels_with_long_axes(rs, small=3)[:3]
    s = make_session(ctx, env, (0, 0, 0), kernels, False)
    assert not s.fused_passes()
    s.set_fused_passes(True)
    assert s.fused_passes()
    s.keep_fields(True)
    assert not s.fused_passes()  # per-value output needs the separate passes
    s.keep_fields(False)
    assert s.fused_passes()
    # 252 cells is still a 256-point grid; 250 is a grid of its own length
    s2 = make_session(ctx, env[:, :252, :250], (0, 0, 0), kernels, True)
    assert not s2.fused_passes()
    s3 = make_session(ctx, env[:, :252, :], (0, 0, 0), kernels, True)
    assert s3.fused_passes()
    s64 = make_session(ctx, env, (0, 0, 0), kernels, True, precision=64)
    assert not s64.fused_passes()
    for x in (s, s2, s3, s64):
        x.close()


def test_fused_equals_separate_on_bracket_workload(ctx):
    """All 96 orientations of the bench workload: fused and separate steps
    find the same lowest feasible corner for every orientation and choose the
    same placements, also with a rejected translation and with an orientation
    subset restricted to a region mask."""
    import spectrapack_amd as sp
    from spectrapack_amd import workloads as W
    w = W.bracket_in_sphere()
    env = ctx.mesh_container(w.container.vertices, w.container.triangles, w.env_first, w.env_shape, w.pitch)
    kernels = ctx.voxelize_objects(w.part.vertices, w.part.triangles, w.catalog, w.pitch)
    sess = {f: make_session(ctx, env, w.env_first, kernels, f) for f in (True, False)}
    assert sess[True].fused_passes() and not sess[False].fused_passes()
    for step in range(5):
        print(f"[fused] bracket step={step}", flush=True)
        if step == 2:
            r = sess[True].step(stamp=False)
            for s in sess.values():
                s.exclude(r["orientation"], r["translation"])
        if step == 3:
            mask = np.zeros((4, 4, 4), np.uint8)
            mask[1:3, :, 2:] = 1
            for s in sess.values():
                s.set_active(list(range(10, 60, 3)), mask)
        if step == 4:
            for s in sess.values():
                s.set_active()
        ra = sess[True].step(stamp=True)
        rb = sess[False].step(stamp=True)
        assert ra["numeric_flags"] == rb["numeric_flags"] == 0
        assert ra["orientation"] == rb["orientation"] and ra["translation"] == rb["translation"], (step, ra, rb)
        assert abs(ra["max_residual"] - rb["max_residual"]) < 0.01 and ra["max_residual"] < 0.05
        assert sess[True].orientation_minima() == sess[False].orientation_minima(), step
    assert np.array_equal(sess[True].environment(), sess[False].environment())


def test_fused_orientation_minima_edge_extents(ctx):
    """Every orientation's lowest feasible corner equals the separate passes'
    for legal extents at the fused pass's edges on a 256 x 256 x 253 window
    (a 256^3 FFT grid): 128 and 129 legal rows (one chunk, and a second chunk
    of one row), 4 legal layers, a single legal column, a kernel that does not
    fit, and a small random kernel; over three occupancies, one of them with a
    rejected corner."""
    rs = np.random.default_rng(77)
    kernels = []
    for sh in ((1, 129, 1), (2, 128, 1), (1, 1, 250), (256, 1, 1), (1, 1, 256), (3, 2, 2)):
        c = (rs.uniform(0, 1, sh[::-1]) < 0.8).astype(np.uint8)
        c.flat[-1] = 1
        kernels.append(dict(first=(0, 0, 0), shape=sh, cells=c))
    for trial, density in enumerate((0.002, 0.01, 0.03)):
        env = (rs.uniform(0, 1, (253, 256, 256)) < density).astype(np.uint8)
        minima = {}
        for fused in (True, False):
            s = make_session(ctx, env, (0, 0, 0), kernels, fused)
            assert s.fused_passes() == fused
            r = s.step(stamp=False)
            if trial == 2:
                s.exclude(r["orientation"], r["translation"])
                r = s.step(stamp=False)
            minima[fused] = (s.orientation_minima(), r["orientation"], r["translation"], r["numeric_flags"])
            s.close()
        print(f"[fused] edge extents trial={trial} minima={minima[True][0]}", flush=True)
        assert minima[True] == minima[False], trial


def test_fused_after_ineffective_incremental_window(ctx):
    """Incremental steps whose update window would not be under half the grid
    (a 128^3 kernel window) run as full steps without overlap fields, which
    is the fused pass on a 256^3 FP32 grid: the same placements and
    per-orientation minima as plain steps on the separate passes."""
    rs = np.random.default_rng(9)
    env = (rs.uniform(0, 1, (256, 256, 256)) < 0.002).astype(np.uint8)
    big = np.zeros((128, 128, 128), np.uint8)
    big[0, 0, 0] = big[-1, -1, -1] = big[0, -1, 0] = big[-1, 0, -1] = 1
    kernels = [dict(first=(0, 0, 0), shape=(128, 128, 128), cells=big),
               dict(first=(1, 0, 0), shape=(2, 2, 2), cells=np.ones((2, 2, 2), np.uint8))]
    inc = make_session(ctx, env, (0, 0, 0), kernels, True)
    inc.set_incremental(True)
    ref = make_session(ctx, env, (0, 0, 0), kernels, False)
    for step in range(3):
        print(f"[fused] ineffective incremental step={step}", flush=True)
        ra = inc.step(stamp=True)
        rb = ref.step(stamp=True)
        assert (ra["orientation"], ra["translation"]) == (rb["orientation"], rb["translation"]), (step, ra, rb)
        assert inc.orientation_minima() == ref.orientation_minima(), step
        assert inc.fused_passes()
    assert np.array_equal(inc.environment(), ref.environment())
    inc.close()
    ref.close()


def test_fused_product_pass_tiles(ctx):
    """The column-major product pass at each radix-16 tile width (8, 16, 32)
    feeds the fused pass the same spectrum: every orientation's lowest
    feasible corner equals the separate passes' on two occupancies."""
    rs = np.random.default_rng(13)
    kernels = kernels_with_long_axes(rs, small=3)
    for trial, density in enumerate((0.003, 0.02)):
        env = (rs.uniform(0, 1, (256, 256, 256)) < density).astype(np.uint8)
        ref = make_session(ctx, env, (0, 0, 0), kernels, False)
        ref.step(stamp=False)
        want = ref.orientation_minima()
        ref.close()
        s = make_session(ctx, env, (0, 0, 0), kernels, True)
        for width in (8, 16, 32):
            s.set_tile(0, 16, width)
            assert s.fused_passes()
            print(f"[fused] product tile width={width} trial={trial}", flush=True)
            s.step(stamp=False)
            assert s.orientation_minima() == want, (trial, width)
        s.close()


def test_product_pass_shares_break_inside_row_tiles(ctx):
    """The column-major product pass (fused path) shares row tiles x active
    orientations out evenly over a grid of resident blocks, so a block's share
    can begin and end inside a row tile; the test pins the tile width 16,
    whose 2032 row tiles at 256^3 (16 per plane, planes kz = 1 .. 127; the
    packed plane of kz = 0 and kz = 128 has its own launch) do not fill whole
    rounds of resident blocks, and checks that the shared grid ran.  With 1, 2, 5 and 7 active
    orientations (shares breaking at different orientations) every active
    orientation keeps the lowest feasible corner it has when all orientations
    run, and the others report none.  The separate passes (row-major product
    pass, one block per row tile) are checked the same way."""
    rs = np.random.default_rng(41)
    kernels = kernels_with_long_axes(rs, small=6)
    env = (rs.uniform(0, 1, (256, 256, 256)) < 0.01).astype(np.uint8)
    for fused in (True, False):
        s = make_session(ctx, env, (0, 0, 0), kernels, fused)
        assert s.fused_passes() == fused
        if fused:
            s.set_tile(0, 16, 16)
        s.step(stamp=False)
        if fused:
            blocks, tiles = s.product_grid()
            assert tiles == 2032 and 0 < blocks < tiles, (blocks, tiles)
        full = s.orientation_minima()
        assert sum(m is not None for m in full) >= 6
        for active in ([4], [0, 8], [1, 2, 3, 6, 7], [8, 7, 5, 4, 2, 1, 0]):
            print(f"[fused] product shares fused={fused} active={active}", flush=True)
            s.set_active(active)
            s.step(stamp=False)
            got = s.orientation_minima()
            for o in range(len(kernels)):
                assert got[o] == (full[o] if o in active else None), (fused, active, o)
        s.close()


def test_row_major_product_pass_shares_on_a_small_grid(ctx, oracle):
    """At 128^3 the row-major product pass (width 16) has 520 row tiles, fewer
    than one round of resident blocks, so its (row tile, orientation) units
    are shared out evenly and shares break inside row tiles.  The step's
    choice equals the oracle's; with 1, 2, 3 and 5 active orientations every
    active orientation keeps the lowest feasible corner it has when all run,
    and the others report none."""
    rs = np.random.default_rng(43)
    kernels = kernels_with_long_axes(rs, small=6)[:6]
    env = (rs.uniform(0, 1, (128, 128, 128)) < 0.01).astype(np.uint8)
    env[:2] = 1
    s = make_session(ctx, env, (0, 0, 0), kernels, False)
    s.set_tile(0, 16, 16)
    print("[fused] row-major shares at 128^3", flush=True)
    r = s.step(stamp=False)
    o, t, _ = oracle.place_step(env, (0, 0, 0), kernels, direct=True)
    assert r["orientation"] == o and tuple(r["translation"]) == tuple(t)
    full = s.orientation_minima()
    assert all(m is not None for m in full)
    for active in ([4], [0, 5], [1, 2, 3], [5, 4, 2, 1, 0]):
        print(f"[fused] row-major shares active={active}", flush=True)
        s.set_active(active)
        s.step(stamp=False)
        got = s.orientation_minima()
        for k in range(len(kernels)):
            assert got[k] == (full[k] if k in active else None), (active, k)
    s.close()
