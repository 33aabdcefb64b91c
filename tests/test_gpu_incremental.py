"""Incremental steps (fields updated by the correlation of each step's new
cells, minima rescanned only when blocked) against the oracle and against
full steps: overlap fields, choices and occupancy bit-exact, with accepted,
rejected and externally stamped proposals."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_session(ctx, env, first, kernels, precision=32, incremental=True):
    import spectrapack_amd as sp
    s = sp.Session(ctx, env.shape[::-1], first, len(kernels), precision=precision)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    s.set_incremental(incremental)
    return s


def random_scene(seed, shape=(48, 44, 40), kmax=7):
    rs = np.random.default_rng(seed)
    first = tuple(int(v) for v in rs.integers(-3, 3, 3))
    env = (rs.uniform(0, 1, shape[::-1]) < 0.03).astype(np.uint8)
    env[:3] = 1
    kernels = []
    for _ in range(6):
        s = tuple(int(v) for v in rs.integers(3, kmax + 1, 3))
        kernels.append(dict(first=tuple(int(v) for v in rs.integers(-4, 4, 3)), shape=s,
                            cells=(rs.uniform(0, 1, s[::-1]) < 0.75).astype(np.uint8)))
    return rs, first, env, kernels


@pytest.mark.parametrize("seed,precision", [(0, 32), (1, 64), (2, 32)])
def test_incremental_matches_oracle_with_verdicts(ctx, oracle, seed, precision):
    rs, first, env, kernels = random_scene(seed)
    s = make_session(ctx, env, first, kernels, precision)
    excluded = []
    for step in range(14):
        print(f"[incremental] seed={seed} fp{precision} step={step}", flush=True)
        o, t, fields = oracle.place_step(env, first, kernels, direct=True, fields=True, excluded=excluded)
        r = s.step(stamp=False)
        assert r["orientation"] == o
        for i in range(len(kernels)):
            g = s.overlap_field(i)
            if fields[i] is None:
                assert g is None
                continue
            # rejected corners are held blocked in the incremental fields
            want = fields[i].copy()
            for eo, et in excluded:
                if eo == i:
                    u = [int(et[a]) + kernels[i]["first"][a] - first[a] for a in range(3)]
                    want[u[2], u[1], u[0]] += 1 << 30
            assert np.array_equal(g, want), (step, i)
        if o < 0:
            break
        assert tuple(r["translation"]) == tuple(t)
        if step % 4 == 3:
            excluded.append((o, tuple(int(v) for v in t)))
            s.exclude(o, t)
        else:
            s.stamp(o, t)
            env = oracle.stamp(env, first, kernels[o], t)
        assert np.array_equal(s.environment(), env)


def test_incremental_equals_full_steps_to_exhaustion(ctx, oracle):
    from spectrapack_amd import workloads as W
    dims = (64.0, 64.0, 64.0)
    first, shape = W.box_window(dims, 1.0)
    env = oracle.box_container(first, shape, 1.0, dims)
    kernels = []
    box = W.cuboid(10.0, 6.0, 4.0)
    for q in oracle.cube_catalog():
        f, sh, c = oracle.voxelize_object(box.vertices, box.triangles, q, 1.0)
        kernels.append(dict(first=f, shape=sh, cells=c))
    inc = make_session(ctx, env, first, kernels, 32, True)
    full = make_session(ctx, env, first, kernels, 32, False)
    steps = 0
    while True:
        a = inc.step(stamp=True)
        b = full.step(stamp=True)
        assert (a["orientation"], a["translation"]) == (b["orientation"], b["translation"]), steps
        assert a["numeric_flags"] == 0
        if a["orientation"] < 0:
            break
        steps += 1
        if steps % 100 == 0:
            print(f"[incremental] box scene step {steps}", flush=True)
    print(f"[incremental] box scene packed {steps} parts", flush=True)
    assert steps >= 400
    assert np.array_equal(inc.environment(), full.environment())


def test_incremental_without_benefit_stays_full(ctx, oracle):
    rs, first, env, kernels = random_scene(5, shape=(20, 18, 16), kmax=7)
    s = make_session(ctx, env, first, kernels, 32)
    for step in range(5):
        print(f"[incremental] small scene step={step}", flush=True)
        o, t, _ = oracle.place_step(env, first, kernels, direct=True)
        r = s.step(stamp=True)
        assert r["orientation"] == o
        if o < 0:
            break
        assert tuple(r["translation"]) == tuple(t)
        env = oracle.stamp(env, first, kernels[o], t)
    assert np.array_equal(s.environment(), env)


def test_incremental_refuses_restrictions(ctx):
    import spectrapack_amd as sp
    _, first, env, kernels = random_scene(6)
    s = make_session(ctx, env, first, kernels, 32)
    with pytest.raises(sp.Failure):
        s.set_active([0, 1])


@pytest.mark.parametrize("sides,precision,side,toggle", [((44, 46, 50), 32, 200, False), ((47, 6, 42), 32, 200, False),
                                                         ((44, 46, 50), 64, 200, False),
                                                         ((44, 46, 50), 32, 232, False),
                                                         pytest.param((44, 46, 50), 32, 200, True, marks=pytest.mark.skip(
                                                             reason="switching the register passes off after the update "
                                                                    "window is set up faulted the GPU (illegal memory "
                                                                    "access), cause not found; the library refuses it"))])
def test_incremental_window_128_equals_full(ctx, sides, precision, side, toggle):
    """FP32 and FP64.  Kernels of 42..64 cells on an axis (2 Kmax - 1 in
    83..127, good lengths 90..128) put that axis of the update window on the
    128-point register transforms (16 x 8); an axis with short kernels keeps
    its LDS Stockham length.  On a side^3 grid (at 232^3 the fields pass 4M
    corners per orientation, so each rescan spreads over several blocks),
    with stamped and rejected corners and then two removed copies: every
    orientation's field (rejected corners held blocked), the choices and the
    occupancy equal full steps'.  toggle: the incremental session turns its
    register passes off after five steps (the window then runs on the LDS
    Stockham passes and writes its own fields for k_apply_update)."""
    rs = np.random.default_rng(sum(sides))
    shape = (side, side, side)
    first = (0, 0, 0)
    # about 80 scattered cells and a floor: large kernels still find free corners
    env = (rs.uniform(0, 1, shape[::-1]) < 1e-5).astype(np.uint8)
    env[:2] = 1
    kernels = []
    for i in range(4):
        s = tuple(int(v) if i == 0 else int(rs.integers(max(1, v // 2), v + 1)) for v in sides)
        c = (rs.uniform(0, 1, s[::-1]) < 0.6).astype(np.uint8)
        c.flat[0] = 1
        kernels.append(dict(first=tuple(int(v) for v in rs.integers(-3, 3, 3)), shape=s, cells=c))
    inc = make_session(ctx, env, first, kernels, precision, True)
    full = make_session(ctx, env, first, kernels, precision, False)
    full.keep_fields(True)
    excluded = []

    def compare(a, b, label):
        assert (a["orientation"], a["translation"]) == (b["orientation"], b["translation"]), label
        assert a["numeric_flags"] == 0
        for i in range(len(kernels)):
            want = full.overlap_field(i)
            got = inc.overlap_field(i)
            if want is None:
                assert got is None
                continue
            for eo, et in excluded:
                if eo == i:
                    u = [int(et[k]) + kernels[i]["first"][k] - first[k] for k in range(3)]
                    want[u[2], u[1], u[0]] += 1 << 30
            assert np.array_equal(got, want), (label, i)

    for step in range(10):
        print(f"[incremental] window 128 sides={sides} fp{precision} step={step}", flush=True)
        if toggle and step == 5:
            inc.set_fast_path(False)
        a = inc.step(stamp=False)
        b = full.step(stamp=False)
        compare(a, b, step)
        o = a["orientation"]
        assert o >= 0, step
        t = tuple(a["translation"])
        if step % 3 == 2:
            excluded.append((o, t))
            inc.exclude(o, t)
            full.exclude(o, t)
        else:
            inc.stamp(o, t)
            full.stamp(o, t)
    live = [c["copy_id"] for c in inc.copies() if not c["removed"]]
    assert len(live) >= 3
    for cid in (live[0], live[2]):
        inc.remove_copy(cid)
        full.remove_copy(cid)
    for step in range(3):
        print(f"[incremental] window 128 sides={sides} fp{precision} after removals step={step}", flush=True)
        a = inc.step(stamp=True)
        b = full.step(stamp=True)
        compare(a, b, ("removed", step))
        assert a["orientation"] >= 0
    assert np.array_equal(inc.environment(), full.environment())
