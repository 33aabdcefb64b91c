"""Ranked candidate pages on the GPU against the oracle.

The proximity field (exact squared distances and exp(-sqrt(d2)/2)) is
bit-exact; the FP64 proximity correlation is within 1e-10 of a direct sum (the
reference's own pin, correlation_tests.cpp:330-348); pages are bit-exact
against the oracle's page given the same proximity values and the direct
binary overlaps (as spectral_tests.cpp:3066-3130 pins its page on the
pipeline's own proximity values), in both modes, after cursors, with rejected
translations and with incremental fields; against the oracle's independent
proximity they agree up to near-ties; the native greedy loop makes the same
placements as the oracle's loop."""
import numpy as np
import pytest

import pages as PG

pytestmark = pytest.mark.gpu

INF = 0x7FFFFFFF


def make_session(ctx, env, first, kernels, incremental=False, precision=32):
    import spectrapack_amd as sp
    s = sp.Session(ctx, env.shape[::-1], first, len(kernels), precision=precision)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    s.set_incremental(incremental)
    return s


def rank_params(env, first):
    ez = env.shape[0]
    return dict(origin_z=0.25, pitch=0.5, cmin=0.25 + 0.5 * first[2], cmax=0.25 + 0.5 * (first[2] + ez))


def max_z_of(kernels):
    return [0.5 * (k["first"][2] + k["shape"][2]) + 0.01 * i for i, k in enumerate(kernels)]


def set_ranking(s, env, first, kernels):
    p = rank_params(env, first)
    s.set_ranking(p["origin_z"], p["pitch"], p["cmin"], p["cmax"], max_z_of(kernels))
    return p


def random_scene(seed, shape=(40, 44, 48), n=6, kmax=7, density=0.03):
    rs = np.random.default_rng(seed)
    first = tuple(int(v) for v in rs.integers(-3, 3, 3))
    env = (rs.uniform(0, 1, shape) < density).astype(np.uint8)
    env[:2] = 1
    kernels = []
    for _ in range(n):
        s = tuple(int(v) for v in rs.integers(2, kmax + 1, 3))
        c = (rs.uniform(0, 1, s[::-1]) < 0.75).astype(np.uint8)
        c.flat[0] = 1
        kernels.append(dict(first=tuple(int(v) for v in rs.integers(-4, 4, 3)), shape=s, cells=c))
    return first, env, kernels


@pytest.mark.parametrize("shape,density", [((29, 50, 37), 0.03), ((6, 5, 300), 0.01), ((4, 600, 8), 0.002),
                                           ((520, 4, 70), 0.001), ((8, 9, 10), 0.0), ((8, 9, 10), 1.0),
                                           ((33, 31, 35), -1.0)])
def test_proximity_field_bit_exact(ctx, oracle, shape, density):
    rs = np.random.default_rng(sum(shape))
    if density < 0:
        env = np.zeros(shape, np.uint8)
        env[3, 17, 30] = 1        # a single occupied cell
    else:
        env = (rs.uniform(size=shape) < density).astype(np.uint8)
    print(f"[ranking] proximity field shape={shape} density={density}", flush=True)
    kernels = [dict(first=(0, 0, 0), shape=(2, 2, 2), cells=np.ones((2, 2, 2), np.uint8))]
    s = make_session(ctx, env, (0, 0, 0), kernels)
    set_ranking(s, env, (0, 0, 0), kernels)
    s.rank()
    vals, sq = s.proximity_field()
    want_v, want_sq = oracle.proximity(env)
    want_sq = np.where(want_sq == np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64(INF), want_sq)
    assert np.array_equal(sq.astype(np.uint64), want_sq)
    assert np.array_equal(vals.view(np.uint64), want_v.view(np.uint64))


def test_proximity_values_within_1e10(ctx, oracle):
    first, env, kernels = random_scene(1)
    s = make_session(ctx, env, first, kernels)
    set_ranking(s, env, first, kernels)
    s.rank()
    field, _ = oracle.proximity(env)
    for o, k in enumerate(kernels):
        got = s.proximity_values(o)
        want = PG.proximity_direct(field, k["cells"])
        err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
        print(f"[ranking] proximity o={o} max rel err {err.max():.2e}", flush=True)
        assert err.max() <= 1e-10


def oracle_page(s, oracle_fields, env, first, kernels, p, o, mode, cursor=None, excluded=()):
    k = kernels[o]
    if oracle_fields[o] is None:
        return []
    tf = [first[a] - k["first"][a] for a in range(3)]
    sc = PG.page_scores(s.proximity_values(o), int(k["cells"].sum()), p["origin_z"], p["pitch"], p["cmin"], p["cmax"],
                        max_z_of(kernels)[o], tf[2])
    return PG.ranked_page(oracle_fields[o], sc, tf, mode=mode, cursor=cursor, excluded=excluded, env=env,
                          kernel_cells=k["cells"])


def same_page(got, want):
    assert [tuple(c["translation"]) for c in got] == [tuple(c["translation"]) for c in want]
    assert [np.float64(c["score"]).view(np.uint64) for c in got] == \
        [np.float64(c["score"]).view(np.uint64) for c in want]
    assert [c["binary_overlap"] for c in got] == [c["binary_overlap"] for c in want]
    assert [c["direct_zero_overlap"] for c in got] == [c["direct_zero_overlap"] for c in want]


@pytest.mark.parametrize("seed,precision", [(2, 32), (3, 64)])
@pytest.mark.parametrize("mode", [0, 1])
def test_pages_bit_exact_given_the_same_proximity(ctx, oracle, seed, precision, mode):
    first, env, kernels = random_scene(seed)
    kernels.append(dict(first=(0, 0, 0), shape=(60, 2, 2), cells=np.ones((2, 2, 60), np.uint8)))  # does not fit
    s = make_session(ctx, env, first, kernels, precision=precision)
    p = set_ranking(s, env, first, kernels)
    _, _, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
    pages = s.ranked_pages(mode=mode)
    assert pages[len(kernels) - 1] == []
    cursors = {}
    for rnd in range(3):
        print(f"[ranking] pages seed={seed} fp{precision} mode={mode} page {rnd}", flush=True)
        for o in pages:
            want = oracle_page(s, fields, env, first, kernels, p, o, mode, cursor=cursors.get(o))
            same_page(pages[o], want)
            if mode == 0:
                assert all(c["direct_zero_overlap"] for c in pages[o])
        cursors = {o: pages[o][-1] for o in pages if pages[o]}
        pages = s.ranked_pages(orientations=sorted(cursors), mode=mode, cursors=cursors)


def test_pages_against_independent_proximity(ctx, oracle):
    first, env, kernels = random_scene(4)
    s = make_session(ctx, env, first, kernels)
    p = set_ranking(s, env, first, kernels)
    _, _, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
    field, _ = oracle.proximity(env)
    pages = s.ranked_pages()
    for o, k in enumerate(kernels):
        tf = [first[a] - k["first"][a] for a in range(3)]
        sc = PG.page_scores(PG.proximity_direct(field, k["cells"]), int(k["cells"].sum()), p["origin_z"], p["pitch"],
                            p["cmin"], p["cmax"], max_z_of(kernels)[o], tf[2])
        full = PG.ranked_page(fields[o], sc, tf, size=PG.PAGE_SIZE + 1)
        want = full[:PG.PAGE_SIZE]
        got = pages[o]
        assert len(got) == len(want)
        for i, (g, w) in enumerate(zip(got, want)):
            assert abs(g["score"] - w["score"]) <= 1e-10 * max(1.0, abs(w["score"]))
            near = [abs(w["score"] - full[j]["score"]) <= 1e-9 for j in (i - 1, i + 1) if 0 <= j < len(full)]
            if not any(near):
                assert tuple(g["translation"]) == tuple(w["translation"]), (o, i)


def test_pages_with_incremental_fields_and_rejections(ctx, oracle):
    first, env, kernels = random_scene(5)
    s = make_session(ctx, env, first, kernels, incremental=True)
    p = set_ranking(s, env, first, kernels)
    excluded = {}
    for step in range(6):
        print(f"[ranking] incremental step {step}", flush=True)
        _, _, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
        pages = s.ranked_pages()
        for o in range(len(kernels)):
            want = oracle_page(s, fields, env, first, kernels, p, o, 0, excluded=excluded.get(o, ()))
            same_page(pages[o], want)
        heads = [(pages[o][0]["score"], o) for o in range(len(kernels)) if pages[o]]
        _, o = min(heads)
        t = pages[o][0]["translation"]
        if step % 3 == 2:
            s.exclude(o, t)
            excluded.setdefault(o, []).append(tuple(t))
        else:
            s.stamp(o, t)
            env = oracle.stamp(env, first, kernels[o], t)


def test_ranked_fill_matches_the_oracle_loop(ctx, oracle):
    first, env, kernels = random_scene(6, shape=(18, 20, 22), n=4, kmax=5, density=0.02)
    s = make_session(ctx, env, first, kernels, incremental=True)
    set_ranking(s, env, first, kernels)
    print("[ranking] native fill", flush=True)
    placed = s.ranked_fill()
    print(f"[ranking] native fill placed {len(placed)}", flush=True)
    assert len(placed) >= 10
    # the oracle loop over oracle pages, with the proximity of a second session
    # kept at the same occupancy
    twin = make_session(ctx, env, first, kernels)
    p = set_ranking(twin, env, first, kernels)
    state = dict(env=env.copy())

    def pages_of(orients, cursors):
        twin.rank()
        _, _, fields = oracle.place_step(state["env"], first, kernels, direct=True, fields=True)
        return {o: oracle_page(twin, fields, state["env"], first, kernels, p, o, 0, cursor=cursors.get(o))
                for o in orients}

    def stamp(o, t):
        twin.stamp(o, t)
        state["env"] = oracle.stamp(state["env"], first, kernels[o], t)

    want = PG.ranked_fill(pages_of, stamp, list(range(len(kernels))))
    assert [(o, tuple(t)) for o, t, _ in placed] == [(o, tuple(t)) for o, t, _ in want]
    assert [np.float64(v).view(np.uint64) for _, _, v in placed] == [np.float64(v).view(np.uint64) for _, _, v in want]
    assert np.array_equal(s.environment(), state["env"])


def test_ranked_fill_with_validation(ctx, oracle):
    first, env, kernels = random_scene(7, shape=(16, 18, 20), n=3, kmax=4)
    s = make_session(ctx, env, first, kernels)
    set_ranking(s, env, first, kernels)
    seen = []

    def validate(o, t):
        seen.append((o, t))
        return len(seen) % 2 == 0          # every other head is rejected

    placed = s.ranked_fill(max_copies=5, validate=validate)
    assert len(placed) == 5
    assert len(seen) == 10


def test_ranking_rejects_large_grids_and_bad_heights(ctx):
    import spectrapack_amd as sp
    s = sp.Session(ctx, (1100, 4, 4), (0, 0, 0), 1)
    s.set_kernel(0, np.ones((1, 1, 1), np.uint8), (1, 1, 1), (0, 0, 0))
    with pytest.raises(sp.Failure):
        s.set_ranking(0.0, 1.0, 0.0, 4.0, [1.0])
    s = sp.Session(ctx, (8, 8, 8), (0, 0, 0), 1)
    s.set_kernel(0, np.ones((1, 1, 1), np.uint8), (1, 1, 1), (0, 0, 0))
    with pytest.raises(sp.Failure):
        s.set_ranking(0.0, 1.0, 2.0, 2.0, [1.0])
    with pytest.raises(sp.Failure):
        s.set_ranking(0.0, 1.0, 0.0, 4.0, [1.0, 2.0])
    # an orientation whose kernel has no occupied cell (spectral_pipeline.cpp:1735-1740)
    s = sp.Session(ctx, (8, 8, 8), (0, 0, 0), 2)
    s.set_environment(np.zeros((8, 8, 8), np.uint8))
    s.set_kernel(0, np.ones((2, 2, 2), np.uint8), (2, 2, 2), (0, 0, 0))
    s.set_kernel(1, np.zeros((2, 2, 2), np.uint8), (2, 2, 2), (0, 0, 0))
    s.set_ranking(0.0, 1.0, 0.0, 8.0, [1.0, 1.0])
    with pytest.raises(sp.Failure) as e:
        s.ranked_pages()
    assert "EMPTY_KERNEL" in str(e.value)


def test_pages_at_full_size_bracket(ctx):
    """256^3 x 96 (L-bracket in the sphere), after a few greedy placements:
    every page is ordered by its key, every ordinary entry has zero overlap by
    a direct count, every score is the pipeline expression of the session's
    own proximity value (bit-exact), the proximity at the page corners is a
    direct sum of the proximity field within 1e-10, and the second page
    continues strictly after the first."""
    import spectrapack_amd as sp
    from spectrapack_amd import workloads as W
    w = W.bracket_in_sphere()
    env = ctx.mesh_container(w.container.vertices, w.container.triangles, w.env_first, w.env_shape, w.pitch)
    kernels = ctx.voxelize_objects(w.part.vertices, w.part.triangles, w.catalog, w.pitch)
    s = make_session(ctx, env, w.env_first, kernels, incremental=True)
    cmin, cmax = W.container_z_range(w)
    max_z = W.oriented_max_z(w.part, w.catalog)
    s.set_ranking(0.0, w.pitch, cmin, cmax, max_z)
    print("[ranking] full size: 3 greedy placements", flush=True)
    placed = s.ranked_fill(max_copies=3)
    assert len(placed) == 3
    pages = s.ranked_pages()
    occ = s.environment()
    field, _ = s.proximity_field()
    first = w.env_first
    checked = 0
    for o, page in pages.items():
        if not page:
            continue
        k = kernels[o]
        kc = np.asarray(k["cells"]) != 0
        mass = int(kc.sum())
        prox = s.proximity_values(o)
        tf = [first[a] - k["first"][a] for a in range(3)]
        keys = []
        for c in page:
            u = [c["translation"][a] - tf[a] for a in range(3)]
            keys.append((c["score"], u[2], u[1], u[0]))
            win = occ[u[2]:u[2] + kc.shape[0], u[1]:u[1] + kc.shape[1], u[0]:u[0] + kc.shape[2]]
            assert not win[kc].any() and c["direct_zero_overlap"]
            want = PG.page_scores(prox[u[2]:u[2] + 1, u[1]:u[1] + 1, u[0]:u[0] + 1], mass, 0.0, w.pitch, cmin, cmax,
                                  max_z[o], tf[2] + u[2])[0, 0, 0]
            assert np.float64(c["score"]).view(np.uint64) == np.float64(want).view(np.uint64)
            fw = field[u[2]:u[2] + kc.shape[0], u[1]:u[1] + kc.shape[1], u[0]:u[0] + kc.shape[2]]
            direct = float(fw[kc].sum())
            assert abs(prox[u[2], u[1], u[0]] - direct) <= 1e-10 * max(1.0, direct)
            checked += 1
        assert keys == sorted(keys)
    print(f"[ranking] full size: {checked} page entries checked", flush=True)
    assert checked >= 32 * 90
    some = [o for o in pages if len(pages[o]) == PG.PAGE_SIZE][:8]
    more = s.ranked_pages(orientations=some, cursors={o: pages[o][-1] for o in some})
    for o in some:
        a, b = pages[o][-1], more[o][0]
        ka = (a["score"], a["translation"][2], a["translation"][1], a["translation"][0])
        kb = (b["score"], b["translation"][2], b["translation"][1], b["translation"][0])
        assert ka < kb


def test_page_queries_reject_bad_arguments(ctx):
    import spectrapack_amd as sp
    first, env, kernels = random_scene(8, shape=(16, 18, 20), n=3, kmax=4)
    s = make_session(ctx, env, first, kernels)
    set_ranking(s, env, first, kernels)
    pages = s.ranked_pages()
    c = dict(pages[0][0])
    with pytest.raises(sp.Failure):
        s.ranked_pages(orientations=[0, 0])                     # listed twice
    with pytest.raises(sp.Failure):
        s.ranked_pages(orientations=[7])                        # no such orientation
    with pytest.raises(sp.Failure):
        s.ranked_pages(mode=2)                                  # no such mode
    with pytest.raises(sp.Failure):
        s.ranked_pages(orientations=[0], cursors={0: dict(c, translation=(10 ** 6, 0, 0))})
    with pytest.raises(sp.Failure):
        s.ranked_pages(orientations=[0], mode=1, cursors={0: dict(c, binary_overlap=0.5)})
    # the session still answers correctly afterwards
    again = s.ranked_pages()
    assert [p["translation"] for p in again[0]] == [p["translation"] for p in pages[0]]
