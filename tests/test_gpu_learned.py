"""Learned proposal on the GPU: the MFMA tile layout, the downsample
(bit-exact), forward outputs and gradients against the numpy oracle (to BF16
tolerance), the Adam update, and the learned step against the restricted
search it stands for (bit-exact), including its fallback."""
import numpy as np
import pytest

import cnn
from helpers import restricted_choice

pytestmark = pytest.mark.gpu


def bf16_bits(x):
    a = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((a + 0x7FFF + ((a >> 16) & 1)) >> 16).astype(np.uint16)


@pytest.fixture
def proposal(ctx):
    import spectrapack_amd as sp
    p = sp.Proposal(ctx, max_batch=4)
    yield p
    p.close()


def params(seed):
    rs = np.random.default_rng(seed)
    P = cnn.unpack(cnn.init_parameters(seed))
    for name in ("b1", "b2", "b3", "b4", "ba", "bo", "bh"):
        P[name][...] = rs.normal(0, 0.1, P[name].shape)
    return cnn.pack(P).astype(np.float32)


def smooth_samples(rs, n):
    """Occupancy-like inputs: a filled bottom layer of random height plus blobs."""
    z = np.arange(64)[:, None, None]
    out = []
    for _ in range(n):
        h = rs.integers(4, 40)
        q = np.where(z < h, 255, 0) + np.zeros((64, 64, 64), np.int64)
        for _ in range(6):
            c = rs.integers(0, 64, 3)
            r = rs.integers(3, 10)
            zz, yy, xx = np.ogrid[:64, :64, :64]
            q[(zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2 < r * r] = rs.integers(60, 256)
        out.append(q.astype(np.uint8))
    return np.stack(out)


def test_mfma_tile_layout(proposal):
    rs = np.random.default_rng(0)
    a = rs.integers(-8, 9, (32, 16)).astype(np.float32)
    b = rs.integers(-8, 9, (16, 32)).astype(np.float32)
    c = proposal.mfma_probe(bf16_bits(a), bf16_bits(b))
    assert np.array_equal(c, a @ b)


def test_downsample_matches_oracle(ctx, proposal):
    import spectrapack_amd as sp
    rs = np.random.default_rng(1)
    env = (rs.uniform(size=(96, 70, 130)) < 0.4).astype(np.uint8)   # [z, y, x]
    s = sp.Session(ctx, (130, 70, 96), (0, 0, 0), 1)
    s.set_environment(env)
    s.set_kernel(0, np.ones((2, 2, 2), np.uint8), (2, 2, 2), (0, 0, 0))
    proposal.reserve(2)
    proposal.record(s, 1)
    assert np.array_equal(proposal.sample(1), cnn.downsample(env))


def test_forward_matches_oracle(proposal):
    rs = np.random.default_rng(2)
    flat = params(2)
    q = smooth_samples(rs, 2)
    desc = rs.normal(size=(2, 32)).astype(np.float32)
    proposal.set_parameters(flat)
    proposal.reserve(3)
    proposal.load_samples(q, first=1)
    logits, heat = proposal.forward([1, 2], desc)
    ref = cnn.forward(flat.astype(np.float64), q, desc.astype(np.float64), emulate=True)
    for got, want in ((logits, ref["logits"]), (heat, ref["heat"])):
        scale = max(1.0, float(np.abs(want).max()))
        err = np.abs(got - want)
        print(f"[forward] max {err.max() / scale:.2e} mean {err.mean() / scale:.2e}", flush=True)
        assert err.max() <= 2e-2 * scale
        assert err.mean() <= 2e-3 * scale


def _train_problem(seed):
    rs = np.random.default_rng(seed)
    flat = params(seed)
    q = smooth_samples(rs, 3)
    desc = rs.normal(size=(3, 32)).astype(np.float32)
    mask = (rs.uniform(size=(3, 96)) < 0.8).astype(np.uint8)
    target = (rs.uniform(size=(3, 96)) * mask).astype(np.float32)
    target /= target.sum(1, keepdims=True)
    region = np.array([int(rs.integers(0, 4096)), -1, int(rs.integers(0, 4096))], np.int32)
    return flat, q, desc, target, mask, region


def test_gradients_match_oracle(proposal):
    flat, q, desc, target, mask, region = _train_problem(3)
    proposal.set_parameters(flat)
    proposal.reserve(3)
    proposal.load_samples(q)
    (lo, lh), g = proposal.train([0, 1, 2], desc, target, mask, region, heat_weight=0.5, update=False,
                                 gradients=True)
    ro, rh, rg, _ = cnn.loss_and_grads(flat.astype(np.float64), q, desc.astype(np.float64), target, mask, region,
                                       0.5, emulate=True)
    assert abs(lo - ro) <= 1e-3 * max(1.0, abs(ro))
    assert abs(lh - rh) <= 1e-3 * max(1.0, abs(rh))
    off = 0
    for name, shape in cnn.LAYOUT:
        n = int(np.prod(shape))
        a, b = g[off:off + n].astype(np.float64), rg[off:off + n]
        off += n
        nb = np.linalg.norm(b)
        rel = np.linalg.norm(a - b) / max(nb, 1e-12)
        print(f"[grad] {name}: |ref| {nb:.3e} rel {rel:.2e}", flush=True)
        if nb < 1e-9:
            assert np.linalg.norm(a) < 1e-6
        else:
            assert rel <= 3e-2, name


def test_adam_update(proposal):
    flat, q, desc, target, mask, region = _train_problem(4)
    proposal.set_parameters(flat)
    proposal.reserve(3)
    proposal.load_samples(q)
    _, g = proposal.train([0, 1, 2], desc, target, mask, region, lr=1e-3, update=True, gradients=True)
    p1 = proposal.parameters()
    # first Adam step: m / (1 - b1) = g, v / (1 - b2) = g^2
    expect = flat - np.float32(1e-3) * g / (np.sqrt(g * g) + np.float32(1e-8))
    assert np.allclose(p1, expect, rtol=0, atol=2e-6)


def test_training_fits_a_small_batch(proposal):
    flat, q, desc, target, mask, region = _train_problem(5)
    region[1] = 77
    proposal.set_parameters(flat)
    proposal.reserve(3)
    proposal.load_samples(q)
    first = None
    for it in range(150):
        (lo, lh), _ = proposal.train([0, 1, 2], desc, target, mask, region, lr=3e-3, update=True)
        first = first if first is not None else lo + lh
    print(f"[fit] loss {first:.3f} -> {lo + lh:.3f}", flush=True)
    assert lo + lh < 0.5 * first


def _box_scene(oracle):
    from spectrapack_amd import workloads as W
    first, shape = W.box_window((64.0, 64.0, 64.0), 1.0)
    env = oracle.box_container(first, shape, 1.0, (64.0, 64.0, 64.0))
    kernels = []
    for q in oracle.cube_catalog():
        f, s, c = oracle.voxelize_object(W.cuboid(10.0, 6.0, 4.0).vertices, W.cuboid(10.0, 6.0, 4.0).triangles, q, 1.0)
        kernels.append(dict(first=f, shape=s, cells=c))
    return first, env, kernels


def _session(ctx, env, first, kernels):
    import spectrapack_amd as sp
    s = sp.Session(ctx, env.shape[::-1], first, len(kernels), precision=32)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    return s


def _top(values, k, valid=None):
    idx = np.arange(values.size)
    if valid is not None:
        idx = idx[valid]
    order = sorted(idx, key=lambda i: (-float(values[i]), i))
    return order[:k]


def test_learned_step_is_the_restricted_search(ctx, oracle, proposal):
    first, env, kernels = _box_scene(oracle)
    s = _session(ctx, env, first, kernels)
    for _ in range(6):
        s.step(stamp=True)
    env = s.environment()
    rs = np.random.default_rng(6)
    desc = rs.normal(size=32).astype(np.float32)
    proposal.set_parameters(params(6))
    proposal.reserve(1)
    proposal.record(s, 0)
    logits, heat = proposal.forward([0], desc[None])
    es = env.shape[::-1]
    fits = np.array([all(k["shape"][a] <= es[a] for a in range(3)) for k in kernels])
    top_o = 8
    for top_r in (64, 700, 4096):
        print(f"[learned] top_r={top_r}", flush=True)
        orients = _top(logits[0, :len(kernels)], top_o, fits)
        regions = _top(heat[0], top_r)
        mask = np.zeros(4096, np.uint8)
        mask[regions] = 1
        mask = mask.reshape(16, 16, 16)
        _, _, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
        o_exp, u_exp, _ = restricted_choice(fields, es, orients, mask)
        r = proposal.step(s, desc, top_orientations=top_o, top_regions=top_r, stamp=False, fallback=False)
        assert r["orientation"] == o_exp
        assert not r["fallback"]
        if o_exp >= 0:
            u = tuple(r["key"][a] - first[a] for a in range(3))
            assert u[::-1] == u_exp
        s.set_active(orients, mask)
        r2 = s.step(stamp=False)
        s.set_active()
        assert (r2["orientation"], tuple(r2["translation"])) == (r["orientation"], tuple(r["translation"]))


def test_learned_step_falls_back_to_the_full_search(ctx, oracle, proposal):
    first, env, kernels = _box_scene(oracle)
    env = env.copy()
    env[:14, :14, :14] = 1          # no corner of region 0 is feasible
    s = _session(ctx, env, first, kernels)
    P = cnn.unpack(params(7).astype(np.float64))
    P["wh"][...] = 0.0              # equal heat logits: region 0 ranks first
    P["bh"][...] = 0.0
    proposal.set_parameters(cnn.pack(P).astype(np.float32))
    desc = np.zeros(32, np.float32)
    r = proposal.step(s, desc, top_orientations=4, top_regions=1, stamp=False, fallback=False)
    assert r["orientation"] == -1 and not r["fallback"]
    r = proposal.step(s, desc, top_orientations=4, top_regions=1, stamp=True, fallback=True)
    assert r["fallback"]
    o, t, _ = oracle.place_step(env, first, kernels, direct=True, fields=False)
    assert r["orientation"] == o and tuple(r["translation"]) == tuple(t)
    env = oracle.stamp(env, first, kernels[o], t)
    assert np.array_equal(s.environment(), env)


def test_learned_step_rejects_small_or_wide_sessions(ctx, oracle, proposal):
    import spectrapack_amd as sp
    s = sp.Session(ctx, (40, 70, 70), (0, 0, 0), 1)
    s.set_environment(np.zeros((70, 70, 40), np.uint8))
    s.set_kernel(0, np.ones((2, 2, 2), np.uint8), (2, 2, 2), (0, 0, 0))
    with pytest.raises(sp.Failure):
        proposal.step(s, np.zeros(32, np.float32))
    s = sp.Session(ctx, (64, 64, 64), (0, 0, 0), 97)
    s.set_environment(np.zeros((64, 64, 64), np.uint8))
    for o in range(97):
        s.set_kernel(o, np.ones((2, 2, 2), np.uint8), (2, 2, 2), (0, 0, 0))
    with pytest.raises(sp.Failure):
        proposal.step(s, np.zeros(32, np.float32))


def test_prior_enters_the_ranking(ctx, oracle, proposal):
    """With a prior set, the learned step ranks orientations by logit +
    weight x prior; with every region allowed (4096) no region mask applies:
    the step is the restricted search over those orientations."""
    first, env, kernels = _box_scene(oracle)
    s = _session(ctx, env, first, kernels)
    for _ in range(4):
        s.step(stamp=True)
    env = s.environment()
    rs = np.random.default_rng(9)
    desc = rs.normal(size=32).astype(np.float32)
    proposal.set_parameters(params(9))
    proposal.reserve(1)
    proposal.record(s, 0)
    logits, _ = proposal.forward([0], desc[None])
    es = env.shape[::-1]
    fits = np.array([all(k["shape"][a] <= es[a] for a in range(3)) for k in kernels])
    prior = rs.normal(size=len(kernels)).astype(np.float32)
    _, _, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
    for weight in (0.0, 0.5, 50.0):
        proposal.set_prior(prior, weight)
        score = logits[0, :len(kernels)] + np.float32(weight) * prior
        orients = _top(score, 3, fits)
        o_exp, u_exp, _ = restricted_choice(fields, es, orients)
        r = proposal.step(s, desc, top_orientations=3, top_regions=4096, stamp=False, fallback=False)
        print(f"[prior] weight={weight} orientations={orients} -> {r['orientation']}", flush=True)
        assert r["orientation"] == o_exp
        if o_exp >= 0:
            assert tuple(r["key"][a] - first[a] for a in range(3))[::-1] == u_exp
    proposal.set_prior([], 0.0)


def test_lazy_step_against_its_restatement(ctx, oracle, proposal):
    """Lazy learned steps (3 orientations searched per step) on the box
    scene, against a host restatement on the oracle's overlap fields: keep
    each orientation's minimum, drop it when the occupancy covers its corner,
    search the 3 best-scored orientations whose dropped minimum lies below the
    lowest kept one, place the lowest kept minimum; the first step, and a
    step with nothing kept, is the full search.  Every step's choice and the
    final occupancy match bit for bit."""
    first, env, kernels = _box_scene(oracle)
    s = _session(ctx, env, first, kernels)
    rs = np.random.default_rng(10)
    desc = rs.normal(size=32).astype(np.float32)
    proposal.set_parameters(params(10))
    proposal.reserve(1)
    prior = rs.normal(size=len(kernels)).astype(np.float32)
    weight = 0.7
    proposal.set_prior(prior, weight)
    proposal.lazy_reset()
    es = env.shape[::-1]
    n = len(kernels)
    fits = [all(k["shape"][a] <= es[a] for a in range(3)) for k in kernels]
    top = 3
    cur = env.copy()
    keys, valid = None, None

    def full(minima):
        return list(minima), [True] * n

    def lowest(keys, valid):
        c = [(keys[o], o) for o in range(n) if valid[o] and keys[o] is not None]
        return min(c) if c else None

    partial = 0
    for step in range(16):
        _, _, fields = oracle.place_step(cur, first, kernels, direct=True, fields=True)
        minima = restricted_choice(fields, es)[2]
        proposal.record(s, 0)
        logits, _ = proposal.forward([0], desc[None])
        if keys is None:
            keys, valid = full(minima)
        else:
            for o in range(n):
                if valid[o] and keys[o] is not None and fields[o][keys[o]] != 0:
                    valid[o] = False
            best = lowest(keys, valid)
            cand = [o for o in range(n) if fits[o] and not valid[o] and (best is None or (keys[o], o) < best)]
            score = logits[0, :n] + np.float32(weight) * prior
            pick = sorted(cand, key=lambda o: (-float(score[o]), o))[:top]
            partial += len(cand) > top
            for o in pick:
                keys[o], valid[o] = minima[o], True
            if lowest(keys, valid) is None:
                keys, valid = full(minima)
        want = lowest(keys, valid)
        r = proposal.lazy_step(s, desc, top, stamp=True)
        print(f"[lazy] step {step}: searched {r['searched']} full {r['used_full']} -> {r['orientation']}", flush=True)
        if want is None:
            assert r["orientation"] == -1
            break
        (uz, uy, ux), o = want
        assert r["orientation"] == o, step
        assert tuple(r["key"][a] - first[a] for a in range(3)) == (ux, uy, uz), step
        cur = oracle.stamp(cur, first, kernels[o], tuple(r["translation"]))
        # a placed minimum is no longer free: the next step drops it
    assert np.array_equal(s.environment(), cur)
    assert partial > 0  # some steps searched only part of the candidates
    proposal.set_prior([], 0.0)
    proposal.lazy_reset()


@pytest.mark.gpu
def test_lazy_step_follows_session_changes(ctx, oracle, proposal):
    """The kept minima hold only while occupancy grows by stamping.  An
    excluded translation, a new environment and another session each make
    the next lazy step run the full search, whose choice is the plain step's
    (an excluded corner is never placed again; a cleared environment gives
    back the first step's corner)."""
    first, env, kernels = _box_scene(oracle)
    s = _session(ctx, env, first, kernels)
    desc = np.random.default_rng(11).normal(size=32).astype(np.float32)
    proposal.set_parameters(params(11))
    proposal.reserve(1)
    proposal.lazy_reset()
    pose = lambda r: (r["orientation"], tuple(r["translation"]))

    r0 = proposal.lazy_step(s, desc, 3, stamp=True)
    assert r0["used_full"] and r0["orientation"] >= 0
    for _ in range(3):
        r = proposal.lazy_step(s, desc, 3, stamp=True)
        assert r["orientation"] >= 0
    r = proposal.lazy_step(s, desc, 3, stamp=False)
    assert not r["used_full"]
    plain = s.step(stamp=False)
    assert pose(r) == pose(plain)

    s.exclude(plain["orientation"], plain["translation"])
    r = proposal.lazy_step(s, desc, 3, stamp=False)
    assert r["used_full"]
    assert pose(r) != pose(plain)
    assert pose(r) == pose(s.step(stamp=False))

    s.set_environment(env)
    r = proposal.lazy_step(s, desc, 3, stamp=False)
    assert r["used_full"]
    assert pose(r) == pose(r0)

    s2 = _session(ctx, env, first, kernels)
    s2.stamp(r0["orientation"], r0["translation"])
    r = proposal.lazy_step(s2, desc, 3, stamp=False)
    assert r["used_full"]
    assert pose(r) == pose(s2.step(stamp=False))
    r = proposal.lazy_step(s, desc, 3, stamp=False)
    assert r["used_full"] and pose(r) == pose(r0)
    s2.close()
    s.close()
    proposal.lazy_reset()
