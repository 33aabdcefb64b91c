"""Pins of the ranked-page oracle (oracle/pages.py) on the CPU.

The page order is checked against a literal sort with the reference's
comparator (spectral_pipeline.cpp:1749-1766) on inputs with forced score ties;
cursor pages against slices of the full order (as spectral_tests.cpp:3066-3130
checks the second page against candidates 32..63); the score against a
hand-computed known answer; the direct proximity sum against the FFT
restatement of correlate_proximity_cpu; the greedy loop against a brute-force
argmin over every feasible (orientation, translation) at each placement."""
import functools

import numpy as np
import pytest

import pages as PG


def reference_order(cands, low_overlap):
    def rank_less(a, b):
        if a["score"] != b["score"]:
            return a["score"] < b["score"]
        return tuple(a["translation"][::-1]) < tuple(b["translation"][::-1])

    def less(a, b):
        if low_overlap and a["binary_overlap"] != b["binary_overlap"]:
            return a["binary_overlap"] < b["binary_overlap"]
        return rank_less(a, b)

    return sorted(cands, key=functools.cmp_to_key(lambda a, b: -1 if less(a, b) else (1 if less(b, a) else 0)))


def all_candidates(overlap, scores, first, low_overlap):
    out = []
    lz, ly, lx = overlap.shape
    for z in range(lz):
        for y in range(ly):
            for x in range(lx):
                if not low_overlap and round(float(overlap[z, y, x])) != 0:
                    continue
                out.append(dict(translation=(first[0] + x, first[1] + y, first[2] + z), score=float(scores[z, y, x]),
                                binary_overlap=float(overlap[z, y, x])))
    return out


def tied_problem(seed, shape=(5, 6, 7)):
    rs = np.random.default_rng(seed)
    overlap = rs.integers(0, 3, shape)
    overlap[rs.uniform(size=shape) < 0.5] = 0
    # few distinct scores: many exact ties, including -0.0 against 0.0
    scores = rs.choice(np.array([-0.25, -0.0, 0.0, 0.125, 0.5, 0.5000000000000001]), shape)
    first = tuple(int(v) for v in rs.integers(-4, 4, 3))
    return overlap, scores, first


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("mode", [0, 1])
def test_page_order_is_the_reference_comparator(seed, mode):
    overlap, scores, first = tied_problem(seed)
    want = reference_order(all_candidates(overlap, scores, first, mode == 1), mode == 1)
    got = PG.ranked_page(overlap, scores, first, mode=mode, size=10 ** 6)
    assert [c["translation"] for c in got] == [c["translation"] for c in want]
    assert [c["score"] for c in got] == [c["score"] for c in want]


@pytest.mark.parametrize("mode", [0, 1])
def test_cursor_pages_are_slices_of_the_order(mode):
    overlap, scores, first = tied_problem(7, shape=(6, 7, 8))
    full = PG.ranked_page(overlap, scores, first, mode=mode, size=10 ** 6)
    assert len(full) > 3 * PG.PAGE_SIZE
    pages, cursor = [], None
    while True:
        p = PG.ranked_page(overlap, scores, first, mode=mode, cursor=cursor)
        if not p:
            break
        pages.extend(p)
        cursor = p[-1]
    assert [c["translation"] for c in pages] == [c["translation"] for c in full]


def test_excluded_translations_never_appear():
    overlap, scores, first = tied_problem(3)
    full = PG.ranked_page(overlap, scores, first, size=10 ** 6)
    drop = [full[0]["translation"], full[5]["translation"]]
    got = PG.ranked_page(overlap, scores, first, excluded=drop, size=10 ** 6)
    assert [c["translation"] for c in got] == [c["translation"] for c in full if c["translation"] not in drop]


def test_page_score_known_answer():
    # origin_z -0.2, pitch 0.375, t_z 3 -> world_z 0.925; object max z 0.6,
    # container z in [0, 1.25] -> top = 1.525 / 1.25 = 1.22; proximity 2.5 over
    # 4 kernel ones -> mean 0.625; score = .65 * 1.22 - .35 * .625 = 0.57425
    s = PG.page_scores(np.full((1, 1, 1), 2.5), 4, -0.2, 0.375, 0.0, 1.25, 0.6, 3)
    assert s[0, 0, 0] == pytest.approx(0.57425, abs=1e-12)
    # one z level up adds .65 * pitch / height; more proximity lowers the score
    s2 = PG.page_scores(np.array([[[2.5]], [[2.5]]]), 4, -0.2, 0.375, 0.0, 1.25, 0.6, 3)
    assert s2[1, 0, 0] - s2[0, 0, 0] == pytest.approx(.65 * .375 / 1.25, abs=1e-12)
    s3 = PG.page_scores(np.full((1, 1, 1), 3.5), 4, -0.2, 0.375, 0.0, 1.25, 0.6, 3)
    assert s[0, 0, 0] - s3[0, 0, 0] == pytest.approx(.35 * 1.0 / 4, abs=1e-12)


def test_proximity_direct_matches_fft_restatement(oracle):
    rs = np.random.default_rng(5)
    occ = (rs.uniform(size=(9, 10, 11)) < 0.1).astype(np.uint8)
    field, _ = oracle.proximity(occ)
    ker = (rs.uniform(size=(3, 4, 2)) < 0.7).astype(np.uint8)
    ker[0, 0, 0] = 1
    direct = PG.proximity_direct(field, ker)
    es, ks = occ.shape[::-1], ker.shape[::-1]
    r = oracle.correlate((es, ks, (0, 0, 0), (0, 0, 0)), field, ker, proximity=True)
    full = r["values"].reshape(r["shape"][2], r["shape"][1], r["shape"][0])
    legal = full[ks[2] - 1:ks[2] - 1 + direct.shape[0], ks[1] - 1:ks[1] - 1 + direct.shape[1],
                 ks[0] - 1:ks[0] - 1 + direct.shape[2]]
    assert np.max(np.abs(legal - direct) / np.maximum(1.0, np.abs(direct))) <= 1e-10


def small_scene(oracle, seed):
    rs = np.random.default_rng(seed)
    env = np.zeros((9, 10, 11), np.uint8)
    env[0] = 1
    env[rs.uniform(size=env.shape) < 0.04] = 1
    kernels = []
    for _ in range(3):
        s = tuple(int(v) for v in rs.integers(2, 4, 3))
        c = (rs.uniform(size=s[::-1]) < 0.8).astype(np.uint8)
        c.flat[0] = 1
        kernels.append(dict(first=tuple(int(v) for v in rs.integers(-2, 2, 3)), shape=s, cells=c))
    return env, (1, -2, 0), kernels


def oracle_pages(oracle, env, first, kernels, params):
    field, _ = oracle.proximity(env)
    _, _, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)

    def pages_of(orients, cursors):
        out = {}
        for o in orients:
            k = kernels[o]
            if fields[o] is None:
                out[o] = []
                continue
            prox = PG.proximity_direct(field, k["cells"])
            tf = [first[a] - k["first"][a] for a in range(3)]
            sc = PG.page_scores(prox, int(k["cells"].sum()), *params, params_max_z(o), tf[2])
            out[o] = PG.ranked_page(fields[o], sc, tf, cursor=cursors.get(o), env=env, kernel_cells=k["cells"])
        return out
    return pages_of


def params_max_z(o):
    return 0.1 * (o + 1)


def test_ranked_fill_takes_the_best_feasible_candidate(oracle):
    env, first, kernels = small_scene(oracle, 1)
    params = (0.0, 1.0, 0.0, 9.0)
    state = dict(env=env.copy())

    def pages_of(orients, cursors):
        return oracle_pages(oracle, state["env"], first, kernels, params)(orients, cursors)

    def stamp(o, t):
        # brute force: the chosen head is the best (score, orientation, z, y, x)
        # over every feasible translation of every orientation
        best = None
        for oo in range(len(kernels)):
            for c in PG.ranked_page(*full_problem(oracle, state["env"], first, kernels, params, oo), size=10 ** 6):
                c = dict(c, orientation=oo)
                if best is None or PG.rank_less(c, best):
                    best = c
        assert (best["orientation"], tuple(best["translation"])) == (o, tuple(t))
        state["env"] = oracle.stamp(state["env"], first, kernels[o], t)

    placed = PG.ranked_fill(pages_of, stamp, list(range(len(kernels))))
    assert len(placed) >= 5
    # maximal: afterwards no orientation has a zero-overlap translation
    _, _, fields = oracle.place_step(state["env"], first, kernels, direct=True, fields=True)
    assert all(f is None or not (f == 0).any() for f in fields)


def full_problem(oracle, env, first, kernels, params, o):
    field, _ = oracle.proximity(env)
    _, _, fields = oracle.place_step(env, first, kernels, direct=True, fields=True)
    k = kernels[o]
    prox = PG.proximity_direct(field, k["cells"])
    tf = [first[a] - k["first"][a] for a in range(3)]
    sc = PG.page_scores(prox, int(k["cells"].sum()), *params, params_max_z(o), tf[2])
    return fields[o], sc, tf
