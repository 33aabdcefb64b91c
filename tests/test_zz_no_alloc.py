"""The allocation-free correlation (sp_hip_correlate_binary_noalloc) against
the allocating entry point on the same inputs: values, shape, first
translation, report and stats identical, and the values equal to the direct
integer oracle.  Two calls in a row, the second smaller, so the second runs
on scratch that still holds the first call's data."""
import numpy as np
import pytest

from helpers import binary_pattern, direct_expected

pytestmark = pytest.mark.gpu


def pair(ctx, oracle, spec, env, ker, precision):
    print(f"[noalloc] spec={spec} precision={precision}", flush=True)
    a = ctx.correlate(spec, env, ker, precision=precision, noalloc=True)
    b = ctx.correlate(spec, env, ker, precision=precision)
    assert np.array_equal(a["values"], b["values"])
    assert np.array_equal(a["translation_first"], b["translation_first"])
    assert np.array_equal(a["shape"], b["shape"])
    # the raw mass is a sum of float atomics, whose order varies from call
    # to call on either entry point; every other report field is exact
    na, nb = dict(a["numeric"]), dict(b["numeric"])
    ma, mb = na.pop("mass_residual"), nb.pop("mass_residual")
    assert na == nb
    assert abs(ma - mb) < 1e-6 and ma < 0.25
    assert a["stats"] == b["stats"]
    first, vals = direct_expected(oracle, spec, env, ker)
    assert np.array_equal(a["translation_first"], first)
    assert np.array_equal(a["values"], vals.astype(np.float64))
    return a


@pytest.mark.parametrize("precision", [0, 32, 64])
def test_noalloc_matches_allocating_call_twice(ctx, oracle, precision):
    rs = np.random.default_rng(11)
    big = ((40, 33, 29), (9, 7, 12), (3, -8, 5), (-2, 4, -6))
    env1 = (rs.uniform(0, 1, (29, 33, 40)) < 0.3).astype(np.uint8)
    ker1 = binary_pattern((9, 7, 12), 5)
    r1 = pair(ctx, oracle, big, env1, ker1, precision)
    small = ((6, 5, 4), (3, 2, 2), (1, 0, -2), (0, 1, 0))
    env2 = binary_pattern((6, 5, 4), 3)
    ker2 = binary_pattern((3, 2, 2), 11)
    r2 = pair(ctx, oracle, small, env2, ker2, precision)
    assert r2["values"].size < r1["values"].size


def test_noalloc_golden_vector(ctx):
    # engine/pack_compute/tests/correlation_tests.cpp:292-308
    spec = ((3, 2, 1), (2, 2, 1), (4, -3, 0), (-1, 1, 0))
    g = ctx.correlate(spec, np.array([1, 0, 1, 0, 1, 1], np.uint8), np.array([1, 0, 1, 1], np.uint8), noalloc=True)
    assert list(g["translation_first"]) == [4, -5, 0] and list(g["shape"]) == [4, 3, 1]
    assert g["values"].tolist() == [1, 1, 1, 1, 0, 2, 2, 2, 0, 0, 1, 1]


def test_noalloc_rejects_what_its_scratch_cannot_hold(ctx):
    import spectrapack_amd as sp
    # FP64 on a 256^3 FFT grid needs far more than the 64 MiB scratch; the
    # call is refused before any device work on the inputs' spectra
    spec = ((250, 250, 250), (4, 4, 4), (0, 0, 0), (0, 0, 0))
    env = np.zeros(250 ** 3, np.uint8)
    env[::7] = 1
    ker = np.ones(64, np.uint8)
    with pytest.raises(sp.Failure) as f:
        ctx.correlate(spec, env, ker, limits=(1 << 40, 0, 1 << 30, 200000000), precision=64, noalloc=True)
    assert f.value.code == "SCRATCH_LIMIT"
    # the same call through the allocating entry point still works afterwards
    small = ((6, 5, 4), (3, 2, 2), (1, 0, -2), (0, 1, 0))
    g = ctx.correlate(small, binary_pattern((6, 5, 4), 3), binary_pattern((3, 2, 2), 11), noalloc=True)
    assert g["values"].size == 8 * 6 * 5
