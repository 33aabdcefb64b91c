"""Pins the oracle of the learned proposal network: BF16 rounding against
torch's conversion, conv taps against scipy.ndimage.correlate, the downsample
against closed forms and a per-bin loop, and the analytic gradients against
central finite differences of the FP64 loss."""
import numpy as np
import pytest
import scipy.ndimage

import cnn


def test_bf16_rounding_matches_torch():
    import torch
    rs = np.random.default_rng(0)
    x = np.concatenate([rs.normal(0, 3, 5000), [1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -1 - 2.0 ** -8, 0.0, 1e-30]])
    x32 = x.astype(np.float32)
    ref = torch.from_numpy(x32).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(cnn.bf16(x32), ref)
    assert cnn.bf16(np.float32(1 + 2.0 ** -8)) == 1.0
    assert cnn.bf16(np.float32(1 + 3 * 2.0 ** -8)) == 1 + 2.0 ** -6


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_taps_match_scipy(stride):
    rs = np.random.default_rng(1 + stride)
    x = rs.normal(size=(1, 8, 8, 8, 3))
    w = rs.normal(size=(2, 27, 3))
    b = rs.normal(size=2)
    y, _ = cnn.conv(x, w, b, stride)
    for co in range(2):
        ref = sum(scipy.ndimage.correlate(x[0, ..., ci], w[co, :, ci].reshape(3, 3, 3), mode="constant", cval=0.0)
                  for ci in range(3)) + b[co]
        ref = ref[::stride, ::stride, ::stride]
        assert np.allclose(y[0, ..., co], ref, atol=1e-12)


def test_col2im_is_adjoint_of_im2col():
    rs = np.random.default_rng(3)
    for stride in (1, 2):
        x = rs.normal(size=(2, 8, 8, 8, 2))
        c = cnn._im2col(x, stride)
        r = rs.normal(size=c.shape)
        assert np.isclose((c * r).sum(), (x * cnn._col2im(r, stride, 8)).sum())


def test_downsample_closed_forms_and_loop():
    assert (cnn.downsample(np.ones((64, 70, 130), np.uint8)) == 255).all()
    assert (cnn.downsample(np.zeros((256, 64, 66), np.uint8)) == 0).all()
    rs = np.random.default_rng(4)
    env = (rs.uniform(size=(128, 128, 128)) < 0.3).astype(np.uint8)
    got = cnn.downsample(env)
    counts = env.reshape(64, 2, 64, 2, 64, 2).sum(axis=(1, 3, 5))
    assert np.array_equal(got, ((255 * counts + 4) // 8).astype(np.uint8))
    env = (rs.uniform(size=(70, 66, 130)) < 0.5).astype(np.uint8)   # [z, y, x]
    got = cnn.downsample(env)
    for bz, by, bx in [(0, 0, 0), (63, 63, 63), (17, 40, 5), (31, 2, 60)]:
        z0, z1 = bz * 70 // 64, (bz + 1) * 70 // 64
        y0, y1 = by * 66 // 64, (by + 1) * 66 // 64
        x0, x1 = bx * 130 // 64, (bx + 1) * 130 // 64
        cnt = int(env[z0:z1, y0:y1, x0:x1].sum())
        cells = (z1 - z0) * (y1 - y0) * (x1 - x0)
        assert got[bz, by, bx] == int(np.floor(255 * cnt / cells + 0.5))


def _problem(seed, B=2):
    rs = np.random.default_rng(seed)
    flat = cnn.init_parameters(seed)
    P = cnn.unpack(flat.copy())
    for name in ("b1", "b2", "b3", "b4", "ba", "bo", "bh"):
        P[name][...] = rs.normal(0, 0.1, P[name].shape)
    flat = cnn.pack(P)
    q = rs.integers(0, 256, size=(B, 64, 64, 64)).astype(np.uint8)
    desc = rs.normal(size=(B, 32))
    mask = (rs.uniform(size=(B, 96)) < 0.8).astype(np.uint8)
    target = rs.uniform(size=(B, 96)) * mask
    target /= target.sum(1, keepdims=True)
    region = np.array([int(rs.integers(0, 4096))] + [-1] * (B - 1))
    return flat, q, desc, target, mask, region


def _fd_errors(flat, q, desc, target, mask, region, hw, seed=6, eps=1e-7):
    """Relative error of the analytic directional derivative per parameter
    group (a random direction inside each group) against central differences.
    eps is small enough that no ReLU changes side."""
    _, _, g, _ = cnn.loss_and_grads(flat, q, desc, target, mask, region, hw, emulate=False)

    def loss(p):
        a, b, _, _ = cnn.loss_and_grads(p, q, desc, target, mask, region, hw, emulate=False)
        return a + hw * b

    rs = np.random.default_rng(seed)
    errs, off = {}, 0
    for name, shape in cnn.LAYOUT:
        n = int(np.prod(shape))
        d = np.zeros_like(flat)
        d[off:off + n] = rs.normal(size=n)
        fd = (loss(flat + eps * d) - loss(flat - eps * d)) / (2 * eps)
        an = float(g @ d)
        errs[name] = abs(fd - an) / max(1.0, abs(an))
        off += n
    return errs


def test_gradients_match_finite_differences():
    errs = _fd_errors(*_problem(5), hw=0.7)
    bad = {k: v for k, v in errs.items() if v > 2e-6}
    assert not bad, bad


def test_planted_sign_error_is_caught():
    """A sign error planted in the oracle's input-gradient scatter fails the
    finite-difference check of the layers below it."""
    orig = cnn._col2im
    cnn._col2im = lambda cols, stride, D: -orig(cols, stride, D)
    try:
        errs = _fd_errors(*_problem(7), hw=1.0)
    finally:
        cnn._col2im = orig
    assert errs["w1"] > 1e-3 and errs["w3"] > 1e-3
    assert errs["wo"] < 2e-6


def test_layout_size():
    assert cnn.PARAMETER_COUNT == 95393 + 64 * 32
