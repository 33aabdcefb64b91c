"""Every tile the autotuner may pick gives the same overlap fields and choice
as the default tiles (which the other step tests hold to the oracle): every
(radix, width) of the register-FFT passes at 128-, 256- and 512-point axes and
every width of the generic passes at other lengths, in FP32 and FP64; and the
autotuned session matches too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def scene(shape, seed):
    rs = np.random.default_rng(seed)
    env = (rs.uniform(0, 1, shape[::-1]) < 0.02).astype(np.uint8)
    env[:, :, : shape[0] // 3] = 1
    kernels = []
    for _ in range(3):
        s = tuple(int(v) for v in rs.integers(30, 60, 3))
        kernels.append(dict(first=tuple(int(v) for v in rs.integers(-5, 5, 3)), shape=s,
                            cells=(rs.uniform(0, 1, s[::-1]) < 0.7).astype(np.uint8)))
    return env, kernels


def run(ctx, env, kernels, precision, tiles=None, autotune=False):
    import spectrapack_amd as sp
    s = sp.Session(ctx, env.shape[::-1], (0, 0, 0), len(kernels), precision=precision)
    s.set_environment(env)
    for i, k in enumerate(kernels):
        s.set_kernel(i, k["cells"], k["shape"], k["first"])
    s.set_autotune(autotune)
    for p, t in (tiles or {}).items():
        s.set_tile(p, *t)
    s.keep_fields(True)
    r = s.step(stamp=False)
    fields = [s.overlap_field(i) for i in range(len(kernels))]
    used = [s.tile(p) for p in range(3)]
    s.close()
    return r, fields, used


@pytest.mark.parametrize("shape", [(128, 128, 128), (256, 256, 256), (512, 256, 256), (256, 512, 256), (256, 256, 512), (96, 90, 100)])
@pytest.mark.parametrize("precision", [32, 64])
def test_every_tile_matches_the_default(ctx, shape, precision):
    import spectrapack_amd as sp
    env, kernels = scene(shape, sum(shape) + precision)
    r0, f0, _ = run(ctx, env, kernels, precision)
    probe = sp.Session(ctx, shape, (0, 0, 0), len(kernels), precision=precision)
    probe.set_environment(env)
    for i, k in enumerate(kernels):
        probe.set_kernel(i, k["cells"], k["shape"], k["first"])
    candidates = {p: probe.tile_candidates(p) for p in range(3)}
    probe.close()
    for p in range(3):
        assert candidates[p], p
        for t in candidates[p]:
            print(f"[tiles] shape={shape} fp{precision} pass={p} tile={t}", flush=True)
            r, f, used = run(ctx, env, kernels, precision, {p: t})
            assert used[p][:2] == t
            assert r["orientation"] == r0["orientation"] and r["translation"] == r0["translation"]
            for a, b in zip(f, f0):
                assert np.array_equal(a, b)
    print(f"[tiles] shape={shape} fp{precision} autotuned", flush=True)
    r, f, used = run(ctx, env, kernels, precision, autotune=True)
    print(f"[tiles] chosen {used}", flush=True)
    assert r["orientation"] == r0["orientation"] and r["translation"] == r0["translation"]
    for a, b in zip(f, f0):
        assert np.array_equal(a, b)
