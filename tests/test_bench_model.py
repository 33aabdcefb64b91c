"""bench.py's byte models and its single-GPU guard (CPU only).

The algorithmic bytes of each step kernel are checked against sums written out
for a small window with two kernels, and the fixed step model against its
closed form at 256^3 x 96 (104.24 GB)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def kernels(*shapes):
    return [dict(shape=s) for s in shapes]


def test_step_model_closed_form():
    # 48 Nc + O (120 Nc + 4 N), Nc = (nx/2 + 1) ny nz, N = nx ny nz
    assert bench.step_model_bytes(256, 256, 256, 96, 32) == 104_239_988_736
    assert bench.step_model_bytes(256, 256, 256, 96, 64) == 2 * 104_239_988_736


def test_kernel_bytes_small_window():
    n = (16, 12, 10)         # FFT grid
    ext = (14, 12, 9)        # environment window
    ks = kernels((3, 2, 4), (5, 12, 2), (15, 1, 1))  # the last one does not fit
    nh = 10 // 2 + 1
    e = 8
    # product pass: each fitting orientation's kernel rows in and legal columns out, the environment spectrum once
    want_mul = nh * 12 * 16 * e + nh * 12 * (3 + 12) * e + nh * 12 * (5 + 10) * e
    assert bench.kernel_bytes("mul_x", n, ks, ext, e) == want_mul
    # the fused path's product pass: kz = 0 and kz = 5 share one of 5 output planes
    want_mul_fused = nh * 12 * 16 * e + (nh * 3 + 5 * 12) * 12 * e + (nh * 5 + 5 * 10) * 12 * e
    assert bench.kernel_bytes("mul_x", n, ks, ext, e, fused=True) == want_mul_fused
    # fused inverse y + z: every plane of every legal column (5 planes)
    assert bench.kernel_bytes("yz_argmin", n, ks, ext, e) == 5 * 12 * 12 * e + 5 * 12 * 10 * e
    # separate inverse y: ny rows in and Ly rows out per legal column; inverse z: the Ly rows once
    assert bench.kernel_bytes("inv_c2c_y", n, ks, ext, e) == nh * (12 + 11) * 12 * e + nh * (12 + 1) * 10 * e
    assert bench.kernel_bytes("c2r_z_argmin", n, ks, ext, e) == nh * 11 * 12 * e + nh * 1 * 10 * e
    assert bench.kernel_bytes("env_c2c_y", n, ks, ext, e) == 2 * nh * 12 * 16 * e
    assert bench.kernel_bytes("unknown", n, ks, ext, e) == 0


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "--gpus 1" in r.stderr
