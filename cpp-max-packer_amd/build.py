"""Build the in-tree libraries: the gfx950 HIP library behind the C ABI, the
C++ host mirror of the reference's pack_compute interface, and its test
driver.  Run as `python cpp-max-packer_amd/build.py` (or via
__graft_entry__.build()).  Outputs stay in the source tree so they travel to
the GPU box with the repository snapshot."""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
PKG = os.path.join(HERE, "spectrapack_amd")
LIB = os.path.join(PKG, "libspectrapack_hip.so")
HOST_LIB = os.path.join(PKG, "libspectrapack_compute.so")
HOST_TEST = os.path.join(PKG, "spectrapack_host_test")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_SOURCES = ["runtime.hip", "correlate.hip", "step.hip", "raster.hip", "learned.hip"]
# Packed FP32 instructions are disabled for all device code; host passes warn
# that they do not know the feature, which is harmless.
HIP_FLAGS = ["-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include")]
PER_FILE_FLAGS = {
    # exact IEEE double arithmetic (no contraction into FMA) for the
    # interval rasteriser and the exact predicates
    "raster.hip": ["-ffp-contract=off"],
}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout[-6000:])
    return r.stdout


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    sources = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    cmds = []
    for s in sources:
        obj = os.path.join(BUILD, s.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + PER_FILE_FLAGS.get(s, []) + ["-c", os.path.join(CSRC, s), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(len(cmds), 4)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # C++20 host mirror of engine/pack_compute over the C ABI (no HIP headers)
    host_srcs = [os.path.join(HOST, f) for f in ("correlation_hip.cpp", "placement.cpp", "exact_validation.cpp",
                                                  "validate_capi.cpp")
                 if os.path.exists(os.path.join(HOST, f))]
    if host_srcs:
        tmp = HOST_LIB + ".tmp"
        _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-I" + os.path.join(HOST, "include"),
              "-I" + os.path.join(ROOT, "include"), "-o", tmp] + host_srcs +
             ["-L" + PKG, "-lspectrapack_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, HOST_LIB)
        test_src = os.path.join(HOST, "host_test.cpp")
        if os.path.exists(test_src):
            tmp = HOST_TEST + ".tmp"
            _run(["g++", "-std=c++20", "-O2", "-I" + os.path.join(HOST, "include"), "-I" + os.path.join(ROOT, "include"),
                  "-o", tmp, test_src, "-L" + PKG, "-lspectrapack_compute", "-lspectrapack_hip",
                  "-Wl,-rpath,$ORIGIN"])
            os.replace(tmp, HOST_TEST)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
