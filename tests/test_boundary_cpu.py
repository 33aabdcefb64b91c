"""CPU checks of the drop-in boundary that need no GPU: the admission
estimate (engine/pack_compute/src/correlation.cpp:596-641) against the
reference's own expected figures, the mirror's result types against the
reference's layout, and the reference solver's sources compiling unchanged
against the mirror headers."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cpp-max-packer_amd")
DRIVER = os.path.join(PKG, "spectrapack_amd", "spectrapack_host_test")
MIRROR = os.path.join(PKG, "host", "include")
REF = "/root/reference/engine"
# 2 sizeof(CorrelationOutcome) + sizeof(vector<double>) + sizeof(OperationControl)
# with g++/libstdc++ on x86-64 (correlation.cpp:67-72)
METADATA = 2 * 120 + 24 + 72

needs_reference = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")


def expected_working_bytes(padded_cells, longest_axis, reserved=0, threads=1):
    """engine/pack_compute/tests/correlation_tests.cpp:208-214."""
    return (reserved + METADATA + 2 * padded_cells * 16 + padded_cells * 8
            + threads * (64 * longest_axis * 16 + (64 << 10)) + 8 * 8)


def test_host_estimate_checks():
    r = subprocess.run([DRIVER, "estimate"], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"metadata bytes {METADATA}" in r.stdout


def test_c_abi_estimate_follows_reference_formula():
    import spectrapack_amd as sp
    one = ((1, 1, 1), (1, 1, 1), (0, 0, 0), (0, 0, 0))
    e = sp.estimate(one, 1, METADATA)
    # correlation_tests.cpp:593: 66,664 accounted bytes besides the metadata
    assert e["working_bytes"] == 66_664 + METADATA
    assert e["padded_cells"] == 1 and e["device_bytes"] > 0
    spec = ((64, 1, 1), (64, 1, 1), (6, -5, 9), (-3, 4, -7))
    for threads in (1, 2, 4, 8):
        assert sp.estimate(spec, threads, METADATA)["working_bytes"] == expected_working_bytes(127, 127,
                                                                                               threads=threads)
    for threads in (0, 9):
        with pytest.raises(sp.Failure) as f:
            sp.estimate(spec, threads, METADATA)
        assert f.value.code == "CPU_THREAD_COUNT_UNSUPPORTED"
    # a padded axis beyond the device FFT: admission bytes still follow the
    # formula; no device figure
    long_line = ((65537, 2, 1), (1, 1, 1), (7, -3, 1), (-4, 2, 0))
    e = sp.estimate(long_line, 8, METADATA)
    assert e["working_bytes"] == expected_working_bytes(65537 * 2, 65537, threads=8) > 512 << 20
    assert e["device_bytes"] == 0


def test_planted_formula_error_is_caught():
    # a dropped output term (the double per padded cell) must disagree with the reference's figure
    planted = METADATA + 2 * 1 * 16 + 64 * 1 * 16 + (64 << 10) + 64
    assert planted != 66_664 + METADATA


@needs_reference
def test_mirror_result_types_have_reference_layout(tmp_path):
    src = tmp_path / "sizes.cpp"
    src.write_text('#include <cstdio>\n#include "spectrapack/compute/correlation.hpp"\n'
                   'int main() { std::printf("%zu %zu %zu %zu\\n", sizeof(spectrapack::compute::CorrelationOutcome), '
                   'sizeof(spectrapack::compute::CorrelationLimits), sizeof(spectrapack::compute::CorrelationSpec), '
                   'sizeof(spectrapack::runtime::OperationControl)); }\n')
    outs = []
    for inc in ([MIRROR], [f"{REF}/pack_compute/include", f"{REF}/common/include"]):
        exe = tmp_path / f"sizes{len(outs)}"
        subprocess.run(["g++", "-std=c++20"] + ["-I" + i for i in inc] + [str(src), "-o", str(exe)], check=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert outs[0] == outs[1], outs
    assert int(outs[0].split()[0]) * 2 + 24 + int(outs[0].split()[3]) == METADATA


@needs_reference
@pytest.mark.parametrize("source", ["gap_fill", "spectral_pipeline", "slab_lattice"])
def test_reference_solver_compiles_against_mirror(source):
    """The pack_solver call sites of correlate_binary_cpu / correlate_proximity_cpu /
    estimate_correlation_cpu compile unchanged when the mirror headers replace
    pack_compute's (the mirror's include directory comes first; pack_compute's
    own include directory is not on the path)."""
    incs = [MIRROR] + [f"{REF}/{d}/include" for d in ("common", "core", "pack_geometry", "pack_io", "pack_solver")]
    cmd = (["g++", "-std=c++20", "-fsyntax-only"] + ["-I" + i for i in incs] + [f"-I{REF}/pack_solver/src",
           f"{REF}/pack_solver/src/{source}.cpp"])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
