"""CPU-side checks of the boundary: the C-ABI library loads and exports every
symbol include/*.h declares (no compute calls without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(prefix="sp_hip_"):
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "sp_hip_correlate_binary" in names and "sp_hip_session_step" in names


def test_library_exports_every_declared_symbol():
    import spectrapack_amd as sp
    lib = ctypes.CDLL(sp.library_path())
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_status_codes_match_reference_strings():
    import spectrapack_amd as sp
    assert sp.status_code(9) == "CORRELATION_NUMERIC"
    assert sp.status_code(8) == "CORRELATION_PADDED_CELL_LIMIT"
    assert sp.status_code(17) == "OPERATION_CANCELLED"


def test_fp32_bound_monotone_and_selects():
    import spectrapack_amd as sp
    n = 256 ** 3
    assert sp.fp32_error_bound(n, n, 1000) < sp.fp32_error_bound(n, n, 10000)
    assert sp.fp32_error_bound(n, n, 20000) < 0.4
    assert sp.fp32_error_bound(512 ** 3, 512 ** 3, 200000) > 0.4


def test_host_library_exports_the_validation_abi():
    from spectrapack_amd import validation
    lib = ctypes.CDLL(validation.LIB_PATH)
    names = declared_symbols("sp_validat")
    assert "sp_validate_placements" in names and "sp_validator_try_add" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
