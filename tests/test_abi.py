"""The C-ABI library loads and exports every function include/*.h declares.

No compute calls: this runs without a GPU.
"""
import glob
import os
import re
import subprocess

import pytest

from hygondoubleblockedfactorizations_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(compat=False):
    """Functions the native headers declare (compat=True: the rocblas_handle
    compatibility header instead)."""
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        if (os.sep + "compat" + os.sep in path) != compat:
            continue
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"typedef[^;]*?\{.*?\}[^;]*;", "", text, flags=re.S)
        text = re.sub(r"typedef[^;]*;", "", text)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name in ("if", "for", "while", "sizeof", "extern"):
                continue
            names.add(name)
    return names


COMPAT_LIB = os.path.join(os.path.dirname(_lib.LIB_PATH), "libmifact_compat.so")


def exported_symbols(path=None):
    out = subprocess.run(["nm", "-D", "--defined-only", path or _lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_loads():
    lib = _lib.load()
    assert lib.mifact_version().startswith(b"mifact")


def test_every_declared_symbol_is_exported():
    declared = declared_functions()
    assert len(declared) >= 40
    missing = declared - exported_symbols()
    assert not missing, sorted(missing)


def test_binding_table_matches_headers():
    assert set(_lib.SIGNATURES) == declared_functions()


def test_reference_entry_points_present():
    # The reference's public driver entry points
    # (include/potrf/double_blocked_potrf.h:9-103,
    #  include/getrf/double_blocked_getrf.h:26-74).
    ref = {
        "double_blocked_potrf", "double_blocked_potrf_dbtrsm",
        "double_blocked_potrf_kblocked_syrk",
        "double_blocked_potrf_kblocked_syrk_adaptive_gemm",
        "double_blocked_potrf_kblocked_syrk_adaptive_gemm_configured",
        "double_blocked_potrf_double", "double_blocked_potrf_double_dbtrsm",
        "double_blocked_potrf_double_kblocked_syrk",
        "double_blocked_potrf_double_kblocked_syrk_adaptive_gemm",
        "double_blocked_potrf_double_kblocked_syrk_adaptive_gemm_configured",
        "double_blocked_getrf", "double_blocked_getrf_double",
        "double_blocked_getrf_batched_pivot",
        "double_blocked_getrf_double_batched_pivot",
        "gemm_nt_update", "gemm_nt_update_double", "gemm_nt_update_m_tiled",
        "gemm_nt_update_m_tiled_double", "syrk_lower_update",
        "syrk_lower_update_double", "syrk_lower_update_kblocked",
        "syrk_lower_update_kblocked_double", "trsm_right_lower_transpose",
        "trsm_right_lower_transpose_double", "double_blocked_trsm",
        "double_blocked_trsm_double",
    }
    assert not (ref - exported_symbols())


def test_no_rocblas_dependency():
    out = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "rocblas" not in out and "rocsolver" not in out


def test_missing_library_fails_loudly(tmp_path):
    saved = _lib._lib
    try:
        _lib._lib = None
        with pytest.raises(_lib.LibraryNotBuilt):
            _lib.load(str(tmp_path / "absent.so"))
    finally:
        _lib._lib = saved


def test_compat_library_exports_reference_prototypes():
    """libmifact_compat.so exports every prototype of the compatibility
    header (the reference's names, taking rocblas_handle); it links rocBLAS
    (for rocblas_get_stream only) while libmifact.so does not, and every
    name it forwards exists in libmifact.so."""
    declared = declared_functions(compat=True)
    assert len(declared) >= 30
    assert not (declared - exported_symbols(COMPAT_LIB))
    assert not (declared - exported_symbols())
    out = subprocess.run(["ldd", COMPAT_LIB], capture_output=True, text=True).stdout
    assert "rocblas" in out
    assert "libmifact.so" not in out           # loaded privately at run time


def test_compat_library_name_not_taken_for_rocblas():
    """rocBLAS looks for its Tensile kernels beside a loaded module whose
    file name ends in "rocblas.so"; a compatibility library so named made a
    later rocBLAS GEMM in the same process abort.  The name must not match."""
    name = os.path.basename(COMPAT_LIB)
    assert "rocblas" not in name
