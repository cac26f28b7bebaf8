"""CPU checks of the boundary: the HIP library loads (no GPU needed to
dlopen it) and exports every function include/elemental_amd.h declares, and
the ctypes binding covers exactly that list."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "elemental_amd.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(elemamd_\w+)\s*\(", src)))


def test_header_declares_the_path():
    names = declared()
    for must in ["elemamd_cholesky_f64", "elemamd_hpd_solve_f64", "elemamd_solve_after_f64", "elemamd_trsm_f64",
                 "elemamd_herk_f64", "elemamd_trrk_f64", "elemamd_trsv_f64", "elemamd_potrf_unb_f64",
                 "elemamd_hpd_solve_mixed_f64"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from elemental_amd import _lib

    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(elemamd_\w+)\b", out))
    assert set(declared()) <= exported


def test_binding_matches_header():
    from elemental_amd import _lib

    assert sorted(_lib.SIGNATURES) == declared()


def test_version_string_without_gpu():
    from elemental_amd import lib

    assert b"gfx950" in lib().elemamd_version()


def test_library_is_gfx950_code_object():
    # the offload bundle embedded in the library names its target
    from elemental_amd import _lib

    data = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
