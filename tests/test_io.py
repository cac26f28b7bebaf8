"""I/O and harness parity on the CPU: elem::Print / elem::Write emit the
reference's text format (io/Print.hpp:15-35: a title line, each row's
entries each followed by a space, an empty line; io/Write.hpp:17-28: the same
in scientific notation), and the Cholesky test driver answers --help with
the reference harness's argument report (tests/lapack-like/Cholesky.cpp:
148-156 through core/imports/choice.hpp:225-270).  The programs are host
code; they load the library but make no device call here."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IO_TEST = os.path.join(ROOT, "tests", "cpp", "io_test")
DRIVER = os.path.join(ROOT, "elemental_amd", "bin", "Cholesky")


def _need(path):
    if not os.path.exists(path):
        pytest.skip(f"{path} is built by __graft_entry__.build()")


def test_print_and_write_format(tmp_path):
    _need(IO_TEST)
    out = tmp_path / "written.txt"
    r = subprocess.run([IO_TEST, str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    # std::ostream's default format: 6 significant digits, %g style
    assert r.stdout == "A\n1 2.5 -0.125 \n-3 1e-07 1.23457e+08 \n\n0.5 -7.25 \n\n"
    assert out.read_text() == ("A written\n1.000000e+00 2.500000e+00 -1.250000e-01 \n"
                               "-3.000000e+00 1.000000e-07 1.234568e+08 \n\n")


def test_driver_help_report():
    _need(DRIVER)
    r = subprocess.run([DRIVER, "--help", "--height", "64", "--uplo", "U"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    expect = ("Optional arguments:\n"
              "  --gridHeight [i,0,0,NOT found]\n    process grid height\n\n"
              "  --uplo [c,L,U,found]\n    upper or lower storage: L/U\n\n"
              "  --height [i,100,64,found]\n    height of matrix\n\n"
              "  --nb [i,96,96,NOT found]\n    algorithmic blocksize\n\n"
              "  --nbLocal [i,32,32,NOT found]\n    local blocksize\n\n"
              "  --correctness [b,1,1,NOT found]\n    test correctness?\n\n"
              "  --print [b,0,0,NOT found]\n    print matrices?\n\n"
              "Out of 0 required arguments, 0 were not specified.\n"
              "Out of 7 optional arguments, 5 were not specified.\n\n")
    assert r.stdout == expect
