"""C-ABI checks that need no GPU: the library loads, exports every function
include/PLA.h declares, keeps the reference's constant values, matches the
MPICH handle ABI, and fails cleanly (an error code, no crash) without a GPU."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from plapack_amd import pla

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_function():
    names = pla.header_functions()
    assert len(names) >= 60
    L = pla.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    # and the binding covers all of them
    assert set(names) <= set(pla._SIGS)


def test_constants_match_reference_values():
    # INCLUDE/PLA_define.h:11-102
    hdr = open(os.path.join(ROOT, "include", "PLA.h")).read()
    expect = {"PLA_LOWER_TRIANGULAR": 52, "PLA_UPPER_TRIANGULAR": 51, "PLA_NO_TRANSPOSE": 21,
              "PLA_TRANSPOSE": 22, "PLA_UNIT_DIAG": 101, "PLA_NONUNIT_DIAG": 102,
              "PLA_SIDE_LEFT": 43, "PLA_SIDE_RIGHT": 44, "PLA_OP_PAN_PAN": 83, "PLA_OP_ALL_ALG": 87,
              "PLA_MATRIX": 13, "PLA_MVECTOR": 12, "PLA_MSCALAR": 14, "PLA_ALL_ROWS": -31,
              "PLA_ALIGN_FIRST": -200, "PLA_MODE_CLOSED": 71, "PLA_MODE_OPEN": 72,
              "PLA_STATE_API_ACTIVE": 73, "PLA_STATE_NORMAL": 74}
    src = "#define PLA_NO_MPI_H\n#include <stdio.h>\n#include \"PLA.h\"\nint main(void){\n"
    for k in expect:
        src += f'printf("{k} %d\\n", (int)({k}));\n'
    src += "return 0;}\n"
    exe = _compile_c(src, "consts", ["gcc"])
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = {l.split()[0]: int(l.split()[1]) for l in out.splitlines()}
    assert got == expect
    for k, v in expect.items():
        if hasattr(pla, k):
            assert getattr(pla, k) == v
    assert "PLA_SUCCESS      0" in hdr


def _compile_c(src, name, cc, extra=()):
    d = os.path.join(ROOT, "tests", "_build")
    os.makedirs(d, exist_ok=True)
    c = os.path.join(d, name + ".c")
    exe = os.path.join(d, name)
    with open(c, "w") as f:
        f.write(src)
    subprocess.run(list(cc) + ["-I", os.path.join(ROOT, "include"), c, "-o", exe] + list(extra),
                   check=True, env=dict(os.environ, MPICH_CC="gcc"))
    return exe


def test_mpich_handle_abi():
    mpicc = shutil.which("mpicc") or "/opt/conda/bin/mpicc"
    if not os.path.exists(mpicc):
        pytest.skip("mpicc not available")
    src = ("#include <mpi.h>\n#include <stdio.h>\nint main(void){printf(\"%d %d %d %d\\n\","
           "(int)MPI_COMM_WORLD,(int)MPI_INT,(int)MPI_DOUBLE,(int)MPI_FLOAT);return 0;}\n")
    exe = _compile_c(src, "mpiabi", [mpicc])
    vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals == [pla.MPI_COMM_WORLD, pla.MPI_INT, pla.MPI_DOUBLE, pla.MPI_FLOAT]


def test_reference_style_driver_compiles_against_header():
    """A driver written like TEST/test_lu.c builds against include/PLA.h and <mpi.h>."""
    mpicc = shutil.which("mpicc") or "/opt/conda/bin/mpicc"
    if not os.path.exists(mpicc):
        pytest.skip("mpicc not available")
    drv = os.path.join(ROOT, "tests", "drivers", "test_lu_driver.c")
    subprocess.run([mpicc, "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    drv], check=True, env=dict(os.environ, MPICH_CC="gcc"))


def test_no_gpu_gives_error_not_crash():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    # run in a child so this process never initialises HIP
    code = ("from plapack_amd import pla\n"
            "rc = pla.lib().PLA_Init(pla.MPI_COMM_WORLD)\n"
            "print(rc, pla.last_error())\n")
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert out.returncode == 0, out.stderr
    rc = int(out.stdout.split()[0])
    assert rc < 0
