"""The reference's own drivers build against include/PLA.h and link against
libpla_mi355x.so unmodified (the drop-in contract), and the constants they
use keep the reference's values.

The driver sources are compiled where they lie under /root/reference (read,
never copied); the tests skip where the reference or mpicc is absent (e.g. on
a GPU box).  Nothing here runs device code: the host-only entry points
(PLA_Abort, PLA_malloc / PLA_calloc / PLA_free, PLA_Comm_1D_to_2D_ratio) are
exercised without a GPU.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
LIBDIR = os.path.join(ROOT, "plapack_amd")
BUILD = os.path.join(ROOT, "tests", "_build")

# each program: its sources under the reference tree (TEST/testall and the
# EXAMPLES makefiles link them this way)
DRIVERS = {
    "test_lu": ["TEST/test_lu.c", "TEST/create_problem.c"],
    "test_chol": ["TEST/test_chol.c", "TEST/create_problem.c"],
    "test_gemm": ["TEST/test_gemm.c", "TEST/create_problem.c"],
    "test_syrk": ["TEST/test_syrk.c", "TEST/create_problem.c"],
    "test_trsm": ["TEST/test_trsm.c", "TEST/create_problem.c"],
    "lu_example": ["EXAMPLES/LU/driver.c", "EXAMPLES/LU/create_problem.c", "EXAMPLES/LU/process_answer.c"],
    "chol_example": ["EXAMPLES/Cholesky/driver.c", "EXAMPLES/Cholesky/create_problem.c"],
}


def _mpicc():
    mpicc = shutil.which("mpicc") or "/opt/conda/bin/mpicc"
    if not os.path.exists(mpicc):
        pytest.skip("mpicc not available")
    return mpicc


def _need_reference():
    if not os.path.isdir(os.path.join(REF, "TEST")):
        pytest.skip("the reference sources are not present")


def _link_flags():
    if not os.path.exists(os.path.join(LIBDIR, "libpla_mi355x.so")):
        pytest.skip("libpla_mi355x.so is not built")
    # the system libstdc++ (needed by the HIP runtime) ahead of MPICH's own
    return ["-lm", "-L/usr/lib/x86_64-linux-gnu", "-Wl,-rpath-link,/usr/lib/x86_64-linux-gnu",
            "-L", LIBDIR, "-lpla_mi355x", "-Wl,-rpath," + LIBDIR, "-Wl,-rpath-link,/opt/rocm/lib",
            "-Wl,-rpath,/usr/lib/x86_64-linux-gnu"]


@pytest.mark.parametrize("name", sorted(DRIVERS))
def test_reference_driver_builds_and_links(name):
    _need_reference()
    mpicc = _mpicc()
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, "ref_" + name)
    srcs = [os.path.join(REF, f) for f in DRIVERS[name]]
    out = subprocess.run([mpicc, "-O1", "-Wimplicit-function-declaration", "-I", os.path.join(ROOT, "include")]
                         + srcs + ["-o", exe] + _link_flags(),
                         capture_output=True, text=True, env=dict(os.environ, MPICH_CC="gcc"))
    log = out.stdout + out.stderr
    assert out.returncode == 0, log
    assert "undefined reference" not in log
    # every PLAPACK routine the driver calls is declared by include/PLA.h
    implicit = set(re.findall(r"implicit declaration of function '(\w+)'", log))
    assert not {f for f in implicit if f.upper().startswith("PLA_")}, implicit
    assert os.path.exists(exe)


def test_constants_equal_the_reference_headers():
    """Every constant the reference defines in PLA_define.h, PLA_Machines.h
    and PLA_Versions.h, printed through the reference's headers and through
    include/PLA.h, has the same value."""
    _need_reference()
    inc = os.path.join(REF, "INCLUDE")
    names = []
    for h in ("PLA_define.h", "PLA_Machines.h", "PLA_Versions.h"):
        for line in open(os.path.join(inc, h)):
            m = re.match(r"\s*#define\s+(\w+)\s+\S", line)
            if m and not m.group(1).endswith("_H"):
                names.append(m.group(1))
    names = sorted(set(names))
    assert len(names) > 60
    prog = "#include <stdio.h>\n@HEAD@\nint main(void){\n" + "".join(
        f'printf("{n} %ld\\n", (long)({n}));\n' for n in names) + "return 0;}\n"
    os.makedirs(BUILD, exist_ok=True)
    results = []
    for tag, head, flags in (("ref", '#include "PLA_Machines.h"\n#include "PLA_define.h"\n#include "PLA_Versions.h"',
                              ["-I", inc]),
                             ("ours", '#define PLA_NO_MPI_H\n#include "PLA.h"', ["-I", os.path.join(ROOT, "include")])):
        c = os.path.join(BUILD, f"consts_{tag}.c")
        exe = c[:-2]
        with open(c, "w") as f:
            f.write(prog.replace("@HEAD@", head))
        subprocess.run(["gcc"] + flags + [c, "-o", exe], check=True)
        results.append(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    assert results[0] == results[1]


_HOST_ONLY = r"""
#define PLA_NO_MPI_H
#include "PLA.h"
#include <string.h>
int main(int argc, char **argv) {
  MPI_Comm c2 = 0;
  if (PLA_Comm_1D_to_2D_ratio(MPI_COMM_WORLD, 2.0, &c2) != PLA_SUCCESS || c2 != MPI_COMM_WORLD) return 2;
  double *p = (double *)PLA_malloc(sizeof(double) * 100);
  int *q = (int *)PLA_calloc(50, sizeof(int));
  if (!p || !q) return 3;
  for (int i = 0; i < 50; i++) if (q[i] != 0) return 4;
  p[99] = 1.0;
  if (PLA_free(p) != NULL) return 5;
  PLA_free(q);
  if (argc > 1 && strcmp(argv[1], "abort") == 0) {
    printf("before abort\n");
    fflush(stdout);
    PLA_Abort("giving up", 42, "driver.c");
    printf("after abort\n");
    return 0;
  }
  printf("ok\n");
  return 0;
}
"""


def test_host_only_entry_points():
    if not os.path.exists(os.path.join(LIBDIR, "libpla_mi355x.so")):
        pytest.skip("libpla_mi355x.so is not built")
    os.makedirs(BUILD, exist_ok=True)
    c = os.path.join(BUILD, "host_only.c")
    exe = c[:-2]
    with open(c, "w") as f:
        f.write(_HOST_ONLY)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe, "-L", LIBDIR, "-lpla_mi355x",
                    "-Wl,-rpath," + LIBDIR, "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    ok = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert ok.returncode == 0 and ok.stdout.strip() == "ok", (ok.returncode, ok.stdout, ok.stderr)
    ab = subprocess.run([exe, "abort"], capture_output=True, text=True, timeout=60)
    assert ab.returncode == 1                      # the reference exits with 0; a failure status here
    assert "before abort" in ab.stdout and "after abort" not in ab.stdout
    assert "giving up" in ab.stderr and "42" in ab.stderr and "driver.c" in ab.stderr
