"""plapack_amd -- PLAPACK's dense factor-and-solve path (PLA_LU, PLA_Chol and
the trsm/gemm/syrk updates beneath them) for AMD Instinct MI355X (gfx950).

The product is the C-ABI library plapack_amd/libpla_mi355x.so
(include/PLA.h).  This package only locates, builds and binds it.
"""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8):
    """Compile the HIP kernels and host layer into libpla_mi355x.so (always rebuilds)."""
    subprocess.run(["make", "-s", "-C", _HERE, "clean"], check=True)
    subprocess.run(["make", "-s", "-j", str(jobs), "-C", _HERE], check=True)


from . import pla  # noqa: E402,F401
