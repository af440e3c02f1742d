"""The reference's own BLAS3 / Cholesky test drivers replayed through the C ABI.

  TEST/test_gemm.c with TEST/input_test_gemm: m, n, k = 20, 30, 10, every
      transa / transb, variants A, B, C (one kernel here), alpha = 1, beta = 0;
  TEST/test_chol.c with TEST/input_test_chol: n = 20, A random + n I, lower;
  TEST/test_syrk.c and TEST/test_trsm.c (interactive drivers without an input
      file): every uplo / trans (/ side / diag) case at the same 20 x 30 size,
      with the drivers' operand shapes, the diagonal shift n of test_trsm.c
      (PLA_Shift) and their alpha = 1, beta = 0.

Operands come from TEST/create_problem.c's stream exactly: srand48(0) once per
process (rank 0), then drand48() * 2 - 1 column by column through every object
the driver creates, in creation order, across cases -- double complex objects
(input_test_gemm's itype 3 cases) draw two numbers per element.  PLA_Obj_fill_drand48(obj, 0, offset) reproduces
that stream bit-exactly on the device.  Results are checked against the oracle
with the usual k * eps bounds (the drivers' own sequential check compares with
PLA_Local_* on the same data)."""
import ctypes

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps

# input_test_gemm, in file order: (itype, variant, m, n, k, itransa, itransb)
_GEMM_INPUT = [(t, v, 20, 30, 10, a, b) for v in (0, 1)
               for t in (1, 3) for a, b in ((0, 0), (0, 1), (1, 0), (1, 1))]
_GEMM_INPUT += [(1, 2, 20, 30, 10, a, b) for a, b in ((0, 0), (0, 1), (1, 0), (1, 1))]
_GEMM_INPUT += [(3, 2, 20, 30, 10, a, b) for a, b in ((0, 0), (1, 0), (1, 1), (0, 1))]


def _gemm_cases():
    """Double cases with the stream offset at which each one's A starts."""
    out, off = [], 0
    for itype, variant, m, n, k, ta, tb in _GEMM_INPUT:
        per = 2 if itype == 3 else 1
        if itype == 1:
            out.append((variant, m, n, k, ta, tb, off))
        off += per * (m * k + k * n + m * n)
    return out


def _filled(s, m, n, offset):
    o = s.matrix(m, n)
    pla.call("PLA_Obj_fill_drand48", o, 0, offset)
    return o


@pytest.mark.parametrize("variant,m,n,k,ta,tb,off", _gemm_cases())
def test_gemm_driver_cases(pla_session, ora, variant, m, n, k, ta, tb, off):
    s = pla_session
    ashape = (k, m) if ta else (m, k)
    bshape = (n, k) if tb else (k, n)
    A = _filled(s, *ashape, off)
    off += m * k
    B = _filled(s, *bshape, off)
    off += k * n
    C = _filled(s, m, n, off)
    a, b, c = ora.uniform(*ashape, 0, off - m * k - k * n), ora.uniform(*bshape, 0, off - k * n), \
        ora.uniform(m, n, 0, off)
    assert np.array_equal(pla.get_contents(A), a) and np.array_equal(pla.get_contents(C), c)
    one, zero = s.scalar(1.0), s.scalar(0.0)
    T = {0: pla.PLA_NO_TRANSPOSE, 1: pla.PLA_TRANSPOSE}
    pla.call("PLA_Gemm", T[ta], T[tb], one, A, B, zero, C)
    ref = ora.gemm("T" if ta else "N", "T" if tb else "N", 1.0, a, b, 0.0, c)
    assert np.abs(pla.get_contents(C) - ref).max() <= 4 * k * EPS * max(1.0, np.abs(ref).max())
    pla.free(A, B, C, one, zero)


def test_chol_driver_case(pla_session, ora):
    """input_test_chol: nb_alg 4, double, n = 20 (the float case follows below)."""
    s = pla_session
    n = 20
    A = _filled(s, n, n, 0)
    a = ora.uniform(n, n, 0) + n * np.eye(n)  # PLA_Shift(A, n)
    pla.set_contents(A, a)
    pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A)
    got = np.tril(pla.get_contents(A))
    ref, info = ora.chol_blocked(a, nb=4)
    assert info == 0
    assert np.abs(got - np.tril(ref)).max() <= 8 * n * EPS * np.abs(ref).max()
    pla.free(A)


@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("trans", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE])
def test_syrk_driver_shapes(pla_session, ora, uplo, trans):
    """test_syrk.c: A is n x k; C is n x n (no transpose) or k x k (transpose)."""
    s = pla_session
    n, k = 20, 30
    A = _filled(s, n, k, 0)
    cn = n if trans == pla.PLA_NO_TRANSPOSE else k
    C = _filled(s, cn, cn, n * k)
    a, c = ora.uniform(n, k, 0), ora.uniform(cn, cn, 0, n * k)
    one, zero = s.scalar(1.0), s.scalar(0.0)
    pla.call("PLA_Syrk", uplo, trans, one, A, zero, C)
    ref = ora.syrk("L" if uplo == pla.PLA_LOWER_TRIANGULAR else "U",
                   "N" if trans == pla.PLA_NO_TRANSPOSE else "T", 1.0, a, 0.0, c)
    got = pla.get_contents(C)
    depth = k if trans == pla.PLA_NO_TRANSPOSE else n
    assert np.abs(got - ref).max() <= 4 * depth * EPS * max(1.0, np.abs(ref).max())
    pla.free(A, C, one, zero)


@pytest.mark.parametrize("side", [pla.PLA_SIDE_LEFT, pla.PLA_SIDE_RIGHT])
@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("trans", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE])
@pytest.mark.parametrize("diag", [pla.PLA_NONUNIT_DIAG, pla.PLA_UNIT_DIAG])
def test_trsm_driver_cases(pla_session, ora, side, uplo, trans, diag):
    """test_trsm.c: A is m x m (left) or n x n (right), random + n I; B m x n."""
    s = pla_session
    m, n = 20, 30
    na = m if side == pla.PLA_SIDE_LEFT else n
    A = _filled(s, na, na, 0)
    a = ora.uniform(na, na, 0) + n * np.eye(na)
    pla.set_contents(A, a)
    B = _filled(s, m, n, na * na)
    b = ora.uniform(m, n, 0, na * na)
    one = s.scalar(1.0)
    pla.call("PLA_Trsm", side, uplo, trans, diag, one, A, B)
    args = ("L" if uplo == pla.PLA_LOWER_TRIANGULAR else "U", "N" if trans == pla.PLA_NO_TRANSPOSE else "T",
            "U" if diag == pla.PLA_UNIT_DIAG else "N", a, b)
    ref = ora.trsm_left(*args) if side == pla.PLA_SIDE_LEFT else ora.trsm_right(*args)
    assert np.abs(pla.get_contents(B) - ref).max() <= 16 * na * EPS * max(1.0, np.abs(ref).max())
    pla.free(A, B, one)


def _zuniform(ora, m, n, offset):
    """A complex m x n object of TEST/create_problem.c: the 2m x n real view
    of its (re, im) pairs is the drand48 stream in column order."""
    u = ora.uniform(2 * m, n, 0, offset)
    return u[0::2] + 1j * u[1::2]


def _gemm_complex_cases():
    """input_test_gemm's double complex (itype 3) cases with their stream offsets."""
    out, off = [], 0
    for itype, variant, m, n, k, ta, tb in _GEMM_INPUT:
        per = 2 if itype == 3 else 1
        if itype == 3:
            out.append((variant, m, n, k, ta, tb, off))
        off += per * (m * k + k * n + m * n)
    return out


@pytest.mark.parametrize("variant,m,n,k,ta,tb,off", _gemm_complex_cases())
def test_gemm_driver_complex_cases(pla_session, ora, variant, m, n, k, ta, tb, off):
    """input_test_gemm's 12 MPI_DOUBLE_COMPLEX cases: PLA_Gemm with alpha = 1,
    beta = 0 on the driver's stream of operands, checked against numpy's
    complex product."""
    s = pla_session
    ashape = (k, m) if ta else (m, k)
    bshape = (n, k) if tb else (k, n)
    objs = []
    for shape in (ashape, bshape, (m, n)):
        o = s.matrix(*shape, pla.MPI_DOUBLE_COMPLEX)
        pla.call("PLA_Obj_fill_drand48", o, 0, off)
        objs.append((o, _zuniform(ora, *shape, off)))
        off += 2 * shape[0] * shape[1]
    (A, a), (B, b), (C, c) = objs
    assert np.array_equal(pla.get_contents(A), a) and np.array_equal(pla.get_contents(C), c)
    one, zero = s.mscalar(dtype=pla.MPI_DOUBLE_COMPLEX), s.mscalar(dtype=pla.MPI_DOUBLE_COMPLEX)
    pla.set_contents(one, np.array([[1.0 + 0j]]))
    pla.set_contents(zero, np.array([[0j]]))
    T = {0: pla.PLA_NO_TRANSPOSE, 1: pla.PLA_TRANSPOSE}
    pla.call("PLA_Gemm", T[ta], T[tb], one, A, B, zero, C)
    ref = (a.T if ta else a) @ (b.T if tb else b)
    assert np.abs(pla.get_contents(C) - ref).max() <= 8 * k * EPS * max(1.0, np.abs(ref).max())
    pla.free(A, B, C, one, zero)


def test_chol_driver_float_case(pla_session):
    """input_test_chol's single-precision case: n = 20, nb_alg 4, after the
    double case the driver's stream continues; create_problem stores each
    draw rounded to float, PLA_Shift adds n, PLA_Chol factors it.  Checked
    against the Cholesky of the same float matrix in float64, rounded."""
    from oracle import oracle as ora
    s = pla_session
    n = 20
    A = s.matrix(n, n, pla.MPI_FLOAT)
    pla.call("PLA_Obj_fill_drand48", A, 0, n * n)      # the double case drew n * n numbers first
    af = ora.uniform(n, n, 0, n * n).astype(np.float32)
    assert np.array_equal(pla.get_contents(A), af)
    d_n = np.array([float(n)])
    pla.call("PLA_Shift", A, pla.MPI_DOUBLE, d_n.ctypes.data_as(ctypes.c_void_p))
    a = af + np.float32(n) * np.eye(n, dtype=np.float32)
    assert np.array_equal(pla.get_contents(A), a)
    pla.call("PLA_Set_device_nb", 4)
    try:
        assert pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A, ok=range(0, 1 << 30)) == 0
    finally:
        pla.call("PLA_Set_device_nb", 0)
    got = pla.get_contents(A)
    ref = np.linalg.cholesky(np.tril(a).astype(np.float64) + np.tril(a, -1).astype(np.float64).T)
    assert got.dtype == np.float32
    assert np.abs(np.tril(got) - ref).max() <= 2 * np.finfo(np.float32).eps * np.abs(ref).max()
    assert np.array_equal(np.triu(got, 1), np.triu(a, 1))
    pla.free(A)
