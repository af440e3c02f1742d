"""GPU parity of the building blocks beneath PLA_LU / PLA_Chol, called through
the C ABI: the drand48 fill (bit-exact against the oracle), PLA_Gemm (FP64
MFMA kernel, every transpose case and ragged edges), PLA_Syrk, PLA_Trsm (both
sides, every uplo / trans / diag) / PLA_Trsv,
PLA_Apply_pivots_to_rows (bit-exact), views and copies.

Tolerances: FP64 results agree with the oracle to c * k * eps * scale, the
usual forward-error bound of a k-term dot product (c = 4)."""
import ctypes

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _obj(s, arr):
    return pla.from_numpy(s, np.asfortranarray(arr))


def test_fill_drand48_bit_exact(pla_session, ora):
    s = pla_session
    A = s.matrix(300, 211)
    pla.call("PLA_Obj_fill_drand48", A, 0, 0)
    assert np.array_equal(pla.get_contents(A), ora.uniform(300, 211, 0))
    pla.call("PLA_Obj_fill_drand48", A, 5, 12345)
    assert np.array_equal(pla.get_contents(A), ora.uniform(300, 211, 5, 12345))
    pla.free(A)


def test_fill_drand48_far_into_stream(pla_session, ora):
    s = pla_session
    n = 8192
    A = s.matrix(n, n)
    pla.call("PLA_Obj_fill_drand48", A, 0, 0)
    left, right = ctypes.c_void_p(), ctypes.c_void_p()
    pla.call("PLA_Obj_vert_split_2", A, -3, ctypes.byref(left), ctypes.byref(right))
    assert pla.dims(right) == (n, 3)
    got = pla.get_contents(right)
    assert np.array_equal(got, ora.uniform(n, 3, 0, skip=n * (n - 3)))
    pla.free(left, right, A)


@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (130, 257, 33), (256, 128, 256),
                                   (77, 5, 300), (300, 1, 129)])
def test_gemm_vs_oracle(pla_session, ora, ta, tb, m, n, k):
    s = pla_session
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    a = rng.uniform(-1, 1, (k, m) if ta else (m, k))
    b = rng.uniform(-1, 1, (n, k) if tb else (k, n))
    c = rng.uniform(-1, 1, (m, n))
    A, B, C = _obj(s, a), _obj(s, b), _obj(s, c)
    al, be = s.scalar(-1.0), s.scalar(0.5)
    pla.call("PLA_Gemm", pla.PLA_TRANSPOSE if ta else pla.PLA_NO_TRANSPOSE,
             pla.PLA_TRANSPOSE if tb else pla.PLA_NO_TRANSPOSE, al, A, B, be, C)
    got = pla.get_contents(C)
    ref = ora.gemm("T" if ta else "N", "T" if tb else "N", -1.0, a, b, 0.5, c)
    assert np.abs(got - ref).max() <= 8 * k * EPS * max(1.0, np.abs(ref).max())
    pla.free(A, B, C, al, be)


def test_gemm_views_leave_outside_untouched(pla_session, ora):
    """Trailing-update shape on views: C = A22 inside a larger matrix."""
    s = pla_session
    rng = np.random.default_rng(1)
    full = rng.uniform(-1, 1, (400, 380))
    F = _obj(s, full)
    # A21 = F[100:, 0:64], A12 = F[0:64, 100:], A22 = F[100:, 100:]
    tl, tr, bl, br = (ctypes.c_void_p() for _ in range(4))
    pla.call("PLA_Obj_split_4", F, 100, 100, ctypes.byref(tl), ctypes.byref(tr), ctypes.byref(bl),
             ctypes.byref(br))
    a21, x1, a12, x2 = (ctypes.c_void_p() for _ in range(4))
    pla.call("PLA_Obj_vert_split_2", bl, 64, ctypes.byref(a21), ctypes.byref(x1))
    pla.call("PLA_Obj_horz_split_2", tr, 64, ctypes.byref(a12), ctypes.byref(x2))
    one, mone = s.scalar(1.0), s.scalar(-1.0)
    pla.call("PLA_Gemm", pla.PLA_NO_TRANSPOSE, pla.PLA_NO_TRANSPOSE, mone, a21, a12, one, br)
    got = pla.get_contents(F)
    ref = full.copy()
    ref[100:, 100:] -= full[100:, :64] @ full[:64, 100:]
    assert np.abs(got - ref).max() <= 4 * 64 * EPS * 2
    assert np.array_equal(got[:100, :], full[:100, :])
    assert np.array_equal(got[:, :100], full[:, :100])
    pla.free(tl, tr, bl, br, a21, x1, a12, x2, one, mone, F)


@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("trans", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE])
@pytest.mark.parametrize("diag", [pla.PLA_UNIT_DIAG, pla.PLA_NONUNIT_DIAG])
@pytest.mark.parametrize("n,nrhs", [(50, 3), (200, 70), (300, 1)])
def test_trsm_left_vs_oracle(pla_session, ora, uplo, trans, diag, n, nrhs):
    s = pla_session
    rng = np.random.default_rng(n + nrhs)
    t = rng.uniform(-1, 1, (n, n)) / n + np.eye(n)
    b = rng.uniform(-1, 1, (n, nrhs))
    T, B, al = _obj(s, t), _obj(s, b), s.scalar(2.0)
    pla.call("PLA_Trsm", pla.PLA_SIDE_LEFT, uplo, trans, diag, al, T, B)
    got = pla.get_contents(B)
    ref = ora.trsm_left("L" if uplo == pla.PLA_LOWER_TRIANGULAR else "U",
                        "N" if trans == pla.PLA_NO_TRANSPOSE else "T",
                        "U" if diag == pla.PLA_UNIT_DIAG else "N", t, b, alpha=2.0)
    assert np.abs(got - ref).max() <= 64 * n * EPS * max(1.0, np.abs(ref).max())
    pla.free(T, B, al)


@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("trans", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE])
@pytest.mark.parametrize("diag", [pla.PLA_UNIT_DIAG, pla.PLA_NONUNIT_DIAG])
@pytest.mark.parametrize("m,n", [(90, 150), (1, 64), (300, 65), (513, 200)])
def test_trsm_right_vs_oracle(pla_session, ora, uplo, trans, diag, m, n):
    """B <- alpha B op(T)^-1: 64-column leaves (one thread per row) under a
    column recursion whose updates are the FP64 GEMM."""
    s = pla_session
    rng = np.random.default_rng(m * 5 + n)
    t = rng.uniform(-1, 1, (n, n)) / n + np.eye(n)
    b = rng.uniform(-1, 1, (m, n))
    T, B, al = _obj(s, t), _obj(s, b), s.scalar(-0.5)
    pla.call("PLA_Trsm", pla.PLA_SIDE_RIGHT, uplo, trans, diag, al, T, B)
    got = pla.get_contents(B)
    ref = ora.trsm_right("L" if uplo == pla.PLA_LOWER_TRIANGULAR else "U",
                         "N" if trans == pla.PLA_NO_TRANSPOSE else "T",
                         "U" if diag == pla.PLA_UNIT_DIAG else "N", t, b, alpha=-0.5)
    assert np.abs(got - ref).max() <= 64 * n * EPS * max(1.0, np.abs(ref).max())
    pla.free(T, B, al)


def test_trsm_right_reads_only_its_triangle(pla_session, ora):
    """NaN in the strict opposite triangle and (unit) diagonal of T is never read."""
    s = pla_session
    rng = np.random.default_rng(8)
    n, m = 130, 70
    t = rng.uniform(-1, 1, (n, n)) / n + np.eye(n)
    b = rng.uniform(-1, 1, (m, n))
    ref = ora.trsm_right("U", "T", "U", t, b)
    t[np.tril(np.ones((n, n), bool), 0)] = np.nan
    T, B = _obj(s, t), _obj(s, b)
    pla.call("PLA_Trsm", pla.PLA_SIDE_RIGHT, pla.PLA_UPPER_TRIANGULAR, pla.PLA_TRANSPOSE, pla.PLA_UNIT_DIAG,
             None, T, B)
    got = pla.get_contents(B)
    assert np.isfinite(got).all()
    assert np.abs(got - ref).max() <= 64 * n * EPS * max(1.0, np.abs(ref).max())
    pla.free(T, B)


@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("trans", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE])
@pytest.mark.parametrize("n,k,alpha,beta", [(1, 1, 1.0, 1.0), (100, 33, 0.75, -2.0), (257, 300, -1.0, 1.0),
                                            (640, 200, 1.0, 0.5), (384, 0, 1.0, 3.0), (130, 64, -1.0, 0.0)])
def test_syrk_vs_oracle(pla_session, ora, uplo, trans, n, k, alpha, beta):
    """PLA_Syrk: all of C scaled by beta, the uplo triangle += alpha op(A) op(A)^T;
    n = 640 / 257 put interior 128 x 128 tiles on the alpha = +-1 fast path."""
    s = pla_session
    rng = np.random.default_rng(n + 3 * k)
    a = rng.uniform(-1, 1, (n, k) if trans == pla.PLA_NO_TRANSPOSE else (k, n))
    c = rng.uniform(-1, 1, (n, n))
    A, C, al, be = _obj(s, a), _obj(s, c), s.scalar(alpha), s.scalar(beta)
    pla.call("PLA_Syrk", uplo, trans, al, A, be, C)
    got = pla.get_contents(C)
    lo = uplo == pla.PLA_LOWER_TRIANGULAR
    ref = ora.syrk("L" if lo else "U", "N" if trans == pla.PLA_NO_TRANSPOSE else "T", alpha, a, beta, c)
    tri = np.tril(np.ones((n, n), bool)) if lo else np.triu(np.ones((n, n), bool))
    assert np.abs(got[tri] - ref[tri]).max() <= 8 * (k + 1) * EPS * max(1.0, np.abs(ref).max())
    assert np.array_equal(got[~tri], ref[~tri])  # beta * C, a single rounding
    pla.free(A, C, al, be)


def test_syrk_on_a_view_leaves_outside_untouched(pla_session, ora):
    """The Cholesky trailing-update shape: C = A22 inside a larger matrix, A = A21."""
    s = pla_session
    rng = np.random.default_rng(9)
    full = rng.uniform(-1, 1, (700, 700))
    F = _obj(s, full)
    tl, tr, bl, br = (ctypes.c_void_p() for _ in range(4))
    pla.call("PLA_Obj_split_4", F, 300, 300, ctypes.byref(tl), ctypes.byref(tr), ctypes.byref(bl),
             ctypes.byref(br))
    mone, one = s.scalar(-1.0), s.scalar(1.0)
    pla.call("PLA_Syrk", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, mone, bl, one, br)
    got = pla.get_contents(F)
    ref = full.copy()
    ref[300:, 300:] = ora.syrk("L", "N", -1.0, full[300:, :300], 1.0, full[300:, 300:])
    assert np.abs(got - ref).max() <= 8 * 300 * EPS * 300
    assert np.array_equal(got[:300, :], full[:300, :])
    assert np.array_equal(np.triu(got[300:, 300:], 1), np.triu(full[300:, 300:], 1))
    pla.free(tl, tr, bl, br, mone, one, F)


def test_syrk_argument_errors(pla_session):
    s = pla_session
    A, C, R = s.matrix(10, 4), s.matrix(10, 10), s.matrix(10, 9)
    one = s.scalar(1.0)
    assert pla.lib().PLA_Syrk(pla.PLA_LOWER_TRIANGULAR, pla.PLA_TRANSPOSE, one, A, one, C) == pla.PLA_ERR_DIMENSIONS
    assert pla.lib().PLA_Syrk(pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, one, A, one, R) == pla.PLA_ERR_DIMENSIONS
    assert pla.lib().PLA_Syrk(12345, pla.PLA_NO_TRANSPOSE, one, A, one, C) == pla.PLA_ERR_INVALID_PARAM
    assert pla.lib().PLA_Syrk(pla.PLA_UPPER_TRIANGULAR, 0, one, A, one, C) == pla.PLA_ERR_INVALID_PARAM
    pla.free(A, C, R, one)


@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("trans", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE])
@pytest.mark.parametrize("diag", [pla.PLA_UNIT_DIAG, pla.PLA_NONUNIT_DIAG])
@pytest.mark.parametrize("n", [1, 63, 333, 2000])
def test_trsv_vs_oracle(pla_session, ora, uplo, trans, diag, n):
    """The sync-free blocked solve (64-row blocks, inverted diagonal blocks)."""
    s = pla_session
    rng = np.random.default_rng(4 + n)
    t = rng.uniform(-1, 1, (n, n)) / n + np.eye(n)
    x = rng.uniform(-1, 1, n)
    T = _obj(s, t)
    X = s.mvector(n)
    pla.set_contents(X, x)
    pla.call("PLA_Trsv", uplo, trans, diag, T, X)
    got = pla.get_contents(X)[:, 0]
    ref = ora.trsm_left("L" if uplo == pla.PLA_LOWER_TRIANGULAR else "U",
                        "N" if trans == pla.PLA_NO_TRANSPOSE else "T",
                        "U" if diag == pla.PLA_UNIT_DIAG else "N", t, x)[:, 0]
    assert np.abs(got - ref).max() <= 64 * n * EPS
    pla.free(T, X)


@pytest.mark.parametrize("m,ncols,npiv", [(50, 7, 20), (3000, 513, 1500), (10, 1, 10)])
def test_apply_pivots_bit_exact(pla_session, ora, m, ncols, npiv):
    s = pla_session
    rng = np.random.default_rng(m)
    a = rng.standard_normal((m, ncols))
    piv = np.array([rng.integers(0, m - i) for i in range(npiv)], dtype=np.int32)
    A = _obj(s, a)
    P = s.mvector(npiv, dtype=pla.MPI_INT)
    pla.set_contents(P, piv)
    pla.call("PLA_Apply_pivots_to_rows", A, P)
    assert np.array_equal(pla.get_contents(A), ora.laswp(a, piv))
    pla.free(A, P)


def test_copy_and_constants(pla_session):
    s = pla_session
    a = np.arange(12.0).reshape(3, 4)
    A = _obj(s, a)
    B = s.matrix(3, 4)
    pla.call("PLA_Copy", A, B)
    assert np.array_equal(pla.get_contents(B), a)
    mone, zero, one = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    pla.call("PLA_Create_constants_conf_to", A, ctypes.byref(mone), ctypes.byref(zero),
             ctypes.byref(one))
    assert pla.get_contents(mone)[0, 0] == -1.0 and pla.get_contents(one)[0, 0] == 1.0
    assert pla.get_contents(zero)[0, 0] == 0.0
    pla.free(A, B, mone, zero, one)


@pytest.mark.parametrize("m,n", [(1, 1), (300, 257), (1000, 3), (3, 1000)])
def test_matrix_norms_vs_numpy(pla_session, m, n):
    s = pla_session
    a = np.random.default_rng(m * n).uniform(-1, 1, (m, n))
    A, nrm = _obj(s, a), s.scalar(0.0)
    pla.call("PLA_Matrix_infinity_norm", A, nrm)
    assert abs(pla.get_contents(nrm)[0, 0] - np.abs(a).sum(axis=1).max()) <= 4 * n * EPS * n
    pla.call("PLA_Matrix_one_norm", A, nrm)
    assert abs(pla.get_contents(nrm)[0, 0] - np.abs(a).sum(axis=0).max()) <= 4 * m * EPS * m
    pla.free(A, nrm)
