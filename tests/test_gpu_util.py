"""GPU tests of the level-1/2 helpers the reference's drivers call around the
factor-and-solve path: PLA_Axpy (BLAS1/PLA_Axpy.c:13), PLA_Nrm2
(BLAS1/PLA_Nrm2.c:13), PLA_Symv / PLA_Hemv (BLAS2/PLA_Symv.c:13,
BLAS2/PLA_Hemv.c:13), PLA_Shift (UTIL/PLA_Shift.c:15) and PLA_Gemm_x
(BLAS3/PLA_Gemm_x.c:13), through the C ABI.

Their semantics are the textbook BLAS ones (daxpy, dnrm2 / dznrm2, dsymv /
zhemv reading one triangle), so numpy is the checker: tolerances are a few
ulps times the length.  The triangle a routine must not read is filled with
NaN, which would poison any result that touched it."""
import ctypes

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu

TYPES = {"d": (pla.MPI_DOUBLE, np.float64, 1e-14), "s": (pla.MPI_FLOAT, np.float32, 2e-6),
         "z": (pla.MPI_DOUBLE_COMPLEX, np.complex128, 1e-14)}


def _obj(s, a, dt):
    a = np.asarray(a)
    o = s.matrix(a.shape[0], a.shape[1] if a.ndim == 2 else 1, dt)
    pla.set_contents(o, a)
    return o


def _scalar(s, v, dt):
    o = s.mscalar(dtype=dt)
    pla.set_contents(o, np.array([[v]]))
    return o


def _rand(rng, shape, npt):
    a = rng.uniform(-1, 1, shape)
    if np.dtype(npt).kind == "c":
        a = a + 1j * rng.uniform(-1, 1, shape)
    return a.astype(npt)


@pytest.mark.parametrize("t", ["d", "s", "z"])
@pytest.mark.parametrize("shape", [(1, 1), (1000, 1), (300, 7), (0, 3)])
def test_axpy(pla_session, t, shape):
    dt, npt, tol = TYPES[t]
    rng = np.random.default_rng(shape[0] + shape[1])
    x, y = _rand(rng, shape, npt), _rand(rng, shape, npt)
    alpha = (-0.75 + 0.5j) if t == "z" else -0.75
    X, Y, A = _obj(pla_session, x, dt), _obj(pla_session, y, dt), _scalar(pla_session, alpha, dt)
    pla.call("PLA_Axpy", A, X, Y)
    got = pla.get_contents(Y)
    want = np.asarray(alpha, dtype=npt) * x + y
    assert got.shape == want.shape
    if want.size:
        assert np.abs(got - want).max() <= tol * 4
    assert np.array_equal(pla.get_contents(X), x)
    pla.free(X, Y, A)


def test_axpy_errors(pla_session):
    s = pla_session
    X, Y = _obj(s, np.zeros((5, 1)), pla.MPI_DOUBLE), _obj(s, np.zeros((6, 1)), pla.MPI_DOUBLE)
    A = _scalar(s, 1.0, pla.MPI_DOUBLE)
    assert pla.lib().PLA_Axpy(A, X, Y) == pla.PLA_ERR_DIMENSIONS
    Z = _obj(s, np.zeros((5, 1), np.complex128), pla.MPI_DOUBLE_COMPLEX)
    assert pla.lib().PLA_Axpy(A, X, Z) == pla.PLA_ERR_INVALID_PARAM
    pla.free(X, Y, A, Z)


@pytest.mark.parametrize("t", ["d", "s", "z"])
@pytest.mark.parametrize("n,row", [(1, False), (777, False), (100000, False), (513, True), (0, False)])
def test_nrm2(pla_session, t, n, row):
    dt, npt, tol = TYPES[t]
    rng = np.random.default_rng(n)
    x = _rand(rng, (n,), npt)
    X = _obj(pla_session, x.reshape(1, -1) if row else x.reshape(-1, 1), dt)
    rt = pla.MPI_FLOAT if t == "s" else pla.MPI_DOUBLE
    al = pla_session.mscalar(dtype=rt)
    pla.call("PLA_Nrm2", X, al)
    got = float(pla.get_contents(al)[0, 0])
    want = float(np.linalg.norm(x.astype(np.complex128 if t == "z" else np.float64)))
    assert abs(got - want) <= tol * max(1.0, want) * max(1, np.sqrt(n)) * 2
    pla.free(X, al)


def test_nrm2_scales_against_overflow_and_underflow(pla_session):
    """dnrm2 scales by the largest component: values whose squares overflow
    or underflow in FP64 still give the norm."""
    for scale in (1e200, 1e-200):
        x = np.array([3.0, -4.0, 0.0, 12.0]) * scale
        X = _obj(pla_session, x.reshape(-1, 1), pla.MPI_DOUBLE)
        al = pla_session.mscalar(dtype=pla.MPI_DOUBLE)
        pla.call("PLA_Nrm2", X, al)
        assert abs(pla.get_contents(al)[0, 0] / (13.0 * scale) - 1.0) < 1e-15
        pla.free(X, al)


@pytest.mark.parametrize("t", ["d", "s", "z"])
@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("n,beta", [(1, 0.0), (65, 0.5), (700, 0.0), (1000, -1.25)])
def test_symv_hemv(pla_session, t, uplo, n, beta):
    dt, npt, tol = TYPES[t]
    rng = np.random.default_rng(n + uplo)
    a = _rand(rng, (n, n), npt)
    full = np.tril(a) + np.tril(a, -1).conj().T          # symmetric / Hermitian from the lower part
    if t == "z":
        full[np.diag_indices(n)] = full[np.diag_indices(n)].real
    stored = full.copy()
    if uplo == pla.PLA_LOWER_TRIANGULAR:
        stored[np.triu_indices(n, 1)] = np.nan
    else:
        stored[np.tril_indices(n, -1)] = np.nan
    if t == "z":  # zhemv ignores the imaginary part of the diagonal
        stored[np.diag_indices(n)] += 7j
    x, y = _rand(rng, (n, 1), npt), _rand(rng, (n, 1), npt)
    alpha = (0.5 - 2j) if t == "z" else 1.5
    betav = complex(beta, 0.25) if (t == "z" and beta != 0.0) else beta
    if beta == 0.0:
        y[:] = np.nan  # beta = 0: y is not read
    A, X, Y = _obj(pla_session, stored, dt), _obj(pla_session, x, dt), _obj(pla_session, y, dt)
    Al, Be = _scalar(pla_session, alpha, dt), _scalar(pla_session, betav, dt)
    pla.call("PLA_Hemv" if t == "z" else "PLA_Symv", uplo, Al, A, X, Be, Y)
    got = pla.get_contents(Y)
    f64 = np.complex128 if t == "z" else np.float64
    want = np.asarray(alpha, f64) * (full.astype(f64) @ x.astype(f64))
    if beta != 0.0:
        want = want + np.asarray(betav, f64) * y.astype(f64)
    assert np.isfinite(got).all()
    assert np.abs(got - want).max() <= tol * 8 * n * max(1.0, np.abs(want).max())
    pla.free(A, X, Y, Al, Be)


def test_symv_datatype_rules(pla_session):
    s = pla_session
    z = _obj(s, np.eye(4, dtype=np.complex128), pla.MPI_DOUBLE_COMPLEX)
    zx = _obj(s, np.ones((4, 1), np.complex128), pla.MPI_DOUBLE_COMPLEX)
    assert pla.lib().PLA_Symv(pla.PLA_LOWER_TRIANGULAR, None, z, zx, None, zx) == pla.PLA_ERR_UNSUPPORTED
    d = _obj(s, np.eye(4), pla.MPI_DOUBLE)
    dx = _obj(s, np.ones((4, 1)), pla.MPI_DOUBLE)
    assert pla.lib().PLA_Hemv(pla.PLA_LOWER_TRIANGULAR, None, d, dx, None, dx) == pla.PLA_ERR_UNSUPPORTED
    pla.free(z, zx, d, dx)


@pytest.mark.parametrize("t,vt", [("d", "d"), ("d", "s"), ("s", "d"), ("s", "s")])
@pytest.mark.parametrize("shape", [(50, 50), (80, 30), (30, 80)])
def test_shift(pla_session, t, vt, shape):
    dt, npt, _ = TYPES[t]
    vdt, vnp, _ = TYPES[vt]
    a = _rand(np.random.default_rng(1), shape, npt)
    A = _obj(pla_session, a, dt)
    v = np.array([2.5], dtype=vnp)
    pla.call("PLA_Shift", A, vdt, v.ctypes.data_as(ctypes.c_void_p))
    want = a.copy()
    k = min(shape)
    want[np.arange(k), np.arange(k)] += npt(v[0])
    assert np.array_equal(pla.get_contents(A), want)
    pla.free(A)


def test_shift_forms_the_reference_spd_input(pla_session):
    """TEST/test_chol.c:121-127: create_problem then PLA_Shift(A, MPI_DOUBLE,
    &d_n) makes the matrix the Cholesky test factors."""
    s = pla_session
    n = 300
    A = s.matrix(n, n)
    pla.call("PLA_Obj_fill_drand48", A, 0, 0)
    a0 = pla.get_contents(A)
    d_n = np.array([float(n)])
    pla.call("PLA_Shift", A, pla.MPI_DOUBLE, d_n.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(pla.get_contents(A), a0 + n * np.eye(n))
    pla.free(A)


def test_gemm_x_versions(pla_session):
    s = pla_session
    rng = np.random.default_rng(3)
    a, b, c = rng.uniform(-1, 1, (70, 40)), rng.uniform(-1, 1, (40, 90)), rng.uniform(-1, 1, (70, 90))
    one, mone = s.scalar(1.0), s.scalar(-1.0)
    for version in (100, 200, 300):
        A, B, C = pla.from_numpy(s, a), pla.from_numpy(s, b), pla.from_numpy(s, c)
        pla.call("PLA_Gemm_x", version, pla.PLA_NO_TRANSPOSE, pla.PLA_NO_TRANSPOSE, mone, A, B, one, C)
        assert np.abs(pla.get_contents(C) - (c - a @ b)).max() < 1e-13
        pla.free(A, B, C)
    A, B, C = pla.from_numpy(s, a), pla.from_numpy(s, b), pla.from_numpy(s, c)
    assert pla.lib().PLA_Gemm_x(7, pla.PLA_NO_TRANSPOSE, pla.PLA_NO_TRANSPOSE, one, A, B, one, C) == \
        pla.PLA_ERR_INVALID_PARAM
    pla.free(A, B, C, one, mone)
