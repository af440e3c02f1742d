"""GPU tests of the MPI_FLOAT / MPI_COMPLEX objects and of the
MPI_DOUBLE_COMPLEX BLAS3, through the C ABI.

Single-precision objects are computed in double precision and rounded once
(PLA_slu / PLA_clu of SOLVERS/LU/PLA_seq_lu.c:50,108 compute in single), so
the bar is the oracle's double-precision result rounded to single: the same
pivots as the oracle's LU of the widened matrix, factors within two single-
precision ulps of its factors, and a scaled residual below 16 in single
precision.  The complex BLAS3 (BLAS3/PLA_Gemm.c, PLA_Trsm.c, PLA_Syrk.c on
MPI_DOUBLE_COMPLEX) are checked against numpy's complex products and
solves, every trans / side / uplo / diag case, with complex alpha and beta."""
import ctypes

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
EPS_F = np.finfo(np.float32).eps
SINGLE = {"s": (pla.MPI_FLOAT, np.float32), "c": (pla.MPI_COMPLEX, np.complex64)}


def _obj(s, a, dt):
    a = np.asarray(a)
    o = s.matrix(a.shape[0], a.shape[1] if a.ndim == 2 else 1, dt)
    pla.set_contents(o, a)
    return o


def _scalar(s, v, dt):
    o = s.mscalar(dtype=dt)
    pla.set_contents(o, np.array([[v]]))
    return o


def _input(ora, t, m, n, seed):
    if t == "s":
        return ora.uniform(m, n, seed).astype(np.float32)
    return ora.uniform_z(m, n, seed).astype(np.complex64)


def test_single_fill_rounds_each_draw(pla_session, ora):
    """TEST/create_problem.c stores drand48() * 2 - 1 into a float (a complex
    float takes two draws): the device fill gives the same floats."""
    s = pla_session
    for t, (dt, npt) in SINGLE.items():
        A = s.matrix(37, 11, dt)
        pla.call("PLA_Obj_fill_drand48", A, 3, 5)
        ref = (ora.uniform(37, 11, 3, 5) if t == "s" else ora.uniform_z(37, 11, 3, 5)).astype(npt)
        assert np.array_equal(pla.get_contents(A), ref)
        pla.free(A)


@pytest.mark.parametrize("t", ["s", "c"])
@pytest.mark.parametrize("m,n,nb", [(64, 64, 32), (300, 300, 64), (1000, 1000, 128), (500, 300, 128),
                                    (300, 500, 128)])
def test_single_lu(pla_session, ora, t, m, n, nb):
    dt, npt = SINGLE[t]
    a = _input(ora, t, m, n, m + n)
    A = _obj(pla_session, a, dt)
    P = pla_session.mvector(min(m, n), dtype=pla.MPI_INT)
    pla.call("PLA_Set_device_nb", nb)
    try:
        pla.call("PLA_LU", A, P)
    finally:
        pla.call("PLA_Set_device_nb", 0)
    lu, piv = pla.get_contents(A), pla.get_contents(P)[:, 0]
    pla.free(A, P)
    assert lu.dtype == npt
    if t == "s":
        lu_o, piv_o = ora.lu_blocked(a.astype(np.float64), nb=nb)
    else:
        lu_o, piv_o = ora.zlu_blocked(a.astype(np.complex128), nb=nb)
    assert np.array_equal(piv, piv_o)
    assert np.abs(lu.astype(lu_o.dtype) - lu_o).max() <= 2 * EPS_F * np.abs(lu_o).max()
    # P A = L U in single precision
    k = min(m, n)
    L = np.tril(lu, -1)[:, :k].astype(np.complex128) + np.eye(m, k)
    U = np.triu(lu)[:k].astype(np.complex128)
    pa = a.astype(np.complex128).copy()
    for j in range(k):
        p = j + piv[j]
        pa[[j, p]] = pa[[p, j]]
    res = np.abs(pa - L @ U).sum(axis=0).max() / (max(m, n) * np.abs(a).sum(axis=0).max() * EPS_F)
    assert res < 16


@pytest.mark.parametrize("t", ["s", "c"])
@pytest.mark.parametrize("n", [100, 777])
def test_single_chol_and_solves(pla_session, ora, t, n):
    s = pla_session
    dt, npt = SINGLE[t]
    rng = np.random.default_rng(n)
    b = rng.uniform(-1, 1, (n, n)) + (1j * rng.uniform(-1, 1, (n, n)) if t == "c" else 0)
    spd = (b @ b.conj().T / n + 0.5 * np.eye(n)).astype(npt)
    A = _obj(s, spd, dt)
    assert pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A, ok=range(0, 1 << 30)) == 0
    got = pla.get_contents(A)
    ref = np.linalg.cholesky(np.tril(spd).astype(np.complex128) + np.tril(spd, -1).astype(np.complex128).conj().T)
    assert np.abs(np.tril(got) - ref).max() <= 2 * EPS_F * np.abs(ref).max()
    assert np.array_equal(np.triu(got, 1), np.triu(spd, 1))
    pla.free(A)
    rhs = rng.uniform(-1, 1, (n, 3)).astype(npt)
    full = np.tril(spd).astype(np.complex128) + np.tril(spd, -1).astype(np.complex128).conj().T
    A, B = _obj(s, spd, dt), _obj(s, rhs, dt)
    pla.call("PLA_Pos_def_solve", A, B)
    x = pla.get_contents(B).astype(np.complex128)
    xr = np.linalg.solve(full, rhs.astype(np.complex128))
    assert np.abs(x - xr).max() <= 64 * EPS_F * np.abs(xr).max() * np.linalg.cond(full)
    pla.free(A, B)
    g = _input(ora, t, n, n, 7)
    A, B, P = _obj(s, g, dt), _obj(s, rhs, dt), s.mvector(n, dtype=pla.MPI_INT)
    pla.call("PLA_General_solve", A, P, B)
    x = pla.get_contents(B).astype(np.complex128)
    gr = g.astype(np.complex128)
    eta = np.abs(rhs - gr @ x).max() / (np.abs(gr).sum(axis=1).max() * np.abs(x).max() + np.abs(rhs).max())
    assert eta <= 16 * n * EPS_F
    pla.free(A, B, P)


@pytest.mark.parametrize("t", ["s", "c"])
def test_single_blas3(pla_session, t):
    s = pla_session
    dt, npt = SINGLE[t]
    rng = np.random.default_rng(5)

    def r(*shape):
        v = rng.uniform(-1, 1, shape) + (1j * rng.uniform(-1, 1, shape) if t == "c" else 0)
        return v.astype(npt)
    a, b, c = r(40, 30), r(30, 50), r(40, 50)
    A, B, C = _obj(s, a, dt), _obj(s, b, dt), _obj(s, c, dt)
    al, be = _scalar(s, 1.5, dt), _scalar(s, -0.5, dt)
    pla.call("PLA_Gemm", pla.PLA_NO_TRANSPOSE, pla.PLA_NO_TRANSPOSE, al, A, B, be, C)
    ref = 1.5 * a.astype(np.complex128) @ b.astype(np.complex128) - 0.5 * c
    assert pla.get_contents(C).dtype == npt
    assert np.abs(pla.get_contents(C) - ref).max() <= 4 * EPS_F * np.abs(ref).max()
    pla.free(A, B, C)
    t_ = r(30, 30) + 30 * np.eye(30, dtype=npt)
    x = r(30, 20)
    T, X = _obj(s, t_, dt), _obj(s, x, dt)
    pla.call("PLA_Trsm", pla.PLA_SIDE_LEFT, pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG,
             al, T, X)
    ref = np.linalg.solve(np.tril(t_).astype(np.complex128), 1.5 * x.astype(np.complex128))
    assert np.abs(pla.get_contents(X) - ref).max() <= 4 * EPS_F * np.abs(ref).max()
    pla.free(T, X)
    sy = r(40, 40)
    A, C = _obj(s, a, dt), _obj(s, sy, dt)
    pla.call("PLA_Syrk", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, al, A, be, C)
    a64 = a.astype(np.complex128)
    want = -0.5 * sy.astype(np.complex128)
    low = np.tril_indices(40)
    want[low] += (1.5 * a64 @ a64.T)[low]
    assert np.abs(pla.get_contents(C) - want).max() <= 4 * EPS_F * np.abs(want).max()
    pla.free(A, C, al, be)


@pytest.mark.parametrize("ta", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE, pla.PLA_CONJUGATE_TRANSPOSE, 24])
@pytest.mark.parametrize("tb", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE, pla.PLA_CONJUGATE_TRANSPOSE, 24])
def test_complex_gemm_all_transposes(pla_session, ta, tb):
    """PLA_CONJUGATE (24) conjugates without transposing."""
    s = pla_session
    rng = np.random.default_rng(ta * 100 + tb)
    m, n, k = 70, 45, 90

    def op(x, t):
        return {21: x, 22: x.T, 23: x.conj().T, 24: x.conj()}[t]

    def r(*shape):
        return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    a = r(*((k, m) if ta in (22, 23) else (m, k)))
    b = r(*((n, k) if tb in (22, 23) else (k, n)))
    c = r(m, n)
    A, B, C = (_obj(s, v, pla.MPI_DOUBLE_COMPLEX) for v in (a, b, c))
    alpha, beta = 0.75 - 1.25j, -0.5 + 0.25j
    Al, Be = _scalar(s, alpha, pla.MPI_DOUBLE_COMPLEX), _scalar(s, beta, pla.MPI_DOUBLE_COMPLEX)
    pla.call("PLA_Gemm", ta, tb, Al, A, B, Be, C)
    ref = alpha * op(a, ta) @ op(b, tb) + beta * c
    assert np.abs(pla.get_contents(C) - ref).max() <= 8 * k * EPS * np.abs(ref).max()
    pla.free(A, B, C, Al, Be)


@pytest.mark.parametrize("side", [pla.PLA_SIDE_LEFT, pla.PLA_SIDE_RIGHT])
@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("trans", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE, pla.PLA_CONJUGATE_TRANSPOSE, 24])
@pytest.mark.parametrize("diag", [pla.PLA_NONUNIT_DIAG, pla.PLA_UNIT_DIAG])
def test_complex_trsm_all_cases(pla_session, side, uplo, trans, diag):
    s = pla_session
    rng = np.random.default_rng(side * 1000 + uplo * 100 + trans + diag)
    m, n = 150, 90
    nt = m if side == pla.PLA_SIDE_LEFT else n
    # A unit triangle with O(1) entries has a condition number exponential in
    # nt; its off-diagonal part is scaled by 1/nt. The stored diagonal (nt)
    # must then be ignored by the solve.
    scale = 1.0 / nt if diag == pla.PLA_UNIT_DIAG else 1.0
    t = scale * (rng.uniform(-1, 1, (nt, nt)) + 1j * rng.uniform(-1, 1, (nt, nt))) + nt * np.eye(nt)
    b = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    tri = np.tril(t) if uplo == pla.PLA_LOWER_TRIANGULAR else np.triu(t)
    if diag == pla.PLA_UNIT_DIAG:
        tri = tri - np.diag(np.diag(tri)) + np.eye(nt)
    optri = {21: tri, 22: tri.T, 23: tri.conj().T, 24: tri.conj()}[trans]
    alpha = 2.0 - 0.5j
    ref = np.linalg.solve(optri, alpha * b) if side == pla.PLA_SIDE_LEFT else \
        np.linalg.solve(optri.T, (alpha * b).T).T
    T, B, Al = _obj(s, t, pla.MPI_DOUBLE_COMPLEX), _obj(s, b, pla.MPI_DOUBLE_COMPLEX), \
        _scalar(s, alpha, pla.MPI_DOUBLE_COMPLEX)
    pla.call("PLA_Trsm", side, uplo, trans, diag, Al, T, B)
    assert np.abs(pla.get_contents(B) - ref).max() <= 64 * nt * EPS * np.abs(ref).max()
    pla.free(T, B, Al)


@pytest.mark.parametrize("uplo", [pla.PLA_LOWER_TRIANGULAR, pla.PLA_UPPER_TRIANGULAR])
@pytest.mark.parametrize("trans", [pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE])
def test_complex_syrk(pla_session, uplo, trans):
    """zsyrk: C <- beta C (all of C, PLA_Syrk_panpan.c:22), then the uplo
    triangle += alpha op(A) op(A)^T (a transpose, not a conjugate)."""
    s = pla_session
    rng = np.random.default_rng(uplo + trans)
    n, k = 80, 60
    a = rng.uniform(-1, 1, (n, k) if trans == pla.PLA_NO_TRANSPOSE else (k, n)) * (1 + 0.5j)
    a = a + 1j * rng.uniform(-1, 1, a.shape)
    c = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    alpha, beta = 1.5 + 0.5j, 0.25 - 1j
    A, C = _obj(s, a, pla.MPI_DOUBLE_COMPLEX), _obj(s, c, pla.MPI_DOUBLE_COMPLEX)
    Al, Be = _scalar(s, alpha, pla.MPI_DOUBLE_COMPLEX), _scalar(s, beta, pla.MPI_DOUBLE_COMPLEX)
    pla.call("PLA_Syrk", uplo, trans, Al, A, Be, C)
    opa = a if trans == pla.PLA_NO_TRANSPOSE else a.T
    prod = alpha * opa @ opa.T
    ref = beta * c
    mask = np.tril(np.ones((n, n), bool)) if uplo == pla.PLA_LOWER_TRIANGULAR else np.triu(np.ones((n, n), bool))
    ref[mask] += prod[mask]
    assert np.abs(pla.get_contents(C) - ref).max() <= 8 * k * EPS * np.abs(ref).max()
    pla.free(A, C, Al, Be)


def test_complex_trsv_and_gemv(pla_session, ora):
    """The LU example's complex path (EXAMPLES/LU/driver.c): b = A x by
    PLA_Gemv, PLA_LU, interchanges, the two PLA_Trsv, then PLA_Axpy and
    PLA_Nrm2 of the error."""
    s = pla_session
    n = 500
    a = ora.uniform_z(n, n, 0)
    xt = ora.uniform_z(n, 1, 1)
    Z = pla.MPI_DOUBLE_COMPLEX
    A, X, Bv = _obj(s, a, Z), _obj(s, xt, Z), s.mvector(n, 1, Z)
    one, zero, mone = _scalar(s, 1.0, Z), _scalar(s, 0.0, Z), _scalar(s, -1.0, Z)
    pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, one, A, X, zero, Bv)
    assert np.abs(pla.get_contents(Bv) - a @ xt).max() <= 8 * n * EPS * np.abs(a @ xt).max()
    P = s.mvector(n, dtype=pla.MPI_INT)
    pla.call("PLA_LU", A, P)
    pla.call("PLA_Apply_pivots_to_rows", Bv, P)
    pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_UNIT_DIAG, A, Bv)
    pla.call("PLA_Trsv", pla.PLA_UPPER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG, A, Bv)
    pla.call("PLA_Axpy", mone, X, Bv)
    nrm = s.mscalar(dtype=pla.MPI_DOUBLE)
    pla.call("PLA_Nrm2", Bv, nrm)
    assert pla.get_contents(nrm)[0, 0] <= 1e-9
    pla.free(A, X, Bv, one, zero, mone, P, nrm)
