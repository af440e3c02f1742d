"""GPU parity of the MPI_DOUBLE_COMPLEX paths through the C ABI: PLA_LU
(SOLVERS/LU/PLA_LU.c on complex data, the default datatype of the reference's
example drivers, EXAMPLES/LU/driver.c:110), PLA_General_solve,
PLA_Apply_pivots_to_rows, PLA_Chol / PLA_Chol_simple (A = L L^H,
SOLVERS/Cholesky/PLA_Chol.c:83-234; EXAMPLES/Cholesky/driver.c:17) and
PLA_Pos_def_solve, against the oracle's complex restatements
(ora_zlu_blocked pinned to LAPACK zgetrf, ora_zchol_blocked to
numpy.linalg.cholesky, in test_oracle.py).

Bar: pivots equal to the oracle's (integers); factors within the reference's
sequential-check tolerance 1e-7 * max|A| of the oracle (in practice ~1e-13 is
asserted); scaled residual ||PA - LU||_1 / (n ||A||_1 eps) < 16; solves to a
normwise backward error below 64 n eps.  Inputs are TEST/create_problem.c's
complex stream (real then imaginary part of each element), generated on the
GPU by PLA_Obj_fill_drand48 and checked bit for bit."""
import ctypes

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _zlu_gpu(s, a, nb):
    pla.call("PLA_Set_device_nb", nb)
    A = pla.from_numpy(s, a)
    m, n = a.shape
    P = s.mvector(min(m, n), dtype=pla.MPI_INT)
    pla.call("PLA_LU", A, P)
    f, piv = pla.get_contents(A), pla.get_contents(P)[:, 0]
    pla.free(A, P)
    pla.call("PLA_Set_device_nb", 0)
    return f, piv


def _scaled_residual(a, f, piv):
    m, n = a.shape
    mn = min(m, n)
    L = np.tril(f[:, :mn], -1) + np.eye(m, mn)
    U = np.triu(f[:mn, :])
    pa = a.copy()
    for i in range(mn):
        if piv[i]:
            pa[[i, i + piv[i]]] = pa[[i + piv[i], i]]
    return np.abs(pa - L @ U).sum(axis=0).max() / (max(m, n) * np.abs(a).sum(axis=0).max() * EPS)


def test_fill_complex_stream_bit_exact(pla_session, ora):
    s = pla_session
    A = s.matrix(301, 77, pla.MPI_DOUBLE_COMPLEX)
    pla.call("PLA_Obj_fill_drand48", A, 0, 0)
    assert np.array_equal(pla.get_contents(A), ora.uniform_z(301, 77, 0))
    pla.call("PLA_Obj_fill_drand48", A, 7, 1001)
    assert np.array_equal(pla.get_contents(A), ora.uniform_z(301, 77, 7, 1001))
    pla.free(A)


@pytest.mark.parametrize("m,n,nb", [(1, 1, 32), (5, 5, 32), (100, 37, 32), (37, 100, 32), (300, 300, 64),
                                    (1000, 1000, 256), (1500, 700, 128), (2049, 2049, 256)])
def test_zlu_vs_oracle(pla_session, ora, m, n, nb):
    a = ora.uniform_z(m, n, seed=m + 7 * n)
    f, piv = _zlu_gpu(pla_session, a, nb)
    f_o, piv_o = ora.zlu_blocked(a, nb=nb)
    assert np.array_equal(piv, piv_o)
    assert np.abs(f - f_o).max() <= 1e-7 * np.abs(a).max()
    assert np.abs(f - f_o).max() <= 1e-11 * np.abs(a).max() * max(m, n)
    assert _scaled_residual(a, f, piv) < 16


def test_zlu_pivot_rule_on_device(pla_session):
    """|re| + |im| picks the pivot (3+3i over 5), the first of a tie (rows 2, 3)."""
    a = np.array([[1.0, 1.0], [5.0, 2.0], [3 + 3j, 0.0], [-6j, 1.0]], dtype=complex)
    _, piv = _zlu_gpu(pla_session, a, 32)
    assert piv[0] == 2


@pytest.mark.parametrize("n,nrhs", [(500, 1), (1200, 3), (4096, 1)])
def test_zgeneral_solve(pla_session, ora, n, nrhs):
    s = pla_session
    A = s.matrix(n, n, pla.MPI_DOUBLE_COMPLEX)
    pla.call("PLA_Obj_fill_drand48", A, 3, 0)
    a = ora.uniform_z(n, n, 3)
    xs = np.random.default_rng(n).uniform(-1, 1, (n, nrhs)) + 1j * np.random.default_rng(n + 1).uniform(-1, 1, (n, nrhs))
    b = a @ xs
    B = pla.from_numpy(s, b)
    P = s.mvector(n, dtype=pla.MPI_INT)
    pla.call("PLA_General_solve", A, P, B)
    x = pla.get_contents(B)
    eta = np.abs(a @ x - b).max() / (np.abs(a).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    assert eta < 64 * n * EPS
    if n <= 1200:
        f_o, piv_o = ora.zlu_blocked(a, nb=256)
        assert np.array_equal(pla.get_contents(P)[:, 0], piv_o)
        x_o = ora.zlu_solve(f_o, piv_o, b)
        assert np.abs(x - x_o).max() <= 1e-9 * np.abs(x_o).max()
    pla.free(A, B, P)


def test_apply_pivots_complex_bit_exact(pla_session):
    s = pla_session
    rng = np.random.default_rng(5)
    a = rng.standard_normal((60, 9)) + 1j * rng.standard_normal((60, 9))
    piv = np.array([3, 0, 57, 1, 0, 2], dtype=np.int32)
    A = pla.from_numpy(s, a)
    P = s.mvector(6, dtype=pla.MPI_INT)
    pla.set_contents(P, piv)
    pla.call("PLA_Apply_pivots_to_rows", A, P)
    ref = a.copy()
    for i, p in enumerate(piv):
        if p:
            ref[[i, i + p]] = ref[[i + p, i]]
    assert np.array_equal(pla.get_contents(A), ref)
    pla.free(A, P)


def test_complex_unsupported_paths_fail_cleanly(pla_session):
    s = pla_session
    A = s.matrix(8, 8, pla.MPI_DOUBLE_COMPLEX)
    P = s.mvector(8, dtype=pla.MPI_INT)
    assert pla.lib().PLA_LU_linpack(A, P) == pla.PLA_ERR_UNSUPPORTED
    pla.free(A, P)


def _hpd(n, seed, delta=0.1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    return x @ x.conj().T / n + delta * np.eye(n)


def _zchol_gpu(s, a, nb, simple=False):
    pla.call("PLA_Set_device_nb", nb)
    A = pla.from_numpy(s, a)
    if simple:
        info = pla.call("PLA_Chol_simple", nb, A, ok=range(0, 1 << 30))
    else:
        info = pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A, ok=range(0, 1 << 30))
    out = pla.get_contents(A)
    pla.free(A)
    pla.call("PLA_Set_device_nb", 0)
    return out, info


@pytest.mark.parametrize("n,nb", [(1, 32), (31, 32), (64, 32), (300, 64), (1000, 256), (2049, 256)])
def test_zchol_vs_oracle(pla_session, ora, n, nb):
    a = _hpd(n, n)
    l, info = _zchol_gpu(pla_session, a, nb)
    lo, info_o = ora.zchol_blocked(a, nb=nb)
    assert info == 0 and info_o == 0
    assert np.abs(np.tril(l) - np.tril(lo)).max() <= 1e-13 * max(1.0, np.abs(a).max()) * np.sqrt(n)
    assert np.array_equal(np.triu(l, 1), np.triu(a, 1))  # the upper triangle is never written
    assert np.all(np.diag(l).imag == 0)


def test_zchol_simple(pla_session, ora):
    a = _hpd(700, 5)
    l, info = _zchol_gpu(pla_session, a, 100, simple=True)
    lo, _ = ora.zchol_blocked(a, nb=100)
    assert info == 0
    assert np.abs(np.tril(l) - np.tril(lo)).max() <= 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("n,nb,bad", [(40, 8, 25), (300, 64, 200), (700, 256, 0)])
def test_zchol_failure_index_and_state(pla_session, ora, n, nb, bad):
    """A non-positive pivot: its 1-based index, blocks before it factored, its
    diagonal block as it was before its factorization."""
    a = _hpd(n, 3)
    a[bad, bad] = -5.0
    l, info = _zchol_gpu(pla_session, a, nb)
    lo, info_o = ora.zchol_blocked(a, nb=nb)
    assert info == info_o and info > 0
    k = ((info - 1) // nb) * nb
    b = min(nb, n - k)
    tol = 1e-12 * np.abs(a).max()
    if k > 0:  # the blocks before it factored
        assert np.abs(np.tril(l[:, :k]) - np.tril(lo[:, :k])).max() <= tol
    # its diagonal block as the earlier steps left it: a - L10 L10^H on the lower triangle
    L10 = np.tril(lo)[k:k + b, :k]
    expect = a[k:k + b, k:k + b] - L10 @ L10.conj().T
    assert np.abs(np.tril(l[k:k + b, k:k + b] - expect)).max() <= tol
    assert np.array_equal(np.triu(l, 1), np.triu(a, 1))


@pytest.mark.parametrize("n,nrhs", [(500, 1), (1500, 4)])
def test_zpos_def_solve(pla_session, n, nrhs):
    s = pla_session
    a = _hpd(n, 9)
    rng = np.random.default_rng(2)
    xs = rng.uniform(-1, 1, (n, nrhs)) + 1j * rng.uniform(-1, 1, (n, nrhs))
    b = a @ xs
    A, B = pla.from_numpy(s, a), pla.from_numpy(s, b)
    pla.call("PLA_Pos_def_solve", A, B)
    x = pla.get_contents(B)
    eta = np.abs(a @ x - b).max() / (np.abs(a).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    assert eta < 64 * n * EPS
    pla.free(A, B)
