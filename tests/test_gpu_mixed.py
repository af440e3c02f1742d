"""GPU tests of PLA_LU_mixed: FP32-storage LU with a low-precision MFMA
trailing update (FP16 / BF16, single or split hi+lo products) and FP64
iterative refinement.

There is no reference counterpart, so the checks are what the mathematics
fixes.  The FP64 backward error is recomputed here on the host (numpy, FP64)
and must be <= the requested tolerance.  The pivots must form a valid
interchange sequence.  For the split precisions the factorization must be
good enough that refinement converges in a few steps.  A well-conditioned
input gives the same pivots as the FP64 oracle.  The FP64 solution agrees
with the exact solution to the conditioning-limited accuracy."""
import ctypes

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu


def _mixed(s, a, b, prec, tol=1e-14, max_iter=30, nb=0, ok=(pla.PLA_SUCCESS,)):
    n, nrhs = b.shape
    pla.call("PLA_Set_device_nb", nb)
    A, B = pla.from_numpy(s, a), pla.from_numpy(s, b)
    X = s.matrix(n, nrhs)
    P = s.mvector(n, dtype=pla.MPI_INT)
    it, be = ctypes.c_int(), ctypes.c_double()
    rc = pla.call("PLA_LU_mixed", A, P, B, X, prec, tol, max_iter, ctypes.byref(it), ctypes.byref(be), ok=ok)
    if len(ok) > 1:
        assert rc == (pla.PLA_SUCCESS if be.value <= tol else pla.PLA_ERR_NOT_CONVERGED)
    x, piv = pla.get_contents(X), pla.get_contents(P)[:, 0]
    a_after = pla.get_contents(A)
    pla.free(A, B, X, P)
    pla.call("PLA_Set_device_nb", 0)
    assert np.array_equal(a_after, a), "A must be left unchanged"
    return x, piv, it.value, be.value


def _eta(a, x, b):
    r = b - a @ x
    an = np.abs(a).sum(axis=1).max()
    return max(np.abs(r[:, j]).max() / (an * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(b.shape[1]))


@pytest.mark.parametrize("prec", [pla.PLA_MIXED_FP16X3, pla.PLA_MIXED_BF16X3])
@pytest.mark.parametrize("n,nrhs,nb", [(300, 1, 64), (1000, 16, 128), (2049, 4, 256), (4096, 1, 512), (5000, 2, 1024)])
def test_split_precisions_reach_fp64_backward_error(pla_session, ora, prec, n, nrhs, nb):
    a = ora.uniform(n, n, seed=n)
    b = np.random.default_rng(n).uniform(-1, 1, (n, nrhs))
    x, piv, it, be = _mixed(pla_session, a, b, prec, nb=nb)
    assert be <= 1e-14, (it, be)
    assert it <= 8
    assert _eta(a, x, b) <= 2e-14  # host FP64 recomputation agrees
    assert piv.min() >= 0 and (piv + np.arange(n)).max() < n


@pytest.mark.parametrize("prec", [pla.PLA_MIXED_FP16, pla.PLA_MIXED_BF16])
def test_single_half_precision_converges_when_well_conditioned(pla_session, ora, prec):
    n = 1500
    a = ora.uniform(n, n, seed=3) + n / 4 * np.eye(n)  # kappa ~ 10: kappa * u_half << 1
    b = np.random.default_rng(1).uniform(-1, 1, (n, 2))
    x, piv, it, be = _mixed(pla_session, a, b, prec, nb=128)
    assert be <= 1e-14, (it, be)
    assert _eta(a, x, b) <= 2e-14
    # diagonally dominant: no interchanges, like the FP64 oracle
    _, piv_o = ora.lu_blocked(a, nb=128)
    assert np.array_equal(piv, piv_o)


def test_solution_accuracy_vs_exact(pla_session, ora):
    n = 1024
    a = ora.uniform(n, n, seed=11)
    xt = np.random.default_rng(2).uniform(-1, 1, (n, 3))
    b = a @ xt
    x, piv, it, be = _mixed(pla_session, a, b, pla.PLA_MIXED_FP16X3, nb=128)
    cond = np.linalg.cond(a, np.inf)
    assert np.abs(x - xt).max() <= 50 * cond * np.finfo(np.float64).eps * np.abs(xt).max()


def test_max_iter_zero_reports_unrefined_error(pla_session, ora):
    n = 512
    a = ora.uniform(n, n, seed=5)
    b = np.random.default_rng(5).uniform(-1, 1, (n, 1))
    x, piv, it, be = _mixed(pla_session, a, b, pla.PLA_MIXED_BF16, max_iter=0, nb=128,
                            ok=(pla.PLA_SUCCESS, pla.PLA_ERR_NOT_CONVERGED))
    assert it == 0
    assert be > 1e-14  # a BF16 update alone is far from FP64 accuracy
    assert abs(be - _eta(a, x, b)) <= 0.05 * be


def test_argument_errors(pla_session):
    s = pla_session
    A, B, X = s.matrix(8, 8), s.matrix(8, 17), s.matrix(8, 17)
    P = s.mvector(8, dtype=pla.MPI_INT)
    it, be = ctypes.c_int(), ctypes.c_double()
    assert pla.lib().PLA_LU_mixed(A, P, B, X, pla.PLA_MIXED_FP16, 1e-14, 5, ctypes.byref(it),
                                  ctypes.byref(be)) < 0  # 17 right-hand sides
    B1, X1 = s.matrix(8, 1), s.matrix(8, 1)
    assert pla.lib().PLA_LU_mixed(A, P, B1, X1, 9, 1e-14, 5, None, None) < 0  # unknown precision
    pla.free(A, B, X, P, B1, X1)


def test_block_size_schedule_reaches_fp64_backward_error(pla_session, ora):
    """A smaller block size for the last steps (PLA_PARAM_NB_TAIL once at most
    PLA_PARAM_NB_SWITCH rows remain; 200 is no multiple of the 32-deep
    low-precision product, so its K is padded) still refines to FP64."""
    n = 3000
    a = ora.uniform(n, n, seed=5)
    b = np.random.default_rng(5).uniform(-1, 1, (n, 2))
    pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_TAIL, 200)
    pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_SWITCH, 1800)
    try:
        x, piv, it, be = _mixed(pla_session, a, b, pla.PLA_MIXED_FP16X3, nb=512)
    finally:
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_TAIL, 0)
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_SWITCH, 0)
    assert be <= 1e-14 and it <= 8, (it, be)
    assert _eta(a, x, b) <= 2e-14
    assert piv.min() >= 0 and (piv + np.arange(n)).max() < n
