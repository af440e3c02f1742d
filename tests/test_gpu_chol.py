"""GPU parity of PLA_Chol (SOLVERS/Cholesky/PLA_Chol.c:13) and
PLA_Pos_def_solve (SOLVERS/Cholesky/PLA_Pos_def.c:44) through the C ABI.

Bar: L within the reference's sequential-check tolerance 1e-6 * max|A|
(PLA_Chol_enter_exit.c:115-132) of the oracle -- in practice 1e-13 is asserted;
the upper triangle is never written; the failure index k+i+1 of a
non-positive pivot (PLA_Chol.c:181-185) equals the oracle's and the exact
first non-positive leading minor, with the failing diagonal block left as it
was before its factorization."""
import os

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _spd(n, seed, delta=1e-2):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1, 1, (n, n))
    return b @ b.T / n + delta * np.eye(n)


def _chol_gpu(s, a, nb=0):
    pla.call("PLA_Set_device_nb", nb)
    A = pla.from_numpy(s, a)
    info = pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A, ok=range(0, 1 << 30))
    out = pla.get_contents(A)
    pla.free(A)
    pla.call("PLA_Set_device_nb", 0)
    return out, info


def test_chol_input_test_chol_case(pla_session):
    """TEST/input_test_chol, double: n=20, nb_alg=4, create_problem + n I."""
    g = np.load(os.path.join(GOLDEN, "chol_n20_shift.npz"))
    a = g["a"]
    l, info = _chol_gpu(pla_session, a, nb=int(g["nb"]))
    assert info == 0
    assert np.abs(np.tril(l) - g["l"]).max() <= 1e-13 * np.abs(a).max()
    assert np.array_equal(np.triu(l, 1), np.triu(a, 1))


@pytest.mark.parametrize("n,nb", [(1, 128), (63, 128), (64, 64), (65, 64), (300, 128), (1000, 128),
                                  (1537, 256), (2048, 512)])
def test_chol_vs_oracle(pla_session, ora, n, nb):
    a = _spd(n, n)
    l, info = _chol_gpu(pla_session, a, nb=nb)
    lo, info_o = ora.chol_blocked(a, nb=nb)
    assert info == 0 and info_o == 0
    assert np.abs(np.tril(l) - np.tril(lo)).max() <= 1e-13 * max(1.0, np.abs(a).max()) * np.sqrt(n)
    assert np.array_equal(np.triu(l, 1), np.triu(a, 1))


@pytest.mark.parametrize("n,nb,bad", [(40, 8, 25), (300, 64, 200), (300, 128, 0), (1000, 256, 777)])
def test_chol_failure_index_and_state(pla_session, ora, n, nb, bad):
    a = _spd(n, 3)
    a[bad, bad] -= 50.0
    dets = [np.linalg.slogdet(a[:k, :k])[0] for k in range(1, min(n, bad + 2) + 1)]
    first_bad = next(k for k, sg in enumerate(dets, start=1) if sg <= 0)
    l, info = _chol_gpu(pla_session, a, nb=nb)
    _, info_o = ora.chol_blocked(a, nb=nb)
    assert info == first_bad == info_o
    k = ((info - 1) // nb) * nb
    b = min(nb, n - k)
    if k > 0:  # blocks before the failure are factored
        ref = np.linalg.cholesky(a[:k, :k])
        assert np.abs(np.tril(l[:k, :k]) - ref).max() < 1e-12
    # the failing diagonal block holds the Schur complement it had on entry
    if k > 0:
        l21 = np.linalg.solve(np.linalg.cholesky(a[:k, :k]), a[:k, k:k + b]).T
        schur = a[k:k + b, k:k + b] - l21 @ l21.T
    else:
        schur = a[:b, :b]
    assert np.abs(np.tril(l[k:k + b, k:k + b]) - np.tril(schur)).max() < 1e-10
    assert np.array_equal(np.triu(l, 1), np.triu(a, 1))


@pytest.mark.parametrize("n,nb_alg", [(20, 4), (300, 32), (1000, 100), (777, 512)])
def test_chol_simple_vs_oracle(pla_session, ora, n, nb_alg):
    """PLA_Chol_simple (SOLVERS/Cholesky/PLA_Chol_simple.c:13) with the
    caller's block size."""
    a = _spd(n, n + 1)
    A = pla.from_numpy(pla_session, a)
    assert pla.call("PLA_Chol_simple", nb_alg, A) == 0
    l = pla.get_contents(A)
    lo, info_o = ora.chol_blocked(a, nb=nb_alg)
    assert info_o == 0
    assert np.abs(np.tril(l) - np.tril(lo)).max() <= 1e-13 * max(1.0, np.abs(a).max()) * np.sqrt(n)
    assert np.array_equal(np.triu(l, 1), np.triu(a, 1))
    pla.free(A)


def test_chol_simple_edge_cases(pla_session, ora):
    s = pla_session
    a = _spd(50, 4)
    A = pla.from_numpy(s, a)
    assert pla.call("PLA_Chol_simple", 0, A) == 0  # nb_alg <= 0: the reference's loop ends at once
    assert np.array_equal(pla.get_contents(A), a)
    a[30, 30] = -1.0
    pla.set_contents(A, a)
    info = pla.call("PLA_Chol_simple", 8, A, ok=range(0, 1 << 30))
    _, info_o = ora.chol_blocked(a, nb=8)
    assert info == info_o and info > 0
    pla.free(A)


def test_chol_upper_is_unsupported(pla_session):
    A = pla.from_numpy(pla_session, np.eye(4))
    assert pla.lib().PLA_Chol(pla.PLA_UPPER_TRIANGULAR, A) < 0
    pla.free(A)


def test_pos_def_solve(pla_session):
    s = pla_session
    n = 900
    a = _spd(n, 9)
    x = np.random.default_rng(1).uniform(-1, 1, (n, 4))
    A, B = pla.from_numpy(s, a), pla.from_numpy(s, a @ x)
    pla.call("PLA_Pos_def_solve", A, B)
    assert np.abs(pla.get_contents(B) - x).max() < 1e-10
    pla.free(A, B)
