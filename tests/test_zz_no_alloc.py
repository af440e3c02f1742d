"""PLA_Chol_no_alloc -- PLA_Chol (SOLVERS/Cholesky/PLA_Chol.c:13) with its
scratch (info words, diagonal-block backup) in a static device array, for
callers that cannot allocate.  It has no counterpart in the reference; the
existing entry point PLA_Chol is its reference here.

Bar: on the same inputs, PLA_Chol_no_alloc returns what PLA_Chol returns and
writes bit-identical factors (the same kernels in the same order), on two
calls in a row, the second with a smaller matrix and a smaller block size
that also fails at a known pivot, so it runs over the scratch the first call
left behind and restores its failing block from that scratch."""
import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu


def _spd(n, seed, delta=1e-2):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1, 1, (n, n))
    return b @ b.T / n + delta * np.eye(n)


def _factor(s, fn, a, nb):
    pla.call("PLA_Set_device_nb", nb)
    try:
        A = pla.from_numpy(s, a)
        info = pla.call(fn, pla.PLA_LOWER_TRIANGULAR, A, ok=range(0, 1 << 30))
        out = pla.get_contents(A)
        pla.free(A)
    finally:
        pla.call("PLA_Set_device_nb", 0)
    return out, info


def test_chol_no_alloc_matches_chol_on_two_calls(pla_session, ora):
    # (n, nb, index of a diagonal entry made negative or None)
    cases = [(1500, 512, None), (700, 128, 450)]
    inputs = []
    for n, nb, bad in cases:
        a = _spd(n, n)
        if bad is not None:
            a[bad, bad] = -1.0
        inputs.append((a, nb, bad))
    refs = [_factor(pla_session, "PLA_Chol", a, nb) for a, nb, _ in inputs]
    outs = [_factor(pla_session, "PLA_Chol_no_alloc", a, nb) for a, nb, _ in inputs]  # two calls in a row
    for (a, nb, bad), (l_ref, info_ref), (l, info) in zip(inputs, refs, outs):
        assert info == info_ref, (a.shape, nb)
        assert np.array_equal(l, l_ref), (a.shape, nb)
        assert np.array_equal(np.triu(l, 1), np.triu(a, 1))  # upper triangle never written
        if bad is None:
            assert info == 0
            lo, info_o = ora.chol_blocked(a, nb=nb)
            assert info_o == 0
            n = a.shape[0]
            assert np.abs(np.tril(l) - np.tril(lo)).max() <= 1e-13 * max(1.0, np.abs(a).max()) * np.sqrt(n)
        else:
            assert info == bad + 1
            # the failing diagonal block holds the Schur complement it had on
            # entry (restored from the scratch's backup)
            k = (bad // nb) * nb
            b = min(nb, a.shape[0] - k)
            l21 = np.linalg.solve(np.linalg.cholesky(a[:k, :k]), a[:k, k:k + b]).T
            schur = a[k:k + b, k:k + b] - l21 @ l21.T
            assert np.abs(np.tril(l[k:k + b, k:k + b]) - np.tril(schur)).max() < 1e-10


def test_chol_no_alloc_rejects_other_datatypes_and_upper(pla_session):
    s = pla_session
    A = pla.from_numpy(s, _spd(8, 1))
    with pytest.raises(pla.PLAError) as e:
        pla.call("PLA_Chol_no_alloc", pla.PLA_UPPER_TRIANGULAR, A)
    assert e.value.code == pla.PLA_ERR_UNSUPPORTED
    pla.free(A)
    Z = pla.from_numpy(s, _spd(8, 1).astype(np.complex128))
    with pytest.raises(pla.PLAError) as e:
        pla.call("PLA_Chol_no_alloc", pla.PLA_LOWER_TRIANGULAR, Z)
    assert e.value.code == pla.PLA_ERR_UNSUPPORTED
    pla.free(Z)
