"""Pins the CPU oracle (oracle/pla_oracle.c) against checks that do not share
its code: glibc's drand48, the committed golden vectors (LAPACK pivots of the
reference's n=2048 seed-0 input, TEST/input_test_chol's case), LAPACK/numpy
factorizations and solves, exact rational elimination on tiny inputs, and the
algebraic invariants each routine must satisfy.  CPU only.
"""
import ctypes
import ctypes.util
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- generator

def test_drand48_matches_glibc(ora):
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (0, 1, 1793, 123456789):
        libc.srand48(seed)
        ref = np.array([libc.drand48() for _ in range(2000)])
        got = ora.drand48_stream(seed, 2000)
        assert np.array_equal(ref, got)


def test_lcg_skip_equals_stepping(ora):
    for skip in (0, 1, 2, 63, 64, 1000, 4097):
        assert np.array_equal(ora.drand48_stream(7, 5, skip=skip),
                              ora.drand48_stream(7, skip + 5)[skip:])


def test_fill_matches_golden_head(ora):
    head = np.load(os.path.join(GOLDEN, "a_n2048_seed0_head.npy"))
    a = ora.uniform(2048, 2048, seed=0)
    assert np.array_equal(a.ravel(order="F")[:16], head)
    # the values quoted for TEST/create_problem.c with seed 0
    assert a[0, 0] == -0.6583439277874206
    assert a[1, 0] == 0.49980396096992763


def test_fill_skip_continues_the_stream(ora):
    a = ora.uniform(7, 5, seed=3)
    b = ora.uniform(7, 3, seed=3, skip=7 * 2)
    assert np.array_equal(a[:, 2:], b)


# ---------------------------------------------------------------- idamax

def test_idamax_first_max(ora):
    x = np.array([1.0, -3.0, 2.0, 3.0, -3.0])
    assert ora.idamax(x) == 1  # ties resolve to the first index
    rng = np.random.default_rng(0)
    for _ in range(50):
        x = rng.standard_normal(rng.integers(1, 100))
        assert ora.idamax(x) == int(np.argmax(np.abs(x)))


# ---------------------------------------------------------------- LU

def _lapack_piv(a):
    _, ipiv = scipy.linalg.lu_factor(a, check_finite=False)
    return (ipiv - np.arange(len(ipiv))).astype(np.int32)


def test_lu_pivots_golden_n2048(ora):
    """n=2048, nb=128, seed 0 -- pivots equal LAPACK's."""
    golden = np.load(os.path.join(GOLDEN, "lu_piv_n2048_seed0.npy"))
    a = ora.uniform(2048, 2048, 0)
    lu, piv = ora.lu_blocked(a, nb=128)
    assert np.array_equal(piv, golden)
    assert ora.lu_scaled_residual(a, lu, piv) < 16


@pytest.mark.parametrize("m,n,nb", [(50, 50, 8), (64, 40, 16), (40, 64, 16), (129, 129, 32),
                                    (300, 300, 128), (1, 1, 4), (5, 1, 4), (1, 5, 4)])
def test_lu_blocked_vs_lapack(ora, m, n, nb):
    a = ora.uniform(m, n, seed=m * 31 + n)
    lu, piv = ora.lu_blocked(a, nb=nb)
    assert np.array_equal(piv, _lapack_piv(a))
    assert ora.lu_scaled_residual(a, lu, piv) < 16
    lu2, piv2 = ora.lu_unblocked(a)
    assert np.array_equal(piv, piv2)
    # reference's own sequential criterion (PLA_LU_enter_exit.c:130-153)
    assert np.abs(lu - lu2).max() <= 1e-7 * np.abs(a).max() * max(m, n)


def _exact_lu_pivots(a):
    """Partial pivoting in exact rational arithmetic (first max on ties)."""
    A = [[Fraction(float(v)) for v in row] for row in np.asarray(a)]
    m, n = len(A), len(A[0])
    piv = []
    for j in range(min(m, n)):
        col = [abs(A[i][j]) for i in range(j, m)]
        p = col.index(max(col))
        piv.append(p)
        A[j], A[j + p] = A[j + p], A[j]
        if A[j][j] != 0:
            for i in range(j + 1, m):
                f = A[i][j] / A[j][j]
                A[i][j] = f
                for c in range(j + 1, n):
                    A[i][c] -= f * A[j][c]
    return np.array(piv, dtype=np.int32)


def test_lu_exact_small_with_ties(ora):
    # integer matrices with deliberate |value| ties: first index wins
    rng = np.random.default_rng(5)
    for _ in range(20):
        n = int(rng.integers(2, 9))
        a = rng.integers(-3, 4, size=(n, n)).astype(np.float64)
        a[rng.integers(0, n), 0] = 3.0
        a[rng.integers(0, n), 0] = -3.0
        if abs(np.linalg.det(a)) < 0.5:
            continue
        _, piv = ora.lu_blocked(a, nb=2)
        assert np.array_equal(piv, _exact_lu_pivots(a))


def test_lu_planted_error_is_caught(ora):
    """A dropped row interchange must fail the residual check."""
    a = ora.uniform(64, 64, 9)
    lu, piv = ora.lu_blocked(a, nb=16)
    bad = piv.copy()
    j = int(np.flatnonzero(bad)[0])
    bad[j] = 0
    assert ora.lu_scaled_residual(a, lu, bad) > 1e3


# ---------------------------------------------------------------- laswp

def test_laswp_is_sequential_interchange(ora):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((20, 7))
    piv = np.array([3, 0, 5, 1, 0, 2], dtype=np.int32)
    got = ora.laswp(a, piv, r0=2)
    ref = a.copy()
    for i, p in enumerate(piv):
        r = 2 + i
        ref[[r, r + p]] = ref[[r + p, r]]
    assert np.array_equal(got, ref)


# ---------------------------------------------------------------- Cholesky

def _spd(n, seed, delta=1e-2):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1, 1, (n, n))
    return b @ b.T / n + delta * np.eye(n)


@pytest.mark.parametrize("n,nb", [(1, 4), (7, 4), (64, 16), (200, 64), (257, 128)])
def test_chol_vs_numpy(ora, n, nb):
    a = _spd(n, n)
    l, info = ora.chol_blocked(a, nb=nb)
    assert info == 0
    ref = np.linalg.cholesky(a)
    assert np.allclose(np.tril(l), ref, rtol=0, atol=1e-12 * n)
    assert np.array_equal(np.triu(l, 1), np.triu(a, 1))  # upper triangle untouched


def test_chol_input_test_chol_fixture(ora):
    """TEST/input_test_chol double case: n=20, nb_alg=4, A = create_problem + n I."""
    g = np.load(os.path.join(GOLDEN, "chol_n20_shift.npz"))
    a = ora.uniform(20, 20, 0) + 20 * np.eye(20)
    assert np.array_equal(a, g["a"])
    l, info = ora.chol_blocked(a, nb=int(g["nb"]))
    assert info == 0
    # reference criterion: max|diff| <= 1e-6 max|A| (PLA_Chol_enter_exit.c:115-132)
    assert np.abs(np.tril(l) - g["l"]).max() <= 1e-6 * np.abs(a).max()


def test_chol_failure_index_matches_leading_minors(ora):
    n = 40
    a = _spd(n, 3)
    a[25, 25] -= 50.0  # the 26th leading minor becomes negative first
    dets = [np.linalg.slogdet(a[:k, :k])[0] for k in range(1, n + 1)]
    first_bad = next(k for k, s in enumerate(dets, start=1) if s <= 0)
    _, info = ora.chol_blocked(a, nb=8)
    assert info == first_bad
    _, info_u = ora.chol_unblocked(a)
    assert info_u == first_bad


# ---------------------------------------------------------------- solves / BLAS

@pytest.mark.parametrize("uplo,trans,diag", [(u, t, d) for u in "LU" for t in "NT" for d in "UN"])
def test_trsm_left_vs_scipy(ora, uplo, trans, diag):
    rng = np.random.default_rng(2)
    n = 37
    t = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    b = rng.standard_normal((n, 3))
    x = ora.trsm_left(uplo, trans, diag, t, b, alpha=2.0)
    ref = scipy.linalg.solve_triangular(t, 2.0 * b, lower=(uplo == "L"), trans=trans,
                                        unit_diagonal=(diag == "U"))
    assert np.allclose(x, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["U", "N"])
def test_trsm_right_vs_scipy(ora, uplo, trans, diag):
    """X op(T) = alpha B  <=>  op(T)^T X^T = alpha B^T."""
    rng = np.random.default_rng(5)
    n, m = 29, 7
    t = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    b = rng.standard_normal((m, n))
    x = ora.trsm_right(uplo, trans, diag, t, b, alpha=-1.5)
    ref = scipy.linalg.solve_triangular(t, -1.5 * b.T, lower=(uplo == "L"),
                                        trans="N" if trans == "T" else "T",
                                        unit_diagonal=(diag == "U")).T
    assert np.allclose(x, ref, rtol=1e-12, atol=1e-12)
    # the strict opposite triangle of T is never read
    t2 = t.copy()
    mask = np.triu(np.ones((n, n), bool), 1) if uplo == "L" else np.tril(np.ones((n, n), bool), -1)
    t2[mask] = np.nan
    if diag == "U":
        np.fill_diagonal(t2, np.nan)
    assert np.array_equal(ora.trsm_right(uplo, trans, diag, t2, b, alpha=-1.5), x)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_syrk_vs_numpy(ora, uplo, trans):
    rng = np.random.default_rng(6)
    n, k = 19, 11
    a = rng.standard_normal((n, k) if trans == "N" else (k, n))
    c = rng.standard_normal((n, n))
    got = ora.syrk(uplo, trans, 0.75, a, -2.0, c)
    opa = a if trans == "N" else a.T
    full = 0.75 * opa @ opa.T - 2.0 * c
    tri = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.allclose(got[tri], full[tri], rtol=1e-13, atol=1e-13)
    # PLA_Syrk_panpan scales all of C by beta before the triangle update
    assert np.array_equal(got[~tri], -2.0 * c[~tri])
    # beta = 0 is a multiply: NaN in C survives in the untouched triangle
    c[~tri] = np.nan
    got0 = ora.syrk(uplo, trans, 1.0, a, 0.0, c)
    assert np.isnan(got0[~tri]).all()


@pytest.mark.parametrize("m,n,nb", [(60, 61, 8), (50, 50, 16), (70, 40, 12)])
def test_lu_linpack_vs_lu_blocked(ora, m, n, nb):
    """The LINPACK variant: U, pivots and the columns right of the square part
    equal PLA_LU's bit for bit; applying each later panel's interchanges to a
    panel's multipliers gives PLA_LU's L; on [A b] the appended column is
    inv(L) P b, so back substitution with U solves A x = b."""
    a = ora.uniform(m, n, 21)
    f1, p1 = ora.lu_blocked(a, nb)
    f2, p2 = ora.lu_linpack(a, nb)
    mn = min(m, n)
    assert np.array_equal(p1, p2)
    assert np.array_equal(np.triu(f1), np.triu(f2))
    L2 = np.tril(f2[:, :mn], -1)
    for k in range(0, mn, nb):
        cols = slice(k, min(k + nb, mn))
        blk = L2[:, cols]
        for j in range(min(k + nb, mn), mn):  # later interchanges, in order
            if p2[j]:
                blk[[j, j + p2[j]]] = blk[[j + p2[j], j]]
        L2[:, cols] = blk
    assert np.array_equal(L2, np.tril(f1[:, :mn], -1))
    if n == m + 1:
        x = scipy.linalg.solve_triangular(np.triu(f2[:, :m]), f2[:, m])
        assert np.abs(a[:, :m] @ x - a[:, m]).max() <= 1e-10 * np.abs(x).max()
    # an interchange into the columns left of a panel would show in L's first block
    if mn > nb and any(p2[nb:]):
        assert not np.array_equal(f1[:, :nb], f2[:, :nb])


@pytest.mark.parametrize("m,n,nb", [(40, 40, 8), (60, 35, 16), (35, 60, 16), (97, 97, 32)])
def test_zlu_blocked_vs_lapack(ora, m, n, nb):
    """MPI_DOUBLE_COMPLEX PLA_LU: izamax pivots (|re| + |im|) and zgeru updates
    agree with LAPACK zgetrf (scipy), which uses the same pivot rule."""
    a = ora.uniform_z(m, n, seed=m + n)
    f, piv = ora.zlu_blocked(a, nb=nb)
    lu, pp = scipy.linalg.lu_factor(a, check_finite=False)
    mn = min(m, n)
    assert np.array_equal(piv, pp[:mn] - np.arange(mn))
    assert np.abs(f - lu).max() <= 1e-12 * np.abs(a).max()


def test_zlu_exact_pivot_rule(ora):
    """|re| + |im| (not the modulus) picks the pivot: 3+3i (6) beats 5 (5) though
    |5| > |3+3i|; a tie keeps the first row."""
    a = np.array([[1.0, 1.0], [5.0, 2.0], [3 + 3j, 0.0], [-6j, 1.0]], dtype=complex)
    _, piv = ora.zlu_blocked(a.copy(), nb=2)
    assert piv[0] == 2  # rows 2 and 3 tie at 6: the first
    b = np.array([[1.0, 0.0], [5.0, 1.0], [3 + 3j, 2.0]], dtype=complex)
    _, piv = ora.zlu_blocked(b, nb=2)
    assert piv[0] == 2


@pytest.mark.parametrize("n,nb", [(30, 8), (64, 16), (97, 32)])
def test_zchol_blocked_vs_numpy(ora, n, nb):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = x @ x.conj().T / n + 0.1 * np.eye(n)
    f, info = ora.zchol_blocked(a, nb=nb)
    assert info == 0
    ref = np.linalg.cholesky(a)
    assert np.abs(np.tril(f) - ref).max() <= 1e-12 * np.abs(a).max()
    assert np.array_equal(np.triu(f, 1), np.triu(a, 1))  # the upper triangle is never written
    assert np.all(np.diag(f).imag == 0)


def test_zchol_failure_index(ora):
    """The first non-positive leading minor (exact signs on a small integer case)."""
    rng = np.random.default_rng(1)
    x = rng.integers(-3, 4, (12, 12)) + 1j * rng.integers(-3, 4, (12, 12))
    a = x @ x.conj().T + np.eye(12)
    a[7, 7] -= 400.0
    minors = [np.linalg.det(a[:k, :k]).real for k in range(1, 13)]
    first = next(k for k, d in enumerate(minors, 1) if d <= 0)
    _, info = ora.zchol_blocked(a, nb=4)
    assert info == first


def test_zlu_solve_and_stream(ora):
    n = 80
    ab = ora.uniform_z(n, n + 1, seed=2)
    # create_problem's complex stream: real then imaginary part of each element
    r = ora.uniform(2 * n, n + 1, 2)
    assert np.array_equal(ab.real, r[0::2]) and np.array_equal(ab.imag, r[1::2])
    f, piv = ora.zlu_blocked(ab[:, :n], nb=16)
    x = ora.zlu_solve(f, piv, ab[:, n])[:, 0]
    assert np.abs(x - np.linalg.solve(ab[:, :n], ab[:, n])).max() <= 1e-10 * np.abs(x).max()


def test_api_multi_vs_numpy_slices(ora):
    """Multi-block axpys: sub-block (i, j) of the local buffer starts at the
    running sums of the sizes and lands at (rd[i], cd[j]); overlapping targets
    receive every add in block order."""
    rng = np.random.default_rng(7)
    a = rng.standard_normal((40, 30))
    buf = rng.standard_normal((12, 9))
    rsz, csz, rd, cd = [5, 0, 7], [4, 5], [30, 2, 10], [20, 1]
    nb, na = ora.api_multi(True, rsz, csz, -0.5, buf, a, rd, cd)
    ref = a.copy()
    ref[30:35, 20:24] += -0.5 * buf[0:5, 0:4]
    ref[30:35, 1:6] += -0.5 * buf[0:5, 4:9]
    ref[10:17, 20:24] += -0.5 * buf[5:12, 0:4]
    ref[10:17, 1:6] += -0.5 * buf[5:12, 4:9]
    assert np.array_equal(na, ref) and np.array_equal(nb, buf)
    nb, na = ora.api_multi(False, rsz, csz, 2.0, buf, a, rd, cd)
    refb = buf.copy()
    refb[0:5, 0:4] += 2.0 * a[30:35, 20:24]
    refb[0:5, 4:9] += 2.0 * a[30:35, 1:6]
    refb[5:12, 0:4] += 2.0 * a[10:17, 20:24]
    refb[5:12, 4:9] += 2.0 * a[10:17, 1:6]
    assert np.array_equal(nb, refb) and np.array_equal(na, a)
    # overlapping row targets: both blocks add into rows 3..5, in order
    nb, na = ora.api_multi(True, [4, 4], [3], 1.0, buf[:8, :3], np.zeros((10, 5)), [2, 3], [1])
    ref = np.zeros((10, 5))
    ref[2:6, 1:4] += buf[0:4, 0:3]
    ref[3:7, 1:4] += buf[4:8, 0:3]
    assert np.array_equal(na, ref)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_gemm_vs_numpy(ora, ta, tb):
    rng = np.random.default_rng(4)
    m, n, k = 13, 17, 9
    a = rng.standard_normal((m, k) if ta == "N" else (k, m))
    b = rng.standard_normal((k, n) if tb == "N" else (n, k))
    c = rng.standard_normal((m, n))
    got = ora.gemm(ta, tb, -1.5, a, b, 0.5, c)
    ref = -1.5 * (a if ta == "N" else a.T) @ (b if tb == "N" else b.T) + 0.5 * c
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)


def test_lu_and_chol_solve(ora):
    a = ora.uniform(100, 100, 11)
    b = a @ np.ones(100)
    lu, piv = ora.lu_blocked(a, nb=32)
    x = ora.lu_solve(lu, piv, b)
    assert np.abs(x[:, 0] - 1).max() < 1e-10
    s = _spd(80, 2)
    l, info = ora.chol_blocked(s, nb=16)
    x = ora.chol_solve(np.tril(l), s @ np.ones(80))
    assert info == 0 and np.abs(x[:, 0] - 1).max() < 1e-10


# ---- E4M3 rounding of the oracle, pinned by the OCP OFP8 specification ----

def test_e4m3_encodings_from_specification(ora):
    # (value, byte): max normal, min normal, min subnormal, 1, -2, 0.1 -> 0.1015625
    cases = [(448.0, 0x7E), (2.0 ** -6, 0x08), (2.0 ** -9, 0x01), (1.0, 0x38), (-2.0, 0xC0),
             (0.1015625, 0x1D), (0.0, 0x00), (240.0, 0x77), (-448.0, 0xFE)]
    for v, byte in cases:
        assert int(ora.e4m3_encode(np.array([v]))[0]) == byte, v
        assert ora.e4m3_decode(np.array([byte], np.uint8))[0] == v


def test_e4m3_rounding_ties_and_subnormals(ora):
    r = ora.e4m3_round
    assert r(0.1) == 0.1015625
    assert r(1.0625) == 1.0 and r(1.1875) == 1.25           # ties to even mantissa
    assert r(0.001) == 2.0 ** -9 and r(0.0009) == 0.0        # subnormal grid 2^-9
    assert r(2.0 ** -10) == 0.0 and r(3 * 2.0 ** -10) == 2 * 2.0 ** -9  # ties to even
    assert r(440.0) == 448.0 and r(-3.3) == -3.25
    # every representable value maps to itself; midpoints go to the even neighbour
    vals = np.unique(np.abs(ora.e4m3_decode(np.arange(256, dtype=np.uint8))))
    vals = vals[np.isfinite(vals)]
    vals = vals[vals <= 448]
    assert np.array_equal(r(vals), vals)
    mids = (vals[:-1] + vals[1:]) / 2
    got = r(mids)
    lo_even = (ora.e4m3_encode(vals[:-1]) & 1) == 0
    assert np.array_equal(got, np.where(lo_even, vals[:-1], vals[1:]))


def test_fp8_tile_quantization_scales(ora):
    rng = np.random.default_rng(0)
    p = rng.uniform(-1, 1, (600, 200)) * np.repeat([3e-3, 17.0, 1e5], [256, 256, 88])[:, None]
    q, ex = ora.fp8_quantize_tiles(p)
    assert q.shape == (600, 256) and ex.tolist() == [-17, -4, 8]
    deq = ora.e4m3_decode(q[:, :200]) * np.exp2(np.repeat(ex, [256, 256, 88]))[:, None]
    assert np.abs(deq - p).max() <= 2.0 ** -4 * np.abs(p).max()     # unit roundoff 2^-4
    assert (q[:, 200:] == 0).all()


def test_fp8_tile_quantization_two_k_tiles_and_clamp(ora):
    """256 x 256 blocks along k too (a 400-column panel in 512-byte rows);
    each block's exponent follows its own maximum; an exponent below -127 (the
    E8M0 range) is clamped, and the block's values then round towards zero."""
    rng = np.random.default_rng(1)
    p = rng.uniform(-1, 1, (300, 400))
    p[:, 256:] *= 1e3
    p[256:, :256] *= 1e-60
    q, ex = ora.fp8_quantize_tiles(p, width=512)
    assert q.shape == (300, 512) and ex.shape == (4,)
    want = [int(np.frexp(np.abs(blk).max() / 448)[1]) for blk in
            (p[:256, :256], p[:256, 256:], p[256:, :256], p[256:, 256:])]
    want = [min(127, max(-127, e)) for e in want]
    assert ex.tolist() == want and ex[2] == -127
    for t, (rs, cs) in enumerate(((slice(0, 256), slice(0, 256)), (slice(0, 256), slice(256, 400)),
                                  (slice(256, 300), slice(256, 400)))):
        e = ex[[0, 1, 3][t]]
        deq = ora.e4m3_decode(q[rs, cs]) * 2.0 ** e
        assert np.abs(deq - p[rs, cs]).max() <= 2.0 ** -4 * np.abs(p[rs, cs]).max()
    assert np.abs(ora.e4m3_decode(q[256:, :256])).max() * 2.0 ** -127 <= 2.0 ** -136
    assert (q[:, 400:] == 0).all()


def test_blas_restatements_match_the_scalar_oracle(ora):
    """The CPU baseline's restatements over the host BLAS (lu_blocked_blas,
    chol_blocked_blas) follow the same algorithms as the scalar oracle: the
    same pivots (and the golden n=2048 vector), factors within rounding, the
    same Cholesky failure index, the upper triangle never written."""
    golden = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lu_piv_n2048_seed0.npy"))
    lu_b, piv_b = ora.lu_blocked_blas(ora.uniform(2048, 2048, 0), nb=128)
    assert np.array_equal(piv_b, golden)
    for m, n, nb in ((300, 300, 64), (700, 500, 128), (500, 700, 128), (1000, 1000, 96)):
        a = ora.uniform(m, n, m + n)
        lu_s, piv_s = ora.lu_blocked(a, nb=nb)
        lu_b, piv_b = ora.lu_blocked_blas(a, nb=nb)
        assert np.array_equal(piv_s, piv_b)
        assert np.abs(lu_s - lu_b).max() <= 1e-10
    rng = np.random.default_rng(4)
    b = rng.uniform(-1, 1, (600, 600))
    spd = b @ b.T / 600 + 1e-2 * np.eye(600)
    l_s, i_s = ora.chol_blocked(spd, nb=128)
    l_b, i_b = ora.chol_blocked_blas(spd, nb=128)
    assert i_s == i_b == 0
    assert np.abs(np.tril(l_s) - np.tril(l_b)).max() <= 1e-12
    assert np.array_equal(np.triu(l_b, 1), np.triu(spd, 1))
    spd[333, 333] = -1.0
    assert ora.chol_blocked(spd, nb=128)[1] == ora.chol_blocked_blas(spd, nb=128)[1] == 334
