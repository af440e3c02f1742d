"""GPU tests of the FP8 Cholesky path and of GMRES-IR.

- The E4M3 panel quantization is byte-exact against the oracle's rounding
  (oracle.e4m3_round, pinned by the OCP specification in test_oracle.py).
- The FP8 MFMA SYRK is bit-exact on inputs whose sums are exact in FP32
  (small E4M3 values, power-of-two tile scales): this pins the MFMA operand and
  result layouts, the lower-triangle masking and the tile scales.
- The symmetric residual (lower triangle only) agrees with an exactly rounded
  host sum and ignores the upper triangle.
- PLA_Chol_mixed reaches FP64 backward error on SPD systems B B^T/n + delta I
  (the configurations' generator), including delta = 4e-4 where the FP8
  factorization needs a diagonal shift and GMRES-IR.
- PLA_LU_mixed falls back to GMRES-IR when classic refinement cannot converge.
There is no reference counterpart: these are what the mathematics fixes.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu

_HOOKS = os.path.join(os.path.dirname(pla.LIB_PATH), "libpla_testhooks.so")


@pytest.fixture(scope="module")
def th(pla_session):
    L = ctypes.CDLL(_HOOKS)
    p, i, l, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
    L.pla_th_fp8_quantize.argtypes = [i, i, p, l, p, p]
    L.pla_th_fp8_syrk.argtypes = [i, p, p, p, l]
    L.pla_th_sym_residual.argtypes = [i, i, p, l, p, l, p, l, p, l, p]
    L.pla_th_sym_inf_norm.argtypes = [i, p, l, ctypes.POINTER(d)]
    L.pla_th_fp8_update.argtypes = [i, i, i, i, p, p, i, i, p, l]
    L.pla_th_mirror_lower.argtypes = [i, p, l]
    L.pla_th_set_fp8_options.argtypes = [i, i, i]
    return L


class _Dev:
    def __init__(self, s, host, dt):
        host = np.asarray(host)
        self.shape = host.shape
        self.obj = s.matrix(host.shape[0], host.shape[1] if host.ndim == 2 else 1, dt)
        ptr = ctypes.c_void_p()
        pla.call("PLA_Obj_device_buffer", self.obj, ctypes.byref(ptr))
        self.ptr = ptr
        pla.set_contents(self.obj, host)

    def get(self):
        return pla.get_contents(self.obj)


def _tiles_panel(rng, R, b):
    scales = np.array([3e-3, 17.0, 1e5, 0.25])[(np.arange(R) // 256) % 4]
    return rng.uniform(-1, 1, (R, b)) * scales[:, None]


def test_fp8_quantize_matches_oracle(pla_session, ora, th):
    s = pla_session
    rng = np.random.default_rng(1)
    for R, b in [(600, 200), (256, 256), (37, 5), (700, 400), (300, 512)]:
        p = _tiles_panel(rng, R, b)
        if b > 256:
            p[:, 256:] *= 1e-3  # the second k tile gets its own exponents
        P = _Dev(s, p, pla.MPI_DOUBLE)
        nt, kq = (R + 255) // 256, 256 * ((b + 255) // 256)
        Q = _Dev(s, np.zeros((R * kq // 4, 1), np.int32), pla.MPI_INT)
        E = _Dev(s, np.zeros((nt * kq // 256, 1), np.int32), pla.MPI_INT)
        assert th.pla_th_fp8_quantize(R, b, P.ptr, R, Q.ptr, E.ptr) == 0
        got = Q.get()[:, 0].view(np.uint8).reshape(R, kq)
        q_ref, ex_ref = ora.fp8_quantize_tiles(p, width=kq)
        assert np.array_equal(E.get()[:, 0], ex_ref)
        assert np.array_equal(got, q_ref), (R, b, np.argwhere(got != q_ref)[:5])
        pla.free(P.obj, Q.obj, E.obj)


def test_fp8_syrk_exact(pla_session, ora, th):
    s = pla_session
    rng = np.random.default_rng(2)
    R = 300
    vals = np.array([0.0, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0, -0.5, -1.0, -2.0, -3.0, -4.0])
    lq = np.zeros((R, 256))
    lq[:, :200] = rng.choice(vals, (R, 200))
    qbytes = ora.e4m3_encode(lq)
    exps = np.array([-2, 3], np.int32)
    c0 = rng.integers(-1000, 1000, (R, R)).astype(np.float32)
    C = _Dev(s, c0, pla.MPI_FLOAT)
    Q = _Dev(s, qbytes.reshape(-1).view(np.int32).reshape(-1, 1), pla.MPI_INT)
    E = _Dev(s, exps.reshape(-1, 1), pla.MPI_INT)
    assert th.pla_th_fp8_syrk(R, Q.ptr, E.ptr, C.ptr, R) == 0
    got = C.get()
    sc = np.exp2(exps[np.arange(R) // 256].astype(np.float64))
    full = c0.astype(np.float64) - (sc[:, None] * sc[None, :]) * (lq @ lq.T)
    lo = np.tril_indices(R)
    up = np.triu_indices(R, 1)
    assert np.array_equal(got[lo].astype(np.float64), full[lo])
    assert np.array_equal(got[up], c0[up])
    pla.free(C.obj, Q.obj, E.obj)


@pytest.mark.parametrize("tri", [True, False])
def test_fp8_update_two_k_tiles_exact(pla_session, ora, th, tri):
    """C -= 2^(e(i,k) + e(j,k)) Lq_i Lq_j^T over a 400-column panel (two k
    tiles, 512-byte rows) whose exponents differ per row tile and per k
    tile; the triangle and the rectangle launches at 256-row offsets.  The
    products and sums are exact in FP32, so the result is bit-exact."""
    s = pla_session
    rng = np.random.default_rng(5 + tri)
    R, b, kq = 900, 400, 512
    vals = np.array([0.0, 0.5, 1.0, 1.5, 2.0, 3.0, -0.5, -1.0, -2.0, -3.0])
    lq = np.zeros((R, kq))
    lq[:, :b] = rng.choice(vals, (R, b))
    exps = rng.integers(-3, 4, (R + 255) // 256 * 2).astype(np.int32)
    Q = _Dev(s, ora.e4m3_encode(lq).reshape(-1).view(np.int32).reshape(-1, 1), pla.MPI_INT)
    E = _Dev(s, exps.reshape(-1, 1), pla.MPI_INT)
    sc = np.exp2(exps.reshape(-1, 2).astype(np.float64))[np.arange(R) // 256]  # R x 2
    full = np.zeros((R, R))
    for kt in range(2):
        blk = lq[:, 256 * kt:256 * (kt + 1)] * sc[:, kt:kt + 1]
        full += blk @ blk.T
    if tri:
        ro = co = 256
        M = N = R - 256
    else:
        ro, co, M, N = 512, 256, R - 512, 256
    c0 = rng.integers(-1000, 1000, (M, N)).astype(np.float32)
    C = _Dev(s, c0, pla.MPI_FLOAT)
    assert th.pla_th_fp8_update(int(tri), M, N, b, Q.ptr, E.ptr, ro, co, C.ptr, M) == 0
    got = C.get().astype(np.float64)
    want = c0.astype(np.float64) - full[ro:ro + M, co:co + N]
    if tri:
        lo, up = np.tril_indices(M), np.triu_indices(M, 1)
        assert np.array_equal(got[lo], want[lo])
        assert np.array_equal(got[up], c0[up].astype(np.float64))
    else:
        assert np.array_equal(got, want)
    pla.free(C.obj, Q.obj, E.obj)


def test_mirror_lower(pla_session, th):
    s = pla_session
    rng = np.random.default_rng(6)
    for n in (1, 63, 64, 200):
        w = rng.standard_normal((n, n)).astype(np.float32)
        W = _Dev(s, w, pla.MPI_FLOAT)
        assert th.pla_th_mirror_lower(n, W.ptr, n) == 0
        got = W.get()
        want = np.tril(w) + np.tril(w, -1).T
        assert np.array_equal(got, want)
        pla.free(W.obj)


def test_sym_residual_lower_only(pla_session, th):
    s = pla_session
    rng = np.random.default_rng(3)
    n, nrhs = 700, 3
    g = rng.uniform(-1, 1, (n, n))
    a = (g + g.T) / 2
    stored = a.copy()
    stored[np.triu_indices(n, 1)] = np.nan           # never read
    x = rng.uniform(-1, 1, (n, nrhs))
    b = a @ x + rng.uniform(-1e-10, 1e-10, (n, nrhs))
    A, X, B = _Dev(s, stored, pla.MPI_DOUBLE), _Dev(s, x, pla.MPI_DOUBLE), _Dev(s, b, pla.MPI_DOUBLE)
    Rr = _Dev(s, np.zeros((n, nrhs)), pla.MPI_DOUBLE)
    rmax = (ctypes.c_double * nrhs)()
    assert th.pla_th_sym_residual(n, nrhs, A.ptr, n, X.ptr, n, B.ptr, n, Rr.ptr, n,
                                  ctypes.cast(rmax, ctypes.c_void_p)) == 0
    r = Rr.get()
    ref = np.array([[math.fsum([b[i, q]] + [-a[i, j] * x[j, q] for j in range(n)]) for q in range(nrhs)]
                    for i in range(0, n, 7)])
    scale = (np.abs(a) @ np.abs(x)).max()
    assert np.abs(r[::7] - ref).max() <= 1e-15 * scale
    assert np.allclose(list(rmax), np.abs(r).max(axis=0), rtol=0, atol=0)
    nrm = ctypes.c_double()
    assert th.pla_th_sym_inf_norm(n, A.ptr, n, ctypes.byref(nrm)) == 0
    assert abs(nrm.value - np.abs(a).sum(axis=1).max()) <= 1e-13 * nrm.value
    pla.free(A.obj, X.obj, B.obj, Rr.obj)


def _spd(rng, n, delta):
    bm = rng.uniform(-1, 1, (n, n))
    return bm @ bm.T / n + delta * np.eye(n)


def _eta(a, x, b):
    an = np.abs(a).sum(axis=1).max()
    r = b - a @ x
    return max(np.abs(r[:, j]).max() / (an * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(b.shape[1]))


def _chol_mixed(s, a, b, tol=1e-14, max_iter=30, upper_nan=True):
    n, nrhs = b.shape
    stored = a.copy()
    if upper_nan:
        stored[np.triu_indices(n, 1)] = np.nan       # only the lower triangle is read
    A, B = pla.from_numpy(s, stored), pla.from_numpy(s, b)
    X = s.matrix(n, nrhs)
    it, be = ctypes.c_int(), ctypes.c_double()
    pla.call("PLA_Chol_mixed", A, B, X, pla.PLA_MIXED_FP8, tol, max_iter, ctypes.byref(it), ctypes.byref(be))
    x = pla.get_contents(X)
    a_after = pla.get_contents(A)
    outer, steps, gcorr, shift = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    pla.call("PLA_Mixed_stats", ctypes.byref(outer), ctypes.byref(steps), ctypes.byref(gcorr), ctypes.byref(shift))
    pla.free(A, B, X)
    assert np.array_equal(a_after, stored, equal_nan=True), "A must be left unchanged"
    return x, it.value, be.value, (outer.value, steps.value, gcorr.value, shift.value)


@pytest.mark.parametrize("n,delta,nrhs", [(300, 4e-2, 1), (1000, 4e-2, 8), (1000, 4e-4, 8), (2600, 4e-2, 4),
                                          (2600, 4e-4, 2)])
def test_chol_fp8_reaches_fp64_backward_error(pla_session, n, delta, nrhs):
    rng = np.random.default_rng(n + int(1 / delta))
    a = _spd(rng, n, delta)
    b = rng.uniform(-1, 1, (n, nrhs))
    x, it, be, (outer, steps, gcorr, shift) = _chol_mixed(pla_session, a, b)
    assert be <= 1e-14, (it, be, steps, shift)
    assert _eta(a, x, b) <= 2e-14
    assert outer == it and steps >= 0 and shift >= 0.0
    x_ref = np.linalg.solve(a, b)
    cond = np.linalg.cond(a, 1)
    assert np.abs(x - x_ref).max() <= 50 * cond * 1.1e-16 * np.abs(x_ref).max()


@pytest.mark.parametrize("nb,masked", [(256, 1), (512, 1), (256, 0)])
def test_chol_fp8_block_sizes_and_streams(pla_session, th, nb, masked):
    """The look-ahead FP8 factorization with either block size, its panel on
    the CU-masked stream pair or beside the update, with a ragged last block."""
    rng = np.random.default_rng(nb + masked)
    n = 1700
    a = _spd(rng, n, 4e-2)
    b = rng.uniform(-1, 1, (n, 3))
    th.pla_th_set_fp8_options(nb, masked, 0)
    try:
        x, it, be, _ = _chol_mixed(pla_session, a, b)
    finally:
        th.pla_th_set_fp8_options(512, 0, 0)
    assert be <= 1e-14 and _eta(a, x, b) <= 2e-14, (it, be)


def test_chol_fp8_not_spd_raises(pla_session):
    n = 300
    a = -np.eye(n)
    b = np.ones((n, 1))
    A, B = pla.from_numpy(pla_session, a), pla.from_numpy(pla_session, b)
    X = pla_session.matrix(n, 1)
    rc = pla.lib().PLA_Chol_mixed(A, B, X, pla.PLA_MIXED_FP8, 1e-14, 5, None, None)
    assert rc < 0 and "positive definite" in pla.last_error()
    assert pla.lib().PLA_Chol_mixed(A, B, X, 1, 1e-14, 5, None, None) < 0   # not an FP8 precision
    pla.free(A, B, X)


def test_lu_mixed_falls_back_to_gmres(pla_session):
    """kappa(A) = 1e4 with FP16 factors: kappa u_half ~ 5 > 1, so classic
    refinement cannot converge; GMRES-IR with the same factors does."""
    rng = np.random.default_rng(7)
    n = 800
    u, _ = np.linalg.qr(rng.standard_normal((n, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (u * np.logspace(0, -4, n)) @ v.T
    b = rng.uniform(-1, 1, (n, 2))
    A, B = pla.from_numpy(pla_session, a), pla.from_numpy(pla_session, b)
    X = pla_session.matrix(n, 2)
    P = pla_session.mvector(n, dtype=pla.MPI_INT)
    it, be = ctypes.c_int(), ctypes.c_double()
    pla.call("PLA_Set_device_nb", 128)
    pla.call("PLA_LU_mixed", A, P, B, X, pla.PLA_MIXED_FP16, 1e-14, 40, ctypes.byref(it), ctypes.byref(be))
    pla.call("PLA_Set_device_nb", 0)
    outer, steps, gcorr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    pla.call("PLA_Mixed_stats", ctypes.byref(outer), ctypes.byref(steps), ctypes.byref(gcorr), None)
    x = pla.get_contents(X)
    pla.free(A, B, X, P)
    assert be.value <= 1e-14, (it.value, be.value, steps.value, gcorr.value)
    assert gcorr.value >= 1 and steps.value >= 1
    assert _eta(a, x, b) <= 2e-14


@pytest.mark.parametrize("n,delta,nb,masked", [(1700, 4e-2, 256, 0), (1700, 4e-2, 512, 1), (1000, 4e-4, 512, 0)])
def test_chol_fp8_side_stream_conversions_identical(pla_session, th, n, delta, nb, masked):
    """The FP8 panel's conversions of the rows below the diagonal block run
    on a side stream ordered by events (A21 before the panel GEMM, X21 after
    it, the main stream joined at the end): the solution is bit-identical to
    the same conversions on the panel stream, with either stream arrangement
    and through a shifted retry (delta = 4e-4)."""
    th.pla_th_set_fp8_side_cvt.argtypes = [ctypes.c_int]
    rng = np.random.default_rng(n + nb + masked)
    a = _spd(rng, n, delta)
    b = rng.uniform(-1, 1, (n, 2))
    out = []
    th.pla_th_set_fp8_options(nb, masked, 0)
    try:
        for side in (0, 1):
            th.pla_th_set_fp8_side_cvt(side)
            x, it, be, stats = _chol_mixed(pla_session, a, b)
            out.append((x, it, be, stats))
    finally:
        th.pla_th_set_fp8_side_cvt(1)
        th.pla_th_set_fp8_options(512, 0, 0)
    (x0, it0, be0, s0), (x1, it1, be1, s1) = out
    assert np.array_equal(x0, x1) and (it0, be0, s0) == (it1, be1, s1)
    assert be1 <= 1e-14 and _eta(a, x1, b) <= 2e-14
