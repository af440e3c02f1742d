"""Full-size runs of the configurations the benchmark is quoted on, through
the C ABI, each checked independently of the library's own refinement loop:

  plain PLA_LU, n = 32768, nb in {128, 256, 512}    (A(i,j) = drand48()*2-1, seed 0)
  plain PLA_Chol, n = 32768                          (A = B B^T / n + 1e-2 I, B from seed 0)
  PLA_LU_mixed, n = 65536: FP16x3 with 16 right-hand sides, BF16x3 with 1
  PLA_Chol_mixed (FP8 E4M3), n = 49152, delta in {4e-2, 4e-4}, 8 right-hand sides
  10,000 batched SPD systems, n_i uniform in [16, 512], 4 right-hand sides, FP64 and FP32 + refinement

Residuals are recomputed from a pristine copy of A with PLA_Gemm / PLA_Gemv
(the FP64 GEMM kernel is pinned against the oracle by test_gpu_blas.py) and
PLA_Matrix_infinity_norm, or in numpy for the batched systems.  The bars:
HPL's scaled residual ||b - A x||_inf / (eps (||A||_inf ||x||_inf + ||b||_inf) n) < 16
for the plain FP64 solves; the requested normwise backward error (1e-14,
recomputed to within 2e-14) for the refined solves.  Right-hand sides of the
refined configurations are drand48()*2-1 from srand48(1)."""
import ctypes
import time

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def _inf_norm(s, A):
    nrm = s.scalar(0.0)
    pla.call("PLA_Matrix_infinity_norm", A, nrm)
    v = float(pla.get_contents(nrm)[0, 0])
    pla.free(nrm)
    return v


def _eta_columns(s, A0, anorm, B, X):
    """max over columns j of ||b_j - A0 x_j||_inf / (||A0|| ||x_j|| + ||b_j||),
    with R = B - A0 X formed by PLA_Gemm on the pristine A0."""
    n, k = pla.dims(B)
    R = s.matrix(n, k)
    one, mone = s.scalar(1.0), s.scalar(-1.0)
    pla.call("PLA_Copy", B, R)
    pla.call("PLA_Gemm", pla.PLA_NO_TRANSPOSE, pla.PLA_NO_TRANSPOSE, mone, A0, X, one, R)
    r, x, b = pla.get_contents(R), pla.get_contents(X), pla.get_contents(B)
    pla.free(R, one, mone)
    eta = np.abs(r).max(axis=0) / (anorm * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))
    return float(eta.max()), x


# ---------------------------------------------------------------------------
# plain FP64 LU, n = 32768

@pytest.fixture(scope="module")
def lu32k(pla_session):
    s, n = pla_session, 32768
    A0 = s.matrix(n, n)
    pla.call("PLA_Obj_fill_drand48", A0, 0, 0)
    ones, b = s.mvector(n), s.mvector(n)
    one, zero = s.scalar(1.0), s.scalar(0.0)
    pla.call("PLA_Obj_set_to_one", ones)
    pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, one, A0, ones, zero, b)   # b = A * 1
    st = {"n": n, "A0": A0, "b": b, "anorm": _inf_norm(s, A0), "piv": None}
    yield st
    pla.free(A0, ones, b, one, zero)


@pytest.mark.parametrize("nb", [128, 256, 512])
def test_plain_lu_n32768(pla_session, lu32k, nb):
    s, n = pla_session, lu32k["n"]
    A, x, r = s.matrix(n, n), s.mvector(n), s.mvector(n)
    P = s.mvector(n, dtype=pla.MPI_INT)
    one, mone = s.scalar(1.0), s.scalar(-1.0)
    pla.call("PLA_Copy", lu32k["A0"], A)
    pla.call("PLA_Set_device_nb", nb)
    try:
        t = time.perf_counter()
        pla.call("PLA_LU", A, P)
        pla.call("PLA_Device_synchronize")
        _log(f"PLA_LU n={n} nb={nb}: {time.perf_counter() - t:.2f} s")
    finally:
        pla.call("PLA_Set_device_nb", 0)
    pla.call("PLA_Copy", lu32k["b"], x)
    pla.call("PLA_Apply_pivots_to_rows", x, P)
    pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_UNIT_DIAG, A, x)
    pla.call("PLA_Trsv", pla.PLA_UPPER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG, A, x)
    pla.call("PLA_Copy", lu32k["b"], r)
    pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, mone, lu32k["A0"], x, one, r)
    xv, rv, bv = (pla.get_contents(o)[:, 0] for o in (x, r, lu32k["b"]))
    piv = pla.get_contents(P)[:, 0]
    pla.free(A, x, r, P, one, mone)
    assert piv.min() >= 0 and (piv + np.arange(n)).max() < n
    eta = np.abs(rv).max() / (lu32k["anorm"] * np.abs(xv).max() + np.abs(bv).max())
    assert eta / (n * EPS) < 16, eta
    assert np.abs(xv - 1.0).max() < 1e-6
    # the pivot sequence is the algorithm's, not the block size's
    if lu32k["piv"] is None:
        lu32k["piv"] = piv
    else:
        assert np.array_equal(piv, lu32k["piv"]), int((piv != lu32k["piv"]).sum())


# ---------------------------------------------------------------------------
# plain FP64 Cholesky, n = 32768

def _spd_full(s, n, seed=0):
    """A = B B^T / n (full, both triangles) with B(i,j) = drand48()*2-1 from
    srand48(seed), formed by PLA_Gemm on the GPU."""
    Bm, A = s.matrix(n, n), s.matrix(n, n)
    pla.call("PLA_Obj_fill_drand48", Bm, seed, 0)
    al, zero = s.scalar(1.0 / n), s.scalar(0.0)
    pla.call("PLA_Gemm", pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE, al, Bm, Bm, zero, A)
    pla.free(Bm, al, zero)
    return A


def _shift(A, delta):
    v = np.array([delta])
    pla.call("PLA_Shift", A, pla.MPI_DOUBLE, v.ctypes.data_as(ctypes.c_void_p))


def test_plain_chol_n32768(pla_session):
    s, n = pla_session, 32768
    # the configuration's block size is autotuned over {128, 256, 512}
    best, ms = ctypes.c_int(), (ctypes.c_double * 4)()
    pla.call("PLA_Autotune_nb", pla.PLA_NB_CHOL, n, ctypes.byref(best), ctypes.cast(ms, ctypes.c_void_p))
    _log(f"PLA_Autotune_nb PLA_Chol n={n}: nb 128/256/512 {ms[0]:.0f}/{ms[1]:.0f}/{ms[2]:.0f} ms -> {best.value}")
    assert best.value in (128, 256, 512)
    A0 = _spd_full(s, n)
    _shift(A0, 1e-2)
    A = s.matrix(n, n)
    pla.call("PLA_Copy", A0, A)
    ones, b, x, r = s.mvector(n), s.mvector(n), s.mvector(n), s.mvector(n)
    one, zero, mone = s.scalar(1.0), s.scalar(0.0), s.scalar(-1.0)
    pla.call("PLA_Obj_set_to_one", ones)
    pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, one, A0, ones, zero, b)   # b = A * 1
    t = time.perf_counter()
    info = pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A, ok=range(0, 1 << 30))
    pla.call("PLA_Device_synchronize")
    _log(f"PLA_Chol n={n}: {time.perf_counter() - t:.2f} s")
    assert info == 0
    pla.call("PLA_Copy", b, x)
    pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG, A, x)
    pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_TRANSPOSE, pla.PLA_NONUNIT_DIAG, A, x)
    pla.call("PLA_Copy", b, r)
    pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, mone, A0, x, one, r)
    anorm = _inf_norm(s, A0)
    xv, rv, bv = (pla.get_contents(o)[:, 0] for o in (x, r, b))
    # the upper triangle is never written: spot-check a strip of it against A0
    top = s.matrix(64, n)
    top_a = s.matrix(64, n)
    for src, dst in ((A, top), (A0, top_a)):
        v = ctypes.c_void_p()
        pla.call("PLA_Obj_horz_split_2", src, 64, ctypes.byref(v), None)
        pla.call("PLA_Copy", v, dst)
        pla.free(v)
    ta, t0 = pla.get_contents(top), pla.get_contents(top_a)
    pla.free(A, A0, ones, b, x, r, one, zero, mone, top, top_a)
    assert np.array_equal(np.triu(ta, 1), np.triu(t0, 1))
    eta = np.abs(rv).max() / (anorm * np.abs(xv).max() + np.abs(bv).max())
    assert eta / (n * EPS) < 16, eta
    assert np.abs(xv - 1.0).max() < 1e-8


# ---------------------------------------------------------------------------
# mixed-precision LU + refinement, n = 65536

@pytest.mark.parametrize("prec,nrhs", [(pla.PLA_MIXED_FP16X3, 16), (pla.PLA_MIXED_BF16X3, 1)])
def test_mixed_lu_n65536(pla_session, prec, nrhs):
    s, n = pla_session, 65536
    A0 = s.matrix(n, n)
    pla.call("PLA_Obj_fill_drand48", A0, 0, 0)
    B, X = s.matrix(n, nrhs), s.matrix(n, nrhs)
    pla.call("PLA_Obj_fill_drand48", B, 1, 0)
    P = s.mvector(n, dtype=pla.MPI_INT)
    it, be = ctypes.c_int(), ctypes.c_double()
    t = time.perf_counter()
    pla.call("PLA_LU_mixed", A0, P, B, X, prec, 1e-14, 30, ctypes.byref(it), ctypes.byref(be))
    pla.call("PLA_Device_synchronize")
    outer, steps, gcorr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    pla.call("PLA_Mixed_stats", ctypes.byref(outer), ctypes.byref(steps), ctypes.byref(gcorr), None)
    _log(f"PLA_LU_mixed n={n} prec={prec} rhs={nrhs}: {time.perf_counter() - t:.2f} s, "
         f"{it.value} corrections, {steps.value} GMRES steps, backward error {be.value:.3e}")
    assert be.value <= 1e-14
    eta, _ = _eta_columns(s, A0, _inf_norm(s, A0), B, X)
    pla.free(A0, B, X, P)
    assert eta <= 2e-14, eta


# ---------------------------------------------------------------------------
# FP8 Cholesky + refinement, n = 49152

@pytest.fixture(scope="module")
def spd49k(pla_session):
    s, n = pla_session, 49152
    Abase = _spd_full(s, n)
    yield {"n": n, "Abase": Abase}
    pla.free(Abase)


@pytest.mark.parametrize("delta", [4e-2, 4e-4])
def test_fp8_chol_n49152(pla_session, spd49k, delta):
    s, n, nrhs = pla_session, spd49k["n"], 8
    A = s.matrix(n, n)
    pla.call("PLA_Copy", spd49k["Abase"], A)
    _shift(A, delta)
    B, X = s.matrix(n, nrhs), s.matrix(n, nrhs)
    pla.call("PLA_Obj_fill_drand48", B, 1, 0)
    it, be = ctypes.c_int(), ctypes.c_double()
    t = time.perf_counter()
    pla.call("PLA_Chol_mixed", A, B, X, pla.PLA_MIXED_FP8, 1e-14, 50, ctypes.byref(it), ctypes.byref(be))
    pla.call("PLA_Device_synchronize")
    outer, steps, gcorr, shift = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    pla.call("PLA_Mixed_stats", ctypes.byref(outer), ctypes.byref(steps), ctypes.byref(gcorr), ctypes.byref(shift))
    _log(f"PLA_Chol_mixed FP8 n={n} delta={delta}: {time.perf_counter() - t:.2f} s, {it.value} corrections, "
         f"{steps.value} GMRES steps, shift {shift.value:.3e}, backward error {be.value:.3e}")
    assert be.value <= 1e-14
    eta, _ = _eta_columns(s, A, _inf_norm(s, A), B, X)
    pla.free(A, B, X)
    assert eta <= 2e-14, eta


# ---------------------------------------------------------------------------
# 10,000 batched uneven SPD systems

def _batched_inputs(count, nrhs, seed=2):
    rng = np.random.default_rng(seed)
    ns = rng.integers(16, 513, count).astype(np.int32)
    a_len = ns.astype(np.int64) ** 2
    b_len = ns.astype(np.int64) * nrhs
    a_off = np.concatenate([[0], np.cumsum(a_len)[:-1]]).astype(np.int64)
    b_off = np.concatenate([[0], np.cumsum(b_len)[:-1]]).astype(np.int64)
    ha = np.empty(int(a_len.sum()))
    hb = rng.uniform(-1.0, 1.0, int(b_len.sum()))
    for i, n in enumerate(ns):
        bm = rng.uniform(-1.0, 1.0, (n, n))
        a = bm @ bm.T / n
        a[np.diag_indices(n)] += 1e-2
        ha[a_off[i]:a_off[i] + a_len[i]] = a.reshape(-1, order="F")
    return ns, a_off, b_off, ha, hb


@pytest.fixture(scope="module")
def batched10k():
    t = time.perf_counter()
    out = _batched_inputs(10000, 4)
    _log(f"batched inputs: {time.perf_counter() - t:.1f} s")
    return out


@pytest.mark.parametrize("mode", [pla.PLA_BATCHED_FP64, pla.PLA_BATCHED_FP32_IR])
def test_batched_10000_systems(pla_session, batched10k, mode):
    s, nrhs = pla_session, 4
    ns, a_off, b_off, ha, hb = batched10k
    count = len(ns)

    def dev(host, dt):
        o = s.mvector(host.size, 1, dt)
        pla.set_contents(o, host)
        p = ctypes.c_void_p()
        pla.call("PLA_Obj_device_buffer", o, ctypes.byref(p))
        return o, p

    A, pa = dev(ha, pla.MPI_DOUBLE)
    B, pb = dev(hb, pla.MPI_DOUBLE)
    info, pi = dev(np.full(count, -99, np.int32), pla.MPI_INT)
    iters, pit = dev(np.zeros(count, np.int32), pla.MPI_INT)
    berr, pbe = dev(np.zeros(count), pla.MPI_DOUBLE)
    plan = ctypes.c_void_p()
    pla.call("PLA_Batched_plan_create", count, ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
             a_off.ctypes.data_as(ctypes.c_void_p), b_off.ctypes.data_as(ctypes.c_void_p), nrhs,
             mode | pla.PLA_BATCHED_AUTOTUNE, 1e-14, 10, pa, pb, pi, pit, pbe, ctypes.byref(plan))
    nbv, wpc = ctypes.c_int(), ctypes.c_int()
    pla.call("PLA_Batched_plan_config", plan, ctypes.byref(nbv), ctypes.byref(wpc))
    pla.set_contents(A, ha)         # autotuning ran on scratch copies; start from pristine inputs
    pla.set_contents(B, hb)
    t = time.perf_counter()
    pla.call("PLA_Batched_plan_execute", plan)
    pla.call("PLA_Device_synchronize")
    _log(f"batched mode {mode}: {count} systems in {1e3 * (time.perf_counter() - t):.1f} ms "
         f"(nb {nbv.value}, {wpc.value} workgroups per CU)")
    pla.call("PLA_Batched_plan_free", ctypes.byref(plan))
    inf = pla.get_contents(info)[:, 0]
    xs = pla.get_contents(B)[:, 0]
    be = pla.get_contents(berr)[:, 0]
    a_after = pla.get_contents(A)[:, 0] if mode == pla.PLA_BATCHED_FP32_IR else None
    pla.free(A, B, info, iters, berr)
    assert (inf == 0).all()
    if a_after is not None:
        assert np.array_equal(a_after, ha), "FP32_IR mode must leave A unchanged"
    worst = 0.0
    for i, n in enumerate(ns):
        a = ha[a_off[i]:a_off[i] + n * n].reshape(n, n, order="F")
        b = hb[b_off[i]:b_off[i] + n * nrhs].reshape(n, nrhs, order="F")
        x = xs[b_off[i]:b_off[i] + n * nrhs].reshape(n, nrhs, order="F")
        r = b - a @ x
        eta = np.abs(r).max(axis=0) / (np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))
        if mode == pla.PLA_BATCHED_FP64:
            worst = max(worst, float(eta.max() / (n * EPS)))
        else:
            worst = max(worst, float(eta.max()))
    if mode == pla.PLA_BATCHED_FP64:
        assert worst < 16, worst                     # HPL's scaled residual, every system
    else:
        assert worst <= 2e-14, worst                 # the requested backward error, every system
        assert be.max() <= 1e-14
