"""GPU parity of PLA_LU (SOLVERS/LU/PLA_LU.c:13) and PLA_General_solve
(SOLVERS/LU/PLA_General.c:67) through the C ABI.

Bar: pivot vectors equal (integers, bit-exact) to the oracle's and, for the
reference's own n=2048 seed-0 case, to the committed LAPACK golden vector;
factors within the reference's sequential-check tolerance 1e-7 * max|A|
(SOLVERS/LU/PLA_LU_enter_exit.c:130-153) of the oracle; scaled residual
||PA - LU||_1 / (n ||A||_1 eps) < 16.  At n=16384 (too large for the CPU
oracle) the size-independent checks are the solve of A x = A*1 and its
normwise backward error."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _lu_gpu(s, a=None, n=None, m=None, seed=0, nb=0):
    """Run PLA_LU on a (host array) or on the drand48 seed matrix of size m x n."""
    pla.call("PLA_Set_device_nb", nb)
    if a is not None:
        A = pla.from_numpy(s, a)
        m, n = a.shape
    else:
        m = m or n
        A = s.matrix(m, n)
        pla.call("PLA_Obj_fill_drand48", A, seed, 0)
    P = s.mvector(min(m, n), dtype=pla.MPI_INT)
    pla.call("PLA_LU", A, P)
    lu, piv = pla.get_contents(A), pla.get_contents(P)[:, 0]
    pla.free(A, P)
    pla.call("PLA_Set_device_nb", 0)
    return lu, piv


def test_lu_golden_n2048_seed0(pla_session, ora):
    """The reference's CPU-runnable case: n=2048, nb=128, A from drand48 seed 0."""
    golden = np.load(os.path.join(GOLDEN, "lu_piv_n2048_seed0.npy"))
    lu, piv = _lu_gpu(pla_session, n=2048, seed=0, nb=128)
    assert np.array_equal(piv, golden)
    a = ora.uniform(2048, 2048, 0)
    assert ora.lu_scaled_residual(a, lu, piv) < 16
    lu_o, piv_o = ora.lu_blocked(a, nb=128)
    assert np.array_equal(piv, piv_o)
    assert np.abs(lu - lu_o).max() <= 1e-7 * np.abs(a).max()


@pytest.mark.parametrize("m,n,nb", [(1, 1, 128), (5, 5, 128), (100, 37, 32), (37, 100, 32),
                                    (300, 300, 64), (1000, 1000, 128), (1000, 1000, 256),
                                    (2049, 2049, 128), (1500, 700, 256), (700, 1500, 512),
                                    (3000, 3000, 1024), (2500, 1100, 1024)])
def test_lu_shapes_vs_oracle(pla_session, ora, m, n, nb):
    a = ora.uniform(m, n, seed=m + 3 * n)
    lu, piv = _lu_gpu(pla_session, a=a, nb=nb)
    lu_o, piv_o = ora.lu_blocked(a, nb=nb)
    assert np.array_equal(piv, piv_o)
    assert np.abs(lu - lu_o).max() <= 1e-7 * np.abs(a).max()
    assert ora.lu_scaled_residual(a, lu, piv) < 16


@pytest.mark.parametrize("n,nb", [(300, 64), (1000, 128), (2049, 256)])
def test_lu_linpack_appended_system(pla_session, ora, n, nb):
    """PLA_LU_linpack on [A b] (SOLVERS/LINPACK/driver.c:108-157): pivots and
    factors against the oracle's LINPACK variant, then PLA_Trsv with U on the
    appended column solves A x = b."""
    s = pla_session
    ab = ora.uniform(n, n + 1, seed=n)
    pla.call("PLA_Set_device_nb", nb)
    A = pla.from_numpy(s, ab)
    P = s.mvector(n, dtype=pla.MPI_INT)
    pla.call("PLA_LU_linpack", A, P)
    pla.call("PLA_Set_device_nb", 0)
    f, piv = pla.get_contents(A), pla.get_contents(P)[:, 0]
    f_o, piv_o = ora.lu_linpack(ab, nb=nb)
    assert np.array_equal(piv, piv_o)
    assert np.abs(f - f_o).max() <= 1e-7 * np.abs(ab).max()
    # back substitution on the appended column, as the LINPACK driver does
    tl, tr, bl, br = (ctypes.c_void_p() for _ in range(4))
    pla.call("PLA_Obj_split_4", A, n, n, ctypes.byref(tl), ctypes.byref(tr), ctypes.byref(bl), ctypes.byref(br))
    pla.call("PLA_Trsv", pla.PLA_UPPER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG, tl, tr)
    x = pla.get_contents(tr)[:, 0]
    a, b = ab[:, :n], ab[:, n]
    eta = np.abs(a @ x - b).max() / (np.abs(a).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    assert eta < 64 * n * np.finfo(np.float64).eps
    pla.free(tl, tr, bl, br, A, P)


def test_lu_zero_column_not_flagged(pla_session, ora):
    """Like the reference, a zero pivot is not detected: PLA_LU returns 0."""
    a = ora.uniform(64, 64, 2)
    a[:, 10] = 0.0
    lu, piv = _lu_gpu(pla_session, a=a, nb=32)
    _, piv_o = ora.lu_blocked(a, nb=32)
    assert np.array_equal(piv[:11], piv_o[:11])
    assert piv[10] == 0  # all-zero column: first index, no interchange


def test_lu_argument_errors(pla_session):
    s = pla_session
    A = s.matrix(10, 10)
    P = s.mvector(9, dtype=pla.MPI_INT)
    rc = pla.lib().PLA_LU(A, P)
    assert rc < 0 and "pivots" in pla.last_error()
    Pd = s.mvector(10, dtype=pla.MPI_DOUBLE)
    assert pla.lib().PLA_LU(A, Pd) < 0
    pla.free(A, P, Pd)


def test_general_solve_multi_rhs(pla_session, ora):
    s = pla_session
    n = 777
    a = ora.uniform(n, n, 4)
    x = np.random.default_rng(0).uniform(-1, 1, (n, 16))
    b = a @ x
    A, B = pla.from_numpy(s, a), pla.from_numpy(s, b)
    P = s.mvector(n, dtype=pla.MPI_INT)
    pla.call("PLA_General_solve", A, P, B)
    got = pla.get_contents(B)
    lu_o, piv_o = ora.lu_blocked(a, nb=128)
    assert np.array_equal(pla.get_contents(P)[:, 0], piv_o)
    assert np.abs(got - x).max() < 1e-9
    pla.free(A, B, P)


@pytest.mark.parametrize("n,nb,repeat", [(16384, 256, 1), (65536, 0, 2)])
def test_lu_solve_large_backward_error(pla_session, n, nb, repeat):
    """Full-path property at sizes the CPU oracle cannot finish quickly.  At
    n=65536 (the library's block size, all streams busy at once) the
    factorization is repeated: the pivots must not change between runs (a
    stream-ordering race shows up as different pivots and a failed solve)."""
    s = pla_session
    A = s.matrix(n, n)
    A0 = s.matrix(n, n)
    pla.call("PLA_Obj_fill_drand48", A, 0, 0)
    pla.call("PLA_Obj_fill_drand48", A0, 0, 0)
    ones, b, x = s.mvector(n), s.mvector(n), s.mvector(n)
    pla.call("PLA_Obj_set_to_one", ones)
    one, zero, mone = s.scalar(1.0), s.scalar(0.0), s.scalar(-1.0)
    pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, one, A0, ones, zero, b)   # b = A*1
    pla.call("PLA_Copy", b, x)
    P = s.mvector(n, dtype=pla.MPI_INT)
    pla.call("PLA_Set_device_nb", nb)
    first = None
    for _ in range(repeat):
        pla.call("PLA_Copy", A0, A)
        pla.call("PLA_LU", A, P)
        piv = pla.get_contents(P)[:, 0]
        if first is None:
            first = piv.copy()
        assert np.array_equal(piv, first)
    pla.call("PLA_Set_device_nb", 0)
    pla.call("PLA_Apply_pivots_to_rows", x, P)
    pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_UNIT_DIAG, A, x)
    pla.call("PLA_Trsv", pla.PLA_UPPER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG, A, x)
    xs = pla.get_contents(x)[:, 0]
    r = s.mvector(n)
    pla.call("PLA_Copy", b, r)
    pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, mone, A0, x, one, r)      # r = b - A x
    rv, bv = pla.get_contents(r)[:, 0], pla.get_contents(b)[:, 0]
    piv = pla.get_contents(P)[:, 0]
    assert piv.min() >= 0 and (piv + np.arange(n)).max() < n
    nrm = s.scalar(0.0)
    pla.call("PLA_Matrix_infinity_norm", A0, nrm)
    anorm = pla.get_contents(nrm)[0, 0]
    assert n / 2 - 4 * np.sqrt(n) < anorm < n / 2 + 8 * np.sqrt(n)
    # HPL's acceptance test: ||Ax-b|| / (eps (||A|| ||x|| + ||b||) n) < 16
    eta = np.abs(rv).max() / (anorm * np.abs(xs).max() + np.abs(bv).max())
    assert eta / (n * np.finfo(np.float64).eps) < 16
    assert np.abs(xs - 1.0).max() < 1e-6
    pla.free(A, A0, ones, b, x, r, P, one, zero, mone, nrm)


def _driver_exe(name="test_lu_driver"):
    """Compile tests/drivers/<name>.c with mpicc against include/PLA.h and the library."""
    mpicc = shutil.which("mpicc") or "/opt/conda/bin/mpicc"
    if not os.path.exists(mpicc):
        pytest.skip("mpicc not available")
    out = os.path.join(ROOT, "tests", "_build")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(out, name)
    libdir = os.path.join(ROOT, "plapack_amd")
    subprocess.run([mpicc, "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "drivers", name + ".c"), "-o", exe, "-lm",
                    # the system libstdc++ (needed by the HIP runtime) ahead of MPICH's own
                    "-L/usr/lib/x86_64-linux-gnu", "-Wl,-rpath-link,/usr/lib/x86_64-linux-gnu",
                    "-L", libdir, "-lpla_mi355x", "-Wl,-rpath," + libdir,
                    "-Wl,-rpath,/usr/lib/x86_64-linux-gnu"], check=True,
                   env=dict(os.environ, MPICH_CC="gcc"))
    return exe


def test_reference_style_c_driver(tmp_path, ora):
    """A TEST/test_lu.c-style C program, host-filled through PLA_Obj_local_buffer."""
    exe = _driver_exe()
    pivfile = str(tmp_path / "piv.bin")
    stdin = "1 1\n128\n128\n1\n1\n0\n1 2048\n-1\n"
    out = subprocess.run([exe, pivfile], input=stdin, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "MFLOPS/node" in out.stdout
    piv = np.fromfile(pivfile, dtype=np.int32)
    assert np.array_equal(piv, np.load(os.path.join(GOLDEN, "lu_piv_n2048_seed0.npy")))
    # TEST/test_lu.c's double-complex datatype (itype 3), n = 1000
    out = subprocess.run([exe, pivfile], input="1 1\n128\n128\n1\n1\n0\n3 1000\n-1\n", capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    _, piv_z = ora.zlu_blocked(ora.uniform_z(1000, 1000, 0), nb=128)  # nb_alg 128 from stdin
    assert np.array_equal(np.fromfile(pivfile, dtype=np.int32), piv_z)
    # a 2 x 1 mesh is refused with an error code
    out = subprocess.run([exe], input="2 1\n128\n128\n1\n1\n0\n-1\n", capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 3 and "1x1" in out.stdout


def test_attach_device_buffer(pla_session, ora):
    """PLA_Obj_attach_buffer (OBJ/PLA_Obj_management.c:15) with a device buffer:
    PLA_LU on an object whose storage is the top-left block of another."""
    s = pla_session
    big = ora.uniform(300, 260, seed=9)
    X = pla.from_numpy(s, big)
    ptr = ctypes.c_void_p()
    pla.call("PLA_Obj_device_buffer", X, ctypes.byref(ptr))
    Y = s.matrix(200, 200)
    P = s.mvector(200, dtype=pla.MPI_INT)
    pla.call("PLA_Obj_attach_buffer", Y, ptr, 300, 1)
    pla.call("PLA_LU", Y, P)
    lu_ref, piv_ref = ora.lu_blocked(big[:200, :200], nb=128)
    out = pla.get_contents(X)
    assert np.array_equal(pla.get_contents(P)[:, 0], piv_ref)
    assert np.allclose(out[:200, :200], lu_ref, atol=1e-10)
    assert np.array_equal(out[200:, :], big[200:, :]) and np.array_equal(out[:, 200:], big[:, 200:])
    pla.free(Y, P)                      # the attached buffer stays X's
    assert np.allclose(pla.get_contents(X)[:200, :200], lu_ref, atol=1e-10)
    host = np.zeros(10)
    Z = s.matrix(10, 1)
    assert pla.lib().PLA_Obj_attach_buffer(Z, host.ctypes.data_as(ctypes.c_void_p), 10, 1) < 0
    assert pla.lib().PLA_Obj_attach_buffer(Z, ptr, 5, 1) < 0     # ldim < rows
    pla.free(X, Z)


@pytest.mark.parametrize("kind,cus,width", [(0, 0, 0), (0, 0, 8), (0, 0, 16), (0, 0, 32), (0, 32, 0),
                                            (1, 0, 0), (1, 32, 0), (1, 32, 8), (1, 32, 16), (1, 64, 32)])
@pytest.mark.parametrize("m,n,nb", [(3000, 3000, 256), (6000, 700, 512), (20000, 100, 64)])
def test_lu_leaf_kinds_widths_and_panel_cus(pla_session, ora, kind, cus, width, m, n, nb):
    """Both panel leaves (rows in registers, rows in LDS), every leaf width and
    reserved panel CUs (CU-masked panel / trailing streams) give the oracle's
    pivots and factors."""
    a = ora.uniform(m, n, seed=7)
    pla.call("PLA_Set_device_param", pla.PLA_PARAM_LEAF_KIND, kind)
    pla.call("PLA_Set_device_param", pla.PLA_PARAM_PANEL_CUS, cus)
    pla.call("PLA_Set_device_param", pla.PLA_PARAM_LEAF_WIDTH, width)
    try:
        lu, piv = _lu_gpu(pla_session, a=a, nb=nb)
    finally:
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_LEAF_KIND, 0)
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_PANEL_CUS, 0)
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_LEAF_WIDTH, 0)
    lu_o, piv_o = ora.lu_blocked(a, nb=nb)
    assert np.array_equal(piv, piv_o)
    assert np.abs(lu - lu_o).max() <= 1e-7 * np.abs(a).max()


def test_panel_param_errors(pla_session):
    for cus in (-32, 8, 48, 100000):
        with pytest.raises(pla.PLAError):
            pla.call("PLA_Set_device_param", pla.PLA_PARAM_PANEL_CUS, cus)
    for w in (-1, 4, 12, 64):
        with pytest.raises(pla.PLAError):
            pla.call("PLA_Set_device_param", pla.PLA_PARAM_LEAF_WIDTH, w)
    for k in (-1, 2):
        with pytest.raises(pla.PLAError):
            pla.call("PLA_Set_device_param", pla.PLA_PARAM_LEAF_KIND, k)
    with pytest.raises(pla.PLAError):
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_TAIL, -1)
    with pytest.raises(pla.PLAError):
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_SWITCH, -1)


@pytest.mark.parametrize("kind", [0, 1])
def test_lu_exchange_timeout_falls_back(pla_session, ora, kind):
    """A panel exchange that gives up (forced from the host with a one-poll
    wait limit) abandons the leaf launch; workgroup 0 then factors the leaf
    alone from the untouched panel, and PLA_LU completes with the same
    pivots and factors, bit for bit, as with the exchange (every leaf uses
    the same arithmetic in the same order).  Some leaves commit anyway (all
    records in time), so both ends of a launch are exercised; the fallback
    count proves the fallback ran.  The mixed LU uses the same leaves."""
    hooks = ctypes.CDLL(os.path.join(os.path.dirname(pla.LIB_PATH), "libpla_testhooks.so"))
    hooks.pla_th_set_leaf_spin_limit.argtypes = [ctypes.c_uint]
    hooks.pla_th_leaf_fallbacks.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
    s = pla_session
    n = 3000
    fb = ctypes.c_ulonglong()
    pla.call("PLA_Set_device_param", pla.PLA_PARAM_LEAF_KIND, kind)
    try:
        lu_ref, piv_ref = _lu_gpu(s, n=n, seed=7, nb=256)
        hooks.pla_th_leaf_fallbacks(ctypes.byref(fb))
        assert fb.value == 0
        hooks.pla_th_set_leaf_spin_limit(1)
        lu, piv = _lu_gpu(s, n=n, seed=7, nb=256)
        hooks.pla_th_leaf_fallbacks(ctypes.byref(fb))
        assert fb.value > 0
        assert np.array_equal(piv, piv_ref)
        assert np.array_equal(lu, lu_ref)
        # a ragged panel and narrow leaves, against the oracle
        a = ora.uniform(1100, 900, 8)
        lu2, piv2 = _lu_gpu(s, a=a, nb=96)
        lu_o, piv_o = ora.lu_blocked(a, nb=96)
        assert np.array_equal(piv2, piv_o)
        assert np.abs(lu2 - lu_o).max() <= 1e-7 * np.abs(a).max()
        # the mixed LU (FP64 panels on the same leaves) reaches FP64 accuracy
        b = np.random.default_rng(8).uniform(-1, 1, (2000, 2))
        am = ora.uniform(2000, 2000, 9)
        A, B = pla.from_numpy(s, am), pla.from_numpy(s, b)
        P, X = s.mvector(2000, dtype=pla.MPI_INT), s.matrix(2000, 2)
        it, be = ctypes.c_int(), ctypes.c_double()
        pla.call("PLA_LU_mixed", A, P, B, X, pla.PLA_MIXED_FP16X3, 1e-14, 30, ctypes.byref(it), ctypes.byref(be))
        x = pla.get_contents(X)
        pla.free(A, B, P, X)
        r = b - am @ x
        eta = (np.abs(r).max(axis=0) / (np.abs(am).sum(axis=1).max() * np.abs(x).max(axis=0)
                                         + np.abs(b).max(axis=0))).max()
        assert eta <= 2e-14
        hooks.pla_th_leaf_fallbacks(ctypes.byref(fb))
        assert fb.value > 0
    finally:
        hooks.pla_th_set_leaf_spin_limit(0)
        pla.call("PLA_Set_device_nb", 0)
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_LEAF_KIND, 0)
    lu, piv = _lu_gpu(s, n=1000, seed=3, nb=128)
    lu_o, piv_o = ora.lu_blocked(ora.uniform(1000, 1000, 3), nb=128)
    assert np.array_equal(piv, piv_o)


@pytest.mark.parametrize("entry", ["PLA_Trsv", "PLA_General_solve", "PLA_Pos_def_solve"])
def test_triangular_solve_timeout_reports_error(pla_session, ora, entry):
    """A triangular-solve block wait that gives up (forced from the host with
    a poll limit of zero: every block that has to wait for a preceding block
    gives up at its first wait, whatever the timing) makes the entry point
    return PLA_ERR_DEVICE instead of a silently wrong solution; the error word
    is cleared, so the next solve and the next LU are correct."""
    hooks = ctypes.CDLL(os.path.join(os.path.dirname(pla.LIB_PATH), "libpla_testhooks.so"))
    hooks.pla_th_trsv_give_up_at_first_wait.argtypes = [ctypes.c_int]
    s = pla_session
    n = 8192
    a = ora.uniform(n, n, 6)
    spd = a @ a.T / n + np.eye(n)
    b = np.random.default_rng(6).uniform(-1, 1, (n, 1))

    def run():
        if entry == "PLA_Trsv":
            A, X = pla.from_numpy(s, np.tril(a) + n * np.eye(n)), pla.from_numpy(s, b)
            args = (pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG, A, X)
            objs, ref = (A, X), np.linalg.solve(np.tril(a) + n * np.eye(n), b)
        elif entry == "PLA_General_solve":
            A, X, P = pla.from_numpy(s, a), pla.from_numpy(s, b), s.mvector(n, dtype=pla.MPI_INT)
            args, objs, ref = (A, P, X), (A, X, P), np.linalg.solve(a, b)
        else:
            A, X = pla.from_numpy(s, spd), pla.from_numpy(s, b)
            args, objs, ref = (A, X), (A, X), np.linalg.solve(spd, b)
        try:
            pla.call(entry, *args)
            return pla.get_contents(X), ref
        finally:
            pla.free(*objs)

    hooks.pla_th_trsv_give_up_at_first_wait(1)
    try:
        with pytest.raises(pla.PLAError) as e:
            run()
        assert e.value.code == pla.PLA_ERR_DEVICE and "timed out" in str(e.value)
    finally:
        hooks.pla_th_trsv_give_up_at_first_wait(0)
    x, ref = run()
    assert np.abs(x - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())
    lu, piv = _lu_gpu(s, n=1000, seed=3, nb=128)
    _, piv_o = ora.lu_blocked(ora.uniform(1000, 1000, 3), nb=128)
    assert np.array_equal(piv, piv_o)


@pytest.mark.parametrize("nrhs", [5, 8, 13, 16])
def test_solve_column_groups_bit_identical(pla_session, ora, nrhs):
    """The untransposed triangular solves split the right-hand sides into
    independent chains of at most 4 columns (the default); 0 keeps every
    column in one chain.  Each column's arithmetic is the same in every
    grouping, so PLA_General_solve (unit lower, then upper) and
    PLA_Pos_def_solve (lower, then lower transposed) give bit-identical
    solutions for every group width, ragged last group and 64-row block
    (n = 1500) included, and they solve the systems."""
    hooks = ctypes.CDLL(os.path.join(os.path.dirname(pla.LIB_PATH), "libpla_testhooks.so"))
    hooks.pla_th_set_trsv_group_cols.argtypes = [ctypes.c_int]
    s = pla_session
    n = 1500
    a = ora.uniform(n, n, 8)
    spd = a @ a.T / n + np.eye(n)
    x = np.random.default_rng(nrhs).uniform(-1, 1, (n, nrhs))
    outs = {}
    try:
        for gc in (0, 1, 2, 4, 8, 16):
            assert hooks.pla_th_set_trsv_group_cols(gc) == pla.PLA_SUCCESS
            A, B, P = pla.from_numpy(s, a), pla.from_numpy(s, a @ x), s.mvector(n, dtype=pla.MPI_INT)
            pla.call("PLA_General_solve", A, P, B)
            xl = pla.get_contents(B)
            pla.free(A, B, P)
            A, B = pla.from_numpy(s, spd), pla.from_numpy(s, spd @ x)
            pla.call("PLA_Pos_def_solve", A, B)
            outs[gc] = (xl, pla.get_contents(B))
            pla.free(A, B)
    finally:
        hooks.pla_th_set_trsv_group_cols(4)
    assert hooks.pla_th_set_trsv_group_cols(3) == pla.PLA_ERR_INVALID_PARAM
    for gc, (xl, xc) in outs.items():
        assert np.array_equal(xl, outs[0][0]), f"LU solve, {gc} columns per chain"
        assert np.array_equal(xc, outs[0][1]), f"Cholesky solve, {gc} columns per chain"
    assert np.abs(outs[4][0] - x).max() < 1e-9
    assert np.abs(outs[4][1] - x).max() < 1e-9


@pytest.mark.parametrize("m,n,nb,tail,switch", [(3000, 3000, 512, 128, 1500), (2500, 1100, 1024, 256, 2000),
                                                (2000, 2000, 256, 64, 100000)])
def test_lu_block_size_schedule(pla_session, ora, m, n, nb, tail, switch):
    """A smaller block size for the last steps (PLA_PARAM_NB_TAIL once at most
    PLA_PARAM_NB_SWITCH rows remain) gives the oracle's pivots and factors."""
    a = ora.uniform(m, n, seed=11)
    pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_TAIL, tail)
    pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_SWITCH, switch)
    try:
        lu, piv = _lu_gpu(pla_session, a=a, nb=nb)
    finally:
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_TAIL, 0)
        pla.call("PLA_Set_device_param", pla.PLA_PARAM_NB_SWITCH, 0)
    lu_o, piv_o = ora.lu_blocked(a, nb=nb)
    assert np.array_equal(piv, piv_o)
    assert np.abs(lu - lu_o).max() <= 1e-7 * np.abs(a).max()
    assert ora.lu_scaled_residual(a, lu, piv) < 16


@pytest.mark.parametrize("fact,n", [(pla.PLA_NB_LU, 1777), (pla.PLA_NB_CHOL, 1901)])
def test_autotune_nb(pla_session, ora, fact, n):
    """PLA_Autotune_nb times the factorization for each candidate block size
    on a scratch matrix, and the winner becomes the block size the device
    uses for that order; the factorization at that order is then still the
    oracle's (pivots, Cholesky values)."""
    s = pla_session
    best = ctypes.c_int()
    ms = (ctypes.c_double * 4)()
    pla.call("PLA_Autotune_nb", fact, n, ctypes.byref(best), ctypes.cast(ms, ctypes.c_void_p))
    assert best.value in (128, 256, 512)
    assert all(ms[i] > 0 for i in range(3)) and ms[3] == 0.0
    assert ms[[128, 256, 512].index(best.value)] == min(ms[:3])
    nb = ctypes.c_int()
    pla.call("PLA_Get_device_nb_for", n, fact, ctypes.byref(nb))
    assert nb.value == best.value
    a = ora.uniform(n, n, 11)
    if fact == pla.PLA_NB_LU:
        lu, piv = _lu_gpu(s, a=a)
        lu_o, piv_o = ora.lu_blocked(a, nb=best.value)
        assert np.array_equal(piv, piv_o)
    else:
        spd = np.tril(a) + n * np.eye(n)
        A = pla.from_numpy(s, spd)
        assert pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A, ok=(0,)) == 0
        L = np.tril(pla.get_contents(A))
        pla.free(A)
        sym = np.tril(spd) + np.tril(spd, -1).T
        assert np.abs(L - np.linalg.cholesky(sym)).max() <= 1e-13 * np.abs(sym).max()
    with pytest.raises(pla.PLAError):
        pla.call("PLA_Autotune_nb", pla.PLA_NB_LU_MIXED, n, ctypes.byref(best), None)
