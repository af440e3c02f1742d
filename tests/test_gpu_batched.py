"""GPU tests of the batched uneven SPD factor-and-solve (PLA_Batched_plan_*,
PLA_Chol_batched): one persistent kernel replayed from a hipGraph.

Each system is checked against the CPU oracle's Cholesky and solve
(oracle/pla_oracle.c, the restatement of SOLVERS/Cholesky/PLA_Chol.c and
PLA_Pos_def.c), with the failure index of a non-SPD system equal to the
oracle's.  The FP32-factor mode must reach the requested FP64 backward error,
recomputed here in numpy.  Graph replays are checked one by one against the
first result."""
import ctypes

import numpy as np
import pytest

from plapack_amd import pla

pytestmark = pytest.mark.gpu

EDGE_SIZES = [1, 2, 3, 15, 16, 17, 31, 32, 33, 47, 63, 64, 65, 100, 127, 128, 129, 200, 255, 256,
              257, 300, 383, 511, 512]


def _spd(rng, n, delta=1e-2):
    b = rng.uniform(-1, 1, (n, n))
    return b @ b.T / n + delta * np.eye(n)


class _Dev:
    """A 1-column library object used as a raw device buffer."""

    def __init__(self, s, host, dt):
        host = np.ascontiguousarray(host)
        self.n = max(host.size, 1)
        self.obj = s.mvector(self.n, 1, dt)
        ptr = ctypes.c_void_p()
        pla.call("PLA_Obj_device_buffer", self.obj, ctypes.byref(ptr))
        self.ptr = ptr
        if host.size:
            pla.set_contents(self.obj, host.reshape(-1))

    def get(self):
        return pla.get_contents(self.obj)[:, 0]

    def free(self):
        pla.free(self.obj)


def _pack(systems, rhs, gap=0):
    """Column-major packing with `gap` unused elements after each block."""
    a_off, b_off, a_parts, b_parts = [], [], [], []
    ao = bo = 0
    for a, b in zip(systems, rhs):
        a_off.append(ao)
        b_off.append(bo)
        a_parts += [a.reshape(-1, order="F"), np.full(gap, 7.0)]
        b_parts += [b.reshape(-1, order="F"), np.full(gap, 7.0)]
        ao += a.size + gap
        bo += b.size + gap
    return (np.concatenate(a_parts), np.concatenate(b_parts), np.array(a_off, np.int64),
            np.array(b_off, np.int64))


def _run(s, systems, rhs, mode, tol=1e-14, max_iter=10, gap=0, offsets=False, replays=1):
    count, nrhs = len(systems), rhs[0].shape[1]
    n = np.array([a.shape[0] for a in systems], np.int32)
    ha, hb, a_off, b_off = _pack(systems, rhs, gap)
    A, B = _Dev(s, ha, pla.MPI_DOUBLE), _Dev(s, hb, pla.MPI_DOUBLE)
    info = _Dev(s, np.full(count, -99, np.int32), pla.MPI_INT)
    iters = _Dev(s, np.zeros(count, np.int32), pla.MPI_INT)
    berr = _Dev(s, np.zeros(count), pla.MPI_DOUBLE)
    plan = ctypes.c_void_p()
    ao = a_off.ctypes.data_as(ctypes.c_void_p) if offsets or gap else None
    bo = b_off.ctypes.data_as(ctypes.c_void_p) if offsets or gap else None
    pla.call("PLA_Batched_plan_create", count, n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ao, bo,
             nrhs, mode, tol, max_iter, A.ptr, B.ptr, info.ptr, iters.ptr, berr.ptr, ctypes.byref(plan))
    outs = []
    for _ in range(replays):
        pla.set_contents(A.obj, ha)
        pla.set_contents(B.obj, hb)
        pla.call("PLA_Batched_plan_execute", plan)
        pla.call("PLA_Device_synchronize")
        outs.append((A.get(), B.get(), info.get().copy(), iters.get().copy(), berr.get().copy()))
    nb, wpc = ctypes.c_int(), ctypes.c_int()
    pla.call("PLA_Batched_plan_config", plan, ctypes.byref(nb), ctypes.byref(wpc))
    pla.call("PLA_Batched_plan_free", ctypes.byref(plan))
    for d in (A, B, info, iters, berr):
        d.free()
    res = []
    for ra, rb, rinfo, rit, rbe in outs:
        per = []
        for i, (a, b) in enumerate(zip(systems, rhs)):
            k = a.shape[0]
            per.append((ra[a_off[i]:a_off[i] + k * k].reshape(k, k, order="F"),
                        rb[b_off[i]:b_off[i] + k * nrhs].reshape(k, nrhs, order="F")))
        res.append((per, rinfo, rit, rbe))
    return res, (nb.value, wpc.value), (ha, hb, a_off, b_off)


def _eta(a, x, b):
    an = np.abs(a).sum(axis=1).max()
    r = b - a @ x
    return max(np.abs(r[:, j]).max() / (an * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(b.shape[1]))


def _systems(seed, sizes, nrhs):
    rng = np.random.default_rng(seed)
    systems = [_spd(rng, k) for k in sizes]
    rhs = [rng.uniform(-1, 1, (k, nrhs)) for k in sizes]
    return systems, rhs


@pytest.mark.parametrize("nrhs", [1, 4, 8])
def test_fp64_matches_oracle(pla_session, ora, nrhs):
    sizes = EDGE_SIZES + list(np.random.default_rng(nrhs).integers(16, 513, 15))
    systems, rhs = _systems(nrhs, sizes, nrhs)
    (res,), cfg, _ = _run(pla_session, systems, rhs, pla.PLA_BATCHED_FP64)
    per, info, _, _ = res
    assert (info == 0).all()
    for (a, b), (l_out, x_out) in zip(zip(systems, rhs), per):
        l_ref, rc = ora.chol_unblocked(a)
        assert rc == 0
        k = a.shape[0]
        lo = np.tril_indices(k)
        up = np.triu_indices(k, 1)
        assert np.allclose(l_out[lo], l_ref[lo], rtol=0, atol=1e-12 * np.abs(l_ref).max()), k
        assert np.array_equal(l_out[up], a[up]), "upper triangle must not be touched"
        x_ref = ora.chol_solve(l_ref, b)
        assert np.abs(x_out - x_ref).max() <= 1e-10 * max(np.abs(x_ref).max(), 1), k
        assert _eta(a, x_out, b) <= 1e-14, k


@pytest.mark.parametrize("nrhs", [1, 4])
def test_fp32_refinement_reaches_fp64_backward_error(pla_session, nrhs):
    sizes = EDGE_SIZES + list(np.random.default_rng(10 + nrhs).integers(16, 513, 15))
    systems, rhs = _systems(10 + nrhs, sizes, nrhs)
    (res,), cfg, _ = _run(pla_session, systems, rhs, pla.PLA_BATCHED_FP32_IR, tol=1e-14)
    per, info, iters, berr = res
    assert (info == 0).all()
    for i, ((a, b), (a_out, x_out)) in enumerate(zip(zip(systems, rhs), per)):
        assert np.array_equal(a_out, a), "FP32_IR leaves A unchanged"
        eta = _eta(a, x_out, b)
        assert berr[i] <= 1e-14 and eta <= 2e-14, (a.shape[0], berr[i], eta)
        assert 1 <= iters[i] <= 6, iters[i]
        assert abs(berr[i] - eta) <= 0.5 * eta + 2e-16  # both are FP64 residuals, summed in different orders


def test_failure_index_and_isolation(pla_session, ora):
    rng = np.random.default_rng(5)
    sizes = [40, 130, 7, 300, 64]
    systems = [_spd(rng, k) for k in sizes]
    systems[1][77, 77] = -5.0          # first non-positive pivot inside block 2
    systems[2][0, 0] = 0.0             # fails at the first pivot
    systems[3][299, 299] = -1.0         # fails at the last pivot
    rhs = [rng.uniform(-1, 1, (k, 2)) for k in sizes]
    for mode in (pla.PLA_BATCHED_FP64, pla.PLA_BATCHED_FP32_IR):
        (res,), _, _ = _run(pla_session, systems, rhs, mode)
        per, info, _, _ = res
        for i, a in enumerate(systems):
            _, rc = ora.chol_unblocked(a)
            assert info[i] == rc, (mode, i, info[i], rc)
            if rc:
                assert np.array_equal(per[i][1], rhs[i]), "a failed system keeps its B"
            else:
                assert _eta(a, per[i][1], rhs[i]) <= 1e-14


def test_caller_offsets_with_gaps(pla_session):
    sizes = [33, 5, 250, 128]
    systems, rhs = _systems(21, sizes, 3)
    (res,), _, (ha, hb, a_off, b_off) = _run(pla_session, systems, rhs, pla.PLA_BATCHED_FP64, gap=5)
    per, info, _, _ = res
    assert (info == 0).all()
    for (a, b), (_, x) in zip(zip(systems, rhs), per):
        assert _eta(a, x, b) <= 1e-14


def test_graph_replays_agree_and_autotune(pla_session):
    sizes = list(np.random.default_rng(3).integers(16, 513, 60))
    systems, rhs = _systems(3, sizes, 4)
    for mode in (pla.PLA_BATCHED_FP64, pla.PLA_BATCHED_FP32_IR):
        res, cfg, _ = _run(pla_session, systems, rhs, mode | pla.PLA_BATCHED_AUTOTUNE, replays=4)
        assert cfg[0] in (16, 32, 64) and cfg[1] in (1, 2)
        first = res[0]
        for other in res[1:]:
            assert np.array_equal(other[1], first[1])
            for (a0, x0), (a1, x1) in zip(first[0], other[0]):
                assert np.array_equal(x0, x1) or mode == pla.PLA_BATCHED_FP32_IR
                if mode == pla.PLA_BATCHED_FP32_IR:  # LDS atomics: order may differ, accuracy may not
                    assert np.abs(x0 - x1).max() <= 1e-12 * max(np.abs(x0).max(), 1)
        for (a, b), (_, x) in zip(zip(systems, rhs), first[0]):
            assert _eta(a, x, b) <= 2e-14


def test_one_shot_and_argument_errors(pla_session):
    s = pla_session
    systems, rhs = _systems(4, [20, 30], 2)
    ha, hb, _, _ = _pack(systems, rhs)
    A, B = _Dev(s, ha, pla.MPI_DOUBLE), _Dev(s, hb, pla.MPI_DOUBLE)
    info = _Dev(s, np.zeros(2, np.int32), pla.MPI_INT)
    n = np.array([20, 30], np.int32)
    npi = n.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    pla.call("PLA_Chol_batched", 2, npi, None, None, 2, pla.PLA_BATCHED_FP64, 0.0, 0, A.ptr, B.ptr,
             info.ptr, None, None)
    x = B.get()
    assert _eta(systems[1], x[40:].reshape(30, 2, order="F"), rhs[1]) <= 1e-14
    L = pla.lib()
    assert L.PLA_Chol_batched(2, npi, None, None, 9, pla.PLA_BATCHED_FP64, 0.0, 0, A.ptr, B.ptr,
                              info.ptr, None, None) < 0          # nrhs > 8
    assert L.PLA_Chol_batched(2, npi, None, None, 2, 7, 0.0, 0, A.ptr, B.ptr, info.ptr, None,
                              None) < 0                            # unknown mode
    bad = np.array([20, -1], np.int32)
    assert L.PLA_Chol_batched(2, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None, 2,
                              pla.PLA_BATCHED_FP64, 0.0, 0, A.ptr, B.ptr, info.ptr, None, None) < 0
    big = np.array([20, 6000], np.int32)                           # does not fit the kernel's LDS
    assert L.PLA_Chol_batched(2, big.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None, 2,
                              pla.PLA_BATCHED_FP64, 0.0, 0, A.ptr, B.ptr, info.ptr, None, None) < 0
    assert "fit" in pla.last_error()
    # an empty batch is a no-op
    pla.call("PLA_Chol_batched", 0, npi, None, None, 2, pla.PLA_BATCHED_FP64, 0.0, 0, A.ptr, B.ptr,
             info.ptr, None, None)
    for d in (A, B, info):
        d.free()
