"""Benchmark: FP64-accurate LU on one MI355X -- mixed precision vs plain FP64,
n = 65536 (the metric's configuration), through the C ABI (libpla_mi355x.so).

mixed step   PLA_LU_mixed: FP32 copy of A, LU with FP64 panels and an FP16x3
             MFMA trailing update, then FP64 iterative refinement of one
             right-hand side until the normwise backward error is <= 1e-14.
plain step   restore A from a pristine HBM copy, PLA_LU (FP64 MFMA trailing
             update), interchanges + unit-lower + upper solves of b
             (SOLVERS/LU/PLA_General.c:67-111).
Both count the LU work 2n^3/3 per step; A is resident in HBM before timing.
`value` is the mixed path's FP64-accurate TFLOP/s (the backward error of
every timed step is checked); `plain_fp64` holds the plain path on the same
matrix, and `speedup_vs_plain` their ratio.  `cholesky` times PLA_Chol at
n = 32768 and the FP8 PLA_Chol_mixed at n = 49152 (n^3/3 work), `batched`
the 10,000-system persistent kernel.

Prints ONE JSON line.  Every roofline object carries the algorithmic work of
its kernel over the kernel's hipEvent-timed launches on the stream that
launches them: MFMA-bound trailing updates in TFLOP/s against the dense
matrix peak of their precision, HBM-bound residuals, triangular solves and
row interchanges in GB/s against 8 TB/s.  The CPU baseline is the oracle's
blocked LU and Cholesky (PLA_LU_right / PLA_Chol_lower_large restated over
the host's multithreaded OpenBLAS) on a bounded sample size.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64-accurate LU TFLOP/s, mixed vs plain FP64, n=65536; batched solves/s"
FP64_MFMA_PEAK = 78.6
FP16_MFMA_PEAK = 2500.0
FP8_MFMA_PEAK = 5000.0
HBM_PEAK_GBS = 8000.0
PREC_NAMES = {1: "FP16", 2: "BF16", 3: "FP16x3", 4: "BF16x3"}
ROLE_LU, ROLE_CHOL, ROLE_MIXED, ROLE_FP8, ROLE_RESID, ROLE_TRSV, ROLE_SWAP = 1, 2, 3, 4, 5, 6, 7

# HBM bytes per launch at the default sizes (rocprofv3 --pmc FETCH_SIZE x 2.0
# + WRITE_SIZE x 1.0, the scale factors measured on a copy of known size);
# profiles/trailing_gemm_traffic_n65536.md and profiles/hbm_kernels_traffic.md
TRAFFIC = {"hgemm": 9.40e9, "dgemm": 3.17e10,
           # n = 65536: the mixed and plain LU's HBM kernels
           "resid_lu": 3.445e10, "trsv_f32": 8.706e9, "swap_f32": 1.711e9, "swap_f64": 4.298e9,
           "trsv_f64": 1.883e10,
           # n = 49152, FP8 Cholesky: per timed update (a triangle launch, and a
           # rectangle launch every other time), residual, 8-column solve
           "fp8_update": 4.05e9, "resid_fp8": 1.966e10, "trsv_fp8": 7.845e9}


def cpu_baseline(n_lu, n_chol):
    """The oracle's blocked LU and Cholesky over the host BLAS, TFLOP/s."""
    from oracle import oracle as ora
    _, threads = ora.host_blas()
    a = ora.uniform(n_lu, n_lu, 0)
    t = time.perf_counter()
    ora.lu_blocked_blas(a, nb=128)
    dt = time.perf_counter() - t
    del a
    out = {"value": round(2.0 / 3.0 * n_lu ** 3 / dt / 1e12, 4), "unit": "TFLOP/s", "cores": threads,
           "kind": "port",
           "sample": f"oracle blocked LU (PLA_LU_right restated, nb=128) over the host's OpenBLAS 0.3.29 "
                     f"({threads} threads) on n={n_lu}, drand48 seed 0: {dt:.1f} s"}
    if n_chol > 0:
        # TEST/test_chol.c:121-127's input: the lower triangle of create_problem's
        # matrix shifted by n on the diagonal (diagonally dominant, so SPD)
        spd = ora.uniform(n_chol, n_chol, 0)
        spd[np.diag_indices(n_chol)] += float(n_chol)
        t = time.perf_counter()
        _, info = ora.chol_blocked_blas(spd, nb=128)
        dt = time.perf_counter() - t
        out["cholesky"] = {"value": round(n_chol ** 3 / 3.0 / dt / 1e12, 4), "unit": "TFLOP/s", "cores": threads,
                           "sample": f"oracle blocked Cholesky (PLA_Chol_lower_large restated, nb=128) over the "
                                     f"same OpenBLAS on n={n_chol}, drand48 seed 0 + n I (TEST/test_chol.c's "
                                     f"input): {dt:.1f} s, info {info}"}
    return out


def timing_query(pla, role):
    launches, ms, fl = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
    pla.call("PLA_Kernel_timing_query", role, ctypes.byref(launches), ctypes.byref(ms), ctypes.byref(fl))
    return launches.value, ms.value, fl.value


def roofline(kernel, launches, ms, flops, peak, mfma_per_flop=1, traffic=None):
    """MFMA roofline: algorithmic flops per second over the dense peak."""
    achieved = flops / (ms * 1e-3) / 1e12 if ms > 0 else None
    out = {"kernel": kernel, "bound": "mfma", "achieved": round(achieved, 3) if achieved else None,
           "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4) if achieved else None,
           "traffic": traffic, "launches": launches, "avg_launch_ms": round(ms / max(1, launches), 4)}
    if mfma_per_flop != 1 and achieved:
        # the hi/lo split runs mfma_per_flop MFMA products per algorithmic flop
        out["mfma_products_per_flop"] = mfma_per_flop
        out["mfma_products_frac"] = round(mfma_per_flop * achieved / peak, 4)
    return out


def hbm_roofline(kernel, launches, ms, nbytes, traffic=None):
    """HBM roofline: algorithmic bytes per second over 8 TB/s."""
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else None
    return {"kernel": kernel, "bound": "hbm", "achieved": round(gbs, 1) if gbs else None, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None, "traffic": traffic,
            "launches": launches, "avg_launch_ms": round(ms / max(1, launches), 4),
            "algorithmic_bytes_per_launch": round(nbytes / max(1, launches))}


def interchange_bytes(piv, n, nb, elem):
    """Algorithmic bytes of the LU's row interchanges outside the panels:
    per step, every row the step's composite permutation moves is read once
    and written once in each of the n - b columns outside the panel."""
    total = 0.0
    for k in range(0, len(piv), nb):
        b = min(nb, len(piv) - k)
        pos = {}
        for j in range(k, k + b):
            p = j + int(piv[j])
            if p != j:
                pos[j], pos[p] = pos.get(p, p), pos.get(j, j)
        moved = sum(1 for i, src in pos.items() if src != i)
        total += moved * (n - b) * 2.0 * elem
    return total


def run_batched(pla, s, count, nrhs, reps, seed=2):
    """Batched uneven SPD solves: `count` systems, n_i uniform in [16, 512],
    A_i = B_i B_i^T / n_i + 1e-2 I (B_i uniform[-1,1]), nrhs uniform[-1,1]
    right-hand sides each, in one persistent kernel replayed from a hipGraph
    (PLA_Batched_plan_*).  FP64 factor-and-solve, then FP32 factors with FP64
    refinement to a backward error <= 1e-14.  Each timed replay starts from
    pristine inputs already in HBM (restored by device copies outside the
    timed region)."""
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
    flops = float((ns.astype(np.float64) ** 3).sum() / 3.0)

    def dev(host, dt):
        o = s.mvector(host.size, 1, dt)
        pla.set_contents(o, host)
        ptr = ctypes.c_void_p()
        pla.call("PLA_Obj_device_buffer", o, ctypes.byref(ptr))
        return o, ptr

    A0, _ = dev(ha, pla.MPI_DOUBLE)
    B0, _ = dev(hb, pla.MPI_DOUBLE)
    A, pa = dev(ha, pla.MPI_DOUBLE)
    B, pb = dev(hb, pla.MPI_DOUBLE)
    del ha
    info, pinfo = dev(np.zeros(count, np.int32), pla.MPI_INT)
    iters, piters = dev(np.zeros(count, np.int32), pla.MPI_INT)
    berr, pberr = dev(np.zeros(count), pla.MPI_DOUBLE)
    nptr = ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    aoff = a_off.ctypes.data_as(ctypes.c_void_p)
    boff = b_off.ctypes.data_as(ctypes.c_void_p)
    out = {"systems": count, "n_range": [16, 512], "rhs": nrhs, "flops_chol": flops}
    for name, mode in (("fp64", pla.PLA_BATCHED_FP64), ("fp32_ir", pla.PLA_BATCHED_FP32_IR)):
        plan = ctypes.c_void_p()
        pla.call("PLA_Batched_plan_create", count, nptr, aoff, boff, nrhs, mode | pla.PLA_BATCHED_AUTOTUNE,
                 1e-14, 10, pa, pb, pinfo, piters, pberr, ctypes.byref(plan))
        nbv, wpc = ctypes.c_int(), ctypes.c_int()
        pla.call("PLA_Batched_plan_config", plan, ctypes.byref(nbv), ctypes.byref(wpc))
        times = []
        for r in range(reps + 1):  # the first replay is warmup
            pla.call("PLA_Copy", A0, A)
            pla.call("PLA_Copy", B0, B)
            pla.call("PLA_Device_synchronize")
            t = time.perf_counter()
            pla.call("PLA_Batched_plan_execute", plan)
            pla.call("PLA_Device_synchronize")
            if r:
                times.append(time.perf_counter() - t)
        pla.call("PLA_Batched_plan_free", ctypes.byref(plan))
        inf = pla.get_contents(info)[:, 0]
        if (inf != 0).any():
            raise RuntimeError(f"batched {name}: {int((inf != 0).sum())} systems reported a failure")
        t = float(np.median(times))
        res = {"solves_per_s": round(count / t, 1), "ms": round(t * 1e3, 3),
               "tflops_chol": round(flops / t / 1e12, 3), "frac_fp64_peak": round(flops / t / 1e12 / FP64_MFMA_PEAK, 4),
               "nb": nbv.value, "workgroups_per_cu": wpc.value}
        if mode == pla.PLA_BATCHED_FP32_IR:
            be = pla.get_contents(berr)[:, 0]
            it = pla.get_contents(iters)[:, 0]
            res["max_backward_error"] = float(be.max())
            res["refinement_steps_max"] = int(it.max())
            if not be.max() <= 1e-14:
                raise RuntimeError(f"batched fp32_ir: backward error {be.max()}")
        out[name] = res
        print(f"batched {name}: {res}", file=sys.stderr, flush=True)
    pla.free(A0, B0, A, B, info, iters, berr)
    return out


def run_cholesky(pla, s, n_plain, n_fp8, deltas, steps):
    """PLA_Chol on SPD A = B B^T / n + 1e-2 I (B from drand48 seed 0) at
    n_plain, and the FP8 PLA_Chol_mixed + FP64 refinement of 8 right-hand
    sides (drand48 seed 1) at n_fp8 for each delta; work n^3/3 per
    factorization, A formed on the GPU outside the timed region; every
    backward error recomputed with PLA_Gemm on the pristine A."""
    out = {}

    def spd(n, delta):
        Bm, A = s.matrix(n, n), s.matrix(n, n)
        pla.call("PLA_Obj_fill_drand48", Bm, 0, 0)
        al, zero = s.scalar(1.0 / n), s.scalar(0.0)
        pla.call("PLA_Gemm", pla.PLA_NO_TRANSPOSE, pla.PLA_TRANSPOSE, al, Bm, Bm, zero, A)
        pla.free(Bm, al, zero)
        if delta:
            v = np.array([delta])
            pla.call("PLA_Shift", A, pla.MPI_DOUBLE, v.ctypes.data_as(ctypes.c_void_p))
        return A

    def eta_cols(A0, B, X):
        n, k = pla.dims(B)
        R = s.matrix(n, k)
        one, mone, nrm = s.scalar(1.0), s.scalar(-1.0), s.scalar(0.0)
        pla.call("PLA_Copy", B, R)
        pla.call("PLA_Gemm", pla.PLA_NO_TRANSPOSE, pla.PLA_NO_TRANSPOSE, mone, A0, X, one, R)
        pla.call("PLA_Matrix_infinity_norm", A0, nrm)
        anorm = float(pla.get_contents(nrm)[0, 0])
        r, x, b = pla.get_contents(R), pla.get_contents(X), pla.get_contents(B)
        pla.free(R, one, mone, nrm)
        return float((np.abs(r).max(axis=0) / (anorm * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))).max())

    if n_plain > 0:
        n = n_plain
        # the block size, autotuned over {128, 256, 512} on a scratch matrix
        # (outside the timed region)
        nbv, tune_ms = ctypes.c_int(), (ctypes.c_double * 4)()
        pla.call("PLA_Autotune_nb", pla.PLA_NB_CHOL, n, ctypes.byref(nbv), ctypes.cast(tune_ms, ctypes.c_void_p))
        A0 = spd(n, 1e-2)
        A = s.matrix(n, n)
        pla.call("PLA_Get_device_nb_for", n, pla.PLA_NB_CHOL, ctypes.byref(nbv))
        pla.call("PLA_Copy", A0, A)
        pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A)       # warmup
        el = 0.0
        pla.call("PLA_Kernel_timing", 1)
        for i in range(steps):
            pla.call("PLA_Copy", A0, A)
            pla.call("PLA_Device_synchronize")
            t = time.perf_counter()
            info = pla.call("PLA_Chol", pla.PLA_LOWER_TRIANGULAR, A, ok=(0,))
            pla.call("PLA_Device_synchronize")
            el += time.perf_counter() - t
        rl = roofline("dgemm_kernel<TRI = lower> + dgemm_kernel<NT> (lower(A22) -= A21 A21^T, FP64)",
                      *timing_query(pla, ROLE_CHOL), FP64_MFMA_PEAK)
        pla.call("PLA_Kernel_timing", 0)
        # solve of A x = A * 1 and its backward error against the pristine A
        ones, b = s.matrix(n, 1), s.matrix(n, 1)
        one, zero = s.scalar(1.0), s.scalar(0.0)
        pla.call("PLA_Obj_set_to_one", ones)
        pla.call("PLA_Gemm", pla.PLA_NO_TRANSPOSE, pla.PLA_NO_TRANSPOSE, one, A0, ones, zero, b)
        x = s.matrix(n, 1)
        pla.call("PLA_Copy", b, x)
        pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG, A, x)
        pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_TRANSPOSE, pla.PLA_NONUNIT_DIAG, A, x)
        eta = eta_cols(A0, b, x)
        pla.free(A0, A, ones, b, x, one, zero)
        ms = el / steps * 1e3
        out["plain_fp64"] = {"n": n, "nb": nbv.value,
                             "nb_autotune_ms": {str(c): round(tune_ms[i], 1) for i, c in enumerate((128, 256, 512))},
                             "value": round(n ** 3 / 3.0 / (ms * 1e-3) / 1e12, 3),
                             "unit": "TFLOP/s", "frac_fp64_peak": round(n ** 3 / 3.0 / (ms * 1e-3) / 1e12 / FP64_MFMA_PEAK, 4),
                             "ms_per_factorization": round(ms, 2), "steps": steps, "roofline": rl,
                             "backward_error_inf": eta, "hpl_scaled_residual": eta / (n * 2.0 ** -52),
                             "data": "A = B B^T / n + 1e-2 I, B(i,j) = drand48()*2-1 from srand48(0), formed by "
                                     "PLA_Gemm; b = A * 1"}
        print(f"cholesky plain: {out['plain_fp64']}", file=sys.stderr, flush=True)
    if n_fp8 > 0:
        n, nrhs = n_fp8, 8
        Abase = spd(n, 0.0)
        A = s.matrix(n, n)
        B, X = s.matrix(n, nrhs), s.matrix(n, nrhs)
        pla.call("PLA_Obj_fill_drand48", B, 1, 0)
        res = {"n": n, "rhs": nrhs, "tile": 256, "runs": []}
        # warmup (untimed): the library's grow-only workspaces are allocated
        # by the first call
        if deltas:
            pla.call("PLA_Copy", Abase, A)
            v = np.array([deltas[0]])
            pla.call("PLA_Shift", A, pla.MPI_DOUBLE, v.ctypes.data_as(ctypes.c_void_p))
            it, be = ctypes.c_int(), ctypes.c_double()
            pla.call("PLA_Chol_mixed", A, B, X, pla.PLA_MIXED_FP8, 1e-14, 50, ctypes.byref(it), ctypes.byref(be))
        for delta in deltas:
            pla.call("PLA_Copy", Abase, A)
            v = np.array([delta])
            pla.call("PLA_Shift", A, pla.MPI_DOUBLE, v.ctypes.data_as(ctypes.c_void_p))
            it, be = ctypes.c_int(), ctypes.c_double()
            pla.call("PLA_Device_synchronize")
            pla.call("PLA_Kernel_timing", 1)
            t = time.perf_counter()
            pla.call("PLA_Chol_mixed", A, B, X, pla.PLA_MIXED_FP8, 1e-14, 50, ctypes.byref(it), ctypes.byref(be))
            pla.call("PLA_Device_synchronize")
            el = time.perf_counter() - t
            dflt = (n == 49152)
            rl = roofline("fp8_update_kernel (lower(W22) -= L21 L21^T, E4M3 with E8M0 tile scales, FP32 accumulate)",
                          *timing_query(pla, ROLE_FP8), FP8_MFMA_PEAK,
                          traffic=TRAFFIC["fp8_update"] if dflt else None)
            lr, msr, br = timing_query(pla, ROLE_RESID)
            lt, mst, bt = timing_query(pla, ROLE_TRSV)
            pla.call("PLA_Kernel_timing", 0)
            outer, steps_g, gcorr, shift = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
            pla.call("PLA_Mixed_stats", ctypes.byref(outer), ctypes.byref(steps_g), ctypes.byref(gcorr),
                     ctypes.byref(shift))
            eta = eta_cols(A, B, X)
            run = {"delta": delta, "value": round(n ** 3 / 3.0 / el / 1e12, 3), "unit": "TFLOP/s (FP64-accurate)",
                   "ms": round(el * 1e3, 1), "backward_error": be.value, "backward_error_recomputed": eta,
                   "corrections": it.value, "gmres_steps": steps_g.value, "gmres_corrections": gcorr.value,
                   "shift": shift.value, "roofline": rl,
                   "hbm_rooflines": [hbm_roofline("resid_partial_kernel<8> (r = b - A x on the full symmetric FP64 copy)",
                                                  lr, msr, br, TRAFFIC["resid_fp8"] if dflt else None),
                                     hbm_roofline("trsv_nt_kernel<float, 4> (solves with L and the mirrored L^T, "
                                                  "8 columns as two 4-column chains)", lt,
                                                  mst, bt, TRAFFIC["trsv_fp8"] if dflt else None)]}
            res["runs"].append(run)
            print(f"cholesky fp8 delta={delta}: {run}", file=sys.stderr, flush=True)
        pla.free(Abase, A, B, X)
        out["fp8_mixed"] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--mode", choices=["mixed", "plain", "none"], default="mixed",
                    help="none: skip the LU paths (profiling the Cholesky / batched sections alone)")
    ap.add_argument("--prec", type=int, default=3, help="1 FP16, 2 BF16, 3 FP16x3, 4 BF16x3")
    ap.add_argument("--tol", type=float, default=1e-14)
    ap.add_argument("--nb", type=int, default=0, help="block size (0: library's choice)")
    ap.add_argument("--plain-steps", type=int, default=1, help="plain FP64 steps timed beside mixed (0: skip)")
    ap.add_argument("--chol-n", type=int, default=32768, help="PLA_Chol size (0: skip)")
    ap.add_argument("--chol-steps", type=int, default=2)
    ap.add_argument("--chol-panel-cus", type=int, default=-1,
                    help="CUs reserved for the Cholesky panel (-1: the library's default)")
    ap.add_argument("--fp8-n", type=int, default=49152, help="FP8 PLA_Chol_mixed size (0: skip)")
    ap.add_argument("--fp8-deltas", type=str, default="4e-2,4e-4")
    ap.add_argument("--cpu-n", type=int, default=16384, help="CPU baseline LU size (0: skip)")
    ap.add_argument("--cpu-chol-n", type=int, default=16384, help="CPU baseline Cholesky size (0: skip)")
    ap.add_argument("--batched", type=int, default=10000, help="batched SPD systems (0: skip)")
    ap.add_argument("--batched-reps", type=int, default=5)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 16384 x 16384 FP64 matrix 3 times (2 GiB read + 2 GiB written each): "
                         "the known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    from plapack_amd import pla

    if args.copy_calibration:
        with pla.Session() as s:
            c_a, c_b = s.matrix(16384, 16384), s.matrix(16384, 16384)
            pla.call("PLA_Obj_fill_drand48", c_a, 0, 0)
            for _ in range(3):
                pla.call("PLA_Copy", c_a, c_b)
            pla.call("PLA_Device_synchronize")
            pla.free(c_a, c_b)
        print(json.dumps({"copy_calibration_bytes": 8 * 16384 * 16384}), flush=True)
        return

    n = args.n if args.mode != "none" else 256
    work = 2.0 / 3.0 * n ** 3
    data = ("synthetic: A(i,j) = drand48()*2-1 from srand48(0), column-major (TEST/create_problem.c), "
            "b = A*1, generated on the GPU")
    default_size = (n == 65536 and args.nb == 0)
    with pla.Session() as s:
        pla.call("PLA_Set_device_nb", args.nb)
        nbv, nb_plain = ctypes.c_int(), ctypes.c_int()
        pla.call("PLA_Get_device_nb_for", n, pla.PLA_NB_LU_MIXED if args.mode == "mixed" else pla.PLA_NB_LU,
                 ctypes.byref(nbv))
        pla.call("PLA_Get_device_nb_for", n, pla.PLA_NB_LU, ctypes.byref(nb_plain))
        A0 = s.matrix(n, n)
        pla.call("PLA_Obj_fill_drand48", A0, 0, 0)
        ones, b, x, r = s.mvector(n), s.mvector(n), s.mvector(n), s.mvector(n)
        P = s.mvector(n, dtype=pla.MPI_INT)
        one, zero, mone, nrm = s.scalar(1.0), s.scalar(0.0), s.scalar(-1.0), s.scalar(0.0)
        pla.call("PLA_Obj_set_to_one", ones)
        pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, one, A0, ones, zero, b)   # b = A*1
        pla.call("PLA_Matrix_infinity_norm", A0, nrm)
        anorm = float(pla.get_contents(nrm)[0, 0])

        def backward_error():
            pla.call("PLA_Copy", b, r)
            pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, mone, A0, x, one, r)
            xv, rv, bv = (pla.get_contents(o)[:, 0] for o in (x, r, b))
            return (float(np.abs(rv).max() / (anorm * np.abs(xv).max() + np.abs(bv).max())),
                    float(np.abs(xv - 1.0).max()))

        def run_plain(steps, warmup):
            A = s.matrix(n, n)

            def step():
                pla.call("PLA_Copy", A0, A)
                pla.call("PLA_LU", A, P)
                pla.call("PLA_Copy", b, x)
                pla.call("PLA_Apply_pivots_to_rows", x, P)
                pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_UNIT_DIAG, A, x)
                pla.call("PLA_Trsv", pla.PLA_UPPER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG,
                         A, x)
            for _ in range(warmup):
                step()
            pla.call("PLA_Device_synchronize")
            pla.call("PLA_Kernel_timing", 1)
            t0 = time.perf_counter()
            for i in range(steps):
                step()
                print(f"plain step {i} done", file=sys.stderr, flush=True)
            pla.call("PLA_Device_synchronize")
            el = time.perf_counter() - t0
            rl = roofline("dgemm_kernel<NN, ROLE 1> (A22 -= A21*A12, FP64)", *timing_query(pla, ROLE_LU),
                          FP64_MFMA_PEAK, traffic=TRAFFIC["dgemm"] if default_size else None)
            ls, mss, cols = timing_query(pla, ROLE_SWAP)
            lt, mst, bt = timing_query(pla, ROLE_TRSV)
            pla.call("PLA_Kernel_timing", 0)
            piv = pla.get_contents(P)[:, 0]
            swap_bytes = interchange_bytes(piv, n, nb_plain.value, 8) * steps
            hbm = [hbm_roofline("panel_swap_kernel<double> (row interchanges outside the panels)", ls, mss,
                                swap_bytes, TRAFFIC["swap_f64"] if default_size else None),
                   hbm_roofline("trsv_nt_kernel<double> (unit-lower and upper solves of b)", lt, mst, bt,
                                TRAFFIC["trsv_f64"] if default_size else None)]
            pla.free(A)
            eta, xerr = backward_error()
            ms = el / steps * 1e3
            return {"value": round(work / (ms * 1e-3) / 1e12, 3), "ms_per_step": round(ms, 2),
                    "frac_fp64_peak": round(work / (ms * 1e-3) / 1e12 / FP64_MFMA_PEAK, 4),
                    "steps": steps, "warmup": warmup, "nb": nb_plain.value, "roofline": rl, "hbm_rooflines": hbm,
                    "backward_error_inf": eta, "hpl_scaled_residual": eta / (n * 2.0 ** -52),
                    "hpl_pass": bool(eta / (n * 2.0 ** -52) < 16),
                    "x_minus_1_inf": xerr}

        if args.mode == "none":
            out = {"metric": METRIC, "value": None, "unit": "TFLOP/s", "n_gpus": 1, "steps": 0, "warmup": 0,
                   "note": "LU paths skipped (--mode none)"}
        elif args.mode == "plain":
            res = run_plain(args.steps, args.warmup)
            out = {"metric": METRIC, "value": res["value"], "unit": "TFLOP/s", "n_gpus": 1,
                   "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
                   "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                   "data": data,
                   "config": {"workload": f"plain FP64 PLA_LU (partial pivoting) + 1-RHS solve, n={n}",
                              "n": n, "nb": nbv.value, "rhs": 1, "template": "1x1"},
                   "roofline": res["roofline"], "hbm_rooflines": res["hbm_rooflines"],
                   "accuracy": {k: res[k] for k in ("backward_error_inf", "hpl_scaled_residual", "hpl_pass",
                                                    "x_minus_1_inf")}}
        else:
            X = s.matrix(n, 1)
            bmat = s.matrix(n, 1)
            pla.call("PLA_Copy", b, bmat)
            it, be = ctypes.c_int(), ctypes.c_double()

            def step():
                pla.call("PLA_LU_mixed", A0, P, bmat, X, args.prec, args.tol, 30, ctypes.byref(it),
                         ctypes.byref(be))
                if not be.value <= args.tol:
                    raise RuntimeError(f"refinement did not reach {args.tol}: {be.value}")
            for _ in range(args.warmup):
                step()
            pla.call("PLA_Device_synchronize")
            pla.call("PLA_Kernel_timing", 1)
            t0 = time.perf_counter()
            errs, iters = [], []
            for i in range(args.steps):
                step()
                errs.append(be.value)
                iters.append(it.value)
                print(f"mixed step {i} done: {it.value} refinement steps, backward error {be.value:.3e}",
                      file=sys.stderr, flush=True)
            pla.call("PLA_Device_synchronize")
            el = time.perf_counter() - t0
            split = 3 if args.prec in (3, 4) else 1
            rl = roofline(f"hgemm_kernel<{PREC_NAMES[args.prec]}> (W22 -= L21*U12, FP32 accumulate)",
                          *timing_query(pla, ROLE_MIXED), FP16_MFMA_PEAK, mfma_per_flop=split,
                          traffic=TRAFFIC["hgemm"] if default_size and args.prec == 3 else None)
            lr, msr, br = timing_query(pla, ROLE_RESID)
            lt, mst, bt = timing_query(pla, ROLE_TRSV)
            ls, mss, cols = timing_query(pla, ROLE_SWAP)
            pla.call("PLA_Kernel_timing", 0)
            piv = pla.get_contents(P)[:, 0]
            swap_bytes = interchange_bytes(piv, n, nbv.value, 4) * args.steps
            dflt = default_size and args.prec == 3
            hbm = [hbm_roofline("resid_partial_kernel (r = b - A x in FP64)", lr, msr, br,
                                TRAFFIC["resid_lu"] if dflt else None),
                   hbm_roofline("trsv_nt_kernel<float> (FP32 L and U solves of the refinement)", lt, mst, bt,
                                TRAFFIC["trsv_f32"] if dflt else None),
                   hbm_roofline("panel_swap_kernel<float> (row interchanges outside the panels)", ls, mss,
                                swap_bytes, TRAFFIC["swap_f32"] if dflt else None)]
            pla.call("PLA_Copy", X, x)
            eta, xerr = backward_error()   # recomputed independently (GEMV on A0)
            ms = el / args.steps * 1e3
            value = work / (ms * 1e-3) / 1e12
            pla.free(X, bmat)
            plain = run_plain(args.plain_steps, 0) if args.plain_steps > 0 else None
            out = {"metric": METRIC, "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1,
                   "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2),
                   "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                   "dtype": f"{PREC_NAMES[args.prec].lower()} MFMA / f32 / f64 (FP64-refined)",
                   "data": data,
                   "config": {"workload": f"mixed-precision PLA_LU_mixed ({PREC_NAMES[args.prec]} MFMA trailing "
                                          f"update, FP64 panels, FP32 factors) + FP64 iterative refinement to "
                                          f"backward error <= {args.tol}, n={n}, 1 RHS",
                              "n": n, "nb": nbv.value, "rhs": 1, "template": "1x1",
                              "precision": PREC_NAMES[args.prec]},
                   "roofline": rl, "hbm_rooflines": hbm,
                   "accuracy": {"backward_error_inf": max(errs), "refinement_steps": iters,
                                "backward_error_recomputed": eta, "x_minus_1_inf": xerr}}
            if plain:
                out["plain_fp64"] = plain
                out["speedup_vs_plain"] = round(value / plain["value"], 3)
        pla.free(A0, ones, b, x, r, P, one, zero, mone, nrm)
        pla.call("PLA_Set_device_nb", 0)
        if args.chol_panel_cus >= 0:
            pla.call("PLA_Set_device_param", pla.PLA_PARAM_CHOL_PANEL_CUS, args.chol_panel_cus)
        if args.chol_n > 0 or args.fp8_n > 0:
            deltas = [float(v) for v in args.fp8_deltas.split(",") if v]
            out["cholesky"] = run_cholesky(pla, s, args.chol_n, args.fp8_n, deltas, args.chol_steps)
        if args.batched > 0:
            out["batched"] = run_batched(pla, s, args.batched, 4, args.batched_reps)

    if args.cpu_n > 0:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_chol_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
