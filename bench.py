"""Benchmark: plain FP64 PLA_LU with partial pivoting plus the 1-RHS solve,
on one MI355X, through the C ABI (libpla_mi355x.so).

One step = restore A from a pristine device copy, PLA_LU(A, pivots), then
apply the interchanges to b = A*1 and do the unit-lower and upper triangular
solves -- the factor-and-solve path of SOLVERS/LU/PLA_General.c:67-111 with
A already resident in HBM.  Throughput counts the LU work 2n^3/3 per step.

Prints ONE JSON line.  The roofline object is the trailing-update GEMM (the
dominant kernel): its algorithmic flops over its hipEvent-timed launches on
the library stream.  The CPU baseline is the oracle's blocked LU (a scalar C
restatement of the reference algorithm) on a bounded sample size.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64-accurate LU TFLOP/s, mixed vs plain FP64, n=65536; batched solves/s"
FP64_MFMA_PEAK_TFLOPS = 78.6


def cpu_baseline(n):
    """Oracle blocked LU on n x n (single thread), in TFLOP/s."""
    from oracle import oracle as ora
    a = ora.uniform(n, n, 0)
    t = time.perf_counter()
    ora.lu_blocked(a, nb=128)
    dt = time.perf_counter() - t
    return {"value": 2.0 / 3.0 * n ** 3 / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"oracle blocked LU (nb=128, scalar C, netlib loop order) on n={n}, "
                      f"drand48 seed 0; {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nb", type=int, default=0, help="block size (0: library's choice)")
    ap.add_argument("--cpu-n", type=int, default=4096, help="oracle sample size (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    from plapack_amd import pla

    n = args.n
    with pla.Session() as s:
        pla.call("PLA_Set_device_nb", args.nb)
        import ctypes
        nbv = ctypes.c_int()
        pla.call("PLA_Get_device_nb", n, ctypes.byref(nbv))
        A0, A = s.matrix(n, n), s.matrix(n, n)
        pla.call("PLA_Obj_fill_drand48", A0, 0, 0)
        ones, b, x, r = s.mvector(n), s.mvector(n), s.mvector(n), s.mvector(n)
        P = s.mvector(n, dtype=pla.MPI_INT)
        one, zero, mone = s.scalar(1.0), s.scalar(0.0), s.scalar(-1.0)
        pla.call("PLA_Obj_set_to_one", ones)
        pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, one, A0, ones, zero, b)

        def step():
            pla.call("PLA_Copy", A0, A)
            pla.call("PLA_LU", A, P)
            pla.call("PLA_Copy", b, x)
            pla.call("PLA_Apply_pivots_to_rows", x, P)
            pla.call("PLA_Trsv", pla.PLA_LOWER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_UNIT_DIAG, A, x)
            pla.call("PLA_Trsv", pla.PLA_UPPER_TRIANGULAR, pla.PLA_NO_TRANSPOSE, pla.PLA_NONUNIT_DIAG,
                     A, x)

        for _ in range(args.warmup):
            step()
        pla.call("PLA_Device_synchronize")
        pla.call("PLA_Kernel_timing", 1)
        t0 = time.perf_counter()
        for i in range(args.steps):
            step()
            print(f"step {i} done", file=sys.stderr, flush=True)
        pla.call("PLA_Device_synchronize")
        elapsed = time.perf_counter() - t0
        launches, kms, kfl = ctypes.c_long(), ctypes.c_double(), ctypes.c_double()
        pla.call("PLA_Kernel_timing_query", 1, ctypes.byref(launches), ctypes.byref(kms),
                 ctypes.byref(kfl))
        pla.call("PLA_Kernel_timing", 0)

        # accuracy of the last step (FP64 on the GPU): normwise backward error
        # eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) and HPL's eta / (n eps)
        pla.call("PLA_Copy", b, r)
        pla.call("PLA_Gemv", pla.PLA_NO_TRANSPOSE, mone, A0, x, one, r)
        nrm = s.scalar(0.0)
        pla.call("PLA_Matrix_infinity_norm", A0, nrm)
        anorm = float(pla.get_contents(nrm)[0, 0])
        xv, rv, bv = (pla.get_contents(o)[:, 0] for o in (x, r, b))
        eta = float(np.abs(rv).max() / (anorm * np.abs(xv).max() + np.abs(bv).max()))
        xerr = float(np.abs(xv - 1.0).max())
        pla.free(A0, A, ones, b, x, r, P, one, zero, mone, nrm)

    ms = elapsed / args.steps * 1e3
    work = 2.0 / 3.0 * n ** 3
    value = work / (ms * 1e-3) / 1e12
    achieved = kfl.value / (kms.value * 1e-3) / 1e12 if kms.value > 0 else None
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: A(i,j) = drand48()*2-1 from srand48(0), column-major "
                "(TEST/create_problem.c), b = A*1, generated on the GPU",
        "config": {"workload": f"plain FP64 PLA_LU (partial pivoting) + 1-RHS solve, n={n}",
                   "n": n, "nb": nbv.value, "rhs": 1, "template": "1x1"},
        "roofline": {"kernel": "dgemm_kernel<NN, ROLE 1> (A22 -= A21*A12)", "bound": "mfma",
                     "achieved": round(achieved, 3) if achieved else None,
                     "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4) if achieved else None,
                     "traffic": None, "launches": launches.value,
                     "avg_launch_ms": round(kms.value / max(1, launches.value), 4)},
        "accuracy": {"backward_error_inf": eta, "hpl_scaled_residual": eta / (n * 2.0 ** -52),
                     "x_minus_1_inf": xerr},
    }
    if args.cpu_n > 0:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
