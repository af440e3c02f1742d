"""Writes the golden vectors under tests/golden/ (run once; outputs are committed).

Independent of the oracle: the inputs come from a pure-Python restatement of
POSIX drand48 (X <- 0x5DEECE66D X + 0xB mod 2^48, srand48(s): X = s<<16 | 0x330E)
in the column-major order of the reference's TEST/create_problem.c:31-52,
and the factorizations from LAPACK through scipy/numpy.

  lu_piv_n2048_seed0.npy   PLA_LU pivots (0-based offsets relative to row j,
                           SOLVERS/LU/PLA_seq_lu.c:33-35) of the n=2048 seed-0
                           matrix, from LAPACK dgetrf: piv[j] = ipiv[j] - 1 - j
  a_n2048_seed0_head.npy   first 16 entries (column-major) of that matrix
  chol_n20_shift.npz       TEST/input_test_chol's double case: n=20, A from
                           create_problem plus n on the diagonal (test_chol.c:119-128),
                           nb_alg=4; L = numpy.linalg.cholesky of the lower-symmetrised A
"""
import os

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
A_MUL, C_ADD, MASK = 0x5DEECE66D, 0xB, (1 << 48) - 1


def drand48_matrix(m, n, seed=0):
    x = ((seed & 0xFFFFFFFF) << 16 | 0x330E) & MASK
    out = np.empty(m * n, dtype=np.float64)
    for t in range(m * n):
        x = (A_MUL * x + C_ADD) & MASK
        out[t] = x / 2.0**48
    return (out * 2.0 - 1.0).reshape((m, n), order="F")


def main():
    a = drand48_matrix(2048, 2048, 0)
    lu, ipiv = scipy.linalg.lu_factor(a, check_finite=False)
    piv = (ipiv - np.arange(len(ipiv))).astype(np.int32)  # scipy ipiv is 0-based
    np.save(os.path.join(HERE, "lu_piv_n2048_seed0.npy"), piv)
    np.save(os.path.join(HERE, "a_n2048_seed0_head.npy"), a.ravel(order="F")[:16].copy())

    n = 20
    a20 = drand48_matrix(n, n, 0) + n * np.eye(n)
    sym = np.tril(a20) + np.tril(a20, -1).T
    L = np.linalg.cholesky(sym)
    np.savez(os.path.join(HERE, "chol_n20_shift.npz"), a=a20, l=L, nb=4)


if __name__ == "__main__":
    main()
