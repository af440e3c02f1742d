"""ctypes front end of the CPU oracle (oracle/pla_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker the HIP library is
compared with; only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline may import this module.  The product library never loads it.

All arrays are numpy float64 in Fortran (column-major) order, int32 pivots in
the reference's encoding (offset of the pivot row relative to row j,
SOLVERS/LU/PLA_seq_lu.c:33-35).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpla_oracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p


def build():
    """Compile the oracle with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_srand48.restype = ctypes.c_uint64
        L.ora_srand48.argtypes = [ctypes.c_long]
        L.ora_drand48.restype = _d
        L.ora_drand48.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.ora_lcg_skip.restype = ctypes.c_uint64
        L.ora_lcg_skip.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.ora_fill_uniform.argtypes = [_i, _i, _p, _i, ctypes.c_long, ctypes.c_int64]
        L.ora_idamax.restype = _i
        L.ora_idamax.argtypes = [_i, _p, _i]
        L.ora_lu_unblocked.argtypes = [_i, _i, _p, _i, _p]
        L.ora_lu_blocked.argtypes = [_i, _i, _p, _i, _i, _p]
        L.ora_lu_linpack.argtypes = [_i, _i, _p, _i, _i, _p]
        L.ora_zlu_blocked.argtypes = [_i, _i, _p, _i, _i, _p]
        L.ora_zlu_solve.argtypes = [_i, _i, _p, _i, _p, _p, _i]
        L.ora_zchol_blocked.restype = _i
        L.ora_zchol_blocked.argtypes = [_i, _p, _i, _i]
        L.ora_laswp.argtypes = [_i, _p, _i, _i, _i, _p]
        L.ora_chol_unblocked.restype = _i
        L.ora_chol_unblocked.argtypes = [_i, _p, _i]
        L.ora_chol_blocked.restype = _i
        L.ora_chol_blocked.argtypes = [_i, _p, _i, _i]
        L.ora_trsv.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_char, _i, _p, _i, _p]
        L.ora_trsm_left.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_char, _i, _i, _d, _p, _i,
                                    _p, _i]
        L.ora_trsm_right.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_char, _i, _i, _d, _p, _i,
                                     _p, _i]
        L.ora_syrk.argtypes = [ctypes.c_char, ctypes.c_char, _i, _i, _d, _p, _i, _d, _p, _i]
        L.ora_api_to_global.argtypes = [_i, _i, _d, _p, _i, _p, _i, _i, _i]
        L.ora_api_from_global.argtypes = [_i, _i, _d, _p, _i, _i, _i, _p, _i]
        L.ora_api_multi.argtypes = [_i, _i, _i, _p, _p, _d, _p, _i, _p, _i, _p, _p]
        L.ora_gemm.argtypes = [ctypes.c_char, ctypes.c_char, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i]
        L.ora_lu_solve.argtypes = [_i, _i, _p, _i, _p, _p, _i]
        L.ora_chol_solve.argtypes = [_i, _i, _p, _i, _p, _i]
        _lib = L
    return _lib


def _f(a):
    a = np.asarray(a, dtype=np.float64)
    return np.asfortranarray(a)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def drand48_stream(seed, count, skip=0):
    """count consecutive drand48() values after srand48(seed) and `skip` draws."""
    L = lib()
    st = ctypes.c_uint64(L.ora_lcg_skip(L.ora_srand48(seed), skip))
    return np.array([L.ora_drand48(ctypes.byref(st)) for _ in range(count)])


def uniform(m, n, seed=0, skip=0):
    """A(i,j) = drand48()*2-1, column-major draw order (TEST/create_problem.c:44-52)."""
    a = np.zeros((m, n), dtype=np.float64, order="F")
    lib().ora_fill_uniform(m, n, _ptr(a), max(m, 1), seed, skip)
    return a


def idamax(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().ora_idamax(len(x), _ptr(x), 1)


def lu_unblocked(a):
    a = _f(a).copy(order="F")
    m, n = a.shape
    piv = np.zeros(min(m, n), dtype=np.int32)
    lib().ora_lu_unblocked(m, n, _ptr(a), max(m, 1), _ptr(piv))
    return a, piv


def lu_blocked(a, nb=128):
    a = _f(a).copy(order="F")
    m, n = a.shape
    piv = np.zeros(min(m, n), dtype=np.int32)
    lib().ora_lu_blocked(m, n, _ptr(a), max(m, 1), nb, _ptr(piv))
    return a, piv


_hblas = None


def host_blas():
    """The host BLAS/LAPACK the CPU baseline calls: the OpenBLAS that scipy
    ships (its Fortran entry points carry a scipy_ prefix), loaded with
    ctypes.  Returns (library, threads)."""
    global _hblas
    if _hblas is None:
        import glob

        import scipy
        libs = glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs", "libscipy_openblas*.so"))
        if not libs:
            raise OSError("scipy's OpenBLAS library was not found")
        L = ctypes.CDLL(libs[0])
        L.scipy_openblas_get_num_threads.restype = ctypes.c_int
        _hblas = L
    return _hblas, _hblas.scipy_openblas_get_num_threads()


def _ri(v):
    return ctypes.byref(ctypes.c_int(v))


def _rd(v):
    return ctypes.byref(ctypes.c_double(v))


def _rc(c):
    return ctypes.c_char_p(c.encode())


def _at(a, i, j):
    """Address of a[i, j] of a column-major array."""
    return ctypes.c_void_p(a.ctypes.data + 8 * (i + j * a.shape[0]))


def lu_blocked_blas(a, nb=128):
    """Blocked right-looking LU with partial pivoting structured as
    PLA_LU_right (SOLVERS/LU/PLA_LU.c:37-212) on a 1x1 template, with the
    flops in a host BLAS/LAPACK as the reference puts them in its Fortran
    BLAS (INCLUDE/PLA_BLAS.h:32-36).  Per step of width b at column k:
      dgetrf on the m-k x b panel       (PLA_LU_right_mv, :144; LAPACK's
                                         partial pivoting picks the first
                                         maximum |a|, as idamax does)
      dlaswp left and right of it       (PLA_Apply_pivots_to_rows on all
                                         columns, :152)
      dtrsm  A12 <- inv(L11) A12        (:174-182)
      dgemm  A22 <- A22 - A21 A12       (:185-186)
    all in place on the column-major array (leading dimension m), in the
    host BLAS's threads.  Returns (factors, pivots) in the reference's
    encoding."""
    L, _ = host_blas()
    a = _f(a).copy(order="F")
    m, n = a.shape
    mn = min(m, n)
    piv = np.zeros(mn, dtype=np.int32)
    ip = np.zeros(max(nb, 1), dtype=np.int32)
    info = ctypes.c_int(0)
    for k in range(0, mn, nb):
        b = min(nb, mn - k)
        L.scipy_dgetrf_(_ri(m - k), _ri(b), _at(a, k, k), _ri(m), _ptr(ip), ctypes.byref(info))
        piv[k:k + b] = ip[:b] - 1 - np.arange(b)          # 1-based panel rows -> offsets from row j
        if k:
            L.scipy_dlaswp_(_ri(k), _at(a, k, 0), _ri(m), _ri(1), _ri(b), _ptr(ip), _ri(1))
        if k + b < n:
            L.scipy_dlaswp_(_ri(n - k - b), _at(a, k, k + b), _ri(m), _ri(1), _ri(b), _ptr(ip), _ri(1))
            L.scipy_dtrsm_(_rc("L"), _rc("L"), _rc("N"), _rc("U"), _ri(b), _ri(n - k - b), _rd(1.0), _at(a, k, k),
                           _ri(m), _at(a, k, k + b), _ri(m))
            if k + b < m:
                L.scipy_dgemm_(_rc("N"), _rc("N"), _ri(m - k - b), _ri(n - k - b), _ri(b), _rd(-1.0),
                               _at(a, k + b, k), _ri(m), _at(a, k, k + b), _ri(m), _rd(1.0), _at(a, k + b, k + b),
                               _ri(m))
    return a, piv


def chol_blocked_blas(a, nb=128):
    """Blocked right-looking lower Cholesky structured as PLA_Chol_lower_large
    (SOLVERS/Cholesky/PLA_Chol.c:83-234), flops in the host BLAS/LAPACK:
    dpotrf on the diagonal block (:178), dtrsm A21 <- A21 inv(L11)^T
    (:189-195), and the lower trailing update (:207-210) as dsyrk on each
    diagonal block plus dgemm below it, so the upper triangle is never
    written.  Returns (factors, info), info the 1-based index of the first
    non-positive pivot (0 if none)."""
    L, _ = host_blas()
    a = _f(a).copy(order="F")
    n = a.shape[0]
    info = ctypes.c_int(0)
    for k in range(0, n, nb):
        b = min(nb, n - k)
        L.scipy_dpotrf_(_rc("L"), _ri(b), _at(a, k, k), _ri(n), ctypes.byref(info))
        if info.value > 0:
            return a, k + info.value
        r = n - k - b
        if r <= 0:
            continue
        L.scipy_dtrsm_(_rc("R"), _rc("L"), _rc("T"), _rc("N"), _ri(r), _ri(b), _rd(1.0), _at(a, k, k), _ri(n),
                       _at(a, k + b, k), _ri(n))
        for j in range(k + b, n, nb):
            w = min(nb, n - j)
            L.scipy_dsyrk_(_rc("L"), _rc("N"), _ri(w), _ri(b), _rd(-1.0), _at(a, j, k), _ri(n), _rd(1.0),
                           _at(a, j, j), _ri(n))
            if j + w < n:
                L.scipy_dgemm_(_rc("N"), _rc("T"), _ri(n - j - w), _ri(w), _ri(b), _rd(-1.0), _at(a, j + w, k),
                               _ri(n), _at(a, j, k), _ri(n), _rd(1.0), _at(a, j + w, j), _ri(n))
    return a, 0


def lu_linpack(a, nb=128):
    """PLA_LU_linpack (ora_lu_linpack): returns (factors, pivots)."""
    a = _f(a).copy(order="F")
    m, n = a.shape
    piv = np.zeros(min(m, n), dtype=np.int32)
    lib().ora_lu_linpack(m, n, _ptr(a), max(m, 1), nb, _ptr(piv))
    return a, piv


def _z(a):
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


def zlu_blocked(a, nb=128):
    """MPI_DOUBLE_COMPLEX PLA_LU (ora_zlu_blocked): returns (factors, pivots)."""
    a = _z(a).copy(order="F")
    m, n = a.shape
    piv = np.zeros(min(m, n), dtype=np.int32)
    lib().ora_zlu_blocked(m, n, _ptr(a), max(m, 1), nb, _ptr(piv))
    return a, piv


def zlu_solve(lu, piv, b):
    lu = _z(lu)
    b = _z(b).copy(order="F")
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    n, nrhs = b.shape
    lib().ora_zlu_solve(n, nrhs, _ptr(lu), max(lu.shape[0], 1), _ptr(np.ascontiguousarray(piv, dtype=np.int32)),
                        _ptr(b), max(n, 1))
    return b


def zchol_blocked(a, nb=128):
    """MPI_DOUBLE_COMPLEX PLA_Chol, lower (ora_zchol_blocked): returns (factors, info)."""
    a = _z(a).copy(order="F")
    info = lib().ora_zchol_blocked(a.shape[0], _ptr(a), max(a.shape[0], 1), nb)
    return a, info


def uniform_z(m, n, seed=0, skip=0):
    """TEST/create_problem.c for MPI_DOUBLE_COMPLEX: drand48()*2-1 for the real
    then the imaginary part of each element, column by column."""
    r = uniform(2 * m, n, seed, skip)
    return np.asfortranarray(r[0::2] + 1j * r[1::2])


def laswp(a, piv, r0=0):
    a = _f(a).copy(order="F")
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    lib().ora_laswp(a.shape[1], _ptr(a), max(a.shape[0], 1), r0, len(piv), _ptr(piv))
    return a


def chol_unblocked(a):
    a = _f(a).copy(order="F")
    info = lib().ora_chol_unblocked(a.shape[0], _ptr(a), max(a.shape[0], 1))
    return a, info


def chol_blocked(a, nb=128):
    a = _f(a).copy(order="F")
    info = lib().ora_chol_blocked(a.shape[0], _ptr(a), max(a.shape[0], 1), nb)
    return a, info


def trsm_left(uplo, trans, diag, t, b, alpha=1.0):
    t = _f(t)
    b = _f(b).copy(order="F")
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    n, nrhs = b.shape
    lib().ora_trsm_left(uplo.encode(), trans.encode(), diag.encode(), n, nrhs, alpha, _ptr(t),
                        max(t.shape[0], 1), _ptr(b), max(n, 1))
    return b


def trsm_right(uplo, trans, diag, t, b, alpha=1.0):
    """B <- alpha B op(T)^-1 (ora_trsm_right)."""
    t = _f(t)
    b = _f(b).copy(order="F")
    m, n = b.shape
    lib().ora_trsm_right(uplo.encode(), trans.encode(), diag.encode(), m, n, alpha, _ptr(t),
                         max(t.shape[0], 1), _ptr(b), max(m, 1))
    return b


def syrk(uplo, trans, alpha, a, beta, c):
    """All of C scaled by beta, then the uplo triangle += alpha op(A) op(A)^T (ora_syrk)."""
    a = _f(a)
    c = _f(c).copy(order="F")
    n = c.shape[0]
    k = a.shape[1] if trans == "N" else a.shape[0]
    lib().ora_syrk(uplo.encode(), trans.encode(), n, k, alpha, _ptr(a), max(a.shape[0], 1), beta, _ptr(c),
                   max(n, 1))
    return c


def api_multi(to_global, rsz, csz, alpha, buf, a, rd, cd):
    """PLA_API_multi_axpy_{matrix_to_global,global_to_matrix} on one process
    (ora_api_multi): returns the updated (buf, a)."""
    buf = _f(buf).copy(order="F")
    a = _f(a).copy(order="F")
    i32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)
    rsz, csz, rd, cd = i32(rsz), i32(csz), i32(rd), i32(cd)
    lib().ora_api_multi(1 if to_global else 0, len(rsz), len(csz), _ptr(rsz), _ptr(csz), alpha, _ptr(buf),
                        max(buf.shape[0], 1), _ptr(a), max(a.shape[0], 1), _ptr(rd), _ptr(cd))
    return buf, a


def gemm(ta, tb, alpha, a, b, beta, c):
    a, b = _f(a), _f(b)
    c = _f(c).copy(order="F")
    m, n = c.shape
    k = a.shape[0] if ta == "T" else a.shape[1]
    lib().ora_gemm(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(a), max(a.shape[0], 1), _ptr(b),
                   max(b.shape[0], 1), beta, _ptr(c), max(m, 1))
    return c


def lu_solve(lu, piv, b):
    lu = _f(lu)
    b = _f(b).copy(order="F")
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    n, nrhs = b.shape
    lib().ora_lu_solve(n, nrhs, _ptr(lu), max(lu.shape[0], 1), _ptr(piv), _ptr(b), max(n, 1))
    return b


def chol_solve(l, b):
    l = _f(l)
    b = _f(b).copy(order="F")
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    n, nrhs = b.shape
    lib().ora_chol_solve(n, nrhs, _ptr(l), max(l.shape[0], 1), _ptr(b), max(n, 1))
    return b


def split_lu(lu, piv):
    """(P, L, U) with P A = L U from the packed factors and reference pivots."""
    m, n = lu.shape
    k = min(m, n)
    L = np.tril(lu[:, :k], -1) + np.eye(m, k)
    U = np.triu(lu[:k, :])
    perm = np.arange(m)
    for j, p in enumerate(piv):
        if p:
            perm[[j, j + p]] = perm[[j + p, j]]
    return perm, L, U


def lu_scaled_residual(a, lu, piv):
    """||P A - L U||_1 / (n ||A||_1 eps), the acceptance measure (< 16)."""
    perm, L, U = split_lu(lu, piv)
    pa = np.asarray(a)[perm, :]
    r = np.abs(pa - L @ U).sum(axis=0).max()
    na = np.abs(np.asarray(a)).sum(axis=0).max()
    n = max(a.shape)
    return r / (n * na * np.finfo(np.float64).eps)


# ---- FP8 (OCP E4M3) reference rounding --------------------------------------
# Restated from the OCP 8-bit floating point specification (OFP8, E4M3):
# 1 sign, 4 exponent (bias 7), 3 mantissa bits; normals 2^(E-7) (1 + m/8) for
# E = 1..15 (E = 15, m = 7 is NaN, so the largest value is 448); subnormals
# 2^-6 (m/8).  Rounding is to nearest, ties to even.  No counterpart in the
# reference; pinned in tests/test_oracle.py by the specification's encodings.

E4M3_MAX = 448.0


def e4m3_round(x):
    """Round to the nearest E4M3 value (ties to even); |x| must be <= 448."""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    if np.any(a > E4M3_MAX):
        raise ValueError("e4m3_round: |x| > 448")
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    e = np.maximum(e, -6.0)                 # subnormal range shares the 2^-9 quantum
    # correct log2 rounding at binade edges
    e = np.where(a >= np.exp2(e + 1), e + 1, e)
    e = np.where((a < np.exp2(e)) & (e > -6), e - 1, e)
    q = np.exp2(e - 3)                      # spacing of the binade
    r = np.rint(a / q) * q                  # exact: a / q is scaled by a power of two
    return np.copysign(r, x)


def e4m3_encode(v):
    """Bytes of E4M3 values (v must be exactly representable)."""
    v = np.asarray(v, dtype=np.float64)
    sign = (np.signbit(v)).astype(np.uint8) << 7
    a = np.abs(v)
    out = np.zeros(v.shape, dtype=np.uint8)
    sub = a < 2.0 ** -6
    m_sub = np.rint(a / 2.0 ** -9).astype(np.int64)
    e = np.floor(np.log2(np.where(sub, 1.0, a))).astype(np.int64)
    m = np.rint((np.where(sub, 1.0, a) / np.exp2(e) - 1.0) * 8).astype(np.int64)
    norm = ((e + 7) << 3) | m
    out = np.where(sub, m_sub, norm).astype(np.uint8) | sign
    if not np.array_equal(e4m3_decode(out), v):
        raise ValueError("e4m3_encode: value not representable")
    return out


def e4m3_decode(b):
    b = np.asarray(b, dtype=np.uint8).astype(np.int64)
    s = np.where(b & 0x80, -1.0, 1.0)
    E = (b >> 3) & 0xF
    m = b & 7
    val = np.where(E == 0, np.exp2(-6.0) * m / 8.0, np.exp2(E - 7.0) * (1.0 + m / 8.0))
    return s * val


def fp8_quantize_tiles(p, tile=256, width=256):
    """Quantization of an R x b panel (b <= width, width a multiple of
    `tile`) in tile x tile blocks: per block the smallest e with
    max|block| <= 448 * 2^e, clamped to [-127, 127] (the range of an E8M0
    scale), values rounded to FP32 then to E4M3.  Returns (bytes R x width,
    row-major; exponents in row-tile-major order, width / tile per row tile)."""
    p = np.asarray(p, dtype=np.float64)
    R, b = p.shape
    q = np.zeros((R, width), dtype=np.uint8)
    exps = []
    for t0 in range(0, R, tile):
        for c0 in range(0, width, tile):
            blk = p[t0:t0 + tile, c0:min(b, c0 + tile)]
            amax = np.abs(blk).max() if blk.size else 0.0
            if amax > 0:
                f, e2 = np.frexp(amax / E4M3_MAX)
                ex = int(e2 - 1 if f == 0.5 else e2)
            else:
                ex = 0
            ex = min(127, max(-127, ex))
            exps.append(ex)
            if blk.size:
                scaled = (blk * 2.0 ** -ex).astype(np.float32).astype(np.float64)
                q[t0:t0 + tile, c0:c0 + blk.shape[1]] = e4m3_encode(e4m3_round(scaled))
    return q, np.array(exps, dtype=np.int32)
