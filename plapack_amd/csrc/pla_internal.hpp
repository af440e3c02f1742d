// pla_internal.hpp -- internal types of libpla_mi355x (not installed).
//
// A PLA_Obj is a view {objtype, shape, offset} onto a reference-counted base
// object that owns one column-major device buffer in HBM, plus an optional
// pinned host mirror for callers that use PLA_Obj_local_buffer.  This is the
// reference's base/view split (INCLUDE/PLA_Obj.h:15-82) on a 1x1 template,
// where every "local" quantity equals the global one.
#pragma once

#define PLA_NO_MPI_H
#include "PLA.h"
#undef min  // PLA.h keeps the reference's min() macro for C drivers

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <initializer_list>
#include <string>

struct pla_base {
  MPI_Datatype datatype;
  int typesize;           // bytes per element
  int length, width;      // base dimensions
  int ldim;               // elements, >= max(1, length)
  void *dev;              // device buffer (HBM)
  void *host;             // pinned mirror or nullptr
  bool host_dirty;        // mirror may hold data newer than the device copy
  int number_of_views;    // reference count (OBJ/PLA_Obj.c:2019)
  int user_buffer;        // dev belongs to the caller (PLA_Obj_attach_buffer)
  PLA_Template templ;
};

struct PLA_Obj_view_struct {
  int objtype;
  int fixed;              // constants from PLA_Create_constants_conf_to
  int length, width;      // view dimensions
  int row0, col0;         // offset of the view's top-left element in the base
  int owner_row, owner_col;
  int proj_onto;
  pla_base *base;
};

struct PLA_Template_struct {
  MPI_Comm comm_all, comm_row, comm_col;
  int nb, zero_or_one, me, myrow, mycol, nprows, npcols, times_used;
};

// A new zero-filled object with like's object type, shape and owners and the
// given datatype (pla_obj.cpp).
extern "C" int pla_obj_create_like(PLA_Obj like, MPI_Datatype datatype, PLA_Obj *new_obj);

namespace pla {

// ---- error reporting ----
int fail(int code, const char *fmt, ...);      // records message, returns code
int hip_fail(hipError_t e, const char *what);  // maps a HIP error
#define PLA_HIP(call)                                         \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return pla::hip_fail(e_, #call);    \
  } while (0)

// ---- library context ----
struct context {
  bool initialized = false;
  int device = 0;
  hipStream_t stream = nullptr;      // main stream (default priority)
  hipStream_t stream_la = nullptr;   // look-ahead stream (default priority)
  hipStream_t stream_tr = nullptr;   // trailing-update stream of the factorizations (default priority)
  int nb_user = 0;                   // set by pla_Environ_set_nb_alg (0 = not set)
  int nb_device = 0;                 // PLA_Set_device_nb (0 = built-in choice)
  int cu_count = 256;
  // tuning parameters (PLA_Set_device_param)
  int lookahead = 1;        // factor panel k+1 beside step k's trailing update
  int leaf_rows = 256;      // target rows per workgroup of the panel leaf
  int panel_cus = 0;        // CUs reserved for the LU panel (0: panel shares every CU)
  int chol_panel_cus = 32;  // CUs reserved for the Cholesky panel (0: panel shares every CU)
  int leaf_width = 0;       // panel-LU leaf width (0: the widest that fits)
  int leaf_kind = 0;        // panel-LU leaf: 0 rows in registers, 1 rows in LDS
  int nb_tail = 0;          // LU block size once at most nb_switch rows remain (0: none)
  int nb_switch = 0;
  unsigned leaf_spin_limit = 400000u;  // bounded wait of a panel-leaf exchange (polls)
  unsigned trsv_spin_limit = 2000000u; // bounded wait of a triangular-solve block (polls; 0: none)
  int trsv_group_cols = 4;  // untransposed solves: columns per independent chain (0: all in one)
  // FP8 Cholesky (PLA_Chol_mixed): block size (256 or 512), whether its panel
  // runs on the CU-masked Cholesky stream pair, and the first diagonal shift
  // tried after an unshifted factorization fails, as 2^fp8_shift_exp times
  // the largest diagonal element
  int fp8_nb = 512;
  int fp8_masked = 0;
  int fp8_shift_exp = -6;
  int fp8_gmres_first = 1;  // measured n = 49152, delta 4e-2: 9 GMRES steps vs 12 classic corrections
  int fp8_side_cvt = 1;     // FP8 panel: the conversions of A21 and X21 on a side stream (off the panel chain)
  // CU-masked stream pair of the LU look-ahead: the panel stream runs on
  // masked_cus CUs (masked_cus / 8 per XCD), the trailing stream on the rest
  hipStream_t stream_panel = nullptr, stream_trail = nullptr;
  int masked_cus = 0;
};
context &ctx();
int ensure_device();                 // lazily initialise device + streams
// Events of the library's stream orderings, created with the streams at
// PLA_Init and destroyed at PLA_Finalize (nullptr outside that window).
enum {
  EV_LU_READY, EV_LU_PANEL, EV_MLU_READY, EV_MLU_PANEL, EV_ZLU_READY, EV_ZLU_PANEL, EV_CHOL_READY,
  EV_CHOL_PANEL, EV_JOIN0, EV_JOIN1, EV_JOIN2, EV_FP8_SIDE, EV_FP8_A21, EV_FP8_X21, EV_FP8_OUT, EV_COUNT
};
hipEvent_t lib_event(int id);

// Streams of a blocked LU with look-ahead.  *sP receives the panel
// factorizations, *sT the interchanges right of the panel and the trailing
// updates; `main` keeps the interchanges left of the panel.  With panel CUs
// reserved (PLA_PARAM_PANEL_CUS) they are a CU-masked pair that never share a
// CU, so the panel leaf's workgroups (which must all be resident at once)
// never wait for trailing-update workgroups to drain; *gcap is then the most
// workgroups a panel-leaf launch may use (0: no cap beyond the device's).
// Without look-ahead both are `main`.  lu_streams_begin makes *sT and *sP
// wait for work already queued on `main`; lu_streams_end makes `main` wait
// for both.
int lu_streams_begin(hipStream_t main, int m, hipStream_t *sP, hipStream_t *sT, int *gcap);
int lu_streams_end(hipStream_t main, hipStream_t sP, hipStream_t sT);
// The same for the Cholesky look-ahead: *sP the panel factorizations, *sT the
// trailing updates.  With PLA_PARAM_CHOL_PANEL_CUS set they are the CU-masked
// pair (created when the parameter is set, never here); otherwise *sP is the
// look-ahead stream and *sT is `main`.  End with lu_streams_end.
int chol_streams_begin(hipStream_t main, hipStream_t *sP, hipStream_t *sT, bool allow_mask = true);
// Creates (or re-creates) the CU-masked stream pair with k panel CUs.
int masked_streams(int k);

// Grow-only device workspaces (reused across calls, freed at PLA_Finalize);
// `zero`: clear the memory when it is (re)allocated.
constexpr int PLA_NUM_WS = 58;
// slots: 0-2 factorizations, 3-4 LU interchanges, 5-9 triangular solves and
// residual, 10-23 mixed LU, 24-28 symmetric residual / inner products,
// 29-33 GMRES, 34-42 mixed Cholesky, 43-44 application-interface staging, 45-48 complex
// (GEMM operand embedding, conjugate transposes, Cholesky info word and block backup,
// the look-ahead panel's operand embedding), 49-50 level-1/2 helpers (pla_util.hip)
enum {
  WS_SRES_PART = 24, WS_SRES_MAX, WS_ONES, WS_NORM_R, WS_MDOT,
  WS_GM_V = 29, WS_GM_W, WS_GM_H, WS_GM_T, WS_GM_Z,
  WS_CH_W = 34, WS_CH_P64, WS_CH_LQ, WS_CH_EXP, WS_CH_INFO, WS_CH_DINVL, WS_CH_DINVLT, WS_CH_D, WS_CH_R,
  WS_API_0 = 43, WS_API_1, WS_ZEMBED, WS_ZCT, WS_ZINFO, WS_ZEMBED_P,
  WS_UTIL = 49, WS_UTIL_NORM,
  WS_ZOPA = 51, WS_ZOPB, WS_ZTRI,  // complex BLAS3 operands (pla_zblas.cpp)
  WS_CH_FULL = 54,                  // full symmetric FP64 copy of A (PLA_Chol_mixed refinement)
  WS_CH_LINV, WS_CH_OUT,            // inverse of L11 and the solved panel (PLA_Chol_mixed)
  WS_LEAF_FB = 57                   // scratch of a panel leaf factored by one workgroup after its exchange gave up
};
void *workspace(size_t bytes, int slot, bool zero = false);
void release_workspaces();

// ---- object helpers ----
bool valid(PLA_Obj obj);
inline size_t elem_offset(PLA_Obj o) {
  return (size_t)o->col0 * (size_t)o->base->ldim + (size_t)o->row0;
}
// Device pointer to the view's top-left element, after syncing a dirty
// host mirror to the device.
int dev_ptr(PLA_Obj o, void **p);
// After the library wrote the object on the device: refresh its mirror.
int written(PLA_Obj o);
int ld(PLA_Obj o);

// Block size a factorization of kind NB_* uses for an n x n problem.
enum { NB_LU = PLA_NB_LU, NB_CHOL = PLA_NB_CHOL, NB_LU_MIXED = PLA_NB_LU_MIXED };
int choose_nb(int n, int kind);
// Width of the LU step at column k of an m-row factorization (mn = min(m, n))
// with block size nb: nb, or PLA_PARAM_NB_TAIL once at most
// PLA_PARAM_NB_SWITCH rows remain; and the number of steps.
int lu_step_width(int m, int mn, int nb, int k);
int lu_step_count(int m, int mn, int nb);

// Device scalar (1x1 object) read back to the host.
int scalar_value(PLA_Obj o, double *v);

// ---- exchange slots of the panel-LU leaf kernel (allocated once, zeroed) ----
struct leaf_ws {
  void *base = nullptr;
  size_t slot_bytes = 0;
  unsigned long long *slots = nullptr;  // tagged exchange words
  int *err = nullptr;                   // timeout flag
  unsigned long long *state = nullptr;  // committed / abandoned word of the current leaf launch,
                                        // then the count of leaves factored by the fallback
};
leaf_ws &leaf_workspace(size_t slot_bytes);
// First of `count` fresh exchange tags for a leaf launch on stream s.
int leaf_tags(hipStream_t s, unsigned count, unsigned *tag0);
// Waits for stream s; returns PLA_ERR_DEVICE (message naming `what`) and
// clears the word if any bounded wait gave up since the last check.
int leaf_check_error(hipStream_t s, const char *what = "panel factorization");
// The device error word shared by the bounded waits of every kernel.
int *leaf_error_word();

// ---- kernel launchers (pla_kernels.hip); all asynchronous on stream s ----
int k_gemm(hipStream_t s, int role, bool ta, bool tb, int m, int n, int k, double alpha,
           const double *A, long lda, const double *B, long ldb, double beta, double *C, long ldc,
           const int *abort_flag);
int k_syrk_lower(hipStream_t s, int role, int n, int k, double alpha, const double *X, long ldx,
                 double beta, double *C, long ldc, const int *abort_flag);
// lower (upper) triangle of the n x n C <- beta C + alpha op(X) op(X)^T,
// op(X) = X (n x k) or, trans, X^T (X k x n)
int k_syrk(hipStream_t s, bool lower, bool trans, int n, int k, double alpha, const double *X, long ldx, double beta,
           double *C, long ldc);
// rows i and i + piv[i] (i < np) of an nrows x ncols block; entries that point
// outside the block are ignored
int k_laswp(hipStream_t s, double *A, long lda, int nrows, int ncols, int np, const int *piv);
int k_trsm_leaf(hipStream_t s, bool lower, bool trans, bool unit, int b, int ncols, const double *T,
                long ldt, double *B, long ldb, const int *abort_flag);
// X <- X inv(op(T)), X nrows x b, T b x b (b <= 64)
int k_trsm_right_leaf(hipStream_t s, bool lower, bool trans, bool unit, int nrows, int b, const double *T, long ldt,
                      double *X, long ldx, const int *abort_flag);
int k_potrf_leaf(hipStream_t s, int b, double *A, long lda, int offset, int *info);
// application-interface transfers of a packed M x N block (leading dimension
// M) to / from its place in an object (maps or offsets r0, c0)
enum { API_F64 = 1, API_F32 = 2, API_I32 = 3, API_Z64 = 4 };
int k_api_scatter_add(hipStream_t s, int elem, int M, int N, const void *alpha, const void *src, void *dst, long ldd,
                      const int *rmap, const int *cmap, int r0, int c0);
int k_api_gather(hipStream_t s, int elem, int M, int N, const void *src, long lds_, void *dst, const int *rmap,
                 const int *cmap, int r0, int c0);
// ---- MPI_DOUBLE_COMPLEX LU (pla_complex.hip, pla_zfactor.cpp) ----
// H (2m x 2n real, ldh) = the real embedding of the complex m x n block A
int k_zembed(hipStream_t s, int m, int n, const void *A, long lda, double *H, long ldh);
int k_zlaswp(hipStream_t s, void *A, long lda, int nrows, int ncols, int np, const int *piv);
int zleaf_width();
int k_zleaf(hipStream_t s, void *P, long lda, int mp, int w, int pw, int pc0, int *piv);
int ztrsm_leaf_size();
int k_ztrsm_leaf(hipStream_t s, bool lower, bool unit, int b, int ncols, const void *T, long ldt, void *B,
                 long ldb);
int k_zconjtrans(hipStream_t s, int m, int n, const void *A, long lda, void *B, long ldb);  // B = A^H
int zpotrf_leaf_size();
int k_zpotrf_leaf(hipStream_t s, int b, void *A, long lda, int offset, int *info);
int k_ztrsm_rlc_leaf(hipStream_t s, int nrows, int b, const void *L, long ldl, void *X, long ldx, const int *info);
int zherk_diag_max();
int k_zherk_diag(hipStream_t s, int n, int k, const void *X, long ldx, void *C, long ldc, const int *info);
int zchol_factor(hipStream_t s, int n, void *a, long lda, int nb, int *info_out);
int zchol_solve(hipStream_t s, int n, int nrhs, const void *l, long lda, void *b, long ldb);
int zlu_factor(hipStream_t s, int m, int n, void *a, long lda, int *piv, int nb);
int zlu_solve(hipStream_t s, int n, int nrhs, const void *lu, long lda, const int *piv, void *b, long ldb);
// complex C -= A B (A m x k, B k x n) on the real embedding, embedded operand
// in workspace slot ws
int zgemm_minus(hipStream_t s, int m, int n, int k, const void *A, long lda, const void *B, long ldb, void *C,
                long ldc, int ws);
// complex B <- inv(T) B, T n x n lower or upper (not transposed)
int ztrsm_left_n(hipStream_t s, bool lower, bool unit, int n, int nrhs, const void *T, long ldt, void *B, long ldb,
                 int ws);
// ---- datatype conversions and complex operand kernels (pla_convert.hip) ----
// dst = src converted (MPI_FLOAT <-> MPI_DOUBLE, MPI_COMPLEX <-> MPI_DOUBLE_COMPLEX), m x n elements
int k_convert2d(hipStream_t s, int m, int n, const void *src, long lds_, MPI_Datatype sdt, void *dst, long ldd,
                MPI_Datatype ddt);
// complex B (mo x no) = alpha op(A): op 0 = A, 1 = A^T, 2 = A^H, 3 = conj(A)
int k_zop(hipStream_t s, int op, int mo, int no, double are, double aim, const void *A, long lda, void *B, long ldb);
// complex C = beta C (beta = 0 clears C)
int k_zscale2d(hipStream_t s, int m, int n, double bre, double bim, void *C, long ldc);
// the uplo triangle (diagonal included) of the complex n x n C += P
int k_ztri_add(hipStream_t s, int n, bool lower, const void *P, long ldp, void *C, long ldc);
// ---- single-precision objects (pla_single.cpp) ----
// True if obj is MPI_FLOAT or MPI_COMPLEX.
bool is_single(PLA_Obj obj);
// A double-precision copy of obj's view (MPI_FLOAT -> MPI_DOUBLE, MPI_COMPLEX
// -> MPI_DOUBLE_COMPLEX, same object type and shape) and the rounding of such a
// copy back into the original view.
int single_promote(PLA_Obj obj, PLA_Obj *wide);
int single_demote(PLA_Obj wide, PLA_Obj obj);
// Complex scalar of a 1 x 1 object (real types give im = 0).
int scalar_value_z(PLA_Obj o, double *re, double *im);
// Runs call() with every single-precision object among *objs replaced by a
// double-precision copy (the pointers are swapped for the call and restored);
// the copies whose `out` flag is set are rounded back when call() returns a
// non-negative value.  Null pointers and other datatypes are left alone.
int on_promoted(std::initializer_list<PLA_Obj *> objs, std::initializer_list<bool> out,
                const std::function<int()> &call);
// ---- MPI_DOUBLE_COMPLEX BLAS3 (pla_zblas.cpp) ----
int zblas_gemm(int transa, int transb, PLA_Obj alpha, PLA_Obj A, PLA_Obj B, PLA_Obj beta, PLA_Obj C);
int zblas_trsm(int side, int uplo, int trans, int diag, PLA_Obj alpha, PLA_Obj A, PLA_Obj B);
int zblas_syrk(int uplo, int trans, PLA_Obj alpha, PLA_Obj A, PLA_Obj beta, PLA_Obj C);
void api_forget(PLA_Obj obj);  // drop a freed object from the open-object list
void api_release();            // free the staging buffers (PLA_Finalize)
int k_copy2d(hipStream_t s, int m, int n, const double *S, long lds_, double *D, long ldd,
             const int *abort_flag);
int k_scale2d(hipStream_t s, int m, int n, double alpha, double *D, long ldd);
int k_fill_drand48(hipStream_t s, int m, int n, double *D, long ldd, uint64_t state0, uint64_t skip);
struct lu_leaf_call {
  double *P;       // leaf top-left
  double *Ppanel;  // panel column 0 at the leaf's first row
  long lda;
  int mp, w;       // rows from the leaf's first row to the panel bottom, leaf width
  int pw, pc0;     // panel width, leaf's first column inside the panel
  int *piv;        // w pivot offsets (relative to each row)
  int *pairs;      // leaf_pairs_stride() ints: count, (dst, src) row pairs from the panel top
  int gcap;        // most workgroups the launch may use (0: the device's limit)
  int W;           // register leaf: compile-time width (8, 16 or 32), c.w <= W
};
// Panel-LU leaf of the kind ctx().leaf_kind selects.
int k_lu_leaf(hipStream_t s, const lu_leaf_call &c);
// The register leaf (pla_leaf.hip): one wave per workgroup, no LDS.
int k_lu_rleaf(hipStream_t s, const lu_leaf_call &c, int W);
// Widest register leaf (32, 16 or 8) for a panel of mp rows within gcap waves
// (0: no cap beyond the kernel's); 0 if none fits.
int rleaf_width_for(int mp, int gcap);
int rleaf_pairs_stride();
// Widest leaf (32, 16 or 8 columns) whose rows fit in LDS when a panel of mp
// rows is spread over at most gcap workgroups; 0 if none fits.
int leaf_width_for(int mp, int gcap);
constexpr int LU_MAX_LEAVES = 128;  // leaves of one panel (nb <= 1024, width >= 8)
// Compose the leaves' (dst, src) row lists of one panel into one list (comp,
// panel_comp_stride() ints) and apply it to the rows of a panel (A = its first
// row) in columns [0, nleft) and [c1, c2).
int k_panel_compose(hipStream_t s, const int *pairs, int nleaves, int *comp);
int k_panel_swap(hipStream_t s, double *A, long lda, int nleft, int c1, int c2, const int *comp);
int leaf_pairs_stride();
int panel_comp_stride();
// infinity norm (max absolute row sum) or one norm (max absolute column sum); synchronous
int k_norm(hipStream_t s, bool inf_norm, int m, int n, const double *A, long lda, double *result);
int leaf_max_width();
int trsm_leaf_size();

// ---- mixed precision and FP64-accurate solves (pla_mixed.hip) ----
int k_panel_swap_f32(hipStream_t s, float *A, long lda, int nleft, int c1, int c2, const int *comp);
// C (FP32) -= A B with A and B in the packed k-slice layout (pla_mixed.hip):
// A rows 0..m-1 (slice stride sa), B columns boff..boff+n-1 (slice stride sb);
// P (optional): also the updated m x n block of C in FP64 (leading dimension ldp)
int k_hgemm(hipStream_t s, int prec, int m, int n, int k, const void *Ah, const void *Al, long sa, const void *Bh,
            const void *Bl, long sb, long boff, float *C, long ldc, double *P = nullptr, long ldp = 0);
int k_cvt_f64_f32(hipStream_t s, int m, int n, const double *A, long lda, float *W, long ldw);
int k_cvt_f32_f64(hipStream_t s, int m, int n, const float *W, long ldw, double *P, long ldp);
int k_panel_out(hipStream_t s, int prec, int rows, int b, int kpad, const double *P, long ldp, void *Lh, void *Ll,
                long sl);
int k_upanel_out(hipStream_t s, int prec, int b, int kpad, int nr, const double *U, long ldu, float *W, long ldw,
                 void *Uh, void *Ul, long su, long col0);
// Sync-free blocked triangular solve op(A) X = X, X n x nrhs (nrhs <= 16).
// Dinv: inverted diagonal blocks (nullptr: computed into a workspace).
int trsv_sf(hipStream_t s, bool lower, bool trans, bool unit, int n, int nrhs, const double *A, long lda,
            double *X, long ldx, double *Dinv);
int trsv_sf_f32(hipStream_t s, bool lower, bool trans, bool unit, int n, int nrhs, const float *A, long lda,
                double *X, long ldx, double *Dinv);
int k_inv_diag_f32(hipStream_t s, bool lower, bool unit, int n, const float *A, long lda, double *Dinv);
int trsv_block_size();
int trsv_max_rhs();
int residual(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *X, long ldx,
             const double *B, long ldb, double *R, long ldr, double *rmax);
int col_amax(hipStream_t s, int n, int nrhs, const double *X, long ldx, double *out);
int k_permute_rows(hipStream_t s, int n, int nrhs, const int *perm, const double *X, long ldx, double *Y,
                   long ldy);
int k_axpy_cols(hipStream_t s, int n, int nrhs, const double *D, long ldd, double *X, long ldx);

// ---- blocked/recursive building blocks (pla_blas.cpp) ----
// B <- op(T)^{-1} B, T n x n triangular, B n x nrhs.
int trsm_left(hipStream_t s, bool lower, bool trans, bool unit, int n, int nrhs, const double *T,
              long ldt, double *B, long ldb, const int *abort_flag);
// X <- X inv(op(T)), X nrows x n, T n x n triangular.
int trsm_right(hipStream_t s, bool lower, bool trans, bool unit, int nrows, int n, const double *T, long ldt,
               double *X, long ldx, const int *abort_flag);
// X <- X inv(L)^T, X nrows x b, L b x b lower non-unit.
int trsm_rltn(hipStream_t s, int nrows, int b, const double *L, long ldl, double *X, long ldx,
              const int *abort_flag);
// Recursive panel LU of an mp x w panel (mp >= w); pivots relative to each row.
// The leaves apply their interchanges to all w columns of the panel and
// append their (dst, src) row pairs to `pairs`; returns the number of leaves
// in *nleaves.
int lu_panel(hipStream_t s, double *P, int mp, int w, long lda, int *piv, int *pairs, int *nleaves, int gcap);
// Recursive lower Cholesky of a b x b block; failures go to *info (1-based,
// offset added).
int potrf_block(hipStream_t s, int b, double *A, long lda, int offset, int *info);

// ---- Cholesky (pla_chol.cpp) ----
constexpr int CHOL_MAX_NB = 512;  // widest diagonal block of chol_factor
// Scratch of one factorization: info[0] first failure index, info[1] the
// trailing stream's snapshot of it; backup: the diagonal block being factored.
struct chol_scratch {
  int *info;
  double *backup;
  size_t backup_elems;
};
// Blocked lower Cholesky with look-ahead; *info_out = 0 or the 1-based index
// of the first non-positive pivot.  scr == nullptr: scratch from the grow-only
// workspaces, trailing updates timed when kernel timing is on; otherwise the
// caller's scratch, no allocation and no timing.
// Block size of a factorization chosen by PLA_Autotune_nb for order n.
void set_tuned_nb(int kind, int n, int nb);
int lu_factor(hipStream_t sMain, int m, int n, double *a, long lda, int *piv, int nb, bool swap_left);
// a(i, i) += v for i < nd
int k_shift_diag(hipStream_t s, int nd, double *a, long lda, double v);
int chol_factor(hipStream_t sM, int n, double *a, long lda, int nb, int *info_out,
                const chol_scratch *scr = nullptr);
// Device address of the no-allocation Cholesky's static scratch on `device`
// (looked up once per device; PLA_Init calls it for the library's device).
int chol_static_scratch(int device, void **ptr);

// ---- kernel timing (hipEvents on the library stream, per role) ----
// role 1: trailing update of PLA_LU (GEMM); role 2: trailing update of PLA_Chol (SYRK)
bool timing_on();
// ---- refinement (pla_krylov.hip, pla_refine.cpp, pla_fp8.hip) ----
int sym_residual(hipStream_t s, int n, int nrhs, const double *A, long lda, const double *X, long ldx,
                 const double *B, long ldb, double *R, long ldr, double *rmax, bool abs_mode = false);
int sym_inf_norm(hipStream_t s, int n, const double *A, long lda, double *out);
int k_fill(hipStream_t s, long count, double v, double *X);
int k_mdot(hipStream_t s, int n, int nrhs, int k, const double *V, long vstride, long ldv, const double *W,
           long ldw, double *out);
int k_maxpy(hipStream_t s, int n, int nrhs, int k, const double *V, long vstride, long ldv, const double *h,
            double alpha, double *W, long ldw);
int k_scale_cols(hipStream_t s, int n, int nrhs, double *W, long ldw, const double *f, bool inv);
int k_inv_diag_f32_op(hipStream_t s, bool lower, bool trans, bool unit, int n, const float *A, long lda,
                      double *Dinv);
int fp8_tile_rows();
int fp8_row_bytes(int b);  // bytes of one quantized row of a b-column panel
// exps: one exponent per 256 x 256 tile, row-tile major (fp8_row_bytes(b) / 256 per row tile)
int k_fp8_quantize(hipStream_t s, int R, int b, const double *P, long ldp, unsigned char *Lq, int *exps);
// C -= Lq[ro..] Lq[co..]^T over k < b (tri: lower triangle of an M x M block, ro == co)
int k_fp8_update(hipStream_t s, bool tri, int M, int N, int b, const unsigned char *Lq, const int *exps, int ro,
                 int co, float *C, long ldc, const int *abort_flag);
int k_fp8_syrk_lower(hipStream_t s, int R, const unsigned char *Lq, const int *exps, float *C, long ldc,
                     const int *abort_flag);
int k_mirror_lower_f32(hipStream_t s, int n, float *W, long ldw);
// S (n x n) <- the full symmetric matrix whose lower triangle A holds
int k_sym_full(hipStream_t s, int n, const double *A, long lda, double *S, long lds);
int k_max_diag(hipStream_t s, int n, const double *A, long lda, double *out);
int k_eye(hipStream_t s, int n, double *T, long ldt);  // T <- I (n x n)
int k_cvt_lower_shift(hipStream_t s, int n, const double *A, long lda, double shift, float *W, long ldw);

// Refinement of X for A X = B with a low-precision factorization M:
// residual(X, B, R, rmax) computes R = B - A X in FP64 (B == nullptr: -A X;
// rmax == nullptr: no norms, asynchronous); precond(Y, Z) computes
// Z = M^{-1} Y.  Classic refinement first; with gmres_fallback, GMRES-IR
// (restart steps, relative tolerance gmres_tol) once a classic step fails
// to halve the backward error.
struct refine_ops {
  int n = 0, nrhs = 0;
  int restart = 50;
  double gmres_tol = 1e-10;
  bool gmres_first = false;  // GMRES-IR from the first correction, no classic steps
  std::function<int(const double *X, long ldx, const double *B, long ldb, double *R, long ldr, double *rmax)>
      residual;
  std::function<int(const double *Y, long ldy, double *Z, long ldz)> precond;
};
struct refine_stats {
  int outer = 0;        // refinement steps (corrections applied)
  int inner = 0;        // GMRES steps in total
  int gmres_outer = 0;  // corrections computed by GMRES
  double eta = INFINITY;
};
int refine(hipStream_t s, const refine_ops &op, double anorm, const double *B, long ldb, double *X, long ldx,
           double tol, int max_iter, bool gmres_fallback, refine_stats *st);
void set_last_refine(const refine_stats &st, double shift);
// PLA_SUCCESS when st.eta <= tol, else PLA_ERR_NOT_CONVERGED with a message
// (a NaN backward error also fails the comparison)
int refine_outcome(const char *who, const refine_stats &st, double tol);

void timing_begin(hipStream_t s, int role);
void timing_release();  // destroys pending timing events, timing off (PLA_Finalize)
void timing_end(hipStream_t s, int role, double flops);

}  // namespace pla
