/*
 * PLA.h -- drop-in PLAPACK R3.0 C API for libpla_mi355x (AMD Instinct
 * MI355X, gfx950), restricted to the dense factor-and-solve path on a 1x1
 * process template.
 *
 * Names, argument order and meaning, and constant values follow the
 * reference headers (INCLUDE/PLA_define.h:11-102, INCLUDE/PLA_Obj.h,
 * INCLUDE/PLA_Template.h) so that drivers such as TEST/test_lu.c compile
 * unmodified.  PLA_Obj buffers live in device memory (HBM); host access
 * through PLA_Obj_local_buffer goes through a pinned host mirror that is
 * synchronised at every library entry and exit (see PLA_Obj_local_buffer).
 *
 * Every entry point below cites the reference routine it replaces.
 * Entry points marked EXTENSION have no counterpart in the reference.
 *
 * Error behaviour: functions return PLA_SUCCESS (0) or a negative error
 * code; PLA_Last_error() gives the message.  No entry point ends the
 * process except PLA_Abort, which a caller invokes itself (the reference's
 * routines abort on bad input, OBJ/PLA_Diagnostics.c:108-129).
 * Not thread-safe, like the reference.
 *
 * Datatypes: every routine computes in MPI_DOUBLE or MPI_DOUBLE_COMPLEX.
 * MPI_FLOAT and MPI_COMPLEX objects are accepted wherever the reference
 * accepts them: they are widened into FP64 temporaries, the FP64 routine
 * runs, and the results are rounded once back into the objects (so they
 * are at least as accurate as single-precision arithmetic would give).
 */
#ifndef PLA_MI355X_H
#define PLA_MI355X_H

#ifdef PLA_NO_MPI_H
#include "PLA_mpi_abi.h"
#else
#include <mpi.h>
#endif

#include <stdint.h>
/* the standard headers the reference's INCLUDE/PLA.h brings in */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>

#ifndef NULL
#define NULL 0
#endif

/* Machine identifiers (INCLUDE/PLA_Machines.h); drivers test
   MANUFACTURE == CRAY, so both must be defined.  The reference's build passes
   MANUFACTURE and MACHINE_TYPE on the command line; a PC running Linux is the
   default here. */
#undef CRAY
#define CRAY  10
#define IBM   20
#define INTEL 30
#define SUN   40
#define PC    50
#define SGI   60
#define HP    70
#define CRAYPVP 100
#define CRAYMPP 101
#define PARAGON 200
#define SP2     300
#define SUNOS   400
#define LINUX   500
#define NT      600
#ifndef MANUFACTURE
#define MANUFACTURE PC
#endif
#ifndef MACHINE_TYPE
#define MACHINE_TYPE LINUX
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Element types of complex objects (INCLUDE/PLA_Misc.h:11-17). */
typedef struct PLA_COMPLEX {
  float real, imaginary;
} PLA_COMPLEX;
typedef struct PLA_DOUBLE_COMPLEX {
  double real, imaginary;
} PLA_DOUBLE_COMPLEX;

/* ---- constants: values identical to INCLUDE/PLA_define.h:11-102 ---- */
#define PLA_OBJTYPE_OFFSET 10
#define PLA_VECTOR   (PLA_OBJTYPE_OFFSET + 1)
#define PLA_MVECTOR  (PLA_OBJTYPE_OFFSET + 2)
#define PLA_MATRIX   (PLA_OBJTYPE_OFFSET + 3)
#define PLA_MSCALAR  (PLA_OBJTYPE_OFFSET + 4)
#define PLA_PVECTOR  (PLA_OBJTYPE_OFFSET + 5)
#define PLA_PMVECTOR (PLA_OBJTYPE_OFFSET + 6)

#define PLA_TRANSPOSE_OFFSET 20
#define PLA_NO_TRANSPOSE        (PLA_TRANSPOSE_OFFSET + 1)
#define PLA_NO_TRANS            (PLA_TRANSPOSE_OFFSET + 1)
#define PLA_TRANSPOSE           (PLA_TRANSPOSE_OFFSET + 2)
#define PLA_TRANS               (PLA_TRANSPOSE_OFFSET + 2)
#define PLA_CONJUGATE_TRANSPOSE (PLA_TRANSPOSE_OFFSET + 3)
#define PLA_CONJ_TRANS          (PLA_TRANSPOSE_OFFSET + 3)
#define PLA_CONJUGATE           (PLA_TRANSPOSE_OFFSET + 4)
#define PLA_CONJ                (PLA_TRANSPOSE_OFFSET + 4)

#define PLA_OWNER_OFFSET -30
#define PLA_ALL_ROWS (PLA_OWNER_OFFSET - 1)
#define PLA_ALL_COLS (PLA_OWNER_OFFSET - 2)

#define PLA_SIDE_OFFSET 40
#define PLA_SIDE_TOP    (PLA_SIDE_OFFSET + 1)
#define PLA_SIDE_BOTTOM (PLA_SIDE_OFFSET + 2)
#define PLA_SIDE_LEFT   (PLA_SIDE_OFFSET + 3)
#define PLA_SIDE_RIGHT  (PLA_SIDE_OFFSET + 4)

#define PLA_SHAPE_OFFSET 50
#define PLA_UPPER_TRIANGULAR        (PLA_SHAPE_OFFSET + 1)
#define PLA_UP_TRIAN                (PLA_SHAPE_OFFSET + 1)
#define PLA_LOWER_TRIANGULAR        (PLA_SHAPE_OFFSET + 2)
#define PLA_LOWER_TRIAN             (PLA_SHAPE_OFFSET + 2)
#define PLA_SHAPE_GENERAL           (PLA_SHAPE_OFFSET + 3)
#define PLA_SHAPE_LOWER_TRAPEZOIDAL (PLA_SHAPE_OFFSET + 4)
#define PLA_SHAPE_LOW_TRAP          (PLA_SHAPE_OFFSET + 4)
#define PLA_SHAPE_UPPER_TRAPEZOIDAL (PLA_SHAPE_OFFSET + 5)
#define PLA_SHAPE_UP_TRAP           (PLA_SHAPE_OFFSET + 5)

#define PLA_FLOW_OFFSET 60
#define PLA_DIR_NOFLOW (PLA_FLOW_OFFSET + 1)
#define PLA_DIR_UP     (PLA_FLOW_OFFSET + 2)
#define PLA_DIR_DOWN   (PLA_FLOW_OFFSET + 3)
#define PLA_DIR_LEFT   (PLA_FLOW_OFFSET + 4)
#define PLA_DIR_RIGHT  (PLA_FLOW_OFFSET + 5)
#define PLA_TEMP_COL   (PLA_FLOW_OFFSET + 6)
#define PLA_TEMP_ROW   (PLA_FLOW_OFFSET + 7)

#define PLA_OP_OFFSET 80
#define PLA_OP_MAT_PAN       (PLA_OP_OFFSET + 1)
#define PLA_OP_PAN_MAT       (PLA_OP_OFFSET + 2)
#define PLA_OP_PAN_PAN       (PLA_OP_OFFSET + 3)
#define PLA_OP_PAN_TRIAN_MAT (PLA_OP_OFFSET + 4)
#define PLA_OP_SYM_PAN_PAN   (PLA_OP_OFFSET + 5)
#define PLA_OP_TRIAN_MAT_PAN (PLA_OP_OFFSET + 6)
#define PLA_OP_ALL_ALG       (PLA_OP_OFFSET + 7)

#define PLA_API_OFFSET 70
#define PLA_MODE_CLOSED      (PLA_API_OFFSET + 1)
#define PLA_MODE_OPEN        (PLA_API_OFFSET + 2)
#define PLA_STATE_API_ACTIVE (PLA_API_OFFSET + 3)
#define PLA_STATE_NORMAL     (PLA_API_OFFSET + 4)

#define PLA_PROJ_OFFSET 90
#define PLA_PROJ_ONTO_COL (PLA_PROJ_OFFSET + 1)
#define PLA_PROJ_ONTO_ROW (PLA_PROJ_OFFSET + 2)

#define PLA_DIAG_OFFSET 100
#define PLA_UNIT_DIAG    (PLA_DIAG_OFFSET + 1)
#define PLA_NONUNIT_DIAG (PLA_DIAG_OFFSET + 2)
#define PLA_ZERO_DIAG    (PLA_DIAG_OFFSET + 3)

#define PLA_DUMMY        0
#define PLA_SUCCESS      0
#define PLA_ALIGN_FIRST  -200
#define PLA_DIM_ALL      -201
#define PLA_INHERIT      -202
#define PLA_UNDEFINED    -203

#ifndef TRUE
#define TRUE 1
#endif
#ifndef FALSE
#define FALSE 0
#endif
#ifndef min
#define min(x, y) ((x) < (y) ? (x) : (y))
#endif
#ifndef dabs
#define dabs(x) ((x) < 0 ? -(x) : (x))
#endif

/* INCLUDE/PLA_Versions.h:11-13 (PLA_Gemm_x versions), INCLUDE/PLA_Misc.h:43-46 */
#define PLA_VERSION_A 100
#define PLA_VERSION_B 200
#define PLA_VERSION_C 300
#define PLA_METHOD_INV     0
#define PLA_METHOD_FACTORS 1
#define PLA_METHOD_STABLE  2
#define PLA_METHOD_TRSM    3

/* Negative error codes (the reference aborts instead). */
#define PLA_ERR_NOT_INITIALIZED -1001
#define PLA_ERR_INVALID_OBJECT  -1002
#define PLA_ERR_INVALID_PARAM   -1003
#define PLA_ERR_UNSUPPORTED     -1004
#define PLA_ERR_DIMENSIONS      -1005
#define PLA_ERR_DEVICE          -1006
#define PLA_ERR_NO_MEMORY       -1007
#define PLA_ERR_MULTI_PROCESS   -1008
#define PLA_ERR_NOT_CONVERGED   -1009  /* refinement ended above its tolerance */

/* ---- opaque handles (INCLUDE/PLA_Obj.h:82, INCLUDE/PLA_Template.h:31) ---- */
typedef struct PLA_Obj_view_struct *PLA_Obj;
typedef struct PLA_Template_struct *PLA_Template;

/* Checking flags (OBJ/PLA_Diagnostics.c:16-19). */
extern int PLA_ERROR_CHECKING;
extern int PLA_CHECK_PARAMETERS;
extern int PLA_CHECK_AGAINST_SEQUENTIAL;
extern int PLA_CHECK_AGAINST_R12;

/* ---- initialisation and templates ---- */
/* TEMPLATE/PLA_Init.c:20.  Only a 1x1 mesh is supported: rows*cols != 1
   returns PLA_ERR_MULTI_PROCESS.  comm2d receives comm1d. */
int PLA_Comm_1D_to_2D(MPI_Comm comm1d, int rows, int cols, MPI_Comm *comm2d);
/* TEMPLATE/PLA_Init.c:38: the mesh whose rows / cols is closest to ratio.
   On one rank that is 1x1 whatever the ratio: comm2d receives comm1d; a
   communicator of more ranks returns PLA_ERR_MULTI_PROCESS. */
int PLA_Comm_1D_to_2D_ratio(MPI_Comm comm1d, double ratio, MPI_Comm *comm2d);
/* TEMPLATE/PLA_Init.c:85.  Also selects and initialises the HIP device. */
int PLA_Init(MPI_Comm comm);
/* TEMPLATE/PLA_Init.c:132 */
int PLA_Finalize(void);
/* TEMPLATE/PLA_Init.c:113 */
int PLA_Initialized(int *initialized);
/* TEMPLATE/PLA_Temp.c:19 */
int PLA_Temp_create(int nb, int zero_or_one, PLA_Template *templ);
/* TEMPLATE/PLA_Temp.c:136 */
int PLA_Temp_free(PLA_Template *templ);
/* TEMPLATE/PLA_Temp.c:176, :269, :169, :197, :205 */
int PLA_Temp_nb(PLA_Template templ, int *nb);
int PLA_Temp_zero_or_one(PLA_Template templ, int *zero_or_one);
int PLA_Temp_comm_all_size(PLA_Template templ, int *nprocs);
int PLA_Temp_comm_all(PLA_Template templ, MPI_Comm *comm);
int PLA_Temp_comm_all_rank(PLA_Template templ, int *rank);

/* TEMPLATE/PLA_Environ.c:88 (lower-case name is the one the reference links) */
int pla_Environ_set_nb_alg(int type, int nb_alg);
/* TEMPLATE/PLA_Environ.c:54 */
int PLA_Environ_nb_alg(int type, PLA_Template templ, int *nb_alg);

/* OBJ/PLA_Diagnostics.c:402, :436 */
int PLA_Set_error_checking(int error, int parameters, int sequential, int r12);
int PLA_Get_error_checking(int *error, int *parameters, int *sequential, int *r12);
/* OBJ/PLA_Diagnostics.c:108.  Called by a program to give up: prints the
   message, file and line to stderr and ends the process with status 1 --
   through MPI_Abort when the program has initialised MPI, else exit(1) (the
   reference exits with status 0).  The library itself never calls it. */
int PLA_Abort(char *message, int line, char *file);

/* OBJ/PLA_Memory.c:18, :51, :68 -- host memory of the reference's drivers:
   malloc / calloc(size * typesize) / free.  PLA_malloc and PLA_calloc return
   NULL when the host is out of memory (the reference exits); PLA_free returns
   NULL. */
void *PLA_malloc(size_t size);
void *PLA_calloc(size_t size, size_t typesize);
void *PLA_free(void *buffer);

/* ---- objects (OBJ/PLA_Obj.c) ---- */
/* :323 */
int PLA_Matrix_create(MPI_Datatype datatype, int global_length, int global_width,
                      PLA_Template templ, int global_align_row, int global_align_col,
                      PLA_Obj *new_obj);
/* :149 */
int PLA_Mvector_create(MPI_Datatype datatype, int global_length, int global_width,
                       PLA_Template templ, int global_align, PLA_Obj *new_obj);
/* :497 */
int PLA_Mscalar_create(MPI_Datatype datatype, int owner_row, int owner_col, int length,
                       int width, PLA_Template templ, PLA_Obj *new_obj);
/* :3985, :3720, :4025 */
int PLA_Matrix_create_conf_to(PLA_Obj obj, PLA_Obj *new_obj);
int PLA_Mvector_create_conf_to(PLA_Obj obj, int global_width, PLA_Obj *new_obj);
int PLA_Mscalar_create_conf_to(PLA_Obj obj, int owner_row, int owner_col, PLA_Obj *new_obj);
/* :4189 -- duplicated 1x1 constants of obj's datatype; NULL skips one */
int PLA_Create_constants_conf_to(PLA_Obj obj, PLA_Obj *minus_one, PLA_Obj *zero, PLA_Obj *one);
/* :803 */
int PLA_Obj_free(PLA_Obj *obj);

/* :831-1430 queries */
int PLA_Obj_objtype(PLA_Obj obj, int *objtype);
int PLA_Obj_datatype(PLA_Obj obj, MPI_Datatype *datatype);
int PLA_Obj_template(PLA_Obj obj, PLA_Template *templ);
int PLA_Obj_global_length(PLA_Obj obj, int *global_length);
int PLA_Obj_global_width(PLA_Obj obj, int *global_width);
int PLA_Obj_owner_row(PLA_Obj obj, int *owner_row);
int PLA_Obj_owner_col(PLA_Obj obj, int *owner_col);
int PLA_Obj_local_length(PLA_Obj obj, int *local_length);
int PLA_Obj_local_width(PLA_Obj obj, int *local_width);
int PLA_Obj_local_ldim(PLA_Obj obj, int *local_ldim);
int PLA_Obj_local_stride(PLA_Obj obj, int *local_stride);
/* :1350.  Returns a HOST pointer into a pinned mirror of the object's
   device buffer.  The mirror is filled from the device on first use, and
   it is treated as possibly written by the caller from then on: every
   library entry that reads the object first copies the mirror to the
   device, and every entry that writes the object copies the result back
   into the mirror before returning. */
int PLA_Obj_local_buffer(PLA_Obj obj, void **local_buffer);
/* :1437, :1545 -- host <-> device copies of the local contents */
int PLA_Obj_get_local_contents(PLA_Obj obj, int trans, int *rows_in_buf, int *cols_in_buf,
                               void *buf, int leading_dim_buf, int stride_buf);
int PLA_Obj_set_local_contents(int trans, int rows_in_buf, int cols_in_buf, void *buf,
                               int leading_dim_buf, int stride_buf, PLA_Obj obj);
/* :1807, :1838, :1870 */
int PLA_Obj_set_to_zero(PLA_Obj obj);
int PLA_Obj_set_to_one(PLA_Obj obj);
int PLA_Obj_set_to_minus_one(PLA_Obj obj);

/* :1981, :2063, :3122, :3156 -- views share the base object */
int PLA_Obj_view_all(PLA_Obj old_obj, PLA_Obj *new_obj);
int PLA_Obj_split_4(PLA_Obj obj, int length, int width, PLA_Obj *upper_left,
                    PLA_Obj *upper_right, PLA_Obj *lower_left, PLA_Obj *lower_right);
int PLA_Obj_horz_split_2(PLA_Obj obj, int length, PLA_Obj *upper, PLA_Obj *lower);
int PLA_Obj_vert_split_2(PLA_Obj obj, int width, PLA_Obj *left, PLA_Obj *right);

/* COPY/PLA_Copy.c:18 -- same-shape copy between any two objects */
int PLA_Copy(PLA_Obj from, PLA_Obj to);

/* ---- application interface (API/PLA_API.c; prototypes as in
   INCLUDE/PLA_API_prototypes.h).  Host buffers are added into objects in HBM
   (*_to_global) or pieces of objects added into host buffers (global_to_*).
   The objects must be PLA_MATRIX or PLA_MVECTOR of MPI_DOUBLE,
   MPI_DOUBLE_COMPLEX, MPI_FLOAT or MPI_INT, opened with PLA_Obj_API_open inside
   PLA_API_begin / PLA_API_end.
   A *_to_global call may return before the add is done (the buffer is
   already copied and may be reused); PLA_Obj_API_sync / _close / PLA_API_end
   complete it.  A global_to_* call returns with the buffer updated.  alpha
   (a pointer to one element of the object's type; NULL means 1) scales the
   added values -- the reference ignores it and always adds with 1.  Vector
   forms address column zero_or_one of the template, as the reference does
   (API/PLA_API.c:247,285); local increments other than 1 are accepted. ---- */
int PLA_API_begin(void);                                   /* API/PLA_API.c:24 */
int PLA_API_end(void);                                     /* API/PLA_API.c:61 */
int PLA_API_state(int *state);                             /* API/PLA_API.c:109 */
int PLA_Obj_API_open(PLA_Obj obj);                         /* API/PLA_API.c:126 */
int PLA_Obj_API_close(PLA_Obj obj);                        /* API/PLA_API.c:166 */
int PLA_Obj_API_mode(PLA_Obj obj, int *mode);              /* API/PLA_API.c:200 */
int PLA_Obj_API_sync(PLA_Obj obj);                         /* API/PLA_API.c:916 */
/* API/PLA_API.c:220 */
int PLA_API_axpy_vector_to_global(int size, void *alpha, void *local_vector, int local_inc, PLA_Obj laobj,
                                  int disp);
/* API/PLA_API.c:257 */
int PLA_API_axpy_global_to_vector(int size, void *alpha, PLA_Obj laobj, int disp, void *local_vector,
                                  int local_inc);
/* API/PLA_API.c:295 */
int PLA_API_multi_axpy_vector_to_global(int nsub, int *sizes, void *alpha, void *local_vector, int local_stride,
                                        PLA_Obj laobj, int *displs);
/* API/PLA_API.c:343 */
int PLA_API_multi_axpy_global_to_vector(int nsub, int *sizes, void *alpha, PLA_Obj laobj, int *displs,
                                        void *local_vector, int local_stride);
/* API/PLA_API.c:390 */
int PLA_API_axpy_matrix_to_global(int m, int n, void *alpha, void *local_buf, int local_ldim, PLA_Obj A,
                                  int disp_row, int disp_col);
/* API/PLA_API.c:593 */
int PLA_API_axpy_global_to_matrix(int m, int n, void *alpha, PLA_Obj A, int disp_row, int disp_col,
                                  void *local_buf, int local_ldim);
/* API/PLA_API.c:797 */
int PLA_API_multi_axpy_matrix_to_global(int nsub_row, int nsub_col, int *size_row, int *size_col, void *alpha,
                                        void *local_matrix, int local_ldim, PLA_Obj obj, int *disp_row,
                                        int *disp_col);
/* API/PLA_API.c:856 */
int PLA_API_multi_axpy_global_to_matrix(int nsub_row, int nsub_col, int *size_row, int *size_col, void *alpha,
                                        PLA_Obj obj, int *disp_row, int *disp_col, void *local_matrix,
                                        int local_ldim);

/* ---- factor and solve path ---- */
/* SOLVERS/LU/PLA_LU.c:13.  A: m x n MPI_DOUBLE, MPI_DOUBLE_COMPLEX, MPI_FLOAT
   or MPI_COMPLEX matrix
   (complex pivots: the first maximum of |re| + |im|, izamax's rule),
   overwritten with L (unit, below the diagonal) and U; pivots: MPI_INT mvector of length
   min(m,n) receiving, for each row j, the 0-based offset of its pivot row
   relative to row j.  Returns PLA_SUCCESS; like the reference, a zero pivot
   is not flagged (the factors then hold inf/nan). */
int PLA_LU(PLA_Obj A, PLA_Obj pivots);
/* SOLVERS/LINPACK/PLA_LU_linpack.c:13: PLA_LU whose interchanges reach only
   the columns from each panel on -- the multipliers of a panel keep the row
   order of their own step (LINPACK's dgefa layout); on [A b] (n x (n+1)) the
   last column ends as inv(L) P b, ready for PLA_Trsv with U */
int PLA_LU_linpack(PLA_Obj A, PLA_Obj pivots);
/* SOLVERS/Cholesky/PLA_Chol.c:13.  MPI_DOUBLE or MPI_FLOAT (A = L L^T), or
   MPI_DOUBLE_COMPLEX (A = L L^H, zero imaginary diagonal).  Lower only (the
   reference aborts on upper).  Returns 0, or the 1-based global index k+i+1 of the first
   non-positive pivot (PLA_Chol.c:181-185); the diagonal block that failed
   is left as it was before its factorisation, as in the reference. */
int PLA_Chol(int uplo, PLA_Obj A);
/* SOLVERS/Cholesky/PLA_Chol_simple.c:13: lower Cholesky with block size
   nb_alg (<= 0: A unchanged); returns 0 or the 1-based index of the first
   non-positive pivot, as PLA_Chol (the reference's variant always returns 0) */
int PLA_Chol_simple(int nb_alg, PLA_Obj A);
/* SOLVERS/LU/PLA_Apply_pivots_new.c:13 -- forward sequential row swaps */
int PLA_Apply_pivots_to_rows(PLA_Obj A, PLA_Obj pivots);
/* BLAS2/PLA_Trsv.c:15 */
int PLA_Trsv(int uplo, int trans, int diag, PLA_Obj A, PLA_Obj x);
/* BLAS3/PLA_Trsm.c:13: B <- alpha op(A)^-1 B (PLA_SIDE_LEFT) or
   alpha B op(A)^-1 (PLA_SIDE_RIGHT), every uplo / trans / diag; real or
   complex (trans also PLA_CONJUGATE_TRANSPOSE and PLA_CONJUGATE) */
int PLA_Trsm(int side, int uplo, int trans, int diag, PLA_Obj alpha, PLA_Obj A, PLA_Obj B);
/* BLAS2/PLA_Gemv.c:13 */
int PLA_Gemv(int transa, PLA_Obj alpha, PLA_Obj A, PLA_Obj x, PLA_Obj beta, PLA_Obj y);
/* BLAS3/PLA_Gemm.c:13: real or complex; for complex operands transa/transb
   also take PLA_CONJUGATE_TRANSPOSE and PLA_CONJUGATE (conj without
   transpose) */
int PLA_Gemm(int transa, int transb, PLA_Obj alpha, PLA_Obj A, PLA_Obj B, PLA_Obj beta,
             PLA_Obj C);
/* BLAS3/PLA_Gemm_x.c:13: the reference's distributed algorithm variants
   (PLA_VERSION_A / _B / _C) of PLA_Gemm; on one GPU every version is
   PLA_Gemm.  An unknown version returns PLA_ERR_INVALID_PARAM. */
int PLA_Gemm_x(int version, int transa, int transb, PLA_Obj alpha, PLA_Obj A, PLA_Obj B, PLA_Obj beta,
               PLA_Obj C);
/* BLAS1/PLA_Axpy.c:13: y <- alpha x + y for x, y of one shape and datatype
   (MPI_DOUBLE, MPI_FLOAT or MPI_DOUBLE_COMPLEX) and alpha a 1x1 object of
   that datatype */
int PLA_Axpy(PLA_Obj alpha, PLA_Obj x, PLA_Obj y);
/* BLAS1/PLA_Nrm2.c:13: alpha <- ||x||_2 of a vector x (one column, or one
   row) of MPI_DOUBLE, MPI_FLOAT or MPI_DOUBLE_COMPLEX; alpha is a 1x1 object
   of the real datatype (MPI_DOUBLE, or MPI_FLOAT for MPI_FLOAT x) */
int PLA_Nrm2(PLA_Obj x, PLA_Obj alpha);
/* BLAS2/PLA_Symv.c:13, BLAS2/PLA_Hemv.c:13: y <- alpha A x + beta y with A
   symmetric (MPI_DOUBLE, MPI_FLOAT) or Hermitian (MPI_DOUBLE_COMPLEX; the
   imaginary part of the diagonal is taken as zero), only the uplo triangle of
   A read; x and y vectors (one column) */
int PLA_Symv(int uplo, PLA_Obj alpha, PLA_Obj A, PLA_Obj x, PLA_Obj beta, PLA_Obj y);
int PLA_Hemv(int uplo, PLA_Obj alpha, PLA_Obj A, PLA_Obj x, PLA_Obj beta, PLA_Obj y);
/* UTIL/PLA_Shift.c:15: add *value (MPI_DOUBLE or MPI_FLOAT, converted to
   obj's datatype) to the diagonal of obj (MPI_DOUBLE or MPI_FLOAT) */
int PLA_Shift(PLA_Obj obj, MPI_Datatype datatype, void *value);
/* BLAS3/PLA_Syrk.c:13: C <- beta C on all of C (PLA_Syrk_panpan.c:22), then
   the uplo triangle += alpha op(A) op(A)^T (op N or T; complex objects give
   the complex-symmetric update, no conjugation) */
int PLA_Syrk(int uplo, int trans, PLA_Obj alpha, PLA_Obj A, PLA_Obj beta, PLA_Obj C);
/* BLAS2/PLA_Matrix_infinity_norm.c:13, BLAS2/PLA_Matrix_one_norm.c:13 --
   alpha (1x1 MPI_DOUBLE) <- max absolute row / column sum of A */
int PLA_Matrix_infinity_norm(PLA_Obj A, PLA_Obj alpha);
int PLA_Matrix_one_norm(PLA_Obj A, PLA_Obj alpha);
/* SOLVERS/LU/PLA_General.c:67 -- factor A, then solve A X = B in place */
int PLA_General_solve(PLA_Obj A, PLA_Obj ipiv, PLA_Obj B);
/* SOLVERS/Cholesky/PLA_Pos_def.c:44 -- factor A = L L^T, then solve in place */
int PLA_Pos_def_solve(PLA_Obj A, PLA_Obj B);

/* ---- EXTENSIONS (no counterpart in the reference) ---- */
/* Mixed-precision factor-and-solve with FP64 iterative refinement.
   A (n x n MPI_DOUBLE) is left unchanged; an FP32 copy is factored with
   partial pivoting (pivots as PLA_LU's) by the blocked algorithm of
   SOLVERS/LU/PLA_LU.c:37-212 with FP64 panels and a low-precision MFMA
   trailing update (FP32 accumulation):
     PLA_MIXED_FP16 / PLA_MIXED_BF16     one product of half-precision operands;
     PLA_MIXED_FP16X3 / PLA_MIXED_BF16X3 operands split into hi + lo halves,
                                          hi*hi + hi*lo + lo*hi.
   X (n x nrhs, nrhs <= 16) receives the solution of A X = B, refined in FP64
   until the normwise backward error
     max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
   is <= tol or max_iter refinement steps were taken.  *iterations and
   *backward_error (either may be NULL) report the outcome.  Returns
   PLA_ERR_NOT_CONVERGED when tol was not reached (or the backward error is
   not a number); X, *iterations and *backward_error then hold the last
   iterate and its figures. */
#define PLA_MIXED_FP16   1
#define PLA_MIXED_BF16   2
#define PLA_MIXED_FP16X3 3
#define PLA_MIXED_BF16X3 4
int PLA_LU_mixed(PLA_Obj A, PLA_Obj pivots, PLA_Obj B, PLA_Obj X, int precision, double tol, int max_iter,
                 int *iterations, double *backward_error);
/* Batched factor-and-solve of `count` independent SPD systems of uneven sizes
   A_i X_i = B_i (the mathematics of PLA_Pos_def_solve, SOLVERS/Cholesky/
   PLA_Pos_def.c:44-99, for thousands of small systems at once).  One
   persistent kernel solves every system; a plan captures it in a hipGraph
   and replays it.
   Storage (all device pointers, HBM): A_i is n[i] x n[i] column-major with
   leading dimension n[i] at A + a_off[i]; B_i is n[i] x nrhs (ld n[i]) at
   B + b_off[i]; NULL offsets mean packed back to back in index order.  Only
   the lower triangle of each A_i is read.  nrhs <= 8.
   mode PLA_BATCHED_FP64:    A_i is overwritten by L_i (lower), B_i by X_i.
   mode PLA_BATCHED_FP32_IR: A_i is left unchanged; FP32 factors, then FP64
     refinement until max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf +
     ||b_j||_inf) <= tol or max_iter steps; B_i is overwritten by X_i.
   OR-ing PLA_BATCHED_AUTOTUNE into mode times every kernel variant (block
   size, workgroups per CU) on scratch copies and keeps the fastest.
   info[i] (device, int): 0, or the 1-based index of the first non-positive
   pivot of system i (then B_i is left unchanged).  iters[i] / berr[i]
   (device, may be NULL): refinement steps and final backward error
   (FP32_IR only).  The plan keeps the pointers: execute replays on them. */
#define PLA_BATCHED_FP64     1
#define PLA_BATCHED_FP32_IR  2
#define PLA_BATCHED_AUTOTUNE 0x100
typedef struct PLA_Batched_plan_struct *PLA_Batched_plan;
int PLA_Batched_plan_create(int count, const int *n, const long *a_off, const long *b_off, int nrhs,
                            int mode, double tol, int max_iter, double *A, double *B, int *info,
                            int *iters, double *berr, PLA_Batched_plan *plan);
/* Enqueue one replay of the plan's graph on the library stream (asynchronous;
   PLA_Device_synchronize waits for it). */
int PLA_Batched_plan_execute(PLA_Batched_plan plan);
/* Autotune an existing plan (see PLA_BATCHED_AUTOTUNE); *ms_best may be NULL. */
int PLA_Batched_plan_autotune(PLA_Batched_plan plan, double *ms_best);
/* The plan's block size and workgroups per CU. */
int PLA_Batched_plan_config(PLA_Batched_plan plan, int *nb, int *workgroups_per_cu);
int PLA_Batched_plan_free(PLA_Batched_plan *plan);
/* create + execute + wait + free */
int PLA_Chol_batched(int count, const int *n, const long *a_off, const long *b_off, int nrhs, int mode,
                     double tol, int max_iter, double *A, double *B, int *info, int *iters, double *berr);
/* OBJ/PLA_Obj_management.c:15 -- attach a caller's DEVICE (HBM) buffer with
   leading dimension ldim to obj, which must view its whole base object.
   Objects here always own a buffer, so the old one is released (the
   reference requires a base without a buffer).  user_buffer != 0: the
   buffer is not freed by PLA_Obj_free. */
int PLA_Obj_attach_buffer(PLA_Obj obj, void *buffer, int local_ldim, int user_buffer);
/* Mixed-precision Cholesky factor-and-solve with FP64 refinement.  Only the
   lower triangle of A (n x n MPI_DOUBLE, symmetric positive definite) is read,
   and A is left unchanged.  An FP32 copy is factored by the blocked algorithm
   of SOLVERS/Cholesky/PLA_Chol.c:83-234 with nb = 256: FP64 diagonal blocks
   and panels, and the trailing lower(A22) -= L21 L21^T on FP8 MFMA --
     PLA_MIXED_FP8: L21 in OCP E4M3 with one power-of-two scale per 256 x 256
                    tile, exact products, FP32 accumulation.
   If the FP8-updated factorization meets a non-positive pivot it is repeated
   for A + shift I with a growing shift (reported by PLA_Mixed_stats).
   X (n x nrhs, nrhs <= 16) is then refined in FP64 against A: classic
   refinement, switching to GMRES-IR preconditioned by the factors when a
   step fails to halve the normwise backward error, until it is <= tol or
   max_iter corrections were applied.  PLA_LU_mixed refines the same way,
   and returns PLA_ERR_NOT_CONVERGED the same way. */
#define PLA_MIXED_FP8    5
int PLA_Chol_mixed(PLA_Obj A, PLA_Obj B, PLA_Obj X, int precision, double tol, int max_iter,
                   int *iterations, double *backward_error);
/* Statistics of the last PLA_LU_mixed / PLA_Chol_mixed call: corrections
   applied, GMRES steps in total, corrections computed by GMRES, and the
   diagonal shift of the factored matrix (0 for LU).  Any pointer may be NULL. */
int PLA_Mixed_stats(int *outer, int *gmres_steps, int *gmres_corrections, double *shift);
/* No counterpart in the reference.  PLA_Chol (SOLVERS/Cholesky/PLA_Chol.c:13)
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc): the same
   kernels and results as PLA_Chol on MPI_DOUBLE, with the factorization's
   scratch (two info words and the backup of the diagonal block being
   factored) in one static device array of 2 MiB + 256 B per device instead
   of library-allocated workspaces.  The block size is PLA_Chol's, capped at
   512 as in PLA_Chol, whose 512 x 512 backup is the array's limit (a block
   the array cannot hold is rejected with PLA_ERR_DIMENSIONS).  Returns 0 or
   the 1-based index of the first non-positive pivot, as PLA_Chol.  The info
   words are cleared at the start of every call; calls that overlap on one
   device share that array and race. */
int PLA_Chol_no_alloc(int uplo, PLA_Obj A);
/* Message of the last error returned by any entry point. */
const char *PLA_Last_error(void);
/* Device pointer of the object's top-left element (HBM).  Drops the host
   mirror, so later writes through the device pointer are authoritative. */
int PLA_Obj_device_buffer(PLA_Obj obj, void **device_buffer);
/* Fill obj with drand48()*2-1 in column-major order, starting `skip` draws
   after srand48(seed) -- the reference's TEST/create_problem.c:31-52 input,
   generated on the device with LCG jump-ahead (bit-identical values). */
int PLA_Obj_fill_drand48(PLA_Obj obj, long seed, int64_t skip);
/* Wait for all device work issued by the library. */
int PLA_Device_synchronize(void);
/* Block size of the device LU/Cholesky.  A value > 0 takes precedence over
   pla_Environ_set_nb_alg; 0 restores the order: the PLA_OP_PAN_PAN entry if
   pla_Environ_set_nb_alg was called, else the built-in choice for n. */
int PLA_Set_device_nb(int nb);
/* Tuning parameters of the device factorizations (an autotuner sets them).
   PLA_PARAM_LOOKAHEAD: 1 (default) factors panel k+1 on a second stream while
   step k's trailing update runs; 0 runs everything on one stream.
   PLA_PARAM_LEAF_ROWS: target rows per workgroup in the panel-LU leaf.
   PLA_PARAM_PANEL_CUS: CUs reserved for the LU panel factorization while the
   trailing update runs on the others (a multiple of 32 -- the same number on
   every shader engine of every XCD -- below the device's CU count; 0 lets the
   panel share every CU).
   PLA_PARAM_LEAF_WIDTH: panel-LU leaf width, 8, 16 or 32 (0: the widest
   that fits).
   PLA_PARAM_LEAF_KIND: 0 (default) keeps the leaf's rows in registers, one
   wave per workgroup and no LDS, so it runs beside the trailing update; 1
   keeps them in LDS, four waves per workgroup.
   PLA_PARAM_NB_TAIL / PLA_PARAM_NB_SWITCH: the LU factorizations use block
   size NB_TAIL (at most the main block size) for the steps where at most
   NB_SWITCH rows remain; NB_TAIL 0 (default) keeps one block size
   throughout. */
#define PLA_PARAM_LOOKAHEAD 1
#define PLA_PARAM_LEAF_ROWS 2
#define PLA_PARAM_PANEL_CUS 3
#define PLA_PARAM_LEAF_WIDTH 4
#define PLA_PARAM_LEAF_KIND 5
#define PLA_PARAM_NB_TAIL 6
#define PLA_PARAM_NB_SWITCH 7
/* PLA_PARAM_CHOL_PANEL_CUS: CUs reserved for the Cholesky panel (diagonal
   block and the solve below it) while the trailing update runs on the others
   (a multiple of 32 below the device's CU count; 0 shares every CU; 32 by
   default, one CU per shader engine).  The CU-masked stream pair is created
   at PLA_Init and when the parameter is set. */
#define PLA_PARAM_CHOL_PANEL_CUS 8
int PLA_Set_device_param(int param, int value);
int PLA_Get_device_nb(int n, int *nb);
/* Block size the device uses for an n x n factorization of the given kind:
   PLA_NB_LU (PLA_LU; what PLA_Get_device_nb returns), PLA_NB_CHOL
   (PLA_Chol), PLA_NB_LU_MIXED (PLA_LU_mixed). */
#define PLA_NB_LU 0
#define PLA_NB_CHOL 1
#define PLA_NB_LU_MIXED 2
int PLA_Get_device_nb_for(int n, int factorization, int *nb);
/* EXTENSION -- the nb_alg knob (TEMPLATE/PLA_Environ.c:88-127) tuned on the
   device: PLA_LU or PLA_Chol (factorization PLA_NB_LU / PLA_NB_CHOL) runs on a
   scratch n x n matrix (the drand48 input of TEST/create_problem.c; for
   Cholesky plus n I) once per candidate block size 128, 256, 512 (and 1024
   for PLA_LU from n = 16384); the fastest becomes the block size
   PLA_Get_device_nb_for and the factorizations use for order n, for the
   process lifetime (PLA_Set_device_nb and pla_Environ_set_nb_alg still take
   precedence).  ms (NULL allowed) receives the four candidates' times in
   milliseconds, 0 where not tried.  Needs 8 n^2 bytes of device memory for
   the duration of the call; the caller's objects are not touched. */
int PLA_Autotune_nb(int factorization, int n, int *best_nb, double *ms);
/* Per-launch hipEvent timing of the trailing-update kernels on the library
   stream (the reference keeps per-operation timers in PLA_TIMINGS,
   OBJ/PLA_Timings.c:17-125).  role 1: PLA_LU's A22 -= A21 A12 GEMM;
   role 2: PLA_Chol's lower(A22) -= A21 A21^T SYRK; role 3: PLA_LU_mixed's
   low-precision trailing GEMM; role 4: PLA_Chol_mixed's FP8 SYRK.  Roles 5-7
   are HBM-bound and count algorithmic bytes instead of flops: role 5 the
   FP64 residual r = b - A x of the refinements (general or symmetric);
   role 6 the sync-free triangular solves (PLA_Trsv, the solves of
   PLA_General_solve / PLA_Pos_def_solve and of the refinements); role 7
   the row interchanges of the LU factorizations, whose unit is the number of
   columns swapped (the rows moved per step are known only on the device).
   PLA_Kernel_timing(on) clears the counters; the query waits for the
   recorded events and returns the launch count, summed kernel milliseconds
   and the summed work. */
int PLA_Kernel_timing(int on);
int PLA_Kernel_timing_query(int role, long *launches, double *total_ms, double *total_flops);

#ifdef __cplusplus
}
#endif
#endif
