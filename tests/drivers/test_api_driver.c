/*
 * test_api_driver.c -- the application-interface path of the reference's
 * example drivers, built against include/PLA.h and MPICH's <mpi.h>, linked
 * with libpla_mi355x.so.
 *
 * As EXAMPLES/LU/create_problem.c:132-183 (FILL_THROUGH_API): A (n x n) is
 * filled through PLA_API_axpy_matrix_to_global in blocks of 10 columns of
 * drand48() * 2 - 1 (srand48(0)), then b through PLA_API_axpy_vector_to_global
 * with the next n numbers.  PLA_General_solve overwrites b with x; as
 * EXAMPLES/LU/process_answer.c:28-82, x is read back into a zeroed local
 * vector with PLA_API_axpy_global_to_vector and PLA_Obj_API_sync.  The driver
 * regenerates A and b on the host, prints the scaled residual
 * ||A x - b||_inf / (||A||_inf ||x||_inf + ||b||_inf) and writes x to argv[2].
 *
 * usage: test_api_driver n [xfile]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mpi.h"
#include "PLA.h"

#define FILL_BLOCK 10

int main(int argc, char *argv[]) {
  MPI_Comm comm = MPI_COMM_NULL;
  PLA_Template templ = NULL;
  PLA_Obj A = NULL, b = NULL, pivots = NULL;
  double d_one = 1.0;
  int n, j, i, rc;

  MPI_Init(&argc, &argv);
  if (argc < 2 || (n = atoi(argv[1])) <= 0) {
    printf("usage: %s n [xfile]\n", argv[0]);
    return 2;
  }
  PLA_Comm_1D_to_2D(MPI_COMM_WORLD, 1, 1, &comm);
  if ((rc = PLA_Init(comm)) != PLA_SUCCESS) {
    printf("PLA_Init failed: %d\n", rc);
    return 3;
  }
  PLA_Temp_create(128, 0, &templ);
  PLA_Matrix_create(MPI_DOUBLE, n, n, templ, PLA_ALIGN_FIRST, PLA_ALIGN_FIRST, &A);
  PLA_Mvector_create(MPI_DOUBLE, n, 1, templ, PLA_ALIGN_FIRST, &b);
  PLA_Mvector_create(MPI_INT, n, 1, templ, PLA_ALIGN_FIRST, &pivots);
  PLA_Obj_set_to_zero(A);
  PLA_Obj_set_to_zero(b);

  PLA_API_begin();
  PLA_Obj_API_open(A);
  PLA_Obj_API_open(b);
  double *loc = (double *)malloc(sizeof(double) * (size_t)n * FILL_BLOCK);
  srand48(0);
  for (j = 0; j < n; j += FILL_BLOCK) {
    int w = n - j < FILL_BLOCK ? n - j : FILL_BLOCK;
    for (i = 0; i < n * w; i++) loc[i] = drand48() * 2.0 - 1.0;
    if ((rc = PLA_API_axpy_matrix_to_global(n, w, &d_one, loc, n, A, 0, j)) != PLA_SUCCESS) {
      printf("axpy_matrix_to_global failed: %d %s\n", rc, PLA_Last_error());
      return 4;
    }
  }
  for (i = 0; i < n; i++) loc[i] = drand48() * 2.0 - 1.0;
  PLA_API_axpy_vector_to_global(n, &d_one, loc, 1, b, 0);
  PLA_Obj_API_close(A);
  PLA_Obj_API_close(b);
  PLA_API_end();

  double t0 = MPI_Wtime();
  if ((rc = PLA_General_solve(A, pivots, b)) != PLA_SUCCESS) {
    printf("PLA_General_solve failed: %d %s\n", rc, PLA_Last_error());
    return 5;
  }
  double t1 = MPI_Wtime();

  double *x = (double *)calloc((size_t)n, sizeof(double));
  PLA_API_begin();
  PLA_Obj_API_open(b);
  PLA_API_axpy_global_to_vector(n, &d_one, b, 0, x, 1);
  PLA_Obj_API_sync(b);
  PLA_Obj_API_close(b);
  PLA_API_end();

  /* residual against A and b regenerated on the host */
  double *a = (double *)malloc(sizeof(double) * (size_t)n * n), *r = (double *)calloc((size_t)n, sizeof(double));
  srand48(0);
  for (j = 0; j < n; j++)
    for (i = 0; i < n; i++) a[(size_t)j * n + i] = drand48() * 2.0 - 1.0;
  double bnorm = 0.0, anorm = 0.0, xnorm = 0.0, rnorm = 0.0;
  for (i = 0; i < n; i++) {
    r[i] = -(drand48() * 2.0 - 1.0);
    if (fabs(r[i]) > bnorm) bnorm = fabs(r[i]);
    if (fabs(x[i]) > xnorm) xnorm = fabs(x[i]);
  }
  for (j = 0; j < n; j++)
    for (i = 0; i < n; i++) r[i] += a[(size_t)j * n + i] * x[j];
  for (i = 0; i < n; i++) {
    double s = 0.0;
    for (j = 0; j < n; j++) s += fabs(a[(size_t)j * n + i]);
    if (s > anorm) anorm = s;
    if (fabs(r[i]) > rnorm) rnorm = fabs(r[i]);
  }
  printf("n %d solve %.3f s scaled residual %.3e\n", n, t1 - t0, rnorm / (anorm * xnorm + bnorm));
  if (argc > 2) {
    FILE *f = fopen(argv[2], "wb");
    if (f) {
      fwrite(x, sizeof(double), (size_t)n, f);
      fclose(f);
    }
  }
  free(loc);
  free(x);
  free(a);
  free(r);
  PLA_Obj_free(&A);
  PLA_Obj_free(&b);
  PLA_Obj_free(&pivots);
  PLA_Temp_free(&templ);
  PLA_Finalize();
  MPI_Finalize();
  return 0;
}
