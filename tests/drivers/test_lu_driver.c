/*
 * test_lu_driver.c -- a PLAPACK driver in the style of the reference's
 * TEST/test_lu.c, built unmodified against include/PLA.h and MPICH's
 * <mpi.h>, linked with libpla_mi355x.so.
 *
 * stdin (as TEST/test_lu.c:40-131): mesh rows cols, nb_distr, nb_alg, error,
 * parameters, sequential, then pairs "datatype n" until datatype -1 (1: double,
 * 3: double complex).
 * The matrix is filled through PLA_Obj_local_buffer on the host exactly as
 * TEST/create_problem.c:31-52 does (srand48(0), drand48()*2-1, column-major);
 * the library synchronises its host mirror with HBM.  For every n the driver
 * prints "n time MFLOPS/node" and writes the pivots to argv[1] (int32) if given.
 */
#include <stdio.h>
#include <stdlib.h>

#include "mpi.h"
#include "PLA.h"

static void create_problem(PLA_Obj A) {
  int m, n, ldim, i, j, per = 1;
  double *buf;
  MPI_Datatype dt;
  PLA_Obj_local_length(A, &m);
  PLA_Obj_local_width(A, &n);
  PLA_Obj_local_ldim(A, &ldim);
  PLA_Obj_datatype(A, &dt);
  if (dt == MPI_DOUBLE_COMPLEX) per = 2; /* real then imaginary part, as TEST/create_problem.c:60-65 */
  PLA_Obj_local_buffer(A, (void **)&buf);
  srand48(0);
  for (j = 0; j < n; j++)
    for (i = 0; i < m * per; i++) buf[(size_t)j * ldim * per + i] = drand48() * 2.0 - 1.0;
}

int main(int argc, char *argv[]) {
  MPI_Comm comm = MPI_COMM_NULL;
  PLA_Template templ = NULL;
  PLA_Obj A = NULL, pivots = NULL;
  int n, nb_distr, nb_alg, error, parameters, sequential, nprows, npcols, itype, rc;
  double time, flops;

  MPI_Init(&argc, &argv);
  if (scanf("%d%d%d%d%d%d%d", &nprows, &npcols, &nb_distr, &nb_alg, &error, &parameters,
            &sequential) != 7)
    return 2;
  pla_Environ_set_nb_alg(PLA_OP_ALL_ALG, nb_alg);
  PLA_Set_error_checking(error, parameters, sequential, FALSE);
  if ((rc = PLA_Comm_1D_to_2D(MPI_COMM_WORLD, nprows, npcols, &comm)) != PLA_SUCCESS) {
    printf("PLA_Comm_1D_to_2D: %d %s\n", rc, PLA_Last_error());
    return 3;
  }
  if ((rc = PLA_Init(comm)) != PLA_SUCCESS) {
    printf("PLA_Init: %d %s\n", rc, PLA_Last_error());
    return 4;
  }
  PLA_Temp_create(nb_distr, 0, &templ);
  while (scanf("%d", &itype) == 1 && itype != -1) {
    if (scanf("%d", &n) != 1) break;
    if (itype != 1 && itype != 3) {
      printf("datatype %d not supported\n", itype);
      continue;
    }
    PLA_Matrix_create(itype == 3 ? MPI_DOUBLE_COMPLEX : MPI_DOUBLE, n, n, templ, PLA_ALIGN_FIRST, PLA_ALIGN_FIRST,
                      &A);
    PLA_Mvector_create(MPI_INT, n, 1, templ, PLA_ALIGN_FIRST, &pivots);
    create_problem(A);
    MPI_Barrier(MPI_COMM_WORLD);
    time = MPI_Wtime();
    rc = PLA_LU(A, pivots);
    MPI_Barrier(MPI_COMM_WORLD);
    time = MPI_Wtime() - time;
    if (rc != PLA_SUCCESS) {
      printf("PLA_LU: %d %s\n", rc, PLA_Last_error());
      return 5;
    }
    flops = 2.0 / 3.0 * n * n * n;
    printf("%d time = %f, MFLOPS/node = %10.4lf \n", n, time, flops / time * 1.0e-6);
    if (argc > 1) {
      int *p;
      FILE *f = fopen(argv[1], "wb");
      PLA_Obj_local_buffer(pivots, (void **)&p);
      if (!f || fwrite(p, sizeof(int), (size_t)n, f) != (size_t)n) return 6;
      fclose(f);
    }
  }
  PLA_Obj_free(&A);
  PLA_Obj_free(&pivots);
  PLA_Temp_free(&templ);
  PLA_Finalize();
  MPI_Finalize();
  return 0;
}
