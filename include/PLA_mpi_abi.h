/*
 * PLA_mpi_abi.h -- the few MPI types and handles the PLAPACK C API names,
 * with the MPICH 3.x ABI values (MPI_Comm and MPI_Datatype are ints).
 *
 * libpla_mi355x runs on a 1x1 process template and calls no MPI routine,
 * so it is built against this header and links no MPI library.  A caller
 * that includes MPICH's own <mpi.h> passes bit-identical handle values;
 * tests/test_abi.py compiles a probe against <mpi.h> to check that.
 * Callers without MPI define PLA_NO_MPI_H before including PLA.h and get
 * this header instead of <mpi.h>.
 */
#ifndef PLA_MPI_ABI_H
#define PLA_MPI_ABI_H

#ifndef MPI_VERSION
typedef int MPI_Comm;
typedef int MPI_Datatype;

#define MPI_COMM_NULL      ((MPI_Comm)0x04000000)
#define MPI_COMM_WORLD     ((MPI_Comm)0x44000000)
#define MPI_COMM_SELF      ((MPI_Comm)0x44000001)
#define MPI_DATATYPE_NULL  ((MPI_Datatype)0x0c000000)
#define MPI_CHAR           ((MPI_Datatype)0x4c000101)
#define MPI_INT            ((MPI_Datatype)0x4c000405)
#define MPI_FLOAT          ((MPI_Datatype)0x4c00040a)
#define MPI_DOUBLE         ((MPI_Datatype)0x4c00080b)
#define MPI_COMPLEX        ((MPI_Datatype)1275070494)
#define MPI_DOUBLE_COMPLEX ((MPI_Datatype)1275072546)
#define MPI_SUCCESS        0
#endif

#endif
