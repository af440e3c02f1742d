// pla_chol_noalloc.hip -- PLA_Chol_no_alloc: PLA_Chol for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// PLA_Chol's only scratch is two info words and the backup of the diagonal
// block being factored (up to CHOL_MAX_NB^2 doubles), taken from grow-only
// workspaces that are allocated on first use.  This variant takes them from
// one file-scope device array instead, of fixed size:
//   256 B (info words) + 512 * 512 * 8 B (block backup) = 2 MiB + 256 B
// per device, held until the process ends.  The array's device address is
// looked up with hipGetSymbolAddress once per device (at PLA_Init for the
// library's device) and only that address is handed to kernels and copies.
// Every call on a device shares its array: the info words are cleared at the
// start of each call, and calls that overlap on one device race.
#include "pla_internal.hpp"

namespace {

constexpr size_t SCRATCH_INFO_BYTES = 256;
constexpr size_t SCRATCH_BACKUP_ELEMS = (size_t)pla::CHOL_MAX_NB * pla::CHOL_MAX_NB;
constexpr size_t SCRATCH_BYTES = SCRATCH_INFO_BYTES + sizeof(double) * SCRATCH_BACKUP_ELEMS;
constexpr int MAX_DEVICES = 64;

void *g_scratch_ptr[MAX_DEVICES];  // device address of `scratch` on each device

}  // namespace

// The scratch itself.  In host code `scratch` names a host-side stand-in, not
// device memory: only g_scratch_ptr[device] may reach kernels and copies.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace pla {

int chol_static_scratch(int device, void **ptr) {
  if (device < 0 || device >= MAX_DEVICES) return fail(PLA_ERR_DEVICE, "device %d out of range", device);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    PLA_HIP(hipGetDevice(&cur));
    if (cur != device) return fail(PLA_ERR_DEVICE, "scratch lookup for device %d while device %d is current",
                                   device, cur);
    void *p = nullptr;
    PLA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(PLA_ERR_DEVICE, "hipGetSymbolAddress returned no address for the Cholesky scratch");
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return PLA_SUCCESS;
}

}  // namespace pla

using namespace pla;

extern "C" {

// PLA_Chol (SOLVERS/Cholesky/PLA_Chol.c:13) with its scratch in the static
// array above: same kernels, same results, no device allocation.
int PLA_Chol_no_alloc(int uplo, PLA_Obj A) {
  if (!ctx().initialized) return fail(PLA_ERR_NOT_INITIALIZED, "PLA_Chol_no_alloc before PLA_Init");
  if (uplo != PLA_LOWER_TRIANGULAR)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Chol_no_alloc: only uplo = PLA_LOWER_TRIANGULAR is implemented");
  if (!valid(A)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Chol_no_alloc: invalid A");
  if (A->base->datatype != MPI_DOUBLE)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Chol_no_alloc: only MPI_DOUBLE matrices are supported");
  const int n = A->length;
  if (A->width != n) return fail(PLA_ERR_DIMENSIONS, "PLA_Chol_no_alloc: A is %d x %d", A->length, A->width);
  // PLA_Chol's block size; chol_factor caps it at CHOL_MAX_NB, whose backup
  // the scratch holds, and rejects a block the scratch cannot hold
  const int nb = choose_nb(n, NB_CHOL);
  if (n == 0) return PLA_SUCCESS;
  double *a;
  int rc = dev_ptr(A, (void **)&a);
  if (rc) return rc;
  void *base = nullptr;
  if ((rc = chol_static_scratch(ctx().device, &base))) return rc;
  const chol_scratch scr{(int *)base, (double *)((char *)base + SCRATCH_INFO_BYTES), SCRATCH_BACKUP_ELEMS};
  int info = 0;
  if ((rc = chol_factor(ctx().stream, n, a, ld(A), nb, &info, &scr))) return rc;
  if ((rc = written(A))) return rc;
  return info;
}

}  // extern "C"
