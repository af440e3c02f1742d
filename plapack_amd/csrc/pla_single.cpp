// pla_single.cpp -- MPI_FLOAT and MPI_COMPLEX objects.
//
// The reference factors and solves single-precision objects with the
// single-precision BLAS (PLA_slu / PLA_clu, SOLVERS/LU/PLA_seq_lu.c:50,108;
// sgemm / ctrsm through INCLUDE/PLA_BLAS.h).  Here an entry point given
// single-precision objects computes on double-precision copies of them
// (MPI_FLOAT -> MPI_DOUBLE, MPI_COMPLEX -> MPI_DOUBLE_COMPLEX) with the FP64
// kernels, and the outputs are rounded back once: the results are the
// double-precision computation's, rounded to single.
#include "pla_internal.hpp"

namespace pla {

bool is_single(PLA_Obj obj) {
  return valid(obj) && (obj->base->datatype == MPI_FLOAT || obj->base->datatype == MPI_COMPLEX);
}

static MPI_Datatype wide_type(MPI_Datatype dt) { return dt == MPI_FLOAT ? MPI_DOUBLE : MPI_DOUBLE_COMPLEX; }

int single_promote(PLA_Obj obj, PLA_Obj *wide) {
  *wide = nullptr;
  int rc = pla_obj_create_like(obj, wide_type(obj->base->datatype), wide);
  if (rc) return rc;
  if (obj->length == 0 || obj->width == 0) return PLA_SUCCESS;
  void *src, *dst;
  if ((rc = dev_ptr(obj, &src)) || (rc = dev_ptr(*wide, &dst))) return rc;
  return k_convert2d(ctx().stream, obj->length, obj->width, src, ld(obj), obj->base->datatype, dst, ld(*wide),
                     (*wide)->base->datatype);
}

int single_demote(PLA_Obj wide, PLA_Obj obj) {
  if (obj->length == 0 || obj->width == 0) return PLA_SUCCESS;
  void *src, *dst;
  int rc;
  if ((rc = dev_ptr(wide, &src)) || (rc = dev_ptr(obj, &dst))) return rc;
  if ((rc = k_convert2d(ctx().stream, obj->length, obj->width, src, ld(wide), wide->base->datatype, dst, ld(obj),
                        obj->base->datatype)))
    return rc;
  return written(obj);
}

int scalar_value_z(PLA_Obj o, double *re, double *im) {
  *im = 0.0;
  if (valid(o) && o->base->datatype == MPI_DOUBLE_COMPLEX) {
    if (o->length < 1 || o->width < 1) return fail(PLA_ERR_DIMENSIONS, "empty scalar object");
    void *p;
    int rc = dev_ptr(o, &p);
    if (rc) return rc;
    double v[2];
    PLA_HIP(hipStreamSynchronize(ctx().stream));
    PLA_HIP(hipMemcpy(v, p, sizeof v, hipMemcpyDeviceToHost));
    *re = v[0];
    *im = v[1];
    return PLA_SUCCESS;
  }
  return scalar_value(o, re);
}

int on_promoted(std::initializer_list<PLA_Obj *> objs, std::initializer_list<bool> out,
                const std::function<int()> &call) {
  struct slot {
    PLA_Obj *p;
    PLA_Obj orig;
    bool out;
  };
  slot sl[8];
  int ns = 0, rc = PLA_SUCCESS;
  auto ob = out.begin();
  for (PLA_Obj *p : objs) {
    const bool o = ob != out.end() ? *ob++ : false;
    if (!p || !is_single(*p)) continue;
    if (ns == 8) {
      rc = fail(PLA_ERR_INVALID_PARAM, "too many single-precision operands");
      break;
    }
    PLA_Obj w = nullptr;
    if ((rc = single_promote(*p, &w))) {
      if (w) PLA_Obj_free(&w);
      break;
    }
    sl[ns++] = slot{p, *p, o};
    *p = w;
  }
  if (!rc) rc = call();
  for (int i = ns - 1; i >= 0; i--) {
    PLA_Obj w = *sl[i].p;
    // a positive return is an index (a Cholesky failure): the outputs hold
    // what the reference leaves in them, so they are rounded back too
    if (rc >= 0 && sl[i].out) {
      const int r2 = single_demote(w, sl[i].orig);
      if (!rc) rc = r2;
    }
    PLA_Obj_free(&w);
    *sl[i].p = sl[i].orig;
  }
  return rc;
}

}  // namespace pla
