// pla_obj.cpp -- PLA_Obj object model on a 1x1 template with HBM buffers.
//
// Reference counterparts: OBJ/PLA_Obj.c (creation :149,:323,:497, queries
// :831-1430, contents :1437,:1545, constants :1807-1900, views and splits
// :1981,:2063,:3122,:3156, free :803) and OBJ/PLA_Obj_management.c
// (reference counting :47-77).  On one process every object's local part
// is the whole object, so local and global sizes coincide.
#include "pla_internal.hpp"

#include <cstdlib>
#include <cstring>

using pla::fail;

namespace pla {

static int typesize_of(MPI_Datatype dt) {
  if (dt == MPI_DOUBLE) return 8;
  if (dt == MPI_FLOAT) return 4;
  if (dt == MPI_INT) return 4;
  if (dt == MPI_DOUBLE_COMPLEX) return 16;
  if (dt == MPI_COMPLEX) return 8;
  if (dt == MPI_CHAR) return 1;
  return 0;
}

bool valid(PLA_Obj o) { return o && o->base && o->base->dev; }

int ld(PLA_Obj o) { return o->base->ldim; }

static size_t base_bytes(const pla_base *b) {
  return (size_t)b->ldim * (size_t)(b->width > 0 ? b->width : 1) * (size_t)b->typesize;
}

// Mirror -> device if the caller may have written the mirror.
static int push_mirror(pla_base *b) {
  if (b->host && b->host_dirty) {
    PLA_HIP(hipMemcpy(b->dev, b->host, base_bytes(b), hipMemcpyHostToDevice));
    b->host_dirty = false;
  }
  return PLA_SUCCESS;
}

int dev_ptr(PLA_Obj o, void **p) {
  if (!valid(o)) return fail(PLA_ERR_INVALID_OBJECT, "invalid object");
  int rc = push_mirror(o->base);
  if (rc) return rc;
  *p = (char *)o->base->dev + elem_offset(o) * (size_t)o->base->typesize;
  return PLA_SUCCESS;
}

int written(PLA_Obj o) {
  pla_base *b = o->base;
  if (!b->host) return PLA_SUCCESS;
  PLA_HIP(hipDeviceSynchronize());
  PLA_HIP(hipMemcpy(b->host, b->dev, base_bytes(b), hipMemcpyDeviceToHost));
  b->host_dirty = false;
  return PLA_SUCCESS;
}

static void detach(pla_base *b) {
  if (--b->number_of_views > 0) return;
  if (b->dev && !b->user_buffer) (void)hipFree(b->dev);
  if (b->host) (void)hipHostFree(b->host);
  free(b);
}

static void free_view(PLA_Obj v) {
  if (!v) return;
  if (v->base) detach(v->base);
  free(v);
}

static int new_object(int objtype, MPI_Datatype dt, int m, int n, PLA_Template templ, int owner_row,
                      int owner_col, PLA_Obj *out) {
  if (!out) return fail(PLA_ERR_INVALID_PARAM, "object pointer is NULL");
  if (*out) {  // the reference frees a non-NULL target first (OBJ/PLA_Obj.c:73-74)
    free_view(*out);
    *out = nullptr;
  }
  if (!ctx().initialized) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  if (!templ) return fail(PLA_ERR_INVALID_PARAM, "invalid template");
  int ts = typesize_of(dt);
  if (!ts) return fail(PLA_ERR_UNSUPPORTED, "unsupported datatype 0x%x", (unsigned)dt);
  if (m < 0 || n < 0) return fail(PLA_ERR_DIMENSIONS, "negative dimensions %d x %d", m, n);
  int rc = ensure_device();
  if (rc) return rc;
  pla_base *b = (pla_base *)calloc(1, sizeof(pla_base));
  PLA_Obj v = (PLA_Obj)calloc(1, sizeof(struct PLA_Obj_view_struct));
  if (!b || !v) {
    free(b);
    free(v);
    return fail(PLA_ERR_NO_MEMORY, "out of host memory");
  }
  b->datatype = dt;
  b->typesize = ts;
  b->length = m;
  b->width = n;
  b->ldim = m > 0 ? m : 1;
  b->templ = templ;
  b->number_of_views = 1;
  size_t bytes = base_bytes(b);
  hipError_t e = hipMalloc(&b->dev, bytes);
  if (e != hipSuccess) {
    free(b);
    free(v);
    return hip_fail(e, "hipMalloc (object buffer)");
  }
  // zero-filled at creation, as PLA_Obj_set_to_zero does in the reference (:370)
  e = hipMemsetAsync(b->dev, 0, bytes, ctx().stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx().stream);
  if (e != hipSuccess) {
    (void)hipFree(b->dev);
    free(b);
    free(v);
    return hip_fail(e, "hipMemset (object buffer)");
  }
  v->objtype = objtype;
  v->length = m;
  v->width = n;
  v->owner_row = owner_row;
  v->owner_col = owner_col;
  v->proj_onto = PLA_PROJ_ONTO_COL;
  v->base = b;
  *out = v;
  return PLA_SUCCESS;
}

static PLA_Obj clone_view(PLA_Obj o) {
  PLA_Obj v = (PLA_Obj)malloc(sizeof(struct PLA_Obj_view_struct));
  if (!v) return nullptr;
  *v = *o;
  v->base->number_of_views++;
  return v;
}

static int set_view(PLA_Obj src, int r0, int c0, int m, int n, PLA_Obj *out) {
  if (!out) return PLA_SUCCESS;  // PLA_DUMMY target
  PLA_Obj v = clone_view(src);   // clone before freeing: *out may be src
  if (!v) return fail(PLA_ERR_NO_MEMORY, "out of host memory");
  v->row0 = src->row0 + r0;
  v->col0 = src->col0 + c0;
  v->length = m;
  v->width = n;
  if (*out) free_view(*out);
  *out = v;
  return PLA_SUCCESS;
}

static int fill_constant(PLA_Obj o, double value) {
  void *p;
  int rc = dev_ptr(o, &p);
  if (rc) return rc;
  pla_base *b = o->base;
  int m = o->length, n = o->width, ts = b->typesize;
  if (m == 0 || n == 0) return PLA_SUCCESS;
  if (value == 0.0) {
    PLA_HIP(hipMemset2DAsync(p, (size_t)b->ldim * ts, 0, (size_t)m * ts, n, ctx().stream));
  } else {
    // small objects (constants, vectors): build on the host, copy once
    size_t cnt = (size_t)m * n;
    char *h = (char *)malloc(cnt * ts);
    if (!h) return fail(PLA_ERR_NO_MEMORY, "out of host memory");
    for (size_t i = 0; i < cnt; i++) {
      char *e = h + i * ts;
      if (b->datatype == MPI_DOUBLE) *(double *)e = value;
      else if (b->datatype == MPI_FLOAT) *(float *)e = (float)value;
      else if (b->datatype == MPI_INT) *(int *)e = (int)value;
      else if (b->datatype == MPI_DOUBLE_COMPLEX) { ((double *)e)[0] = value; ((double *)e)[1] = 0; }
      else if (b->datatype == MPI_COMPLEX) { ((float *)e)[0] = (float)value; ((float *)e)[1] = 0; }
      else *(char *)e = (char)value;
    }
    hipError_t e = hipMemcpy2D(p, (size_t)b->ldim * ts, h, (size_t)m * ts, (size_t)m * ts, n,
                               hipMemcpyHostToDevice);
    free(h);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy2D (constant)");
  }
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  return written(o);
}

int scalar_value(PLA_Obj o, double *v) {
  if (!valid(o)) return fail(PLA_ERR_INVALID_OBJECT, "invalid scalar object");
  if (o->length < 1 || o->width < 1) return fail(PLA_ERR_DIMENSIONS, "empty scalar object");
  void *p;
  int rc = dev_ptr(o, &p);
  if (rc) return rc;
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  if (o->base->datatype == MPI_DOUBLE) {
    PLA_HIP(hipMemcpy(v, p, sizeof(double), hipMemcpyDeviceToHost));
  } else if (o->base->datatype == MPI_FLOAT) {
    float f;
    PLA_HIP(hipMemcpy(&f, p, sizeof(float), hipMemcpyDeviceToHost));
    *v = f;
  } else if (o->base->datatype == MPI_INT) {
    int i;
    PLA_HIP(hipMemcpy(&i, p, sizeof(int), hipMemcpyDeviceToHost));
    *v = i;
  } else {
    return fail(PLA_ERR_UNSUPPORTED, "scalar of unsupported datatype");
  }
  return PLA_SUCCESS;
}

}  // namespace pla

using namespace pla;

extern "C" {

int pla_obj_create_like(PLA_Obj like, MPI_Datatype datatype, PLA_Obj *new_obj) {
  if (!valid(like)) return fail(PLA_ERR_INVALID_OBJECT, "invalid object");
  return new_object(like->objtype, datatype, like->length, like->width, like->base->templ, like->owner_row,
                    like->owner_col, new_obj);
}

int PLA_Matrix_create(MPI_Datatype datatype, int global_length, int global_width,
                      PLA_Template templ, int global_align_row, int global_align_col,
                      PLA_Obj *new_obj) {
  (void)global_align_row;
  (void)global_align_col;  // one process: every alignment is the first element
  return new_object(PLA_MATRIX, datatype, global_length, global_width, templ, 0, 0, new_obj);
}

int PLA_Mvector_create(MPI_Datatype datatype, int global_length, int global_width,
                       PLA_Template templ, int global_align, PLA_Obj *new_obj) {
  (void)global_align;
  return new_object(PLA_MVECTOR, datatype, global_length, global_width, templ, 0, 0, new_obj);
}

int PLA_Mscalar_create(MPI_Datatype datatype, int owner_row, int owner_col, int length,
                       int width, PLA_Template templ, PLA_Obj *new_obj) {
  return new_object(PLA_MSCALAR, datatype, length, width, templ, owner_row, owner_col, new_obj);
}

int PLA_Matrix_create_conf_to(PLA_Obj obj, PLA_Obj *new_obj) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Matrix_create_conf_to: invalid obj");
  return new_object(PLA_MATRIX, obj->base->datatype, obj->length, obj->width, obj->base->templ,
                    0, 0, new_obj);
}

int PLA_Mvector_create_conf_to(PLA_Obj obj, int global_width, PLA_Obj *new_obj) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Mvector_create_conf_to: invalid obj");
  return new_object(PLA_MVECTOR, obj->base->datatype, obj->length, global_width,
                    obj->base->templ, 0, 0, new_obj);
}

int PLA_Mscalar_create_conf_to(PLA_Obj obj, int owner_row, int owner_col, PLA_Obj *new_obj) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Mscalar_create_conf_to: invalid obj");
  return new_object(PLA_MSCALAR, obj->base->datatype, obj->length, obj->width, obj->base->templ,
                    owner_row, owner_col, new_obj);
}

int PLA_Create_constants_conf_to(PLA_Obj obj, PLA_Obj *minus_one, PLA_Obj *zero, PLA_Obj *one) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Create_constants_conf_to: invalid obj");
  MPI_Datatype dt = obj->base->datatype;
  PLA_Template t = obj->base->templ;
  struct { PLA_Obj *p; double v; } want[3] = {{minus_one, -1.0}, {zero, 0.0}, {one, 1.0}};
  for (auto &w : want) {
    if (!w.p) continue;
    int rc = new_object(PLA_MSCALAR, dt, 1, 1, t, PLA_ALL_ROWS, PLA_ALL_COLS, w.p);
    if (!rc) rc = fill_constant(*w.p, w.v);
    if (rc) return rc;
    (*w.p)->fixed = 1;
  }
  return PLA_SUCCESS;
}

int PLA_Obj_free(PLA_Obj *obj) {
  if (!obj) return PLA_SUCCESS;
  if (*obj) pla::api_forget(*obj);
  free_view(*obj);
  *obj = nullptr;
  return PLA_SUCCESS;
}

#define QUERY(name, type, expr)                                                     \
  int name(PLA_Obj obj, type *out) {                                                \
    if (!valid(obj) || !out) return fail(PLA_ERR_INVALID_OBJECT, #name ": invalid"); \
    *out = (expr);                                                                  \
    return PLA_SUCCESS;                                                             \
  }

QUERY(PLA_Obj_objtype, int, obj->objtype)
QUERY(PLA_Obj_datatype, MPI_Datatype, obj->base->datatype)
QUERY(PLA_Obj_template, PLA_Template, obj->base->templ)
QUERY(PLA_Obj_global_length, int, obj->length)
QUERY(PLA_Obj_global_width, int, obj->width)
QUERY(PLA_Obj_owner_row, int, obj->owner_row)
QUERY(PLA_Obj_owner_col, int, obj->owner_col)
QUERY(PLA_Obj_local_length, int, obj->length)
QUERY(PLA_Obj_local_width, int, obj->width)
QUERY(PLA_Obj_local_ldim, int, obj->base->ldim)
QUERY(PLA_Obj_local_stride, int, 1)
#undef QUERY

int PLA_Obj_local_buffer(PLA_Obj obj, void **local_buffer) {
  if (!valid(obj) || !local_buffer) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_local_buffer");
  pla_base *b = obj->base;
  if (!b->host) {
    size_t bytes = base_bytes(b);
    PLA_HIP(hipHostMalloc(&b->host, bytes, hipHostMallocDefault));
    PLA_HIP(hipDeviceSynchronize());
    PLA_HIP(hipMemcpy(b->host, b->dev, bytes, hipMemcpyDeviceToHost));
  }
  b->host_dirty = true;  // the caller may write through the pointer from now on
  *local_buffer = (char *)b->host + elem_offset(obj) * (size_t)b->typesize;
  return PLA_SUCCESS;
}

int PLA_Obj_device_buffer(PLA_Obj obj, void **device_buffer) {
  if (!valid(obj) || !device_buffer) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_device_buffer");
  pla_base *b = obj->base;
  int rc = dev_ptr(obj, device_buffer);
  if (rc) return rc;
  if (b->host) {
    (void)hipHostFree(b->host);
    b->host = nullptr;
    b->host_dirty = false;
  }
  return PLA_SUCCESS;
}

// OBJ/PLA_Obj_management.c:15-42.  The reference attaches a buffer to a base
// object that has none; every object created here already owns an HBM
// buffer, so attaching replaces it (the old buffer and any host mirror are
// released).  `buffer` is a device pointer; with user_buffer != 0 it stays
// the caller's and PLA_Obj_free leaves it alone.
int PLA_Obj_attach_buffer(PLA_Obj obj, void *buffer, int local_ldim, int user_buffer) {
  if (!valid(obj) || !buffer) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_attach_buffer: invalid argument");
  pla_base *b = obj->base;
  if (obj->row0 || obj->col0 || obj->length != b->length || obj->width != b->width)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Obj_attach_buffer: obj must view its whole base object");
  if (obj->fixed) return fail(PLA_ERR_INVALID_PARAM, "PLA_Obj_attach_buffer: obj is a constant");
  if (local_ldim < (b->length > 0 ? b->length : 1))
    return fail(PLA_ERR_DIMENSIONS, "PLA_Obj_attach_buffer: ldim %d < %d rows", local_ldim, b->length);
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, buffer) != hipSuccess || attr.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Obj_attach_buffer: buffer is not device (HBM) memory");
  }
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  if (b->dev && !b->user_buffer) (void)hipFree(b->dev);
  if (b->host) (void)hipHostFree(b->host);
  b->host = nullptr;
  b->host_dirty = false;
  b->dev = buffer;
  b->ldim = local_ldim;
  b->user_buffer = user_buffer ? 1 : 0;
  return PLA_SUCCESS;
}

int PLA_Obj_get_local_contents(PLA_Obj obj, int trans, int *rows_in_buf, int *cols_in_buf,
                               void *buf, int leading_dim_buf, int stride_buf) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_get_local_contents: invalid obj");
  if (trans != PLA_NO_TRANSPOSE)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Obj_get_local_contents: only PLA_NO_TRANSPOSE");
  if (stride_buf != 1)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Obj_get_local_contents: only stride 1");
  int m = obj->length, n = obj->width, ts = obj->base->typesize;
  if (rows_in_buf) *rows_in_buf = m;
  if (cols_in_buf) *cols_in_buf = n;
  if (m == 0 || n == 0) return PLA_SUCCESS;
  if (leading_dim_buf < m) return fail(PLA_ERR_INVALID_PARAM, "leading_dim_buf < rows");
  void *p;
  int rc = dev_ptr(obj, &p);
  if (rc) return rc;
  PLA_HIP(hipDeviceSynchronize());
  PLA_HIP(hipMemcpy2D(buf, (size_t)leading_dim_buf * ts, p, (size_t)obj->base->ldim * ts,
                      (size_t)m * ts, n, hipMemcpyDeviceToHost));
  return PLA_SUCCESS;
}

int PLA_Obj_set_local_contents(int trans, int rows_in_buf, int cols_in_buf, void *buf,
                               int leading_dim_buf, int stride_buf, PLA_Obj obj) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_set_local_contents: invalid obj");
  if (trans != PLA_NO_TRANSPOSE)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Obj_set_local_contents: only PLA_NO_TRANSPOSE");
  if (stride_buf != 1)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Obj_set_local_contents: only stride 1");
  int m = obj->length, n = obj->width, ts = obj->base->typesize;
  if (rows_in_buf != m || cols_in_buf != n)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Obj_set_local_contents: %d x %d buffer for %d x %d object",
                rows_in_buf, cols_in_buf, m, n);
  if (m == 0 || n == 0) return PLA_SUCCESS;
  if (leading_dim_buf < m) return fail(PLA_ERR_INVALID_PARAM, "leading_dim_buf < rows");
  void *p;
  int rc = dev_ptr(obj, &p);
  if (rc) return rc;
  PLA_HIP(hipDeviceSynchronize());
  PLA_HIP(hipMemcpy2D(p, (size_t)obj->base->ldim * ts, buf, (size_t)leading_dim_buf * ts,
                      (size_t)m * ts, n, hipMemcpyHostToDevice));
  return written(obj);
}

int PLA_Obj_fill_drand48(PLA_Obj obj, long seed, int64_t skip) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_fill_drand48: invalid obj");
  // MPI_FLOAT / MPI_COMPLEX: the same draws, each rounded to single as the
  // assignment buff_A[...] = drand48() * 2.0 - 1.0 does (TEST/create_problem.c:35-43)
  if (is_single(obj)) return on_promoted({&obj}, {true}, [&] { return PLA_Obj_fill_drand48(obj, seed, skip); });
  const bool cplx = obj->base->datatype == MPI_DOUBLE_COMPLEX;
  if (obj->base->datatype != MPI_DOUBLE && !cplx)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Obj_fill_drand48: only MPI_DOUBLE, MPI_FLOAT and complex objects");
  if (skip < 0) return fail(PLA_ERR_INVALID_PARAM, "PLA_Obj_fill_drand48: negative skip");
  void *p;
  int rc = dev_ptr(obj, &p);
  if (rc) return rc;
  const uint64_t state0 = ((((uint64_t)(uint32_t)seed) << 16) | 0x330EULL) & ((1ULL << 48) - 1);
  // a complex element is two consecutive draws, real part first (TEST/create_problem.c:60-65):
  // the stream fills the 2m x n real view of the doubles
  const int mr = cplx ? 2 * obj->length : obj->length;
  const long ldr = cplx ? 2L * ld(obj) : ld(obj);
  if ((rc = k_fill_drand48(ctx().stream, mr, obj->width, (double *)p, ldr, state0, (uint64_t)skip))) return rc;
  return written(obj);
}

int PLA_Obj_set_to_zero(PLA_Obj obj) { return fill_constant(obj, 0.0); }
int PLA_Obj_set_to_one(PLA_Obj obj) { return fill_constant(obj, 1.0); }
int PLA_Obj_set_to_minus_one(PLA_Obj obj) { return fill_constant(obj, -1.0); }

int PLA_Obj_view_all(PLA_Obj old_obj, PLA_Obj *new_obj) {
  if (!valid(old_obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_view_all: invalid obj");
  return set_view(old_obj, 0, 0, old_obj->length, old_obj->width, new_obj);
}

// OBJ/PLA_Obj.c:2063-2090: the signs of length/width choose which block they size.
int PLA_Obj_split_4(PLA_Obj obj, int length, int width, PLA_Obj *upper_left,
                    PLA_Obj *upper_right, PLA_Obj *lower_left, PLA_Obj *lower_right) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_split_4: invalid obj");
  int m = obj->length, n = obj->width;
  int top = length >= 0 ? length : m + length;
  int left = width >= 0 ? width : n + width;
  if (top > m) top = m;
  if (left > n) left = n;
  if (top < 0 || left < 0) return fail(PLA_ERR_DIMENSIONS, "PLA_Obj_split_4: split exceeds object");
  // clone the source first: an output may alias obj (e.g. &A_BR)
  PLA_Obj src = clone_view(obj);
  if (!src) return fail(PLA_ERR_NO_MEMORY, "out of host memory");
  int rc = set_view(src, 0, 0, top, left, upper_left);
  if (!rc) rc = set_view(src, 0, left, top, n - left, upper_right);
  if (!rc) rc = set_view(src, top, 0, m - top, left, lower_left);
  if (!rc) rc = set_view(src, top, left, m - top, n - left, lower_right);
  free_view(src);
  return rc;
}

int PLA_Obj_horz_split_2(PLA_Obj obj, int length, PLA_Obj *upper, PLA_Obj *lower) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_horz_split_2: invalid obj");
  return PLA_Obj_split_4(obj, length, obj->width, upper, nullptr, lower, nullptr);
}

int PLA_Obj_vert_split_2(PLA_Obj obj, int width, PLA_Obj *left, PLA_Obj *right) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_vert_split_2: invalid obj");
  return PLA_Obj_split_4(obj, obj->length, width, left, right, nullptr, nullptr);
}

int PLA_Copy(PLA_Obj from, PLA_Obj to) {
  if (!valid(from) || !valid(to)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Copy: invalid object");
  if (from->length != to->length || from->width != to->width)
    return fail(PLA_ERR_DIMENSIONS, "PLA_Copy: %d x %d into %d x %d", from->length, from->width,
                to->length, to->width);
  if (from->base->datatype != to->base->datatype)
    return fail(PLA_ERR_UNSUPPORTED, "PLA_Copy: datatype conversion not supported");
  int m = from->length, n = from->width, ts = from->base->typesize;
  if (m == 0 || n == 0) return PLA_SUCCESS;
  void *ps, *pd;
  int rc = dev_ptr(from, &ps);
  if (!rc) rc = dev_ptr(to, &pd);
  if (rc) return rc;
  PLA_HIP(hipMemcpy2DAsync(pd, (size_t)to->base->ldim * ts, ps, (size_t)from->base->ldim * ts,
                           (size_t)m * ts, n, hipMemcpyDeviceToDevice, ctx().stream));
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  return written(to);
}

}  // extern "C"
