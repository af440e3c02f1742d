// pla_api.cpp -- the application interface (PLA_API_*): host buffers added
// into global objects in HBM and pieces of global objects added into host
// buffers.
//
// Reference: API/PLA_API.c (begin/end/state :24-121, open/close/mode
// :126-215, vector axpys :220-385, matrix axpys :390-792, multi-block axpys
// :797-911, sync :916) and API/PLA_API_util.c (open-object list :1495-1926,
// PLA_Obj_add_to_local_contents :1519).
//
// On one process every block is local, so the reference's put / request
// machinery (PLA_API_put, the request buffers) has nothing to do; what
// remains is a host <-> HBM transfer.  A transfer goes through a pinned
// staging slot in one DMA copy per chunk, and a kernel moves the packed chunk
// between staging memory in HBM and its place in the object (scatter-add or
// gather).  Two slots alternate, so the host packs chunk c + 1 (or adds
// chunk c - 1 into the caller's buffer) while the copy engine moves chunk c.
//
// Completion: a *_to_global call returns once the caller's buffer has been
// copied into pinned memory (the buffer may be reused at once); the add into
// the object is complete at PLA_Obj_API_sync / PLA_Obj_API_close /
// PLA_API_end, as in the reference.  A global_to_* call returns with the
// caller's buffer updated.
//
// alpha: the reference passes alpha down but never applies it
// (PLA_Obj_add_to_local_contents adds the buffer unscaled); here alpha
// scales the added values, the meaning the prototypes document
// (INCLUDE/PLA_API_prototypes.h).  The two agree for alpha = 1, the value
// every reference caller passes.
#include "pla_internal.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

using namespace pla;

namespace {

struct api_ctx {
  bool active = false;
  std::vector<PLA_Obj> open;   // open objects (by handle, as the reference's list)
  void *pinned[2] = {nullptr, nullptr};
  size_t pinned_bytes = 0;     // capacity of each pinned slot
  hipEvent_t done[2] = {nullptr, nullptr};  // last copy that used slot i
  int next = 0;                // slot of the next chunk
};

api_ctx &api() {
  static api_ctx a;
  return a;
}

constexpr size_t API_STAGE_MIN = 16u << 20;  // bytes per staging slot

int elem_kind(MPI_Datatype dt) {
  if (dt == MPI_DOUBLE) return API_F64;
  if (dt == MPI_FLOAT) return API_F32;
  if (dt == MPI_INT) return API_I32;
  if (dt == MPI_DOUBLE_COMPLEX) return API_Z64;
  return 0;
}

bool is_open(PLA_Obj o) {
  const auto &v = api().open;
  return std::find(v.begin(), v.end(), o) != v.end();
}

// Both staging slots hold at least `bytes` (pinned host + HBM).
int ensure_staging(size_t bytes) {
  api_ctx &a = api();
  if (!a.done[0]) {
    PLA_HIP(hipEventCreateWithFlags(&a.done[0], hipEventDisableTiming));
    PLA_HIP(hipEventCreateWithFlags(&a.done[1], hipEventDisableTiming));
  }
  if (bytes < API_STAGE_MIN) bytes = API_STAGE_MIN;
  if (a.pinned_bytes < bytes) {
    PLA_HIP(hipStreamSynchronize(ctx().stream));
    for (int i = 0; i < 2; i++) {
      if (a.pinned[i]) (void)hipHostFree(a.pinned[i]);
      a.pinned[i] = nullptr;
    }
    a.pinned_bytes = 0;
    for (int i = 0; i < 2; i++) PLA_HIP(hipHostMalloc(&a.pinned[i], bytes, hipHostMallocDefault));
    a.pinned_bytes = bytes;
  }
  for (int i = 0; i < 2; i++)
    if (!workspace(a.pinned_bytes, WS_API_0 + i)) return PLA_ERR_NO_MEMORY;
  return PLA_SUCCESS;
}

// A grid of sub-blocks: the rsz[0] x csz[0] block at the top-left of the
// local buffer belongs at (rdisp[0], cdisp[0]) of the object, the next
// rsz[1] rows at rdisp[1], and so on; the same for columns.  M x N is the
// whole local block.
struct blocks {
  std::vector<int> rsz, rdisp, csz, cdisp;
  int M = 0, N = 0;
};

int check_target(PLA_Obj A, const char *who) {
  if (!api().active) return fail(PLA_ERR_INVALID_PARAM, "%s: API not active (PLA_API_begin)", who);
  if (!valid(A)) return fail(PLA_ERR_INVALID_OBJECT, "%s: invalid object", who);
  if (!is_open(A)) return fail(PLA_ERR_INVALID_PARAM, "%s: object not open (PLA_Obj_API_open)", who);
  if (A->objtype != PLA_MATRIX && A->objtype != PLA_MVECTOR)
    return fail(PLA_ERR_INVALID_PARAM, "%s: Illegal target object type", who);
  if (!elem_kind(A->base->datatype))
    return fail(PLA_ERR_UNSUPPORTED, "%s: only MPI_DOUBLE, MPI_DOUBLE_COMPLEX, MPI_FLOAT and MPI_INT objects", who);
  return PLA_SUCCESS;
}

int check_blocks(PLA_Obj A, const blocks &b, const char *who) {
  for (size_t i = 0; i < b.rsz.size(); i++)
    if (b.rsz[i] < 0 || b.rdisp[i] < 0 || (long)b.rdisp[i] + b.rsz[i] > A->length)
      return fail(PLA_ERR_DIMENSIONS, "%s: rows %d..%d of a %d-row object", who, b.rdisp[i],
                  b.rdisp[i] + b.rsz[i] - 1, A->length);
  for (size_t j = 0; j < b.csz.size(); j++)
    if (b.csz[j] < 0 || b.cdisp[j] < 0 || (long)b.cdisp[j] + b.csz[j] > A->width)
      return fail(PLA_ERR_DIMENSIONS, "%s: columns %d..%d of a %d-column object", who, b.cdisp[j],
                  b.cdisp[j] + b.csz[j] - 1, A->width);
  return PLA_SUCCESS;
}

// Do the non-empty intervals [disp, disp + size) overlap?
bool overlapping(const std::vector<int> &sz, const std::vector<int> &disp) {
  std::vector<std::pair<int, int>> iv;
  for (size_t i = 0; i < sz.size(); i++)
    if (sz[i] > 0) iv.emplace_back(disp[i], disp[i] + sz[i]);
  std::sort(iv.begin(), iv.end());
  for (size_t i = 1; i < iv.size(); i++)
    if (iv[i].first < iv[i - 1].second) return true;
  return false;
}

// Global index of each packed row (column).
std::vector<int> index_map(const std::vector<int> &sz, const std::vector<int> &disp) {
  std::vector<int> m;
  for (size_t i = 0; i < sz.size(); i++)
    for (int r = 0; r < sz[i]; r++) m.push_back(disp[i] + r);
  return m;
}

// Columns per chunk: the packed data plus both index maps fit one slot.
int chunk_cols(int M, int N, int ts) {
  const size_t per_col = (size_t)M * ts + sizeof(int);
  const size_t fixed = (size_t)M * sizeof(int) + 1024;  // row map and alignment padding
  size_t c = (api().pinned_bytes - fixed) / per_col;
  if (c < 1) c = 1;
  return (int)std::min<size_t>(c, (size_t)N);
}

// Packed layout of a staging slot: data (M x nc, leading dimension M), then
// the row map (M ints), then the column map (nc ints), each 256-byte aligned.
struct slot_layout {
  size_t rmap_off, cmap_off, bytes;
};
slot_layout layout(int M, int nc, int ts) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  slot_layout l;
  l.rmap_off = up((size_t)M * nc * ts);
  l.cmap_off = up(l.rmap_off + (size_t)M * sizeof(int));
  l.bytes = l.cmap_off + (size_t)nc * sizeof(int);
  return l;
}

// Host buffer (column j at buf + j * ldim, element i at stride `inc`) added
// into the object: alpha times the packed M x N block, packed row i at
// global row rmap[i], packed column j at cmap[j].
int to_global(PLA_Obj A, int M, int N, const void *alpha, const void *buf, int ldim, int inc,
              const std::vector<int> &rmap, const std::vector<int> &cmap) {
  if (M <= 0 || N <= 0) return PLA_SUCCESS;
  const int ts = A->base->typesize, kind = elem_kind(A->base->datatype);
  int rc;
  if ((rc = ensure_staging(layout(M, 1, ts).bytes))) return rc;
  // a host mirror the caller may have written is pushed by dev_ptr: let the
  // adds already queued on the object finish first
  if (A->base->host && A->base->host_dirty) PLA_HIP(hipStreamSynchronize(ctx().stream));
  void *dst;
  if ((rc = dev_ptr(A, &dst))) return rc;
  api_ctx &a = api();
  hipStream_t s = ctx().stream;
  const int cc = chunk_cols(M, N, ts);
  for (int j0 = 0; j0 < N; j0 += cc) {
    const int nc = std::min(cc, N - j0);
    const int slot = a.next;
    a.next ^= 1;
    PLA_HIP(hipEventSynchronize(a.done[slot]));  // the copy that last read this slot is done
    char *h = (char *)a.pinned[slot];
    const slot_layout L = layout(M, nc, ts);
    for (int j = 0; j < nc; j++) {
      const char *src = (const char *)buf + (size_t)(j0 + j) * ldim * ts;
      char *d = h + (size_t)j * M * ts;
      if (inc == 1) {
        memcpy(d, src, (size_t)M * ts);
      } else {
        for (int i = 0; i < M; i++) memcpy(d + (size_t)i * ts, src + (size_t)i * inc * ts, ts);
      }
    }
    memcpy(h + L.rmap_off, rmap.data(), (size_t)M * sizeof(int));
    memcpy(h + L.cmap_off, cmap.data() + j0, (size_t)nc * sizeof(int));
    char *dv = (char *)workspace(a.pinned_bytes, WS_API_0 + slot);
    PLA_HIP(hipMemcpyAsync(dv, h, L.bytes, hipMemcpyHostToDevice, s));
    PLA_HIP(hipEventRecord(a.done[slot], s));
    if ((rc = k_api_scatter_add(s, kind, M, nc, alpha, dv, dst, ld(A), (const int *)(dv + L.rmap_off),
                                (const int *)(dv + L.cmap_off), 0, 0)))
      return rc;
  }
  return PLA_SUCCESS;
}

struct zpair {
  double re, im;
};
inline zpair operator*(zpair a, zpair b) {
#pragma clang fp contract(off)
  return zpair{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline zpair &operator+=(zpair &a, zpair b) {
  a.re += b.re;
  a.im += b.im;
  return a;
}

template <typename T>
void add_scaled(int M, int nc, const void *alpha, const char *packed, void *buf, int ldim, int inc, int j0) {
#pragma clang fp contract(off)  // alpha * x rounded, then added, as on the device
  const T al = *(const T *)alpha;
  const T *p = (const T *)packed;
  for (int j = 0; j < nc; j++) {
    T *c = (T *)buf + (size_t)(j0 + j) * ldim;
    const T *q = p + (size_t)j * M;
    for (int i = 0; i < M; i++) c[(size_t)i * inc] += al * q[i];
  }
}

// The host buffer += alpha times the gathered M x N block.
int from_global(PLA_Obj A, int M, int N, const void *alpha, void *buf, int ldim, int inc,
                const std::vector<int> &rmap, const std::vector<int> &cmap) {
  if (M <= 0 || N <= 0) return PLA_SUCCESS;
  const int ts = A->base->typesize, kind = elem_kind(A->base->datatype);
  int rc;
  if ((rc = ensure_staging(layout(M, 1, ts).bytes))) return rc;
  if (A->base->host && A->base->host_dirty) PLA_HIP(hipStreamSynchronize(ctx().stream));
  void *src;
  if ((rc = dev_ptr(A, &src))) return rc;
  api_ctx &a = api();
  hipStream_t s = ctx().stream;
  const int cc = chunk_cols(M, N, ts);
  // Queue chunk c + 1 (maps up, gather, data down) before adding chunk c into
  // the caller's buffer on the host.
  struct pending {
    int slot, j0, nc;
  };
  auto issue = [&](int j0, pending *p) -> int {
    const int nc = std::min(cc, N - j0);
    const int slot = a.next;
    a.next ^= 1;
    PLA_HIP(hipEventSynchronize(a.done[slot]));
    char *h = (char *)a.pinned[slot];
    const slot_layout L = layout(M, nc, ts);
    memcpy(h + L.rmap_off, rmap.data(), (size_t)M * sizeof(int));
    memcpy(h + L.cmap_off, cmap.data() + j0, (size_t)nc * sizeof(int));
    char *dv = (char *)workspace(a.pinned_bytes, WS_API_0 + slot);
    PLA_HIP(hipMemcpyAsync(dv + L.rmap_off, h + L.rmap_off, L.bytes - L.rmap_off, hipMemcpyHostToDevice, s));
    int r = k_api_gather(s, kind, M, nc, src, ld(A), dv, (const int *)(dv + L.rmap_off),
                         (const int *)(dv + L.cmap_off), 0, 0);
    if (r) return r;
    PLA_HIP(hipMemcpyAsync(h, dv, (size_t)M * nc * ts, hipMemcpyDeviceToHost, s));
    PLA_HIP(hipEventRecord(a.done[slot], s));
    *p = pending{slot, j0, nc};
    return PLA_SUCCESS;
  };
  pending cur;
  if ((rc = issue(0, &cur))) return rc;
  for (;;) {
    pending nxt{-1, 0, 0};
    const int jn = cur.j0 + cur.nc;
    if (jn < N && (rc = issue(jn, &nxt))) return rc;
    PLA_HIP(hipEventSynchronize(a.done[cur.slot]));
    const char *h = (const char *)a.pinned[cur.slot];
    if (kind == API_F64)
      add_scaled<double>(M, cur.nc, alpha, h, buf, ldim, inc, cur.j0);
    else if (kind == API_F32)
      add_scaled<float>(M, cur.nc, alpha, h, buf, ldim, inc, cur.j0);
    else if (kind == API_Z64)
      add_scaled<zpair>(M, cur.nc, alpha, h, buf, ldim, inc, cur.j0);
    else
      add_scaled<int>(M, cur.nc, alpha, h, buf, ldim, inc, cur.j0);
    if (nxt.slot < 0) break;
    cur = nxt;
  }
  return PLA_SUCCESS;
}

const void *one_of(PLA_Obj A) {
  static const double d1 = 1.0;
  static const float f1 = 1.0f;
  static const int i1 = 1;
  static const zpair z1 = {1.0, 0.0};
  const int k = elem_kind(A->base->datatype);
  return k == API_F64 ? (const void *)&d1
         : k == API_F32 ? (const void *)&f1
         : k == API_Z64 ? (const void *)&z1
                        : (const void *)&i1;
}

// One rectangular block or a grid of sub-blocks, in either direction.  The
// sub-blocks go in one pass when their row intervals and their column
// intervals are each pairwise disjoint (no element of the object is touched
// twice); otherwise block by block, in the reference's order.
int transfer(bool to_obj, PLA_Obj A, const blocks &b, const void *alpha, void *buf, int ldim, int inc,
             const char *who) {
  int rc;
  if ((rc = check_target(A, who)) || (rc = check_blocks(A, b, who))) return rc;
  if (b.M <= 0 || b.N <= 0) return PLA_SUCCESS;
  if (!buf) return fail(PLA_ERR_INVALID_PARAM, "%s: NULL local buffer", who);
  if (inc < 1) return fail(PLA_ERR_INVALID_PARAM, "%s: local increment %d", who, inc);
  if ((long)ldim < (long)(b.M - 1) * inc + 1 && b.N > 1)
    return fail(PLA_ERR_INVALID_PARAM, "%s: local leading dimension %d < %d rows", who, ldim, b.M);
  if (!alpha) alpha = one_of(A);
  const bool split = overlapping(b.rsz, b.rdisp) || overlapping(b.csz, b.cdisp);
  if (!split) {
    const std::vector<int> rmap = index_map(b.rsz, b.rdisp), cmap = index_map(b.csz, b.cdisp);
    return to_obj ? to_global(A, b.M, b.N, alpha, buf, ldim, inc, rmap, cmap)
                  : from_global(A, b.M, b.N, alpha, buf, ldim, inc, rmap, cmap);
  }
  const int ts = A->base->typesize;
  int li = 0;
  for (size_t i = 0; i < b.rsz.size(); i++) {
    int lj = 0;
    for (size_t j = 0; j < b.csz.size(); j++) {
      std::vector<int> rmap(b.rsz[i]), cmap(b.csz[j]);
      for (int r = 0; r < b.rsz[i]; r++) rmap[r] = b.rdisp[i] + r;
      for (int c = 0; c < b.csz[j]; c++) cmap[c] = b.cdisp[j] + c;
      char *sub = (char *)buf + ((size_t)lj * ldim + (size_t)li * inc) * ts;
      rc = to_obj ? to_global(A, b.rsz[i], b.csz[j], alpha, sub, ldim, inc, rmap, cmap)
                  : from_global(A, b.rsz[i], b.csz[j], alpha, sub, ldim, inc, rmap, cmap);
      if (rc) return rc;
      lj += b.csz[j];
    }
    li += b.rsz[i];
  }
  return PLA_SUCCESS;
}

blocks one_block(int m, int n, int dr, int dc) {
  blocks b;
  b.rsz = {m};
  b.rdisp = {dr};
  b.csz = {n};
  b.cdisp = {dc};
  b.M = m;
  b.N = n;
  return b;
}

int grid_blocks(int nr, int nc, const int *rsz, const int *csz, const int *rdisp, const int *cdisp, blocks *b,
                const char *who) {
  if (nr < 0 || nc < 0) return fail(PLA_ERR_INVALID_PARAM, "%s: %d x %d sub-blocks", who, nr, nc);
  if ((nr && (!rsz || !rdisp)) || (nc && (!csz || !cdisp)))
    return fail(PLA_ERR_INVALID_PARAM, "%s: NULL size or displacement array", who);
  b->rsz.assign(rsz, rsz + nr);
  b->rdisp.assign(rdisp, rdisp + nr);
  b->csz.assign(csz, csz + nc);
  b->cdisp.assign(cdisp, cdisp + nc);
  long M = 0, N = 0;
  for (int i = 0; i < nr; i++) M += rsz[i] > 0 ? rsz[i] : 0;
  for (int j = 0; j < nc; j++) N += csz[j] > 0 ? csz[j] : 0;
  if (M > INT32_MAX || N > INT32_MAX) return fail(PLA_ERR_DIMENSIONS, "%s: block too large", who);
  b->M = (int)M;
  b->N = (int)N;
  return PLA_SUCCESS;
}

// The column a vector transfer addresses: the reference passes the template's
// zero_or_one as the column displacement (API/PLA_API.c:247, :285).
int vector_col(PLA_Obj laobj) {
  return (valid(laobj) && laobj->base->templ) ? laobj->base->templ->zero_or_one : 0;
}

}  // namespace

namespace pla {

void api_forget(PLA_Obj obj) {
  auto &v = api().open;
  v.erase(std::remove(v.begin(), v.end(), obj), v.end());
}

void api_release() {
  api_ctx &a = api();
  if (ctx().stream) (void)hipStreamSynchronize(ctx().stream);
  for (int i = 0; i < 2; i++) {
    if (a.pinned[i]) (void)hipHostFree(a.pinned[i]);
    a.pinned[i] = nullptr;
    if (a.done[i]) (void)hipEventDestroy(a.done[i]);
    a.done[i] = nullptr;
  }
  a.pinned_bytes = 0;
  a.next = 0;
  a.open.clear();
  a.active = false;
}

}  // namespace pla

extern "C" {

int PLA_API_begin(void) {
  if (!ctx().initialized) return fail(PLA_ERR_NOT_INITIALIZED, "PLA_API_begin: PLAPACK not yet initialized");
  if (api().active) return fail(PLA_ERR_INVALID_PARAM, "PLA_API_begin: API already active");
  int rc = ensure_device();
  if (rc) return rc;
  api().active = true;
  return PLA_SUCCESS;
}

int PLA_API_end(void) {
  api_ctx &a = api();
  if (!a.active) return fail(PLA_ERR_INVALID_PARAM, "PLA_API_end: API not active");
  if (!a.open.empty())
    return fail(PLA_ERR_INVALID_PARAM, "PLA_API_end: Cannot end API without closing all objects (%zu open)",
                a.open.size());
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  a.active = false;
  return PLA_SUCCESS;
}

int PLA_API_state(int *state) {
  if (!state) return fail(PLA_ERR_INVALID_PARAM, "PLA_API_state: NULL state");
  *state = api().active ? PLA_STATE_API_ACTIVE : PLA_STATE_NORMAL;
  return PLA_SUCCESS;
}

int PLA_Obj_API_open(PLA_Obj obj) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_API_open: invalid object");
  if (!api().active) return fail(PLA_ERR_INVALID_PARAM, "PLA_Obj_API_open: API not active");
  if (is_open(obj)) return fail(PLA_ERR_INVALID_PARAM, "PLA_Obj_API_open: object already open");
  api().open.push_back(obj);
  return PLA_SUCCESS;
}

// Completes the object's pending adds (and refreshes a host mirror).
int PLA_Obj_API_sync(PLA_Obj obj) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_API_sync: invalid object");
  PLA_HIP(hipStreamSynchronize(ctx().stream));
  return written(obj);
}

int PLA_Obj_API_close(PLA_Obj obj) {
  if (!valid(obj)) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_API_close: invalid object");
  if (!is_open(obj)) return fail(PLA_ERR_INVALID_PARAM, "PLA_Obj_API_close: object not open");
  int rc = PLA_Obj_API_sync(obj);
  if (rc) return rc;
  api_forget(obj);
  return PLA_SUCCESS;
}

int PLA_Obj_API_mode(PLA_Obj obj, int *mode) {
  if (!valid(obj) || !mode) return fail(PLA_ERR_INVALID_OBJECT, "PLA_Obj_API_mode: invalid argument");
  *mode = is_open(obj) ? PLA_MODE_OPEN : PLA_MODE_CLOSED;
  return PLA_SUCCESS;
}

int PLA_API_axpy_matrix_to_global(int m, int n, void *alpha, void *local_buf, int local_ldim, PLA_Obj A,
                                  int disp_row, int disp_col) {
  return transfer(true, A, one_block(m, n, disp_row, disp_col), alpha, local_buf, local_ldim, 1,
                  "PLA_API_axpy_matrix_to_global");
}

int PLA_API_axpy_global_to_matrix(int m, int n, void *alpha, PLA_Obj A, int disp_row, int disp_col,
                                  void *local_buf, int local_ldim) {
  return transfer(false, A, one_block(m, n, disp_row, disp_col), alpha, local_buf, local_ldim, 1,
                  "PLA_API_axpy_global_to_matrix");
}

int PLA_API_axpy_vector_to_global(int size, void *alpha, void *local_vector, int local_inc, PLA_Obj laobj,
                                  int disp) {
  return transfer(true, laobj, one_block(size, 1, disp, vector_col(laobj)), alpha, local_vector,
                  size > 0 ? (size - 1) * local_inc + 1 : 1, local_inc, "PLA_API_axpy_vector_to_global");
}

int PLA_API_axpy_global_to_vector(int size, void *alpha, PLA_Obj laobj, int disp, void *local_vector,
                                  int local_inc) {
  return transfer(false, laobj, one_block(size, 1, disp, vector_col(laobj)), alpha, local_vector,
                  size > 0 ? (size - 1) * local_inc + 1 : 1, local_inc, "PLA_API_axpy_global_to_vector");
}

int PLA_API_multi_axpy_vector_to_global(int nsub, int *sizes, void *alpha, void *local_vector, int local_stride,
                                        PLA_Obj laobj, int *displs) {
  const char *who = "PLA_API_multi_axpy_vector_to_global";
  const int one = 1, col = vector_col(laobj);
  blocks b;
  int rc = grid_blocks(nsub, 1, sizes, &one, displs, &col, &b, who);
  if (rc) return rc;
  return transfer(true, laobj, b, alpha, local_vector, b.M > 0 ? (b.M - 1) * local_stride + 1 : 1, local_stride,
                  who);
}

int PLA_API_multi_axpy_global_to_vector(int nsub, int *sizes, void *alpha, PLA_Obj laobj, int *displs,
                                        void *local_vector, int local_stride) {
  const char *who = "PLA_API_multi_axpy_global_to_vector";
  const int one = 1, col = vector_col(laobj);
  blocks b;
  int rc = grid_blocks(nsub, 1, sizes, &one, displs, &col, &b, who);
  if (rc) return rc;
  return transfer(false, laobj, b, alpha, local_vector, b.M > 0 ? (b.M - 1) * local_stride + 1 : 1, local_stride,
                  who);
}

int PLA_API_multi_axpy_matrix_to_global(int nsub_row, int nsub_col, int *size_row, int *size_col, void *alpha,
                                        void *local_matrix, int local_ldim, PLA_Obj obj, int *disp_row,
                                        int *disp_col) {
  const char *who = "PLA_API_multi_axpy_matrix_to_global";
  blocks b;
  int rc = grid_blocks(nsub_row, nsub_col, size_row, size_col, disp_row, disp_col, &b, who);
  if (rc) return rc;
  return transfer(true, obj, b, alpha, local_matrix, local_ldim, 1, who);
}

int PLA_API_multi_axpy_global_to_matrix(int nsub_row, int nsub_col, int *size_row, int *size_col, void *alpha,
                                        PLA_Obj obj, int *disp_row, int *disp_col, void *local_matrix,
                                        int local_ldim) {
  const char *who = "PLA_API_multi_axpy_global_to_matrix";
  blocks b;
  int rc = grid_blocks(nsub_row, nsub_col, size_row, size_col, disp_row, disp_col, &b, who);
  if (rc) return rc;
  return transfer(false, obj, b, alpha, local_matrix, local_ldim, 1, who);
}

}  // extern "C"
