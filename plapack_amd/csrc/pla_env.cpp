// pla_env.cpp -- initialisation, 1x1 template, block-size environment,
// checking flags, error reporting and device workspaces.
//
// Reference counterparts: TEMPLATE/PLA_Init.c, TEMPLATE/PLA_Temp.c,
// TEMPLATE/PLA_Environ.c, OBJ/PLA_Diagnostics.c.
#include "pla_internal.hpp"

#include <map>
#include <utility>

#include <dlfcn.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

int PLA_ERROR_CHECKING = 1;            // OBJ/PLA_Diagnostics.c:16-19
int PLA_CHECK_PARAMETERS = 1;
int PLA_CHECK_AGAINST_SEQUENTIAL = 0;
int PLA_CHECK_AGAINST_R12 = 0;

namespace pla {

static std::string g_last_error;
static context g_ctx;
static hipEvent_t g_events[EV_COUNT];
context &ctx() { return g_ctx; }

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? PLA_ERR_NO_MEMORY : PLA_ERR_DEVICE, "%s: %s", what,
              hipGetErrorString(e));
}

int ensure_device() {
  context &c = g_ctx;
  if (c.stream) return PLA_SUCCESS;
  int count = 0;
  PLA_HIP(hipGetDeviceCount(&count));
  if (count < 1) return fail(PLA_ERR_DEVICE, "no HIP device visible");
  PLA_HIP(hipGetDevice(&c.device));
  PLA_HIP(hipSetDevice(c.device));
  int cus = 0;
  PLA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
  c.cu_count = cus > 0 ? cus : 256;
  PLA_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  PLA_HIP(hipStreamCreateWithFlags(&c.stream_la, hipStreamNonBlocking));
  PLA_HIP(hipStreamCreateWithFlags(&c.stream_tr, hipStreamNonBlocking));
  for (auto &e : g_events)
    if (!e) PLA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (c.chol_panel_cus > 0 && c.chol_panel_cus < c.cu_count) {
    const int rc = masked_streams(c.chol_panel_cus);
    if (rc) return rc;
  }
  // the no-allocation Cholesky's static scratch: its address on this device
  void *scratch_ptr = nullptr;
  return chol_static_scratch(c.device, &scratch_ptr);
}

hipEvent_t lib_event(int id) { return id >= 0 && id < EV_COUNT ? g_events[id] : nullptr; }

// Destroys every stream and event the library created (PLA_Finalize); the
// next PLA_Init creates them again.
static void release_streams(context &c) {
  hipStream_t *streams[] = {&c.stream, &c.stream_la, &c.stream_tr, &c.stream_panel, &c.stream_trail};
  for (hipStream_t *sp : streams)
    if (*sp) (void)hipStreamSynchronize(*sp);
  for (auto &e : g_events) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  for (hipStream_t *sp : streams) {
    if (*sp) (void)hipStreamDestroy(*sp);
    *sp = nullptr;
  }
  c.masked_cus = 0;
}

struct ws_slot {
  void *p = nullptr;
  size_t bytes = 0;
};
static ws_slot g_ws[PLA_NUM_WS];

void *workspace(size_t bytes, int slot, bool zero) {
  if (slot < 0 || slot >= PLA_NUM_WS) return nullptr;
  ws_slot &w = g_ws[slot];
  if (w.bytes >= bytes && w.p) return w.p;
  if (w.p) {
    // in-flight kernels may still use the old buffer
    if (hipDeviceSynchronize() != hipSuccess) return nullptr;
    (void)hipFree(w.p);
    w.p = nullptr;
    w.bytes = 0;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  if (hipMalloc(&w.p, want) != hipSuccess) {
    w.p = nullptr;
    fail(PLA_ERR_NO_MEMORY, "workspace allocation of %zu bytes failed", want);
    return nullptr;
  }
  // the library's streams do not synchronise with the null stream: finish the
  // clear before any of them can touch the buffer
  if (zero && (hipMemset(w.p, 0, want) != hipSuccess || hipDeviceSynchronize() != hipSuccess)) {
    (void)hipFree(w.p);
    w.p = nullptr;
    fail(PLA_ERR_DEVICE, "workspace initialisation failed");
    return nullptr;
  }
  w.bytes = want;
  return w.p;
}

// CU-masked stream pair for the LU look-ahead.  Logical CU i of a mask lives
// on XCD i % 8 and shader engine (i / 8) % 4 of it, so bits [0, k) with k a
// multiple of 32 give every shader engine of every XCD k / 32 CUs.  Uneven
// counts per shader engine are avoided: the panel leaf needs all of its
// workgroups resident at once, and workgroups are spread evenly over the
// shader engines whatever CUs each one has free.
static int make_masked_streams(context &c, int k) {
  if (c.masked_cus == k && c.stream_panel) return PLA_SUCCESS;
  if (c.stream_panel) {
    PLA_HIP(hipStreamSynchronize(c.stream_panel));
    PLA_HIP(hipStreamSynchronize(c.stream_trail));
    PLA_HIP(hipStreamDestroy(c.stream_panel));
    PLA_HIP(hipStreamDestroy(c.stream_trail));
    c.stream_panel = c.stream_trail = nullptr;
    c.masked_cus = 0;
  }
  const int words = (c.cu_count + 31) / 32;
  std::vector<uint32_t> mp(words, 0u), mt(words, 0u);
  for (int i = 0; i < c.cu_count; i++) {
    if (i < k) mp[i / 32] |= 1u << (i % 32);
    else mt[i / 32] |= 1u << (i % 32);
  }
  PLA_HIP(hipExtStreamCreateWithCUMask(&c.stream_panel, (uint32_t)words, mp.data()));
  PLA_HIP(hipExtStreamCreateWithCUMask(&c.stream_trail, (uint32_t)words, mt.data()));
  c.masked_cus = k;
  return PLA_SUCCESS;
}

int lu_streams_begin(hipStream_t main, int m, hipStream_t *sP, hipStream_t *sT, int *gcap) {
  context &c = g_ctx;
  *gcap = 0;
  if (!c.lookahead) {
    *sP = *sT = main;
    return PLA_SUCCESS;
  }
  if (!g_events[EV_JOIN0]) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  int k = c.panel_cus;
  if (k > 0 && c.cu_count % 32 == 0 && k < c.cu_count) {
    // enough CUs that the first (tallest) panel's leaf rows fit in LDS
    while (k + 32 < c.cu_count && (c.leaf_kind ? leaf_width_for(m, k) : rleaf_width_for(m, 4 * k)) == 0) k += 32;
    int rc = make_masked_streams(c, k);
    if (rc) return rc;
    *sP = c.stream_panel;
    *sT = c.stream_trail;
    *gcap = k;
  } else {
    *sP = c.stream_la;
    *sT = c.stream_tr;
  }
  PLA_HIP(hipEventRecord(g_events[EV_JOIN0], main));
  PLA_HIP(hipStreamWaitEvent(*sP, g_events[EV_JOIN0], 0));
  if (*sT != main) PLA_HIP(hipStreamWaitEvent(*sT, g_events[EV_JOIN0], 0));
  return PLA_SUCCESS;
}

int masked_streams(int k) { return make_masked_streams(g_ctx, k); }

int chol_streams_begin(hipStream_t main, hipStream_t *sP, hipStream_t *sT, bool allow_mask) {
  context &c = g_ctx;
  if (!c.lookahead) {
    *sP = *sT = main;
    return PLA_SUCCESS;
  }
  if (!g_events[EV_JOIN0]) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  if (allow_mask && c.chol_panel_cus > 0 && c.masked_cus == c.chol_panel_cus && c.stream_panel) {
    *sP = c.stream_panel;
    *sT = c.stream_trail;
  } else {
    *sP = c.stream_la;
    *sT = main;
  }
  PLA_HIP(hipEventRecord(g_events[EV_JOIN0], main));
  PLA_HIP(hipStreamWaitEvent(*sP, g_events[EV_JOIN0], 0));
  if (*sT != main) PLA_HIP(hipStreamWaitEvent(*sT, g_events[EV_JOIN0], 0));
  return PLA_SUCCESS;
}

int lu_streams_end(hipStream_t main, hipStream_t sP, hipStream_t sT) {
  if (sP != main) {
    PLA_HIP(hipEventRecord(g_events[EV_JOIN1], sP));
    PLA_HIP(hipStreamWaitEvent(main, g_events[EV_JOIN1], 0));
  }
  if (sT != main) {
    PLA_HIP(hipEventRecord(g_events[EV_JOIN2], sT));
    PLA_HIP(hipStreamWaitEvent(main, g_events[EV_JOIN2], 0));
  }
  return PLA_SUCCESS;
}

static leaf_ws g_leaf;

leaf_ws &leaf_workspace(size_t slot_bytes) {
  if (g_leaf.base && g_leaf.slot_bytes >= slot_bytes) return g_leaf;
  if (g_leaf.base) {
    if (hipDeviceSynchronize() != hipSuccess) return g_leaf;
    (void)hipFree(g_leaf.base);
    g_leaf = leaf_ws();
  }
  const size_t total = slot_bytes + 256;
  void *p = nullptr;
  if (hipMalloc(&p, total) != hipSuccess) {
    fail(PLA_ERR_NO_MEMORY, "panel exchange workspace allocation failed");
    return g_leaf;
  }
  if (hipMemset(p, 0, total) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(p);
    fail(PLA_ERR_DEVICE, "panel exchange workspace initialisation failed");
    return g_leaf;
  }
  g_leaf.base = p;
  g_leaf.slot_bytes = slot_bytes;
  g_leaf.slots = (unsigned long long *)p;
  g_leaf.err = (int *)((char *)p + slot_bytes);
  g_leaf.state = (unsigned long long *)((char *)p + slot_bytes + 64);
  return g_leaf;
}

// Exchange tags of the panel leaves: one counter for every leaf kernel (they
// share the slot words), so a stale word never carries a current tag.
static unsigned g_leaf_tag = 1;

int leaf_tags(hipStream_t s, unsigned count, unsigned *tag0) {
  if (!g_leaf.base) return fail(PLA_ERR_INVALID_PARAM, "panel exchange workspace not allocated");
  if (g_leaf_tag > 0xF0000000u) {  // tags only grow; restart far below wrap-around
    PLA_HIP(hipMemsetAsync(g_leaf.slots, 0, g_leaf.slot_bytes, s));
    PLA_HIP(hipMemsetAsync(g_leaf.state, 0, sizeof(unsigned long long), s));
    g_leaf_tag = 1;
  }
  *tag0 = g_leaf_tag;
  g_leaf_tag += count;
  return PLA_SUCCESS;
}

int *leaf_error_word() {
  leaf_ws &w = leaf_workspace(g_leaf.base ? g_leaf.slot_bytes : 4096);
  return w.err;
}

int leaf_check_error(hipStream_t s, const char *what) {
  if (!g_leaf.base) return PLA_SUCCESS;
  int h = 0;
  PLA_HIP(hipMemcpyAsync(&h, g_leaf.err, sizeof(int), hipMemcpyDeviceToHost, s));
  PLA_HIP(hipStreamSynchronize(s));
  if (h) {
    PLA_HIP(hipMemsetAsync(g_leaf.err, 0, sizeof(int), s));
    PLA_HIP(hipStreamSynchronize(s));
    return fail(PLA_ERR_DEVICE, "%s: a cross-workgroup exchange timed out", what);
  }
  return PLA_SUCCESS;
}

void release_workspaces() {
  for (auto &w : g_ws) {
    if (w.p) (void)hipFree(w.p);
    w.p = nullptr;
    w.bytes = 0;
  }
  if (g_leaf.base) (void)hipFree(g_leaf.base);
  g_leaf = leaf_ws();
}

// Block sizes measured by PLA_Autotune_nb, per (factorization, n), for the
// process lifetime.
static std::map<std::pair<int, int>, int> g_tuned_nb;

void set_tuned_nb(int kind, int n, int nb) { g_tuned_nb[{kind, n}] = nb; }

int choose_nb(int n, int kind) {
  const context &c = g_ctx;
  if (c.nb_device > 0) return c.nb_device;
  if (c.nb_user > 0) return c.nb_user;
  const auto t = g_tuned_nb.find({kind, n});
  if (t != g_tuned_nb.end()) return t->second;
  // Measured on MI355X.  Plain FP64 LU: n=32768 nb 256 / 384 / 512 -> 25.8 /
  // 27.1 / 27.9 TF/s; n=65536 nb 512 / 1024 -> 42.6 / 45.1 TF/s (the
  // trailing GEMM runs faster at K = 1024).  Mixed LU n=65536: nb 512 / 1024
  // -> 101.3 / 92.1 TF/s (the wider FP64 panel lengthens the critical path).
  if (kind == NB_LU && n >= 49152) return 1024;
  if (n >= 16384) return 512;
  if (n >= 4096) return 256;
  return 128;
}

int lu_step_width(int m, int mn, int nb, int k) {
  const context &c = g_ctx;
  int w = nb;
  if (c.nb_tail > 0 && m - k <= c.nb_switch) w = c.nb_tail < nb ? c.nb_tail : nb;  // PLA_PARAM_NB_TAIL / _SWITCH
  return mn - k < w ? mn - k : w;
}

int lu_step_count(int m, int mn, int nb) {
  int steps = 0;
  for (int k = 0; k < mn; k += lu_step_width(m, mn, nb, k)) steps++;
  return steps > 0 ? steps : 1;
}

}  // namespace pla

using pla::fail;

static int g_initialized = 0;
static MPI_Comm g_comm = MPI_COMM_NULL;
static int g_nb_table[6] = {128, 128, 128, 128, 128, 128};  // TEMPLATE/PLA_Environ.c:37-44

// If the calling process has initialised an MPI library, ask it for the
// communicator size (1x1 template only).  No MPI library is linked.
static int mpi_comm_size(MPI_Comm comm, int *size) {
  *size = 1;
  typedef int (*initialized_fn)(int *);
  typedef int (*size_fn)(MPI_Comm, int *);
  initialized_fn initd = (initialized_fn)dlsym(RTLD_DEFAULT, "MPI_Initialized");
  size_fn sz = (size_fn)dlsym(RTLD_DEFAULT, "MPI_Comm_size");
  if (!initd || !sz) return 0;
  int flag = 0;
  if (initd(&flag) != 0 || !flag) return 0;
  if (comm == MPI_COMM_NULL) return 0;
  return sz(comm, size);
}

extern "C" {

const char *PLA_Last_error(void) { return pla::g_last_error.c_str(); }

int PLA_Comm_1D_to_2D(MPI_Comm comm1d, int rows, int cols, MPI_Comm *comm2d) {
  if (!comm2d) return fail(PLA_ERR_INVALID_PARAM, "PLA_Comm_1D_to_2D: comm2d is NULL");
  if (rows * cols != 1)
    return fail(PLA_ERR_MULTI_PROCESS,
                "PLA_Comm_1D_to_2D: only a 1x1 mesh is supported (asked for %d x %d)", rows, cols);
  int size = 1;
  mpi_comm_size(comm1d, &size);
  if (size != 1)
    return fail(PLA_ERR_MULTI_PROCESS, "PLA_Comm_1D_to_2D: communicator has %d ranks", size);
  *comm2d = comm1d;
  return PLA_SUCCESS;
}

// TEMPLATE/PLA_Init.c:38-83 picks the divisor of the rank count closest to
// sqrt(ratio * ranks) as the row count; with one rank that is 1 x 1.
int PLA_Comm_1D_to_2D_ratio(MPI_Comm comm1d, double ratio, MPI_Comm *comm2d) {
  (void)ratio;
  if (!comm2d) return fail(PLA_ERR_INVALID_PARAM, "PLA_Comm_1D_to_2D_ratio: comm2d is NULL");
  return PLA_Comm_1D_to_2D(comm1d, 1, 1, comm2d);
}

int PLA_Init(MPI_Comm comm) {
  if (g_initialized) return fail(PLA_ERR_INVALID_PARAM, "PLA_Init called more than once");
  int size = 1;
  mpi_comm_size(comm, &size);
  if (size != 1)
    return fail(PLA_ERR_MULTI_PROCESS, "PLA_Init: communicator has %d ranks, need 1", size);
  int rc = pla::ensure_device();
  if (rc) return rc;
  g_comm = comm;
  g_initialized = 1;
  pla::ctx().initialized = true;
  return PLA_SUCCESS;
}

int PLA_Finalize(void) {
  if (!g_initialized) return fail(PLA_ERR_NOT_INITIALIZED, "PLAPACK not yet initialized");
  pla::context &c = pla::ctx();
  pla::api_release();
  pla::timing_release();
  pla::release_streams(c);
  pla::release_workspaces();
  g_initialized = 0;
  c.initialized = false;
  return PLA_SUCCESS;
}

int PLA_Initialized(int *initialized) {
  if (initialized) *initialized = g_initialized;
  return PLA_SUCCESS;
}

int PLA_Temp_create(int nb, int zero_or_one, PLA_Template *templ) {
  if (!templ) return fail(PLA_ERR_INVALID_PARAM, "PLA_Temp_create: templ is NULL");
  if (!g_initialized) return fail(PLA_ERR_NOT_INITIALIZED, "PLA_Temp_create before PLA_Init");
  if (nb <= 0) return fail(PLA_ERR_INVALID_PARAM, "PLA_Temp_create: nb = %d", nb);
  if (zero_or_one != 0 && zero_or_one != 1)
    return fail(PLA_ERR_INVALID_PARAM, "Error in PLA_temp_create: zero_or_one");
  PLA_Template t = (PLA_Template)calloc(1, sizeof(struct PLA_Template_struct));
  if (!t) return fail(PLA_ERR_NO_MEMORY, "PLA_Temp_create: out of host memory");
  t->comm_all = t->comm_row = t->comm_col = g_comm;
  t->nb = nb;
  t->zero_or_one = zero_or_one;
  t->nprows = t->npcols = 1;
  t->times_used = 1;
  *templ = t;
  return PLA_SUCCESS;
}

int PLA_Temp_free(PLA_Template *templ) {
  if (!templ || !*templ) return PLA_SUCCESS;
  if (--(*templ)->times_used <= 0) free(*templ);
  *templ = NULL;
  return PLA_SUCCESS;
}

int PLA_Temp_nb(PLA_Template templ, int *nb) {
  if (!templ || !nb) return fail(PLA_ERR_INVALID_PARAM, "PLA_Temp_nb: NULL argument");
  *nb = templ->nb;
  return PLA_SUCCESS;
}
int PLA_Temp_zero_or_one(PLA_Template templ, int *zero_or_one) {
  if (!templ || !zero_or_one) return fail(PLA_ERR_INVALID_PARAM, "PLA_Temp_zero_or_one: NULL");
  *zero_or_one = templ->zero_or_one;
  return PLA_SUCCESS;
}
int PLA_Temp_comm_all_size(PLA_Template templ, int *nprocs) {
  if (!templ || !nprocs) return fail(PLA_ERR_INVALID_PARAM, "PLA_Temp_comm_all_size: NULL");
  *nprocs = 1;
  return PLA_SUCCESS;
}
int PLA_Temp_comm_all(PLA_Template templ, MPI_Comm *comm) {
  if (!templ || !comm) return fail(PLA_ERR_INVALID_PARAM, "PLA_Temp_comm_all: NULL");
  *comm = templ->comm_all;
  return PLA_SUCCESS;
}
int PLA_Temp_comm_all_rank(PLA_Template templ, int *rank) {
  if (!templ || !rank) return fail(PLA_ERR_INVALID_PARAM, "PLA_Temp_comm_all_rank: NULL");
  *rank = 0;
  return PLA_SUCCESS;
}

int pla_Environ_set_nb_alg(int type, int nb_alg) {
  if (nb_alg <= 0) return fail(PLA_ERR_INVALID_PARAM, "pla_Environ_set_nb_alg: nb_alg = %d", nb_alg);
  if (type == PLA_OP_ALL_ALG) {
    for (int &v : g_nb_table) v = nb_alg;
  } else if (type >= PLA_OP_MAT_PAN && type <= PLA_OP_TRIAN_MAT_PAN) {
    g_nb_table[type - PLA_OP_MAT_PAN] = nb_alg;
  } else {
    return fail(PLA_ERR_INVALID_PARAM, "Invalid operation type");
  }
  // the factorisations use the PAN_PAN entry (SOLVERS/LU/PLA_LU.c:26)
  pla::ctx().nb_user = g_nb_table[PLA_OP_PAN_PAN - PLA_OP_MAT_PAN];
  return PLA_SUCCESS;
}

int PLA_Environ_nb_alg(int type, PLA_Template templ, int *nb_alg) {
  (void)templ;
  if (!nb_alg) return fail(PLA_ERR_INVALID_PARAM, "PLA_Environ_nb_alg: NULL");
  if (type == PLA_OP_ALL_ALG) type = PLA_OP_PAN_PAN;
  if (type < PLA_OP_MAT_PAN || type > PLA_OP_TRIAN_MAT_PAN)
    return fail(PLA_ERR_INVALID_PARAM, "Invalid operation type");
  *nb_alg = g_nb_table[type - PLA_OP_MAT_PAN];
  return PLA_SUCCESS;
}

int PLA_Set_error_checking(int error, int parameters, int sequential, int r12) {
  if (error >= 0) PLA_ERROR_CHECKING = (error > 0);
  if (parameters >= 0) PLA_CHECK_PARAMETERS = (parameters > 0);
  if (sequential >= 0) PLA_CHECK_AGAINST_SEQUENTIAL = (sequential > 0);
  if (r12 >= 0) PLA_CHECK_AGAINST_R12 = (r12 > 0);
  return PLA_SUCCESS;
}

int PLA_Get_error_checking(int *error, int *parameters, int *sequential, int *r12) {
  if (error) *error = PLA_ERROR_CHECKING;
  if (parameters) *parameters = PLA_CHECK_PARAMETERS;
  if (sequential) *sequential = PLA_CHECK_AGAINST_SEQUENTIAL;
  if (r12) *r12 = PLA_CHECK_AGAINST_R12;
  return PLA_SUCCESS;
}

int PLA_Device_synchronize(void) {
  if (!pla::ctx().stream) return PLA_SUCCESS;
  PLA_HIP(hipDeviceSynchronize());
  return PLA_SUCCESS;
}

int PLA_Set_device_nb(int nb) {
  if (nb < 0) return fail(PLA_ERR_INVALID_PARAM, "PLA_Set_device_nb: nb = %d", nb);
  pla::ctx().nb_device = nb;
  return PLA_SUCCESS;
}

int PLA_Set_device_param(int param, int value) {
  pla::context &c = pla::ctx();
  switch (param) {
    case PLA_PARAM_LOOKAHEAD:
      if (value < 0 || value > 1) return fail(PLA_ERR_INVALID_PARAM, "look-ahead must be 0 or 1");
      c.lookahead = value;
      return PLA_SUCCESS;
    case PLA_PARAM_LEAF_ROWS:
      if (value < 32 || value > 1024) return fail(PLA_ERR_INVALID_PARAM, "leaf rows must be in [32, 1024]");
      c.leaf_rows = value;
      return PLA_SUCCESS;
    case PLA_PARAM_PANEL_CUS:
      if (value < 0 || value % 32 != 0 || (value > 0 && value >= c.cu_count))
        return fail(PLA_ERR_INVALID_PARAM, "panel CUs must be a multiple of 32 below %d", c.cu_count);
      c.panel_cus = value;
      return PLA_SUCCESS;
    case PLA_PARAM_CHOL_PANEL_CUS:
      if (value < 0 || value % 32 != 0 || (value > 0 && value >= c.cu_count))
        return fail(PLA_ERR_INVALID_PARAM, "panel CUs must be a multiple of 32 below %d", c.cu_count);
      c.chol_panel_cus = value;
      // the masked pair is made here, so the factorization itself creates nothing
      if (value > 0 && c.initialized) return pla::masked_streams(value);
      return PLA_SUCCESS;
    case PLA_PARAM_LEAF_KIND:
      if (value != 0 && value != 1) return fail(PLA_ERR_INVALID_PARAM, "leaf kind must be 0 or 1");
      c.leaf_kind = value;
      return PLA_SUCCESS;
    case PLA_PARAM_NB_TAIL:
      if (value < 0) return fail(PLA_ERR_INVALID_PARAM, "tail block size %d", value);
      c.nb_tail = value;
      return PLA_SUCCESS;
    case PLA_PARAM_NB_SWITCH:
      if (value < 0) return fail(PLA_ERR_INVALID_PARAM, "switch rows %d", value);
      c.nb_switch = value;
      return PLA_SUCCESS;
    case PLA_PARAM_LEAF_WIDTH:
      if (value != 0 && value != 8 && value != 16 && value != 32)
        return fail(PLA_ERR_INVALID_PARAM, "leaf width must be 0, 8, 16 or 32");
      c.leaf_width = value;
      return PLA_SUCCESS;
    default:
      return fail(PLA_ERR_INVALID_PARAM, "unknown device parameter %d", param);
  }
}

int PLA_Get_device_nb_for(int n, int factorization, int *nb) {
  if (!nb) return fail(PLA_ERR_INVALID_PARAM, "PLA_Get_device_nb_for: NULL");
  if (factorization < PLA_NB_LU || factorization > PLA_NB_LU_MIXED)
    return fail(PLA_ERR_INVALID_PARAM, "PLA_Get_device_nb_for: factorization %d", factorization);
  *nb = pla::choose_nb(n, factorization);
  return PLA_SUCCESS;
}

int PLA_Get_device_nb(int n, int *nb) {
  if (!nb) return fail(PLA_ERR_INVALID_PARAM, "PLA_Get_device_nb: NULL");
  *nb = pla::choose_nb(n, pla::NB_LU);
  return PLA_SUCCESS;
}

}  // extern "C"
