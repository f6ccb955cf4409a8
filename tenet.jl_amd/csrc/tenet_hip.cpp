// Host runtime of libtenet_hip.so: contexts, the stream-ordered memory pool,
// the contraction planner (labels -> offset tables), and the orchestration
// of QR / SVD / gate update / transfer sweeps on top of the HIP kernels.
// Every exported symbol is declared in include/tenet_hip.h.
#include "../../include/tenet_hip.h"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <tuple>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace tenet {
static thread_local std::string g_err;
// library log (tenet_set_log): 0 off, 1 errors, 2 errors and resource events
static int g_log_level = 0;
static FILE* g_log_file = nullptr;
static std::mutex g_log_mu;
void log_line(int level, const std::string& m) {
  if (level > g_log_level) return;
  std::lock_guard<std::mutex> g(g_log_mu);
  FILE* f = g_log_file ? g_log_file : stderr;
  std::fprintf(f, "[tenet_hip] %s\n", m.c_str());
  std::fflush(f);
}
void set_error(const std::string& m) {
  g_err = m;
  if (g_log_level >= 1) log_line(1, "error: " + m);
}
}  // namespace tenet

using namespace tenet;

#define CHECK_ARG(cond, msg)      \
  do {                            \
    if (!(cond)) {                \
      set_error(msg);             \
      return TENET_ERR_ARG;       \
    }                             \
  } while (0)
#define TRY(expr)              \
  do {                         \
    int _st = (expr);          \
    if (_st != TENET_OK) return _st; \
  } while (0)

struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// A contraction whose matricised operands are ordinary column-major
// matrices: op(A) (M x K), op(B) (K x N), C (M x N, unit stride along M),
// and, with batch labels, each operand's batch entries evenly spaced.  The
// base ops are N or T (the caller's conjugation turns T into C; a
// conjugated operand with op N is not a plain GEMM).
struct LibLayout {
  bool ok = false;
  rocblas_operation opA = rocblas_operation_none, opB = rocblas_operation_none;
  int64_t lda = 0, ldb = 0, ldc = 0;
  int64_t sA = 0, sB = 0, sC = 0;  // batch strides (elements)
};

struct tenet_plan_s {
  tenet_ctx ctx = nullptr;
  int dtype = 0, prec = 0;
  int64_t M = 0, N = 0, K = 0, batch = 0;
  int a_mfast = 0, b_nfast = 0;
  int64_t* tables = nullptr;
  GemmTables t{};
  size_t ws_bytes = 0;
  bool cached = false;
  // spans (elements) each operand must cover, for pool bounds checks
  int64_t spanA = 0, spanB = 0, spanC = 0;
  // a plain (strided-batched) GEMM of at least LIB_MIN_FLOPS in the native
  // precision runs in the vendor library (not under graph capture)
  LibLayout lib;
};
constexpr double LIB_MIN_FLOPS = 1e8;

struct tenet_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;  // the capture stream: replays go there when launched with NULL
};

// Workspace of tenet_svd_thin_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  Host code never passes `scratch`
// itself to a kernel (on the host that name is only a stand-in for the
// symbol): tenet_create looks up each device's address once, with that
// device current, and the entry point hands that address to the kernels.
static __device__ __align__(256) unsigned char scratch[TENET_SVD_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
static hipError_t g_scratch_err[kMaxDevices] = {};
static bool g_scratch_done[kMaxDevices] = {};

// Device address of `scratch` on `device` (the current device), looked up
// on the first call for that device; the lookup's status is kept with it.
static hipError_t scratch_lookup(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_done[device]) {
    void* p = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? p : nullptr;
    if (g_scratch_err[device] != hipSuccess) (void)hipGetLastError();
    g_scratch_done[device] = true;
  }
  return g_scratch_err[device];
}

struct tenet_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::map<size_t, std::vector<void*>> free_blocks;
  std::map<uintptr_t, size_t> live;  // base -> class size
  // blocks freed by tenet_free_async: reusable once their event completes
  struct PendingFree {
    hipEvent_t ev;
    void* ptr;
    size_t cls;
  };
  std::vector<PendingFree> pending;
  std::vector<hipEvent_t> spare_events;
  size_t pool_bytes = 0;
  std::unordered_map<hipStream_t, Workspace> ws;
  std::unordered_map<hipStream_t, Workspace> ws_gemm;
  std::unordered_map<std::string, tenet_plan> plans;
  std::list<std::string> plan_lru;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int32_t* host_flag = nullptr;  // pinned
  int64_t* host_k = nullptr;     // pinned
  // device status word set by captured work (bit 0: a Jacobi SVD inside a
  // graph ended unconverged); read and cleared by tenet_graph_status
  int32_t* dev_status = nullptr;
  // per-kernel timing (tenet_stats_*)
  int stat_family = 0;
  std::vector<hipEvent_t> stat_events;  // pairs
  size_t stat_used = 0;
  int64_t stat_launches = 0, stat_sampled = 0, stat_every = 17;
  bool svd_precond = true;  // QR-preconditioned Jacobi for complex128 bond matrices (tenet_svd_set_precondition)
  // grouped-launch tile costs per (tile, K step) relative to 64 x 64, per dtype (0 = unset)
  double tile_cost[2][3] = {{0, 0, 0}, {0, 0, 0}};
  double stat_ms = 0.0, stat_flops = 0.0, stat_bytes = 0.0;
  // pinned landing slots for the per-sweep convergence flags of grouped SVD
  // runs (free list; a run takes one for its duration)
  struct FlagSlot {
    int32_t* host = nullptr;
    size_t bytes = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
  };
  std::vector<FlagSlot*> flag_slots;
  // library GEMM handle (created on first use; calls on it hold blas_mu)
  rocblas_handle blas = nullptr;
  std::mutex blas_mu;
  // pinned staging ring for small per-call table uploads (a copy from
  // pageable memory would wait for the stream to drain first)
  static constexpr int UP_RING = 8;
  char* up_host[UP_RING] = {};
  size_t up_cap[UP_RING] = {};
  hipEvent_t up_ev[UP_RING] = {};
  bool up_used[UP_RING] = {};
  int up_next = 0;
};
// Every stat_every-th launch of the tracked family is bracketed by events
// (default 17: an odd stride, so the two alternating phases of a Jacobi round
// are both sampled); the timed total is scaled by launches / sampled.

// Drain recorded event pairs into the running totals (synchronises).
static int stats_drain(tenet_ctx ctx) {
  for (size_t i = 0; i + 1 < ctx->stat_used; i += 2) {
    TENET_HIP_CHECK(hipEventSynchronize(ctx->stat_events[i + 1]));
    float ms = 0.f;
    TENET_HIP_CHECK(hipEventElapsedTime(&ms, ctx->stat_events[i], ctx->stat_events[i + 1]));
    ctx->stat_ms += ms;
  }
  ctx->stat_used = 0;
  return TENET_OK;
}

// RAII bracket around one launch of a tracked family.
struct StatScope {
  tenet_ctx ctx;
  hipStream_t s;
  bool on = false;
  // launch_records: the sampled pair is handed to the launch
  // (start_event / stop_event), which stamps it in the dispatch itself
  // (hipExtLaunchKernelGGL), so the sample is the kernel's own duration
  bool by_launch = false, recorded = false;
  StatScope(tenet_ctx c, hipStream_t st, int family, double flops, double bytes, bool launch_records = false)
      : ctx(c), s(st) {
    if (!ctx || ctx->stat_family != family || family == 0) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive) return;
    ctx->stat_launches += 1;
    ctx->stat_flops += flops;
    ctx->stat_bytes += bytes;
    if ((ctx->stat_launches - 1) % ctx->stat_every != 0) return;
    if (ctx->stat_used + 2 > ctx->stat_events.size()) {
      if (ctx->stat_events.size() < 8192) {
        for (int i = 0; i < 512; ++i) {
          hipEvent_t e;
          if (hipEventCreate(&e) != hipSuccess) return;
          ctx->stat_events.push_back(e);
        }
      } else if (stats_drain(ctx) != TENET_OK) {
        return;
      }
    }
    if (launch_records) {
      by_launch = true;
    } else if (hipEventRecord(ctx->stat_events[ctx->stat_used], s) != hipSuccess) {
      return;
    }
    on = true;
    ctx->stat_sampled += 1;
  }
  hipEvent_t start_event() const { return on && by_launch ? ctx->stat_events[ctx->stat_used] : nullptr; }
  hipEvent_t stop_event() const { return on && by_launch ? ctx->stat_events[ctx->stat_used + 1] : nullptr; }
  // the launch with start_event / stop_event succeeded
  void launched() { recorded = true; }
  ~StatScope() {
    if (!on) return;
    if (by_launch) {
      if (recorded)
        ctx->stat_used += 2;
      else
        ctx->stat_sampled -= 1;
      return;
    }
    if (hipEventRecord(ctx->stat_events[ctx->stat_used + 1], s) == hipSuccess) ctx->stat_used += 2;
  }
};

// Take work back out of a family's running totals: launches counted with
// problems that, as it turned out, had already finished (their kernels
// returned at once).  Applies exactly when StatScope counted them.
static void stats_retract(tenet_ctx ctx, int family, hipStream_t s, double flops, double bytes) {
  if (!ctx || family == 0 || ctx->stat_family != family || (flops == 0.0 && bytes == 0.0)) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive) return;
  ctx->stat_flops -= flops;
  ctx->stat_bytes -= bytes;
}

// Queue an upload of host bytes through the context's pinned ring: the bytes
// are copied into a slot at once (the caller's buffer may go right after), the
// slot is reused once its copy has completed.  Not for captured streams (a
// graph's copy node would read the slot at every replay).
static int upload_async(tenet_ctx ctx, hipStream_t s, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return TENET_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const int slot = ctx->up_next;
  ctx->up_next = (slot + 1) % tenet_ctx_s::UP_RING;
  if (!ctx->up_ev[slot]) TENET_HIP_CHECK(hipEventCreateWithFlags(&ctx->up_ev[slot], hipEventDisableTiming));
  if (ctx->up_used[slot]) TENET_HIP_CHECK(hipEventSynchronize(ctx->up_ev[slot]));
  if (ctx->up_cap[slot] < bytes) {
    if (ctx->up_host[slot]) TENET_HIP_CHECK(hipHostFree(ctx->up_host[slot]));
    ctx->up_host[slot] = nullptr;
    ctx->up_cap[slot] = 0;
    const size_t cap = std::max<size_t>((bytes + 65535) / 65536 * 65536, 65536);
    TENET_HIP_CHECK(hipHostMalloc((void**)&ctx->up_host[slot], cap, hipHostMallocDefault));
    ctx->up_cap[slot] = cap;
  }
  std::memcpy(ctx->up_host[slot], src, bytes);
  TENET_HIP_CHECK(hipMemcpyAsync(dst, ctx->up_host[slot], bytes, hipMemcpyHostToDevice, s));
  TENET_HIP_CHECK(hipEventRecord(ctx->up_ev[slot], s));
  ctx->up_used[slot] = true;
  return TENET_OK;
}

// launch_svd_round with its timing sample stamped by the dispatch
static int svd_round_timed(tenet_ctx ctx, hipStream_t s, double flops, double bytes, const SvdBatch& B, int64_t nprob,
                           int max_items, int round, int phase, int mode) {
  StatScope st(ctx, s, TENET_KSTAT_SVD_ROUND, flops, bytes, true);
  const int rc = launch_svd_round(B, nprob, max_items, round, phase, mode, s, st.start_event(), st.stop_event());
  if (rc == TENET_OK) st.launched();
  return rc;
}

static const size_t ELEM[2] = {16, 8};

static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ---------------------------------------------------------------------------
// device guard / workspace
// ---------------------------------------------------------------------------
static int use_device(tenet_ctx ctx) {
  if (!ctx) {
    set_error("null context");
    return TENET_ERR_ARG;
  }
  TENET_HIP_CHECK(hipSetDevice(ctx->device));
  return TENET_OK;
}

static hipStream_t pick(tenet_ctx ctx, tenet_stream s) { return s ? (hipStream_t)s : ctx->stream; }

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st == hipStreamCaptureStatusActive;
}

// Stream-ordered scratch: calls on one stream reuse the same region in order.
// `which` 0: the API call's own scratch; 1: contraction scratch (split packs),
// so an orchestrating call and the contractions it issues never overlap.
static int ws_get(tenet_ctx ctx, hipStream_t s, size_t bytes, void** out, int which = 0) {
  std::lock_guard<std::mutex> g(ctx->mu);
  Workspace& w = which ? ctx->ws_gemm[s] : ctx->ws[s];
  if (w.bytes < bytes) {
    if (capturing(s)) {
      set_error("workspace would grow during graph capture: run the sequence once before capturing");
      return TENET_ERR_OOM;
    }
    const size_t nb = align_up(std::max(bytes, w.bytes + w.bytes / 2), 1 << 20);
    if (g_log_level >= 2)
      log_line(2, "workspace " + std::to_string(which) + " grows to " + std::to_string(nb) + " bytes");
    TENET_HIP_CHECK(hipStreamSynchronize(s));
    if (w.ptr) TENET_HIP_CHECK(hipFree(w.ptr));
    w.ptr = nullptr;
    w.bytes = 0;
    if (hipMalloc(&w.ptr, nb) != hipSuccess) {
      (void)hipGetLastError();
      w.ptr = nullptr;
      set_error("workspace allocation failed");
      return TENET_ERR_OOM;
    }
    w.bytes = nb;
  }
  *out = w.ptr;
  return TENET_OK;
}

struct Bump {
  char* base;
  size_t off = 0;
  explicit Bump(void* p) : base((char*)p) {}
  void* take(size_t b) {
    void* p = base + off;
    off = align_up(off + b, 256);
    return p;
  }
};

// ---------------------------------------------------------------------------
// pool bounds check: if ptr lies in a pool block, [ptr, ptr+bytes) must fit
// ---------------------------------------------------------------------------
static int check_span(tenet_ctx ctx, const void* ptr, size_t bytes, const char* what) {
  if (!ptr) {
    set_error(std::string("null device pointer: ") + what);
    return TENET_ERR_ARG;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  const uintptr_t p = (uintptr_t)ptr;
  auto it = ctx->live.upper_bound(p);
  if (it == ctx->live.begin()) return TENET_OK;
  --it;
  if (p >= it->first && p < it->first + it->second) {
    if (p + bytes > it->first + it->second) {
      set_error(std::string("operand exceeds its allocation: ") + what);
      return TENET_ERR_SHAPE;
    }
  }
  return TENET_OK;
}

// ---------------------------------------------------------------------------
// library / context
// ---------------------------------------------------------------------------
extern "C" const char* tenet_version(void) { return "tenet_hip 0.1.0 (gfx950)"; }
extern "C" const char* tenet_last_error(void) { return g_err.c_str(); }

extern "C" int tenet_device_count(int* count) {
  CHECK_ARG(count, "null count");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return TENET_OK;
}

extern "C" int tenet_create(int device, tenet_ctx* out) {
  CHECK_ARG(out, "null output");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    set_error("no HIP device visible");
    return TENET_ERR_NODEVICE;
  }
  if (device < 0 || device >= n) {
    set_error("device index out of range");
    return TENET_ERR_ARG;
  }
  hipDeviceProp_t prop;
  TENET_HIP_CHECK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("device is not gfx950: ") + prop.gcnArchName);
    return TENET_ERR_NODEVICE;
  }
  TENET_HIP_CHECK(hipSetDevice(device));
  if (device >= kMaxDevices) {
    set_error("device index above the library's per-device table");
    return TENET_ERR_ARG;
  }
  // this device's copy of the allocation-free SVD's scratch array; a failed
  // lookup is reported by tenet_svd_thin_noalloc, the only user
  (void)scratch_lookup(device);
  auto* c = new tenet_ctx_s();
  c->device = device;
  TENET_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  TENET_HIP_CHECK(hipEventCreate(&c->ev0));
  TENET_HIP_CHECK(hipEventCreate(&c->ev1));
  TENET_HIP_CHECK(hipHostMalloc((void**)&c->host_flag, 4096, hipHostMallocDefault));
  TENET_HIP_CHECK(hipHostMalloc((void**)&c->host_k, 64, hipHostMallocDefault));
  TENET_HIP_CHECK(hipMalloc((void**)&c->dev_status, 256));
  TENET_HIP_CHECK(hipMemset(c->dev_status, 0, 256));
  *out = c;
  return TENET_OK;
}

extern "C" int tenet_destroy(tenet_ctx ctx) {
  if (!ctx) return TENET_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : ctx->plans) {
    if (kv.second->tables) (void)hipFree(kv.second->tables);
    delete kv.second;
  }
  for (auto& kv : ctx->free_blocks)
    for (void* p : kv.second) (void)hipFree(p);
  for (auto& kv : ctx->live) (void)hipFree((void*)kv.first);
  for (auto& pf : ctx->pending) {
    (void)hipFree(pf.ptr);
    (void)hipEventDestroy(pf.ev);
  }
  for (hipEvent_t e : ctx->spare_events) (void)hipEventDestroy(e);
  for (auto& kv : ctx->ws)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  for (auto& kv : ctx->ws_gemm)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  (void)hipHostFree(ctx->host_flag);
  (void)hipHostFree(ctx->host_k);
  if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
  for (int i = 0; i < tenet_ctx_s::UP_RING; ++i) {
    if (ctx->up_host[i]) (void)hipHostFree(ctx->up_host[i]);
    if (ctx->up_ev[i]) (void)hipEventDestroy(ctx->up_ev[i]);
  }
  for (auto* f : ctx->flag_slots) {
    (void)hipHostFree(f->host);
    for (hipEvent_t e : f->ev) (void)hipEventDestroy(e);
    delete f;
  }
  if (ctx->dev_status) (void)hipFree(ctx->dev_status);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  for (hipEvent_t e : ctx->stat_events) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return TENET_OK;
}

extern "C" int tenet_stream_create(tenet_ctx ctx, tenet_stream* s) {
  TRY(use_device(ctx));
  CHECK_ARG(s, "null stream out");
  hipStream_t st;
  TENET_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *s = (tenet_stream)st;
  return TENET_OK;
}
extern "C" int tenet_stream_destroy(tenet_ctx ctx, tenet_stream s) {
  TRY(use_device(ctx));
  if (!s) return TENET_OK;
  TENET_HIP_CHECK(hipStreamSynchronize((hipStream_t)s));
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    for (auto* mp : {&ctx->ws, &ctx->ws_gemm}) {
      auto it = mp->find((hipStream_t)s);
      if (it != mp->end()) {
        if (it->second.ptr) (void)hipFree(it->second.ptr);
        mp->erase(it);
      }
    }
  }
  TENET_HIP_CHECK(hipStreamDestroy((hipStream_t)s));
  return TENET_OK;
}
extern "C" int tenet_stream_sync(tenet_ctx ctx, tenet_stream s) {
  TRY(use_device(ctx));
  TENET_HIP_CHECK(hipStreamSynchronize(pick(ctx, s)));
  return TENET_OK;
}
extern "C" int tenet_device_sync(tenet_ctx ctx) {
  TRY(use_device(ctx));
  TENET_HIP_CHECK(hipDeviceSynchronize());
  return TENET_OK;
}

// ---------------------------------------------------------------------------
// events and pinned host staging (asynchronous request loops)
// ---------------------------------------------------------------------------
extern "C" int tenet_event_create(tenet_ctx ctx, tenet_event* ev) {
  TRY(use_device(ctx));
  CHECK_ARG(ev, "null event out");
  hipEvent_t e;
  TENET_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  *ev = (tenet_event)e;
  return TENET_OK;
}
extern "C" int tenet_event_destroy(tenet_ctx ctx, tenet_event ev) {
  TRY(use_device(ctx));
  if (ev) TENET_HIP_CHECK(hipEventDestroy((hipEvent_t)ev));
  return TENET_OK;
}
extern "C" int tenet_event_record(tenet_ctx ctx, tenet_event ev, tenet_stream stream) {
  TRY(use_device(ctx));
  CHECK_ARG(ev, "null event");
  TENET_HIP_CHECK(hipEventRecord((hipEvent_t)ev, pick(ctx, stream)));
  return TENET_OK;
}
extern "C" int tenet_event_query(tenet_ctx ctx, tenet_event ev, int* done) {
  TRY(use_device(ctx));
  CHECK_ARG(ev && done, "bad argument");
  const hipError_t e = hipEventQuery((hipEvent_t)ev);
  if (e == hipSuccess) {
    *done = 1;
    return TENET_OK;
  }
  if (e == hipErrorNotReady) {
    *done = 0;
    return TENET_OK;
  }
  TENET_HIP_CHECK(e);
  return TENET_OK;
}
extern "C" int tenet_stream_wait_event(tenet_ctx ctx, tenet_stream stream, tenet_event ev) {
  TRY(use_device(ctx));
  CHECK_ARG(ev, "null event");
  TENET_HIP_CHECK(hipStreamWaitEvent(pick(ctx, stream), (hipEvent_t)ev, 0));
  return TENET_OK;
}
extern "C" int tenet_host_alloc(tenet_ctx ctx, size_t bytes, void** hptr) {
  TRY(use_device(ctx));
  CHECK_ARG(hptr, "null pointer out");
  TENET_HIP_CHECK(hipHostMalloc(hptr, std::max<size_t>(bytes, 1), hipHostMallocDefault));
  return TENET_OK;
}
extern "C" int tenet_host_free(tenet_ctx ctx, void* hptr) {
  TRY(use_device(ctx));
  if (hptr) TENET_HIP_CHECK(hipHostFree(hptr));
  return TENET_OK;
}
extern "C" int tenet_memcpy_async(tenet_ctx ctx, void* dst, const void* src, size_t bytes, int to_host,
                                  tenet_stream stream) {
  TRY(use_device(ctx));
  if (bytes == 0) return TENET_OK;
  TRY(check_span(ctx, to_host ? src : dst, bytes, to_host ? "async d2h src" : "async h2d dst"));
  TENET_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice,
                                 pick(ctx, stream)));
  return TENET_OK;
}

// ---------------------------------------------------------------------------
// memory pool
// ---------------------------------------------------------------------------
static size_t size_class(size_t b) {
  if (b <= 256) return 256;
  if (b <= (1u << 20)) {
    size_t c = 256;
    while (c < b) c <<= 1;
    return c;
  }
  return align_up(b, 1 << 20);
}

extern "C" int tenet_malloc(tenet_ctx ctx, size_t bytes, void** dptr) {
  TRY(use_device(ctx));
  CHECK_ARG(dptr, "null output");
  const size_t cls = size_class(std::max<size_t>(bytes, 1));
  std::lock_guard<std::mutex> g(ctx->mu);
  // blocks whose stream-ordered free has completed become reusable
  for (size_t i = 0; i < ctx->pending.size();) {
    auto& pf = ctx->pending[i];
    if (hipEventQuery(pf.ev) == hipSuccess) {
      ctx->free_blocks[pf.cls].push_back(pf.ptr);
      ctx->spare_events.push_back(pf.ev);
      pf = ctx->pending.back();
      ctx->pending.pop_back();
    } else {
      (void)hipGetLastError();  // hipErrorNotReady is not an error here
      ++i;
    }
  }
  auto it = ctx->free_blocks.find(cls);
  if (it != ctx->free_blocks.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    ctx->live[(uintptr_t)p] = cls;
    *dptr = p;
    return TENET_OK;
  }
  void* p = nullptr;
  if (hipMalloc(&p, cls) != hipSuccess) {
    (void)hipGetLastError();
    // release cached blocks and retry once
    (void)hipDeviceSynchronize();
    for (auto& kv : ctx->free_blocks)
      for (void* q : kv.second) {
        (void)hipFree(q);
        ctx->pool_bytes -= kv.first;
      }
    ctx->free_blocks.clear();
    if (hipMalloc(&p, cls) != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation failed");
      return TENET_ERR_OOM;
    }
  }
  ctx->pool_bytes += cls;
  ctx->live[(uintptr_t)p] = cls;
  *dptr = p;
  return TENET_OK;
}

extern "C" int tenet_free(tenet_ctx ctx, void* dptr) {
  if (!ctx) return TENET_ERR_ARG;
  if (!dptr) return TENET_OK;
  std::lock_guard<std::mutex> g(ctx->mu);
  auto it = ctx->live.find((uintptr_t)dptr);
  if (it == ctx->live.end()) {
    set_error("tenet_free: pointer not from tenet_malloc");
    return TENET_ERR_ARG;
  }
  ctx->free_blocks[it->second].push_back(dptr);
  ctx->live.erase(it);
  return TENET_OK;
}

extern "C" int tenet_free_async(tenet_ctx ctx, void* dptr, tenet_stream stream) {
  TRY(use_device(ctx));
  if (!dptr) return TENET_OK;
  hipStream_t s = pick(ctx, stream);
  std::lock_guard<std::mutex> g(ctx->mu);
  auto it = ctx->live.find((uintptr_t)dptr);
  if (it == ctx->live.end()) {
    set_error("tenet_free_async: pointer not from tenet_malloc");
    return TENET_ERR_ARG;
  }
  hipEvent_t ev = nullptr;
  if (!ctx->spare_events.empty()) {
    ev = ctx->spare_events.back();
    ctx->spare_events.pop_back();
  } else {
    TENET_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  TENET_HIP_CHECK(hipEventRecord(ev, s));
  ctx->pending.push_back({ev, dptr, it->second});
  ctx->live.erase(it);
  return TENET_OK;
}

extern "C" int tenet_memcpy_h2d(tenet_ctx ctx, void* dst, const void* src, size_t bytes, tenet_stream s) {
  TRY(use_device(ctx));
  if (bytes == 0) return TENET_OK;
  TRY(check_span(ctx, dst, bytes, "h2d dst"));
  hipStream_t st = pick(ctx, s);
  TENET_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  TENET_HIP_CHECK(hipStreamSynchronize(st));  // pageable source: complete before returning
  return TENET_OK;
}
extern "C" int tenet_memcpy_d2h(tenet_ctx ctx, void* dst, const void* src, size_t bytes, tenet_stream s) {
  TRY(use_device(ctx));
  if (bytes == 0) return TENET_OK;
  TRY(check_span(ctx, src, bytes, "d2h src"));
  hipStream_t st = pick(ctx, s);
  TENET_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
  TENET_HIP_CHECK(hipStreamSynchronize(st));
  return TENET_OK;
}
extern "C" int tenet_memcpy_d2d(tenet_ctx ctx, void* dst, const void* src, size_t bytes, tenet_stream s) {
  TRY(use_device(ctx));
  if (bytes == 0) return TENET_OK;
  TRY(check_span(ctx, dst, bytes, "d2d dst"));
  TRY(check_span(ctx, src, bytes, "d2d src"));
  TENET_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, pick(ctx, s)));
  return TENET_OK;
}
extern "C" int tenet_memset_zero(tenet_ctx ctx, void* dst, size_t bytes, tenet_stream s) {
  TRY(use_device(ctx));
  TRY(check_span(ctx, dst, bytes, "memset dst"));
  return launch_fill_zero(dst, bytes, pick(ctx, s));
}
extern "C" int tenet_mem_info(tenet_ctx ctx, size_t* free_b, size_t* total_b, size_t* pool_b) {
  TRY(use_device(ctx));
  size_t f = 0, t = 0;
  TENET_HIP_CHECK(hipMemGetInfo(&f, &t));
  if (free_b) *free_b = f;
  if (total_b) *total_b = t;
  if (pool_b) *pool_b = ctx->pool_bytes;
  return TENET_OK;
}

extern "C" int tenet_set_log(int level, const char* path) {
  CHECK_ARG(level >= 0 && level <= 2, "log level must be 0, 1 or 2");
  std::lock_guard<std::mutex> g(g_log_mu);
  if (g_log_file) {
    std::fclose(g_log_file);
    g_log_file = nullptr;
  }
  if (path && *path) {
    g_log_file = std::fopen(path, "a");
    if (!g_log_file) {
      g_log_level = 0;
      g_err = std::string("tenet_set_log: cannot open ") + path;
      return TENET_ERR_ARG;
    }
  }
  g_log_level = level;
  return TENET_OK;
}

// ---------------------------------------------------------------------------
// contraction planner
// ---------------------------------------------------------------------------
struct LabelInfo {
  int64_t label, extent, sa, sb, sc;
  bool ina, inb, inc;
};

static std::string plan_key(int dtype, int prec, const tenet_tdesc* a, const tenet_tdesc* b, const tenet_tdesc* c) {
  std::string k;
  k.reserve(1024);
  auto put = [&](int64_t v) { k.append(reinterpret_cast<const char*>(&v), sizeof v); };
  put(dtype);
  put(prec);
  for (const tenet_tdesc* d : {a, b, c}) {
    put(d->rank);
    for (int i = 0; i < d->rank; ++i) {
      put(d->extent[i]);
      put(d->stride[i]);
      put(d->label[i]);
    }
  }
  return k;
}

static void radix_table(const std::vector<LabelInfo*>& ls, int which, std::vector<int64_t>& out) {
  int64_t total = 1;
  for (auto* l : ls) total *= l->extent;
  out.assign((size_t)total, 0);
  int64_t span = 1;
  for (auto* l : ls) {
    const int64_t st = which == 0 ? l->sa : (which == 1 ? l->sb : l->sc);
    for (int64_t i = 0; i < total; ++i) out[(size_t)i] += ((i / span) % l->extent) * st;
    span *= l->extent;
  }
}

// Label analysis of one contraction: the nine offset tables (rowA, kA, bA,
// kB, colB, bB, rowC, colC, bC) and the GEMM shape, on the host.
struct HostPlan {
  int64_t M = 0, N = 0, K = 0, batch = 0;
  int a_mfast = 0, b_nfast = 0;
  int64_t spanA = 0, spanB = 0, spanC = 0;
  std::vector<int64_t> v[9];
};

static LibLayout plain_layout(const HostPlan& hp) {
  LibLayout L;
  if (hp.M < 1 || hp.N < 1 || hp.K < 1 || hp.batch < 1) return L;
  auto affine = [](const std::vector<int64_t>& v, int64_t* st) {
    if (v.empty() || v[0] != 0) return false;
    *st = v.size() > 1 ? v[1] : 0;
    for (size_t i = 1; i < v.size(); ++i)
      if (v[i] != (int64_t)i * *st) return false;
    return true;
  };
  int64_t sM, sKa, sKb, sN, sMc, sNc, bA, bB, bC;
  if (!affine(hp.v[0], &sM) || !affine(hp.v[1], &sKa) || !affine(hp.v[3], &sKb) || !affine(hp.v[4], &sN) ||
      !affine(hp.v[6], &sMc) || !affine(hp.v[7], &sNc) || !affine(hp.v[2], &bA) || !affine(hp.v[5], &bB) ||
      !affine(hp.v[8], &bC))
    return L;
  if (bA < 0 || bB < 0 || bC < 0) return L;
  if (hp.M == 1) sM = sMc = 1;
  if (hp.K == 1) sKa = sKb = 1;
  if (hp.N == 1) sN = 1;
  if (sMc != 1) return L;
  L.ldc = hp.N == 1 ? hp.M : sNc;
  if (L.ldc < hp.M) return L;
  if (sM == 1 && (hp.K == 1 || sKa >= hp.M)) {
    L.opA = rocblas_operation_none;
    L.lda = hp.K == 1 ? hp.M : sKa;
  } else if (sKa == 1 && (hp.M == 1 || sM >= hp.K)) {
    L.opA = rocblas_operation_transpose;
    L.lda = hp.M == 1 ? hp.K : sM;
  } else {
    return L;
  }
  if (sKb == 1 && (hp.N == 1 || sN >= hp.K)) {
    L.opB = rocblas_operation_none;
    L.ldb = hp.N == 1 ? hp.K : sN;
  } else if (sN == 1 && (hp.K == 1 || sKb >= hp.N)) {
    L.opB = rocblas_operation_transpose;
    L.ldb = hp.K == 1 ? hp.N : sKb;
  } else {
    return L;
  }
  L.sA = bA;
  L.sB = bB;
  L.sC = bC;
  L.ok = L.lda <= INT32_MAX && L.ldb <= INT32_MAX && L.ldc <= INT32_MAX && hp.M <= INT32_MAX &&
         hp.N <= INT32_MAX && hp.K <= INT32_MAX && hp.batch <= INT32_MAX;
  return L;
}

// the op of an operand after the caller's conjugation (false: not plain)
static bool lib_op(rocblas_operation base, bool conj, rocblas_operation* op) {
  if (!conj) {
    *op = base;
    return true;
  }
  if (base == rocblas_operation_none) return false;
  *op = rocblas_operation_conjugate_transpose;
  return true;
}

static int analyse_plan(const tenet_tdesc* a, const tenet_tdesc* b, const tenet_tdesc* c, HostPlan& hp) {
  for (const tenet_tdesc* d : {a, b, c}) {
    if (d->rank < 0 || d->rank > TENET_MAX_RANK) {
      set_error("rank out of range");
      return TENET_ERR_SHAPE;
    }
    for (int i = 0; i < d->rank; ++i) {
      if (d->extent[i] < 0) {
        set_error("negative extent");
        return TENET_ERR_SHAPE;
      }
      for (int j = 0; j < i; ++j)
        if (d->label[i] == d->label[j]) {
          set_error("repeated label within one tensor");
          return TENET_ERR_SHAPE;
        }
    }
  }
  std::vector<LabelInfo> L;
  auto find = [&](int64_t lab) -> LabelInfo* {
    for (auto& x : L)
      if (x.label == lab) return &x;
    return nullptr;
  };
  for (int i = 0; i < a->rank; ++i) L.push_back({a->label[i], a->extent[i], a->stride[i], 0, 0, true, false, false});
  for (int i = 0; i < b->rank; ++i) {
    LabelInfo* x = find(b->label[i]);
    if (x) {
      if (x->extent != b->extent[i]) {
        set_error("extent mismatch on a shared label");
        return TENET_ERR_SHAPE;
      }
      x->inb = true;
      x->sb = b->stride[i];
    } else {
      L.push_back({b->label[i], b->extent[i], 0, b->stride[i], 0, false, true, false});
    }
  }
  for (int i = 0; i < c->rank; ++i) {
    LabelInfo* x = find(c->label[i]);
    if (!x) {
      set_error("output label absent from both operands");
      return TENET_ERR_SHAPE;
    }
    if (x->extent != c->extent[i]) {
      set_error("extent mismatch on an output label");
      return TENET_ERR_SHAPE;
    }
    x->inc = true;
    x->sc = c->stride[i];
  }
  std::vector<LabelInfo*> fa, fb, kk, bt;
  for (auto& x : L) {
    if (x.ina && x.inb) {
      (x.inc ? bt : kk).push_back(&x);
    } else if (x.ina) {
      if (!x.inc) {
        set_error("label of A alone must appear in the output");
        return TENET_ERR_SHAPE;
      }
      fa.push_back(&x);
    } else {
      if (!x.inc) {
        set_error("label of B alone must appear in the output");
        return TENET_ERR_SHAPE;
      }
      fb.push_back(&x);
    }
  }
  auto by = [](int w) {
    return [w](const LabelInfo* x, const LabelInfo* y) {
      const int64_t sx = std::llabs(w == 0 ? x->sa : (w == 1 ? x->sb : x->sc));
      const int64_t sy = std::llabs(w == 0 ? y->sa : (w == 1 ? y->sb : y->sc));
      return sx < sy;
    };
  };
  std::stable_sort(fa.begin(), fa.end(), by(0));
  std::stable_sort(fb.begin(), fb.end(), by(1));
  std::stable_sort(kk.begin(), kk.end(), by(0));
  std::stable_sort(bt.begin(), bt.end(), by(2));

  radix_table(fa, 0, hp.v[0]);  // rowA
  radix_table(kk, 0, hp.v[1]);  // kA
  radix_table(bt, 0, hp.v[2]);  // bA
  radix_table(kk, 1, hp.v[3]);  // kB
  radix_table(fb, 1, hp.v[4]);  // colB
  radix_table(bt, 1, hp.v[5]);  // bB
  radix_table(fa, 2, hp.v[6]);  // rowC
  radix_table(fb, 2, hp.v[7]);  // colC
  radix_table(bt, 2, hp.v[8]);  // bC
  hp.M = (int64_t)hp.v[0].size();
  hp.N = (int64_t)hp.v[4].size();
  hp.K = (int64_t)hp.v[1].size();
  hp.batch = (int64_t)hp.v[2].size();
  hp.a_mfast = (!fa.empty() && fa[0]->sa == 1) ? 1 : 0;
  hp.b_nfast = (!fb.empty() && fb[0]->sb == 1) ? 1 : 0;
  auto span = [](const tenet_tdesc* d) {
    int64_t s = 0;
    for (int i = 0; i < d->rank; ++i) {
      if (d->extent[i] == 0) return (int64_t)0;
      s += (d->extent[i] - 1) * std::llabs(d->stride[i]);
    }
    return s + 1;
  };
  hp.spanA = span(a);
  hp.spanB = span(b);
  hp.spanC = span(c);
  return TENET_OK;
}

static void set_tables(GemmTables& t, int64_t* T, const size_t* off) {
  t.rowA = T + off[0];
  t.kA = T + off[1];
  t.bA = T + off[2];
  t.kB = T + off[3];
  t.colB = T + off[4];
  t.bB = T + off[5];
  t.rowC = T + off[6];
  t.colC = T + off[7];
  t.bC = T + off[8];
}

static int build_plan(tenet_ctx ctx, int dtype, int prec, const tenet_tdesc* a, const tenet_tdesc* b,
                      const tenet_tdesc* c, tenet_plan p) {
  HostPlan hp;
  TRY(analyse_plan(a, b, c, hp));
  p->ctx = ctx;
  p->dtype = dtype;
  p->prec = prec;
  p->M = hp.M;
  p->N = hp.N;
  p->K = hp.K;
  p->batch = hp.batch;
  p->a_mfast = hp.a_mfast;
  p->b_nfast = hp.b_nfast;
  p->spanA = hp.spanA;
  p->spanB = hp.spanB;
  p->spanC = hp.spanC;
  if (prec == TENET_PREC_NATIVE && 8.0 * hp.M * hp.N * hp.K * hp.batch >= LIB_MIN_FLOPS)
    p->lib = plain_layout(hp);
  std::vector<int64_t> host;
  size_t off[9];
  for (int i = 0; i < 9; ++i) {
    off[i] = host.size();
    host.insert(host.end(), hp.v[i].begin(), hp.v[i].end());
  }
  if (host.empty()) host.push_back(0);
  TENET_HIP_CHECK(hipMalloc((void**)&p->tables, host.size() * sizeof(int64_t)));
  TENET_HIP_CHECK(hipMemcpy(p->tables, host.data(), host.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  set_tables(p->t, p->tables, off);
  GemmLaunch g{};
  g.M = p->M;
  g.N = p->N;
  g.K = p->K;
  g.batch = p->batch;
  p->ws_bytes = gemm_workspace_bytes(g, dtype, prec);
  return TENET_OK;
}

static int valid_dtype_prec(int dtype, int prec) {
  if (dtype != TENET_C128 && dtype != TENET_C64) {
    set_error("unknown dtype");
    return TENET_ERR_ARG;
  }
  if (prec < TENET_PREC_NATIVE || prec > TENET_PREC_SPLIT_I8_FAST) {
    set_error("unknown precision");
    return TENET_ERR_ARG;
  }
  return TENET_OK;
}

extern "C" int tenet_einsum_plan(tenet_ctx ctx, int dtype, int prec, const tenet_tdesc* a, const tenet_tdesc* b,
                                 const tenet_tdesc* c, tenet_plan* plan) {
  TRY(use_device(ctx));
  CHECK_ARG(a && b && c && plan, "null argument");
  TRY(valid_dtype_prec(dtype, prec));
  auto* p = new tenet_plan_s();
  int st = build_plan(ctx, dtype, prec, a, b, c, p);
  if (st != TENET_OK) {
    if (p->tables) (void)hipFree(p->tables);
    delete p;
    return st;
  }
  *plan = p;
  return TENET_OK;
}

// The context's rocBLAS handle bound to stream s (the caller holds blas_mu).
static int blas_bind(tenet_ctx ctx, hipStream_t s) {
  if (!ctx->blas && rocblas_create_handle(&ctx->blas) != rocblas_status_success) {
    ctx->blas = nullptr;
    set_error("rocBLAS handle creation failed");
    return TENET_ERR_HIP;
  }
  if (rocblas_set_stream(ctx->blas, s) != rocblas_status_success) {
    set_error("rocBLAS: cannot bind the stream");
    return TENET_ERR_HIP;
  }
  return TENET_OK;
}

// One plain (strided-batched) GEMM through rocBLAS: C = alpha op(A) op(B) + beta C.
static int lib_gemm(tenet_ctx ctx, hipStream_t s, int dtype, rocblas_operation opA, rocblas_operation opB,
                    int64_t M, int64_t N, int64_t K, const double* alpha, const void* A, int64_t lda, int64_t sA,
                    const void* B, int64_t ldb, int64_t sB, const double* beta, void* C, int64_t ldc, int64_t sC,
                    int64_t batch) {
  std::lock_guard<std::mutex> lk(ctx->blas_mu);
  TRY(blas_bind(ctx, s));
  rocblas_status st;
  if (dtype == TENET_C128) {
    const rocblas_double_complex al = {alpha[0], alpha[1]}, be = {beta[0], beta[1]};
    st = rocblas_zgemm_strided_batched(ctx->blas, opA, opB, (rocblas_int)M, (rocblas_int)N, (rocblas_int)K, &al,
                                       (const rocblas_double_complex*)A, (rocblas_int)lda, sA,
                                       (const rocblas_double_complex*)B, (rocblas_int)ldb, sB, &be,
                                       (rocblas_double_complex*)C, (rocblas_int)ldc, sC, (rocblas_int)batch);
  } else {
    const rocblas_float_complex al = {(float)alpha[0], (float)alpha[1]}, be = {(float)beta[0], (float)beta[1]};
    st = rocblas_cgemm_strided_batched(ctx->blas, opA, opB, (rocblas_int)M, (rocblas_int)N, (rocblas_int)K, &al,
                                       (const rocblas_float_complex*)A, (rocblas_int)lda, sA,
                                       (const rocblas_float_complex*)B, (rocblas_int)ldb, sB, &be,
                                       (rocblas_float_complex*)C, (rocblas_int)ldc, sC, (rocblas_int)batch);
  }
  if (st != rocblas_status_success) {
    set_error(std::string("rocBLAS GEMM failed: ") + rocblas_status_to_string(st));
    return TENET_ERR_HIP;
  }
  return TENET_OK;
}

static int exec_plan(tenet_plan p, hipStream_t s, const void* A, const void* B, void* C, int conj_a, int conj_b,
                     const double* alpha, const double* beta) {
  tenet_ctx ctx = p->ctx;
  const size_t es = ELEM[p->dtype];
  if (p->M * p->N * p->batch == 0) return TENET_OK;
  TRY(check_span(ctx, C, (size_t)p->spanC * es, "C"));
  if (p->K > 0) {
    TRY(check_span(ctx, A, (size_t)p->spanA * es, "A"));
    TRY(check_span(ctx, B, (size_t)p->spanB * es, "B"));
  }
  GemmLaunch g{};
  g.A = A;
  g.B = B;
  g.C = C;
  g.t = p->t;
  g.M = p->M;
  g.N = p->N;
  g.K = p->K;
  g.batch = p->batch;
  g.conjA = conj_a;
  g.conjB = conj_b;
  g.a_mfast = p->a_mfast;
  g.b_nfast = p->b_nfast;
  g.alpha_re = alpha ? alpha[0] : 1.0;
  g.alpha_im = alpha ? alpha[1] : 0.0;
  g.beta_re = beta ? beta[0] : 0.0;
  g.beta_im = beta ? beta[1] : 0.0;
  rocblas_operation oa, ob;
  if (p->lib.ok && p->K > 0 && lib_op(p->lib.opA, conj_a != 0, &oa) && lib_op(p->lib.opB, conj_b != 0, &ob) &&
      !capturing(s)) {
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    return lib_gemm(ctx, s, p->dtype, oa, ob, p->M, p->N, p->K, alpha ? alpha : one, A, p->lib.lda, p->lib.sA, B,
                    p->lib.ldb, p->lib.sB, beta ? beta : zero, C, p->lib.ldc, p->lib.sC, p->batch);
  }
  void* ws = nullptr;
  if (p->ws_bytes) TRY(ws_get(ctx, s, p->ws_bytes, &ws, 1));
  StatScope st(ctx, s, TENET_KSTAT_GEMM, 8.0 * p->M * p->N * p->K * p->batch,
               (double)es * ((double)p->M * p->K + (double)p->K * p->N + (double)p->M * p->N) * p->batch);
  if (p->dtype == TENET_C128) return launch_gemm_c128(g, p->prec, s, ws, p->ws_bytes);
  return launch_gemm_c64(g, p->prec, s, ws, p->ws_bytes);
}

extern "C" int tenet_einsum_exec(tenet_plan plan, tenet_stream stream, const void* A, const void* B, void* C,
                                 int conj_a, int conj_b, const double alpha[2], const double beta[2]) {
  CHECK_ARG(plan, "null plan");
  TRY(use_device(plan->ctx));
  return exec_plan(plan, pick(plan->ctx, stream), A, B, C, conj_a, conj_b, alpha, beta);
}

extern "C" int tenet_plan_destroy(tenet_plan plan) {
  if (!plan) return TENET_OK;
  if (plan->cached) return TENET_OK;  // owned by the context cache
  (void)hipSetDevice(plan->ctx->device);
  if (plan->tables) (void)hipFree(plan->tables);
  delete plan;
  return TENET_OK;
}

static int cached_plan(tenet_ctx ctx, hipStream_t s, int dtype, int prec, const tenet_tdesc* a, const tenet_tdesc* b,
                       const tenet_tdesc* c, tenet_plan* out) {
  const std::string key = plan_key(dtype, prec, a, b, c);
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) {
      *out = it->second;
      return TENET_OK;
    }
  }
  if (capturing(s)) {
    set_error("contraction plan not cached: run the sequence once before graph capture");
    return TENET_ERR_UNSUPPORTED;
  }
  auto* p = new tenet_plan_s();
  int st = build_plan(ctx, dtype, prec, a, b, c, p);
  if (st != TENET_OK) {
    if (p->tables) (void)hipFree(p->tables);
    delete p;
    return st;
  }
  p->cached = true;
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->plans[key] = p;
  ctx->plan_lru.push_back(key);
  if (ctx->plan_lru.size() > 8192) {
    // evict the oldest entry; plans are only freed after the device drains
    (void)hipDeviceSynchronize();
    const std::string old = ctx->plan_lru.front();
    ctx->plan_lru.pop_front();
    auto it = ctx->plans.find(old);
    if (it != ctx->plans.end()) {
      if (it->second->tables) (void)hipFree(it->second->tables);
      delete it->second;
      ctx->plans.erase(it);
    }
  }
  *out = p;
  return TENET_OK;
}

extern "C" int tenet_einsum(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, const tenet_tdesc* a,
                            const void* A, const tenet_tdesc* b, const void* B, const tenet_tdesc* c, void* C,
                            int conj_a, int conj_b) {
  TRY(use_device(ctx));
  CHECK_ARG(a && b && c, "null descriptor");
  TRY(valid_dtype_prec(dtype, prec));
  hipStream_t s = pick(ctx, stream);
  tenet_plan p = nullptr;
  TRY(cached_plan(ctx, s, dtype, prec, a, b, c, &p));
  return exec_plan(p, s, A, B, C, conj_a, conj_b, nullptr, nullptr);
}

// column-major contiguous descriptor helper
static tenet_tdesc mkdesc(std::initializer_list<int64_t> ext, std::initializer_list<int64_t> lab) {
  tenet_tdesc d{};
  d.rank = (int32_t)ext.size();
  int i = 0;
  int64_t st = 1;
  auto li = lab.begin();
  for (int64_t e : ext) {
    d.extent[i] = e;
    d.stride[i] = st;
    d.label[i] = *li++;
    st *= e;
    ++i;
  }
  return d;
}

static int einsum_internal(tenet_ctx ctx, hipStream_t s, int dtype, int prec, const tenet_tdesc& a, const void* A,
                           const tenet_tdesc& b, const void* B, const tenet_tdesc& c, void* C, int conj_a,
                           int conj_b) {
  tenet_plan p = nullptr;
  TRY(cached_plan(ctx, s, dtype, prec, &a, &b, &c, &p));
  return exec_plan(p, s, A, B, C, conj_a, conj_b, nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// QR
// ---------------------------------------------------------------------------
static size_t qr_ws_bytes(int64_t m, int64_t n) {
  const int64_t k = std::min(m, n);
  return std::max(align_up((size_t)m * n * 16 + (size_t)m * k * 16, 256) + align_up((size_t)k * 16, 256) + 256,
                  qr_blocked_ws_bytes(m, n) + 256);
}

static int qr_run_inner(hipStream_t s, int dtype, int flags, int64_t m, int64_t n, const void* A, int64_t lda,
                        void* Q, int64_t ldq, void* R, int64_t ldr, Bump& b, int64_t k);

static int qr_run(tenet_ctx ctx, hipStream_t s, int dtype, int flags, int64_t m, int64_t n, const void* A,
                  int64_t lda, void* Q, int64_t ldq, void* R, int64_t ldr, void* wsp) {
  Bump b(wsp);
  const int64_t k = std::min(m, n);
  const double p = (double)std::max(m, n), q = (double)k;
  const double fl = 4.0 * (4.0 * p * q * q - 4.0 * q * q * q / 3.0);
  const double by = (double)ELEM[dtype] * ((double)m * n + (double)m * k + (double)k * n);
  if (qr_small_fits(m, n)) {
    // one-launch QR reading A in place: the timing sample is that kernel, stamped by its dispatch
    StatScope st(ctx, s, TENET_KSTAT_QR, fl, by, true);
    const int rc =
        launch_qr_small(A, lda, m, n, Q, ldq, R, ldr, dtype, flags, nullptr, s, st.start_event(), st.stop_event());
    if (rc == TENET_OK) st.launched();
    return rc;
  }
  StatScope st(ctx, s, TENET_KSTAT_QR, fl, by);
  return qr_run_inner(s, dtype, flags, m, n, A, lda, Q, ldq, R, ldr, b, k);
}

static int qr_run_inner(hipStream_t s, int dtype, int flags, int64_t m, int64_t n, const void* A, int64_t lda,
                        void* Q, int64_t ldq, void* R, int64_t ldr, Bump& b, int64_t k) {
  double* W = (double*)b.take((size_t)m * n * 16 + (size_t)m * k * 16);
  double* tau = (double*)b.take((size_t)k * 16);
  TRY(launch_pack_c128(A, dtype, m, n, lda, (flags & TENET_QR_IN_ADJ) ? 1 : 0, W, s));
  if (qr_blocked_fits(m)) {
    return launch_qr_blocked(W, m, n, Q, ldq, R, ldr, dtype, flags & (TENET_QR_Q_ADJ | TENET_QR_R_ADJ), s);
  }
  TRY(launch_qr_factor(W, tau, m, n, 1, s));
  TRY(launch_qr_extract(W, tau, m, n, 1, Q, ldq, R, ldr, dtype, flags & (TENET_QR_Q_ADJ | TENET_QR_R_ADJ), s));
  return TENET_OK;
}

extern "C" int tenet_qr_thin_ex(tenet_ctx ctx, tenet_stream stream, int dtype, int flags, int64_t m, int64_t n,
                                const void* A, int64_t lda, void* Q, int64_t ldq, void* R, int64_t ldr) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(m > 0 && n > 0, "empty matrix");
  CHECK_ARG((flags & ~7) == 0, "unknown QR flags");
  const int64_t k = std::min(m, n);
  const bool in_adj = flags & TENET_QR_IN_ADJ, q_adj = flags & TENET_QR_Q_ADJ, r_adj = flags & TENET_QR_R_ADJ;
  CHECK_ARG(lda >= (in_adj ? n : m), "lda too small");
  CHECK_ARG(ldq >= (q_adj ? k : m), "ldq too small");
  CHECK_ARG(ldr >= (r_adj ? n : k), "ldr too small");
  const size_t es = ELEM[dtype];
  TRY(check_span(ctx, A, (in_adj ? ((size_t)(m - 1) * lda + n) : ((size_t)(n - 1) * lda + m)) * es, "A"));
  TRY(check_span(ctx, Q, (q_adj ? ((size_t)(m - 1) * ldq + k) : ((size_t)(k - 1) * ldq + m)) * es, "Q"));
  TRY(check_span(ctx, R, (r_adj ? ((size_t)(k - 1) * ldr + n) : ((size_t)(n - 1) * ldr + k)) * es, "R"));
  hipStream_t s = pick(ctx, stream);
  void* w = nullptr;
  TRY(ws_get(ctx, s, qr_ws_bytes(m, n), &w));
  return qr_run(ctx, s, dtype, flags, m, n, A, lda, Q, ldq, R, ldr, w);
}

static bool spans_overlap(const void* a, size_t na, const void* b, size_t nb) {
  const char* pa = static_cast<const char*>(a);
  const char* pb = static_cast<const char*>(b);
  return pa < pb + nb && pb < pa + na;
}

extern "C" int tenet_qr_absorb(tenet_ctx ctx, tenet_stream stream, int dtype, int flags, int64_t m, int64_t n,
                               const void* A, int64_t lda, void* Q, int64_t ldq, int side, int64_t nb,
                               const void* B, int64_t ldb, void* Bout, int64_t ldbo) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(m > 0 && n > 0 && nb > 0, "empty matrix");
  CHECK_ARG((flags & ~(TENET_QR_IN_ADJ | TENET_QR_Q_ADJ)) == 0, "flags: only TENET_QR_IN_ADJ and TENET_QR_Q_ADJ");
  CHECK_ARG(side == TENET_ABSORB_R_B || side == TENET_ABSORB_B_RH, "unknown absorb side");
  CHECK_ARG(nb <= INT32_MAX / 64, "nb too large");
  if (!qr_small_fits(m, n)) {
    set_error("tenet_qr_absorb: m, n <= 64 only (use tenet_qr_thin_ex and tenet_einsum)");
    return TENET_ERR_UNSUPPORTED;
  }
  const int64_t k = std::min(m, n);
  const bool in_adj = flags & TENET_QR_IN_ADJ, q_adj = flags & TENET_QR_Q_ADJ;
  const bool rb = side == TENET_ABSORB_R_B;
  CHECK_ARG(lda >= (in_adj ? n : m), "lda too small");
  CHECK_ARG(ldq >= (q_adj ? k : m), "ldq too small");
  CHECK_ARG(ldb >= (rb ? n : nb), "ldb too small");
  CHECK_ARG(ldbo >= (rb ? k : nb), "ldbo too small");
  const size_t es = ELEM[dtype];
  const size_t sa = (in_adj ? ((size_t)(m - 1) * lda + n) : ((size_t)(n - 1) * lda + m)) * es;
  const size_t sq = (q_adj ? ((size_t)(m - 1) * ldq + k) : ((size_t)(k - 1) * ldq + m)) * es;
  const size_t sb = (rb ? ((size_t)(nb - 1) * ldb + n) : ((size_t)(n - 1) * ldb + nb)) * es;
  const size_t so = (rb ? ((size_t)(nb - 1) * ldbo + k) : ((size_t)(k - 1) * ldbo + nb)) * es;
  TRY(check_span(ctx, A, sa, "A"));
  TRY(check_span(ctx, Q, sq, "Q"));
  TRY(check_span(ctx, B, sb, "B"));
  TRY(check_span(ctx, Bout, so, "Bout"));
  CHECK_ARG(!spans_overlap(Bout, so, A, sa) && !spans_overlap(Bout, so, Q, sq) && !spans_overlap(Bout, so, B, sb),
            "Bout overlaps A, Q or B");
  CHECK_ARG(!spans_overlap(Q, sq, B, sb), "Q overlaps B (B is read while Q is written)");
  hipStream_t s = pick(ctx, stream);
  const double p = (double)std::max(m, n), q = (double)k;
  const double fl = 4.0 * (4.0 * p * q * q - 4.0 * q * q * q / 3.0) + 8.0 * (double)nb * q * (double)n;
  const double by = (double)es * ((double)m * n + (double)m * k + 2.0 * (double)nb * n);
  const QrAbsorbArgs abs{B, Bout, ldb, ldbo, nb, side};
  StatScope st(ctx, s, TENET_KSTAT_QR, fl, by, true);
  const int rc =
      launch_qr_small(A, lda, m, n, Q, ldq, nullptr, 0, dtype, flags, &abs, s, st.start_event(), st.stop_event());
  if (rc == TENET_OK) st.launched();
  return rc;
}

extern "C" int tenet_qr_thin(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t m, int64_t n, const void* A,
                             int64_t lda, void* Q, int64_t ldq, void* R, int64_t ldr) {
  return tenet_qr_thin_ex(ctx, stream, dtype, 0, m, n, A, lda, Q, ldq, R, ldr);
}

// ---------------------------------------------------------------------------
// SVD (one-sided block Jacobi)
// ---------------------------------------------------------------------------
struct SvdShape {
  int64_t nr, ncols, b, nblk, ncp, q, npairs;
  bool transposed;
};

static SvdShape svd_shape(int64_t m, int64_t n, bool allow_split, bool force_adj = false) {
  SvdShape sh{};
  sh.transposed = force_adj || m < n;
  sh.nr = std::max(m, n);
  sh.ncols = std::min(m, n);
  if (sh.ncols <= 32) {
    sh.b = std::max<int64_t>(1, (sh.ncols + 1) / 2);
    sh.nblk = 2;
  } else {
    sh.b = 16;
    sh.nblk = (sh.ncols + 15) / 16;
    if (sh.nblk % 2) sh.nblk += 1;
  }
  sh.ncp = sh.nblk * sh.b;
  sh.npairs = sh.nblk / 2;
  sh.q = 1;
  if (allow_split) sh.q = std::min<int64_t>(16, std::max<int64_t>(1, (sh.nr + 63) / 64));
  return sh;
}

static size_t svd_ws_bytes(const SvdShape& sh) {
  const int64_t n2 = 2 * sh.b;
  size_t t = 0;
  t += align_up((size_t)sh.nr * sh.ncp * 16, 256);
  t += align_up((size_t)sh.ncp * sh.ncp * 16, 256);
  t += align_up((size_t)sh.npairs * sh.q * n2 * n2 * 16, 256);
  t += align_up((size_t)sh.ncp * 8, 256);
  t += align_up((size_t)sh.ncp * 4, 256);
  t += 256;  // flags
  t += 256;  // k / disc fallback
  return t + 256;
}

// Sweeps a captured SVD runs (it cannot stop at convergence; rounds of a
// converged matrix return at once): Vidal-gauge bond matrices at chi = 256
// have needed up to ~20 of the eager limit of 40.
constexpr int CAPTURED_SWEEPS = 30;

static double svd_tol(int64_t nr) { return 1e-14 * std::max(1.0, std::sqrt((double)nr / 16.0)); }

static int svd_run(tenet_ctx ctx, hipStream_t s, int dtype, int64_t m, int64_t n, const void* A, int64_t lda, void* U,
                   int64_t ldu, double* S, void* Vh, int64_t ldvh, int64_t maxdim, double thr, int64_t* k_dev,
                   double* disc_dev, void* wsp, bool allow_split, bool force_adj = false) {
  const SvdShape sh = svd_shape(m, n, allow_split, force_adj);
  Bump bm(wsp);
  SvdProblem P{};
  P.X = (double*)bm.take((size_t)sh.nr * sh.ncp * 16);
  P.V = (double*)bm.take((size_t)sh.ncp * sh.ncp * 16);
  P.Gpart = (double*)bm.take((size_t)sh.npairs * sh.q * (2 * sh.b) * (2 * sh.b) * 16);
  P.sigma = (double*)bm.take((size_t)sh.ncp * 8);
  P.perm = (int32_t*)bm.take((size_t)sh.ncp * 4);
  P.flag = (int32_t*)bm.take(256);
  P.fro2 = (double*)(P.flag + 8);
  int64_t* kfall = (int64_t*)bm.take(256);
  P.nr = (int32_t)sh.nr;
  P.ncols = (int32_t)sh.ncols;
  P.ncp = (int32_t)sh.ncp;
  P.b = (int32_t)sh.b;
  P.nblk = (int32_t)sh.nblk;
  P.q = (int32_t)sh.q;
  P.m = (int32_t)m;
  P.n = (int32_t)n;
  P.transposed = sh.transposed ? 1 : 0;
  P.accum = 1;
  P.A = A;
  P.lda = lda;
  P.U = U;
  P.ldu = ldu;
  P.Vh = Vh;
  P.ldvh = ldvh;
  P.S = S;
  P.maxdim = maxdim;
  P.threshold = thr;
  P.k_dev = k_dev ? k_dev : kfall;
  P.disc_dev = disc_dev ? disc_dev : (double*)(kfall + 1);
  P.tol = svd_tol(sh.nr);
  SvdBatch B{};
  B.arr = nullptr;
  B.one = P;
  TRY(launch_svd_init(B, 1, sh.nr * sh.ncp + sh.ncp * sh.ncp, (int)sh.ncp, dtype, s));
  const bool sync_ok = !capturing(s);
  const int max_sweeps = sync_ok ? 40 : CAPTURED_SWEEPS;
  bool converged = false;
  // one sweep = a full-inner round on the fixed pairs (2i, 2i+1), then the
  // circle rounds with cross-block inner steps (every column pair once)
  // algorithmic work of one round: Gram of every pair over nr rows, and the
  // rotation applied to the pair's nr rows of X and ncp rows of V
  const double n2 = 2.0 * sh.b;
  const double gram_fl = 8.0 * sh.npairs * n2 * n2 * sh.nr;
  const double apply_fl = 8.0 * sh.npairs * n2 * n2 * (sh.nr + sh.ncp);
  const double xv_bytes = 16.0 * sh.npairs * n2 * (sh.nr + sh.ncp);
  auto round = [&](int r, int mode) -> int {
    if (sh.q == 1)
      return svd_round_timed(ctx, s, gram_fl + apply_fl, 16.0 * sh.npairs * n2 * sh.nr + 2.0 * xv_bytes, B, 1,
                             (int)sh.npairs, r, 2, mode);
    TRY(svd_round_timed(ctx, s, gram_fl, 16.0 * sh.npairs * n2 * sh.nr, B, 1, (int)(sh.npairs * sh.q), r, 0, mode));
    return svd_round_timed(ctx, s, apply_fl, 2.0 * xv_bytes, B, 1, (int)(sh.npairs * sh.q), r, 1, mode);
  };
  for (int sw = 0; sw < max_sweeps && !converged; ++sw) {
    TRY(round(0, 0));
    if (sh.nblk > 2)
      for (int r = 0; r < sh.nblk - 1; ++r) TRY(round(r, 1));
    TRY(launch_svd_sweep_end(B, 1, s));
    if (sync_ok) {
      TENET_HIP_CHECK(hipMemcpyAsync(ctx->host_flag, P.flag, 16, hipMemcpyDeviceToHost, s));
      TENET_HIP_CHECK(hipStreamSynchronize(s));
      converged = ctx->host_flag[1] != 0;
    }
  }
  TRY(launch_svd_finalize(B, 1, (int)sh.ncp, (int64_t)(m + n) * sh.ncols, dtype, s));
  // under capture the sweep count is fixed (rounds of converged problems
  // return at once); whether it sufficed is known only on the device
  if (!sync_ok) TRY(launch_svd_status(B, 1, ctx->dev_status, s));
  if (sync_ok && !converged) {
    set_error("Jacobi SVD did not converge within the sweep limit");
    return TENET_ERR_NOCONV;
  }
  return TENET_OK;
}

// ---- QR-preconditioned SVD -----------------------------------------------
// For the tall form (p x q, p = max(m, n), q = min(m, n)) take the thin
// Householder QR P = Q R, then run the one-sided Jacobi on R^H (Drmac and
// Veselic, SIAM J. Matrix Anal. Appl. 29 (2008): the rows of R are graded,
// so the sweeps converge in roughly half as many rounds, with the same
// relative accuracy).  SVD(R) = U_R S V_R^H with the Jacobi on X = R^H gives
//   m >= n (A = Q R):        U = Q U_R,          V^H = V_R^H
//   m <  n (A^H = Q R):      U = (V_R^H)^H,      V^H = (Q U_R)^H
// Preconditioning is available for every form with q >= 64.  Up to 1024
// rows (the blocked QR) it runs first; above, the QR is the single-workgroup
// kernel (~20x the plain sweeps' time on a Gaussian 1500 x 300), so the
// plain sweeps run first and the preconditioned path only when they do not
// converge — strongly graded matrices (sigma over 12 decades, mixing
// singular vectors) need more sweeps than the limit without it.
static bool svd_precond_applies(tenet_ctx ctx, int dtype, int64_t m, int64_t n) {
  const int64_t q = std::min(m, n);
  return ctx->svd_precond && dtype == TENET_C128 && q >= 64;
}

static size_t svd_precond_ws_bytes(int64_t m, int64_t n) {
  const int64_t p = std::max(m, n), q = std::min(m, n);
  return 2 * align_up((size_t)p * q * 16, 256) + 3 * align_up((size_t)q * q * 16, 256) +
         align_up(qr_ws_bytes(p, q), 256) + svd_ws_bytes(svd_shape(q, q, true, true)) + align_up((size_t)q * 8, 256) +
         align_up((size_t)q * 4, 256) + 1024;
}

// workspace any SVD path of an m x n matrix may need
static size_t svd_any_ws_bytes(tenet_ctx ctx, int dtype, int64_t m, int64_t n) {
  size_t b = svd_ws_bytes(svd_shape(m, n, true));
  if (svd_precond_applies(ctx, dtype, m, n)) b = std::max(b, svd_precond_ws_bytes(m, n));
  return b;
}

static tenet_tdesc mkdesc2(int64_t rows, int64_t cols, int64_t ld, int64_t lr, int64_t lc) {
  tenet_tdesc d{};
  d.rank = 2;
  d.extent[0] = rows;
  d.extent[1] = cols;
  d.stride[0] = 1;
  d.stride[1] = ld;
  d.label[0] = lr;
  d.label[1] = lc;
  return d;
}

static int svd_any(tenet_ctx ctx, hipStream_t s, int dtype, int64_t m, int64_t n, const void* A, int64_t lda, void* U,
                   int64_t ldu, double* S, void* Vh, int64_t ldvh, int64_t maxdim, double thr, int64_t* k_dev,
                   double* disc_dev, void* wsp) {
  if (!svd_precond_applies(ctx, dtype, m, n))
    return svd_run(ctx, s, dtype, m, n, A, lda, U, ldu, S, Vh, ldvh, maxdim, thr, k_dev, disc_dev, wsp, true);
  if (!qr_blocked_fits(std::max(m, n))) {
    // the Jacobi works on its own copy of A, so a retry starts from the untouched input
    const int st0 = svd_run(ctx, s, dtype, m, n, A, lda, U, ldu, S, Vh, ldvh, maxdim, thr, k_dev, disc_dev, wsp, true);
    if (st0 != TENET_ERR_NOCONV) return st0;
  }
  const int64_t p = std::max(m, n), q = std::min(m, n);
  const bool tall = m >= n;
  Bump bm(wsp);
  void* Q = bm.take((size_t)p * q * 16);
  void* R = bm.take((size_t)q * q * 16);
  void* UR = bm.take((size_t)q * q * 16);
  void* VRh = bm.take((size_t)q * q * 16);
  void* qr_ws = bm.take(align_up(qr_ws_bytes(p, q), 256));
  void* svd_ws = bm.take(svd_ws_bytes(svd_shape(q, q, true, true)));
  void* Tp = bm.take((size_t)p * q * 16);
  double* nrm = (double*)bm.take((size_t)q * 8);
  int32_t* perm = (int32_t*)bm.take((size_t)q * 4);
  // T = A (tall) or A^H (wide), its columns by descending norm: T[:, perm] = Q R
  // (the ordering makes R's rows decrease in norm as column pivoting would:
  // the Jacobi on R^H of TEBD bond matrices then needs ~30% fewer sweeps)
  TRY(launch_pre1_order_pack(A, lda, (int)p, (int)q, tall ? 0 : 1, TENET_C128, nrm, perm, (double*)Tp, s));
  TRY(qr_run(ctx, s, TENET_C128, 0, p, q, Tp, p, Q, p, R, q, qr_ws));
  int st = svd_run(ctx, s, TENET_C128, q, q, R, q, UR, q, S, VRh, q, maxdim, thr, k_dev, disc_dev, svd_ws, true, true);
  if (st != TENET_OK && st != TENET_ERR_NOCONV) return st;
  if (tall) {
    // A[:, perm] = Q U_R S V_R^H: U = Q U_R, V^H[:, perm[c]] = V_R^H[:, c]
    const tenet_tdesc dQ = mkdesc2(p, q, p, 1, 2), dU = mkdesc2(q, q, q, 2, 3), dO = mkdesc2(m, q, ldu, 1, 3);
    TRY(einsum_internal(ctx, s, TENET_C128, TENET_PREC_NATIVE, dQ, Q, dU, UR, dO, U, 0, 0));
    TRY(launch_perm_scatter((const double*)VRh, (int)q, (int)q, q, (double*)Vh, ldvh, perm, 0, s));
    return st;
  }
  // A^H[:, perm] = Q U_R S V_R^H: V^H = (Q U_R)^H, i.e.
  // V^H[j, c] = sum_i conj(U_R[i, j]) conj(Q[c, i]), and U[perm[c], j] = conj(V_R^H[j, c])
  const tenet_tdesc dUR = mkdesc2(q, q, q, 1, 2), dQ = mkdesc2(p, q, p, 3, 1), dVh = mkdesc2(q, n, ldvh, 2, 3);
  TRY(einsum_internal(ctx, s, TENET_C128, TENET_PREC_NATIVE, dUR, UR, dQ, Q, dVh, Vh, 1, 1));
  TRY(launch_perm_scatter((const double*)VRh, (int)q, (int)q, q, (double*)U, ldu, perm, 1, s));
  return st;
}

extern "C" int tenet_svd_set_precondition(tenet_ctx ctx, int on) {
  TRY(use_device(ctx));
  ctx->svd_precond = on != 0;
  return TENET_OK;
}

extern "C" int tenet_svd_thin(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t m, int64_t n, const void* A,
                              int64_t lda, void* U, int64_t ldu, double* S, void* Vh, int64_t ldvh, int64_t maxdim,
                              double threshold, int64_t* k_dev, double* discarded_dev) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(m > 0 && n > 0, "empty matrix");
  const int64_t k = std::min(m, n);
  CHECK_ARG(lda >= m && ldu >= m && ldvh >= k, "leading dimension too small");
  CHECK_ARG(k <= 4096, "min(m, n) above 4096 is not supported");
  const size_t es = ELEM[dtype];
  TRY(check_span(ctx, A, ((size_t)(n - 1) * lda + m) * es, "A"));
  TRY(check_span(ctx, U, ((size_t)(k - 1) * ldu + m) * es, "U"));
  TRY(check_span(ctx, Vh, ((size_t)(n - 1) * ldvh + k) * es, "Vh"));
  if (S) TRY(check_span(ctx, S, (size_t)k * 8, "S"));
  hipStream_t s = pick(ctx, stream);
  void* w = nullptr;
  TRY(ws_get(ctx, s, svd_any_ws_bytes(ctx, dtype, m, n), &w));
  return svd_any(ctx, s, dtype, m, n, A, lda, U, ldu, S, Vh, ldvh, maxdim, threshold, k_dev, discarded_dev, w);
}

extern "C" int tenet_svd_thin_noalloc(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t m, int64_t n,
                                      const void* A, int64_t lda, void* U, int64_t ldu, double* S, void* Vh,
                                      int64_t ldvh, int64_t maxdim, double threshold, int64_t* k_dev,
                                      double* discarded_dev) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(m > 0 && n > 0, "empty matrix");
  const int64_t k = std::min(m, n);
  CHECK_ARG(lda >= m && ldu >= m && ldvh >= k, "leading dimension too small");
  const size_t es = ELEM[dtype];
  TRY(check_span(ctx, A, ((size_t)(n - 1) * lda + m) * es, "A"));
  TRY(check_span(ctx, U, ((size_t)(k - 1) * ldu + m) * es, "U"));
  TRY(check_span(ctx, Vh, ((size_t)(n - 1) * ldvh + k) * es, "Vh"));
  if (S) TRY(check_span(ctx, S, (size_t)k * 8, "S"));
  // the plain Jacobi path (no QR preconditioning: its contractions build
  // plans, which allocate); its whole workspace must fit the scratch array
  const size_t need = k <= 4096 ? svd_ws_bytes(svd_shape(m, n, true)) : SIZE_MAX;
  if (need > TENET_SVD_NOALLOC_SCRATCH_BYTES) {
    set_error("tenet_svd_thin_noalloc: a " + std::to_string(m) + " x " + std::to_string(n) + " SVD needs " +
              std::to_string(need) + " bytes of workspace, above the " +
              std::to_string((size_t)TENET_SVD_NOALLOC_SCRATCH_BYTES) +
              "-byte scratch array (tenet_svd_thin allocates instead)");
    return TENET_ERR_UNSUPPORTED;
  }
  const hipError_t e = scratch_lookup(ctx->device);
  if (e != hipSuccess) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return TENET_ERR_HIP;
  }
  return svd_run(ctx, pick(ctx, stream), dtype, m, n, A, lda, U, ldu, S, Vh, ldvh, maxdim, threshold, k_dev,
                 discarded_dev, g_scratch_ptr[ctx->device], true);
}

// workspace of one tenet_simple_update with bond matrix m x n
static size_t simple_update_ws_bytes(tenet_ctx ctx, int dtype, int64_t m, int64_t n) {
  const size_t es = ELEM[dtype];
  const int64_t kmax = std::min(m, n);
  return 2 * align_up((size_t)m * n * es, 256) + align_up((size_t)m * kmax * es, 256) +
         align_up((size_t)kmax * n * es, 256) + align_up((size_t)kmax * 8, 256) + 512 +
         align_up(svd_any_ws_bytes(ctx, dtype, m, n), 256);
}

extern "C" int tenet_workspace_size(tenet_ctx ctx, int op, int dtype, int64_t m, int64_t n, size_t* bytes) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(bytes && m > 0 && n > 0, "bad argument");
  switch (op) {
    case TENET_OP_QR:
      *bytes = qr_ws_bytes(m, n);
      return TENET_OK;
    case TENET_OP_SVD:
      *bytes = svd_any_ws_bytes(ctx, dtype, m, n);
      return TENET_OK;
    case TENET_OP_SIMPLE_UPDATE:
      *bytes = simple_update_ws_bytes(ctx, dtype, m, n);
      return TENET_OK;
    default:
      set_error("tenet_workspace_size: unknown op");
      return TENET_ERR_ARG;
  }
}

extern "C" int tenet_workspace_reserve(tenet_ctx ctx, tenet_stream stream, size_t bytes) {
  TRY(use_device(ctx));
  void* w = nullptr;
  return ws_get(ctx, pick(ctx, stream), bytes, &w);
}

// ---------------------------------------------------------------------------
// fused gate update
// ---------------------------------------------------------------------------
extern "C" int tenet_simple_update(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t chil, int64_t d1,
                                   int64_t chi, int64_t d2, int64_t chir, const void* A, const void* B, const void* G,
                                   void* Aout, void* Bout, double* S_out, int64_t maxdim, double threshold, int absorb,
                                   int64_t* k_dev, double* disc_dev, int64_t* k_host) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, prec));
  CHECK_ARG(chil > 0 && d1 > 0 && chi > 0 && d2 > 0 && chir > 0, "non-positive extent");
  CHECK_ARG(absorb >= TENET_ABSORB_NONE && absorb <= TENET_ABSORB_RIGHT, "bad absorb mode");
  hipStream_t s = pick(ctx, stream);
  const size_t es = ELEM[dtype];
  const int64_t m = chil * d1, n = d2 * chir, kmax = std::min(m, n);
  CHECK_ARG(kmax <= 4096, "bond matrix too large");
  TRY(check_span(ctx, A, (size_t)chil * d1 * chi * es, "A"));
  TRY(check_span(ctx, B, (size_t)chi * d2 * chir * es, "B"));
  TRY(check_span(ctx, G, (size_t)d1 * d2 * d1 * d2 * es, "G"));
  TRY(check_span(ctx, Aout, (size_t)m * kmax * es, "Aout"));
  TRY(check_span(ctx, Bout, (size_t)kmax * n * es, "Bout"));
  if (S_out) TRY(check_span(ctx, S_out, (size_t)kmax * 8, "S_out"));
  const size_t svd_b = align_up(svd_any_ws_bytes(ctx, dtype, m, n), 256);
  const size_t theta_b = align_up((size_t)m * n * es, 256);
  const size_t need = simple_update_ws_bytes(ctx, dtype, m, n);
  void* w = nullptr;
  TRY(ws_get(ctx, s, need, &w));
  Bump bm(w);
  void* theta = bm.take(theta_b);
  void* theta2 = bm.take(theta_b);
  void* U = bm.take((size_t)m * kmax * es);
  void* Vh = bm.take((size_t)kmax * n * es);
  double* S = S_out ? S_out : (double*)bm.take((size_t)kmax * 8);
  int64_t* kd = (int64_t*)bm.take(256);
  int64_t* kuse = k_dev ? k_dev : kd;
  double* duse = disc_dev ? disc_dev : (double*)(kd + 1);
  void* svd_ws = bm.take(svd_b);
  // theta(l, ia, ib, r) = A(l, ia, c) B(c, ib, r)
  const tenet_tdesc dA = mkdesc({chil, d1, chi}, {1, 2, 3});
  const tenet_tdesc dB = mkdesc({chi, d2, chir}, {3, 4, 5});
  const tenet_tdesc dT = mkdesc({chil, d1, d2, chir}, {1, 2, 4, 5});
  TRY(einsum_internal(ctx, s, dtype, prec, dA, A, dB, B, dT, theta, 0, 0));
  // theta2(l, oa, ob, r) = theta(l, ia, ib, r) G(oa, ob, ia, ib)
  const tenet_tdesc dG = mkdesc({d1, d2, d1, d2}, {6, 7, 2, 4});
  const tenet_tdesc dT2 = mkdesc({chil, d1, d2, chir}, {1, 6, 7, 5});
  const int gprec = (prec == TENET_PREC_SPLIT_BF16 || prec == TENET_PREC_SPLIT_FP16) ? TENET_PREC_NATIVE : prec;
  TRY(einsum_internal(ctx, s, dtype, gprec, dT, theta, dG, G, dT2, theta2, 0, 0));
  int st = svd_any(ctx, s, dtype, m, n, theta2, m, U, m, S, Vh, kmax, maxdim, threshold, kuse, duse, svd_ws);
  if (st != TENET_OK && st != TENET_ERR_NOCONV) return st;
  TRY(launch_absorb(U, m, Vh, kmax, n, S, kuse, absorb, Aout, Bout, kmax, dtype, s));
  if (k_host) {
    TENET_HIP_CHECK(hipMemcpyAsync(ctx->host_k, kuse, 8, hipMemcpyDeviceToHost, s));
    TENET_HIP_CHECK(hipStreamSynchronize(s));
    *k_host = ctx->host_k[0];
  }
  return st;
}

// ---------------------------------------------------------------------------
// transfer-matrix sweeps
// ---------------------------------------------------------------------------
extern "C" int tenet_overlap(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t n, int64_t d,
                             const int64_t* chi_a, const void* const* a_sites, const int64_t* chi_b,
                             const void* const* b_sites, double* out) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, prec));
  CHECK_ARG(n >= 1 && d >= 1 && chi_a && chi_b && a_sites && b_sites && out, "bad argument");
  CHECK_ARG(chi_a[0] == 1 && chi_a[n] == 1 && chi_b[0] == 1 && chi_b[n] == 1, "open boundary bonds must be 1");
  hipStream_t s = pick(ctx, stream);
  const size_t es = ELEM[dtype];
  for (int64_t i = 0; i < n; ++i) {
    TRY(check_span(ctx, a_sites[i], (size_t)chi_a[i] * d * chi_a[i + 1] * es, "a site"));
    TRY(check_span(ctx, b_sites[i], (size_t)chi_b[i] * d * chi_b[i + 1] * es, "b site"));
  }
  if (prec == TENET_PREC_NATIVE && overlap_small_fits(n, d, chi_a, chi_b)) {
    // small bonds: the whole chain in one workgroup, the environment in LDS
    OverlapTab tab{};
    tab.n = (int32_t)n;
    tab.d = (int32_t)d;
    for (int64_t i = 0; i <= n; ++i) {
      tab.chi_a[i] = chi_a[i];
      tab.chi_b[i] = chi_b[i];
    }
    for (int64_t i = 0; i < n; ++i) {
      tab.a[i] = (uint64_t)(uintptr_t)a_sites[i];
      tab.b[i] = (uint64_t)(uintptr_t)b_sites[i];
    }
    return launch_overlap_small(tab, out, dtype, s);
  }
  int64_t max_env = 1, max_t = 1;
  for (int64_t i = 0; i <= n; ++i) max_env = std::max(max_env, chi_a[i] * chi_b[i]);
  for (int64_t i = 0; i < n; ++i) max_t = std::max(max_t, chi_b[i] * d * chi_a[i + 1]);
  void* w = nullptr;
  TRY(ws_get(ctx, s, 2 * align_up((size_t)max_env * es, 256) + align_up((size_t)max_t * es, 256) + 256, &w));
  Bump bm(w);
  void* env[2] = {bm.take((size_t)max_env * es), bm.take((size_t)max_env * es)};
  void* T = bm.take((size_t)max_t * es);
  TRY(launch_fill_zero(out, 3 * sizeof(double), s));
  int cur = 0;
  // site 0: env(ra, rb) = sum_s A(1, s, ra) conj(B(1, s, rb))
  {
    const tenet_tdesc dA = mkdesc({1, d, chi_a[1]}, {100, 3, 4});
    const tenet_tdesc dB = mkdesc({1, d, chi_b[1]}, {100, 3, 5});
    const tenet_tdesc dE = mkdesc({1, chi_a[1], chi_b[1]}, {100, 4, 5});
    TRY(einsum_internal(ctx, s, dtype, prec, dA, a_sites[0], dB, b_sites[0], dE, env[cur], 0, 1));
    TRY(launch_rescale_log(env[cur], chi_a[1] * chi_b[1], out, dtype, s));
  }
  for (int64_t i = 1; i < n; ++i) {
    const int64_t la = chi_a[i], lb = chi_b[i], ra = chi_a[i + 1], rb = chi_b[i + 1];
    // T(lb, s, ra) = env(la, lb) A(la, s, ra)
    const tenet_tdesc dE = mkdesc({la, lb}, {1, 2});
    const tenet_tdesc dA = mkdesc({la, d, ra}, {1, 3, 4});
    const tenet_tdesc dT = mkdesc({lb, d, ra}, {2, 3, 4});
    TRY(einsum_internal(ctx, s, dtype, prec, dE, env[cur], dA, a_sites[i], dT, T, 0, 0));
    // env'(ra, rb) = T(lb, s, ra) conj(B(lb, s, rb))
    const tenet_tdesc dB = mkdesc({lb, d, rb}, {2, 3, 5});
    const tenet_tdesc dE2 = mkdesc({ra, rb}, {4, 5});
    TRY(einsum_internal(ctx, s, dtype, prec, dT, T, dB, b_sites[i], dE2, env[cur ^ 1], 0, 1));
    cur ^= 1;
    TRY(launch_rescale_log(env[cur], ra * rb, out, dtype, s));
  }
  return launch_write_scalar_from(env[cur], out, dtype, s);
}

extern "C" int tenet_amplitudes(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t nreq, const int64_t* n_dev,
                                const uint64_t* site_ptrs_dev, const uint64_t* chis_dev, const uint64_t* bits_dev,
                                int64_t d, double* out) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(nreq >= 0 && d >= 1, "bad argument");
  if (nreq == 0) return TENET_OK;
  CHECK_ARG(n_dev && site_ptrs_dev && chis_dev && bits_dev && out, "null pointer");
  hipStream_t s = pick(ctx, stream);
  StatScope st(ctx, s, TENET_KSTAT_AMPLITUDE, 0.0, 0.0);
  return launch_amplitudes(nreq, n_dev, site_ptrs_dev, chis_dev, bits_dev, d, out, dtype, 2048, s);
}

extern "C" int tenet_amplitudes_sites(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t nreq, int64_t n_max,
                                      int64_t chi_max, const int64_t* n_dev, const uint64_t* site_ptrs_dev,
                                      const uint64_t* chis_dev, const uint64_t* bits_dev, int64_t d, double* out) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(nreq >= 0 && d >= 1 && n_max >= 1 && chi_max >= 1, "bad argument");
  if (nreq == 0) return TENET_OK;
  CHECK_ARG(n_dev && site_ptrs_dev && chis_dev && bits_dev && out, "null pointer");
  CHECK_ARG(nreq <= 65535, "more than 65535 requests in one call");
  hipStream_t s = pick(ctx, stream);
  void* w = nullptr;
  const size_t vb = align_up((size_t)nreq * 4 * chi_max * sizeof(double), 256);
  TRY(ws_get(ctx, s, vb + align_up((size_t)nreq * sizeof(double), 256), &w));
  AmpSiteArgs a{};
  a.n = n_dev;
  a.sites = site_ptrs_dev;
  a.chis = chis_dev;
  a.bits = bits_dev;
  a.d = d;
  a.chi_max = chi_max;
  a.n_max = n_max;
  a.vbuf = static_cast<double*>(w);
  a.logs = reinterpret_cast<double*>(static_cast<char*>(w) + vb);
  a.out = out;
  for (int64_t i = 0; i < n_max; ++i) {
    StatScope st(ctx, s, TENET_KSTAT_AMPLITUDE, 0.0, 0.0, true);
    TRY(launch_amp_site(a, nreq, i, dtype, s, st.start_event(), st.stop_event()));
    st.launched();
  }
  return TENET_OK;
}

// Transfer-matrix sweep over sites [start, start + w): env <- conj(A) (env .
// P_j A) with P_j the Pauli letter j (NULL letters: identity, the site is
// used directly), renormalised after every site with the scale's log added
// to out[2].  Leaves the final environment in *env_final (workspace).
static int env_sweep(tenet_ctx ctx, hipStream_t s, int dtype, int prec, int64_t start, int64_t w, int64_t d,
                     const int64_t* chi, const void* const* sites, const void* env_l, const char* paulis,
                     double* out, void** env_final, void* const* env_each = nullptr, double* log_each = nullptr) {
  const size_t es = ELEM[dtype];
  int64_t max_env = 1, max_site = 1;
  for (int64_t j = 0; j <= w; ++j) max_env = std::max(max_env, chi[start + j] * chi[start + j]);
  for (int64_t j = 0; j < w; ++j) max_site = std::max(max_site, chi[start + j] * d * chi[start + j + 1]);
  void* wsp = nullptr;
  TRY(ws_get(ctx, s, 2 * align_up((size_t)max_env * es, 256) + 2 * align_up((size_t)max_site * es, 256) + 256, &wsp));
  Bump bm(wsp);
  void* env[2] = {bm.take((size_t)max_env * es), bm.take((size_t)max_env * es)};
  void* PA = bm.take((size_t)max_site * es);
  void* T = bm.take((size_t)max_site * es);
  TRY(launch_fill_zero(out, 3 * sizeof(double), s));
  int cur = 0;
  const int64_t c0 = chi[start];
  if (env_l) {
    TENET_HIP_CHECK(hipMemcpyAsync(env[cur], env_l, (size_t)c0 * c0 * es, hipMemcpyDeviceToDevice, s));
  } else {
    TRY(launch_set_identity(env[cur], c0, dtype, s));
  }
  for (int64_t j = 0; j < w; ++j) {
    const int64_t i = start + j, l = chi[i], r = chi[i + 1];
    const void* op_site = sites[i];
    if (paulis && paulis[j] != 'I') {
      TRY(launch_apply_pauli(sites[i], l, r, paulis[j], PA, dtype, s));
      op_site = PA;
    }
    // T(a, o, r') = env(a, a') PA(a', o, r')
    const tenet_tdesc dE = mkdesc({l, l}, {1, 2});
    const tenet_tdesc dP = mkdesc({l, d, r}, {2, 3, 4});
    const tenet_tdesc dT = mkdesc({l, d, r}, {1, 3, 4});
    TRY(einsum_internal(ctx, s, dtype, prec, dE, env[cur], dP, op_site, dT, T, 0, 0));
    // env'(r, r') = conj(A(a, o, r)) T(a, o, r')
    const tenet_tdesc dA = mkdesc({l, d, r}, {1, 3, 5});
    const tenet_tdesc dE2 = mkdesc({r, r}, {5, 4});
    TRY(einsum_internal(ctx, s, dtype, prec, dA, sites[i], dT, T, dE2, env[cur ^ 1], 1, 0));
    cur ^= 1;
    TRY(launch_rescale_log(env[cur], r * r, out, dtype, s));
    if (env_each && env_each[j + 1])
      TENET_HIP_CHECK(hipMemcpyAsync(env_each[j + 1], env[cur], (size_t)r * r * es, hipMemcpyDeviceToDevice, s));
    if (log_each)
      TENET_HIP_CHECK(hipMemcpyAsync(log_each + j + 1, out + 2, sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  *env_final = env[cur];
  return TENET_OK;
}

extern "C" int tenet_expect_pauli(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t start, int64_t w,
                                  int64_t d, const int64_t* chi, const void* const* sites, const void* env_l,
                                  const char* paulis, double* out) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, prec));
  CHECK_ARG(d == 2, "Pauli strings need d = 2");
  CHECK_ARG(w >= 1 && start >= 0 && chi && sites && paulis && out, "bad argument");
  for (int64_t j = 0; j < w; ++j) {
    const char c = paulis[j];
    CHECK_ARG(c == 'I' || c == 'X' || c == 'Y' || c == 'Z', "Pauli letters must be I, X, Y or Z");
  }
  hipStream_t s = pick(ctx, stream);
  void* env = nullptr;
  TRY(env_sweep(ctx, s, dtype, prec, start, w, d, chi, sites, env_l, paulis, out, &env));
  return launch_trace_diag(env, chi[start + w], out, dtype, s);
}

extern "C" int tenet_left_envs(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t upto, int64_t d,
                               const int64_t* chi, const void* const* sites, void* const* env_outs, double* logs) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, prec));
  CHECK_ARG(upto >= 0 && d >= 1 && chi && sites && env_outs && logs, "bad argument");
  hipStream_t s = pick(ctx, stream);
  const size_t es = ELEM[dtype];
  for (int64_t j = 0; j <= upto; ++j)
    if (env_outs[j]) TRY(check_span(ctx, env_outs[j], (size_t)chi[j] * chi[j] * es, "env_outs"));
  TRY(check_span(ctx, logs, (size_t)(upto + 1) * sizeof(double), "logs"));
  TRY(launch_fill_zero(logs, sizeof(double), s));
  if (env_outs[0]) TRY(launch_set_identity(env_outs[0], chi[0], dtype, s));
  if (upto == 0) return TENET_OK;
  void* wsp = nullptr;
  // the sweep's running (re, im, log) triple lives in its own small buffer
  TRY(tenet_malloc(ctx, 64, &wsp));
  void* env = nullptr;
  const int st = env_sweep(ctx, s, dtype, prec, 0, upto, d, chi, sites, nullptr, nullptr, (double*)wsp, &env,
                           env_outs, logs);
  // the triple is consumed by copies already on the stream; keep it until they ran
  TENET_HIP_CHECK(hipStreamSynchronize(s));
  (void)tenet_free(ctx, wsp);
  return st;
}

extern "C" int tenet_left_env(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t upto, int64_t d,
                              const int64_t* chi, const void* const* sites, void* env_out, double* out) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, prec));
  CHECK_ARG(upto >= 0 && d >= 1 && chi && sites && env_out && out, "bad argument");
  hipStream_t s = pick(ctx, stream);
  const size_t es = ELEM[dtype];
  TRY(check_span(ctx, env_out, (size_t)chi[upto] * chi[upto] * es, "env_out"));
  if (upto == 0) {
    TRY(launch_fill_zero(out, 3 * sizeof(double), s));
    return launch_set_identity(env_out, chi[0], dtype, s);
  }
  void* env = nullptr;
  TRY(env_sweep(ctx, s, dtype, prec, 0, upto, d, chi, sites, nullptr, nullptr, out, &env));
  TENET_HIP_CHECK(hipMemcpyAsync(env_out, env, (size_t)chi[upto] * chi[upto] * es, hipMemcpyDeviceToDevice, s));
  return TENET_OK;
}

// ---------------------------------------------------------------------------
// grouped launches
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// grouped contractions: plan (tables, tile shapes, K chunks, unit queue) once,
// launch many times
// ---------------------------------------------------------------------------
struct tenet_gplan_s {
  tenet_ctx ctx = nullptr;
  int dtype = 0, prec = 0;
  int64_t count = 0, nunits = 0, grid = 0;
  void* dev = nullptr;  // one allocation: tables, problems, units, counters, partials, pointer staging
  GroupedProblem* probs = nullptr;
  GroupedUnit* units = nullptr;
  uint32_t* counters = nullptr;
  void* partials = nullptr;
  // operand pointers of tenet_grouped_exec: a ring of device tables filled
  // from pinned host slots, each slot reused once the event recorded after
  // its last launch has completed, so an exec returns without waiting
  static constexpr int RING = 8;
  uint64_t* ptrs_dev = nullptr;  // RING tables of 3 * count
  uint64_t* ptrs_pinned = nullptr;
  hipEvent_t ring_ev[RING] = {};
  bool ring_used[RING] = {};
  int ring_next = 0;
  std::vector<int64_t> spanA, spanB, spanC;
  double flops = 0.0, bytes = 0.0;
  int64_t shape_count[4] = {0, 0, 0, 0};  // 64, 32, 16 tiles; GROUPED_SMALLK
  int64_t split_tiles = 0;
  // Problems that are plain column-major GEMMs (every operand a matrix with
  // unit stride along rows or columns, no batch label), grouped by shape:
  // tenet_grouped_exec runs each group as one batched library GEMM (rocBLAS)
  // and the persistent kernel takes the other problems (units_rest).  The
  // device-table exec (capturable) keeps every problem in the kernel.
  struct LibGroup {
    rocblas_operation opA, opB;
    int64_t M, N, K, lda, ldb, ldc;
    std::vector<int64_t> idx;
    int64_t arr_off;  // first pointer of its A array after the 3 * count table (B, C follow)
  };
  std::vector<LibGroup> lib;
  int64_t nlib = 0;
  GroupedUnit* units_rest = nullptr;
  int64_t nunits_rest = 0;
  double flops_rest = 0.0, bytes_rest = 0.0;
  int64_t slot_elems = 0;  // pointers per ring slot: 3 * count + 3 * nlib
};

// A grouped problem as a plain GEMM (no batch label), with its conjugation.
static bool plain_gemm(const HostPlan& hp, bool conjA, bool conjB, rocblas_operation* opA, int64_t* lda,
                       rocblas_operation* opB, int64_t* ldb, int64_t* ldc) {
  if (hp.batch != 1) return false;
  const LibLayout L = plain_layout(hp);
  if (!L.ok || !lib_op(L.opA, conjA, opA) || !lib_op(L.opB, conjB, opB)) return false;
  *lda = L.lda;
  *ldb = L.ldb;
  *ldc = L.ldc;
  return true;
}

static const int TILE_EDGE[3] = {64, 32, 16};

// Relative cost of one (tile, 16-wide K step) per tile shape.  Measured by
// tenet_autotune_tiles; until then a model: the three shapes run 64, 16 and
// 4 MFMA blocks per K step on four waves, with a fixed per-step overhead.
static void tile_costs(tenet_ctx ctx, int dtype, double c[3]) {
  const int d = dtype == TENET_C128 ? 0 : 1;
  if (ctx->tile_cost[d][0] > 0) {
    for (int i = 0; i < 3; ++i) c[i] = ctx->tile_cost[d][i];
    return;
  }
  c[0] = 1.0;
  c[1] = 0.40;
  c[2] = 0.25;
}

static int grouped_plan_build(tenet_ctx ctx, int dtype, int prec, int64_t count, const tenet_tdesc* descs,
                              int force_shape, tenet_gplan g, const int32_t* conj = nullptr) {
  g->ctx = ctx;
  g->dtype = dtype;
  g->prec = prec;
  g->count = count;
  const size_t es = ELEM[dtype];
  std::vector<HostPlan> hps((size_t)count);
  std::vector<GroupedProblem> probs((size_t)count);
  g->spanA.resize((size_t)count);
  g->spanB.resize((size_t)count);
  g->spanC.resize((size_t)count);
  double c[3];
  tile_costs(ctx, dtype, c);
  size_t table_elems = 0;
  double total_cost = 0.0;
  std::vector<double> tile_step_cost((size_t)count);
  for (int64_t i = 0; i < count; ++i) {
    HostPlan& hp = hps[(size_t)i];
    TRY(analyse_plan(&descs[3 * i], &descs[3 * i + 1], &descs[3 * i + 2], hp));
    for (auto& v : hp.v) table_elems += v.size();
    GroupedProblem& q = probs[(size_t)i];
    q.M = hp.M;
    q.N = hp.N;
    q.K = hp.K;
    q.batch = hp.batch;
    q.a_mfast = hp.a_mfast;
    q.b_nfast = hp.b_nfast;
    q.conjA = conj ? (conj[2 * i] != 0) : 0;
    q.conjB = conj ? (conj[2 * i + 1] != 0) : 0;
    g->spanA[(size_t)i] = hp.spanA;
    g->spanB[(size_t)i] = hp.spanB;
    g->spanC[(size_t)i] = hp.spanC;
    g->flops += 8.0 * hp.M * hp.N * hp.K * hp.batch;
    g->bytes += (double)es * ((double)hp.M * hp.K + (double)hp.K * hp.N + (double)hp.M * hp.N) * hp.batch;
    if (force_shape < 0 && hp.K <= 8 && std::min(hp.M, hp.N) < 16) {
      // a tiny contracted range and a thin output: one thread per output
      // element, the element order following C's unit-stride class
      q.shape = GROUPED_SMALLK;
      g->shape_count[GROUPED_SMALLK] += 1;
      const int64_t total = hp.M * hp.N * hp.batch;
      q.tiles_m = (total + SMALLK_ELEMS - 1) / SMALLK_ELEMS;
      q.tiles_n = 1;
      auto stride_of = [](const std::vector<int64_t>& v) -> int64_t {
        return v.size() > 1 ? std::llabs(v[1] - v[0]) : INT64_MAX;
      };
      const int64_t sm = hp.M > 1 ? stride_of(hp.v[6]) : INT64_MAX;
      const int64_t sn = hp.N > 1 ? stride_of(hp.v[7]) : INT64_MAX;
      const int64_t sz = hp.batch > 1 ? stride_of(hp.v[8]) : INT64_MAX;
      q.order = (sm <= sn && sm <= sz) ? 0 : (sn <= sz ? 1 : 2);
      tile_step_cost[(size_t)i] = c[0];
      if (total > 0) total_cost += c[0] * (double)q.tiles_m;
      continue;
    }
    q.order = 0;
    // tile shape with the least modelled time for this problem
    const int64_t ksteps = std::max<int64_t>(1, (hp.K + 15) / 16);
    int best = 0;
    double best_t = 0.0;
    for (int v = 0; v < 3; ++v) {
      if (force_shape >= 0 && v != force_shape) continue;
      const int64_t e = TILE_EDGE[v];
      const double t = c[v] * (double)((hp.M + e - 1) / e) * (double)((hp.N + e - 1) / e) * hp.batch * ksteps;
      if (force_shape >= 0 || t < best_t || best_t == 0.0) {
        best = v;
        best_t = t;
      }
    }
    q.shape = best;
    g->shape_count[best] += 1;
    q.tiles_m = (hp.M + TILE_EDGE[best] - 1) / TILE_EDGE[best];
    q.tiles_n = (hp.N + TILE_EDGE[best] - 1) / TILE_EDGE[best];
    tile_step_cost[(size_t)i] = c[best];
    if (hp.M * hp.N * hp.batch > 0) total_cost += best_t;
  }
  // plain GEMMs grouped by shape; a group with at least LIB_MIN_FLOPS of work
  // goes to the library
  std::vector<char> in_lib((size_t)count, 0);
  if (force_shape < 0 && prec == TENET_PREC_NATIVE) {
    std::map<std::tuple<int, int, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t>, std::vector<int64_t>> groups;
    for (int64_t i = 0; i < count; ++i) {
      const GroupedProblem& q = probs[(size_t)i];
      if (q.shape == GROUPED_SMALLK) continue;
      rocblas_operation oa, ob;
      int64_t lda, ldb, ldc;
      if (!plain_gemm(hps[(size_t)i], q.conjA != 0, q.conjB != 0, &oa, &lda, &ob, &ldb, &ldc)) continue;
      groups[std::make_tuple((int)oa, (int)ob, q.M, q.N, q.K, lda, ldb, ldc)].push_back(i);
    }
    for (auto& kv : groups) {
      const auto& key = kv.first;
      const double fl = 8.0 * std::get<2>(key) * std::get<3>(key) * std::get<4>(key) * (double)kv.second.size();
      if (fl < LIB_MIN_FLOPS || std::get<2>(key) < 16 || std::get<3>(key) < 16 || std::get<4>(key) < 16) continue;
      tenet_gplan_s::LibGroup lg;
      lg.opA = (rocblas_operation)std::get<0>(key);
      lg.opB = (rocblas_operation)std::get<1>(key);
      lg.M = std::get<2>(key);
      lg.N = std::get<3>(key);
      lg.K = std::get<4>(key);
      lg.lda = std::get<5>(key);
      lg.ldb = std::get<6>(key);
      lg.ldc = std::get<7>(key);
      lg.idx = kv.second;
      lg.arr_off = 3 * count + 3 * g->nlib;
      g->nlib += (int64_t)lg.idx.size();
      for (int64_t i : lg.idx) in_lib[(size_t)i] = 1;
      g->lib.push_back(std::move(lg));
    }
  }
  g->slot_elems = 3 * count + 3 * g->nlib;
  for (int64_t i = 0; i < count; ++i) {
    if (in_lib[(size_t)i]) continue;
    const HostPlan& hp = hps[(size_t)i];
    g->flops_rest += 8.0 * hp.M * hp.N * hp.K * hp.batch;
    g->bytes_rest += (double)es * ((double)hp.M * hp.K + (double)hp.K * hp.N + (double)hp.M * hp.N) * hp.batch;
  }
  // persistent grid: resident workgroups on every CU
  int dev = 0, ncu = 256;
  TENET_HIP_CHECK(hipGetDevice(&dev));
  TENET_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int occ = grouped_occupancy(dtype);
  g->grid = (int64_t)ncu * std::max(1, occ);
  // K chunking: no unit above ~1/4 of one workgroup's fair share (with four
  // or more tiles per workgroup the LPT queue balances them whole; chunks
  // cost partial sums through memory)
  const double unit_cap = std::max(total_cost / ((double)g->grid * 4.0), 16.0 * c[0]);
  int64_t part_elems = 0, ctr_elems = 0, nunits = 0;
  for (int64_t i = 0; i < count; ++i) {
    GroupedProblem& q = probs[(size_t)i];
    if (q.shape == GROUPED_SMALLK) {
      q.kchunks = 1;
      q.kc = q.K;
      q.part_base = -1;
      q.ctr_base = -1;
      nunits += q.tiles_m;
      continue;
    }
    const int64_t ksteps = std::max<int64_t>(1, (q.K + 15) / 16);
    const double per_tile = tile_step_cost[(size_t)i] * ksteps;
    int64_t ch = 1;
    if (per_tile > 1.5 * unit_cap) ch = std::min<int64_t>((int64_t)std::ceil(per_tile / unit_cap), 256);
    ch = std::min<int64_t>(ch, ksteps);
    const int64_t kcs = (ksteps + ch - 1) / ch;  // K steps per chunk
    ch = (ksteps + kcs - 1) / kcs;
    q.kchunks = (int32_t)ch;
    q.kc = kcs * 16;
    const int64_t tiles = q.tiles_m * q.tiles_n * q.batch;
    q.part_base = -1;
    q.ctr_base = -1;
    if (ch > 1) {
      const int64_t e = TILE_EDGE[q.shape];
      q.part_base = part_elems;
      part_elems += tiles * ch * e * e;
      q.ctr_base = ctr_elems;
      ctr_elems += tiles;
      g->split_tiles += tiles;
    }
    nunits += tiles * ch;
  }
  CHECK_ARG(nunits < (int64_t)0xffffffffLL, "too many work units for one grouped launch");
  g->nunits = nunits;
  // unit queue, most expensive first (chunks of one tile stay adjacent)
  std::vector<GroupedUnit> units;
  units.reserve((size_t)nunits);
  std::vector<std::pair<double, int64_t>> order;
  order.reserve((size_t)count);
  for (int64_t i = 0; i < count; ++i) {
    const GroupedProblem& q = probs[(size_t)i];
    order.push_back({tile_step_cost[(size_t)i] * (double)((q.K + 15) / 16) / q.kchunks, i});
  }
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<double, int64_t>& x, const std::pair<double, int64_t>& y) { return x.first > y.first; });
  std::vector<GroupedUnit> units_rest;
  for (auto& o : order) {
    const GroupedProblem& q = probs[(size_t)o.second];
    const int64_t tiles = q.shape == GROUPED_SMALLK ? q.tiles_m : q.tiles_m * q.tiles_n * q.batch;
    for (int64_t t = 0; t < tiles; ++t)
      for (int32_t ch = 0; ch < q.kchunks; ++ch) {
        units.push_back({(int32_t)o.second, ch, t});
        if (!in_lib[(size_t)o.second]) units_rest.push_back({(int32_t)o.second, ch, t});
      }
  }
  g->nunits_rest = (int64_t)units_rest.size();
  // one device allocation
  const size_t rs = dtype == TENET_C128 ? 8 : 4;  // partial accumulators in the MFMA precision
  const size_t b_tab = align_up(std::max<size_t>(table_elems, 1) * 8, 256);
  const size_t b_prob = align_up((size_t)count * sizeof(GroupedProblem), 256);
  const size_t b_unit = align_up(std::max<size_t>(units.size(), 1) * sizeof(GroupedUnit), 256);
  const size_t b_unit_rest = g->nlib ? align_up(std::max<size_t>(units_rest.size(), 1) * sizeof(GroupedUnit), 256) : 0;
  const size_t b_ctr = align_up((size_t)(64 + ctr_elems) * 4, 256);
  const size_t b_part = align_up((size_t)std::max<int64_t>(part_elems, 1) * 2 * rs, 256);
  const size_t b_ptr = align_up((size_t)g->slot_elems * 8, 256) * tenet_gplan_s::RING;
  // from the context's caching pool: plans are built per batch shape, and a
  // device allocation per plan would cost more than the plan's analysis
  TRY(tenet_malloc(ctx, b_tab + b_prob + b_unit + b_ctr + b_part + b_ptr + b_unit_rest, &g->dev));
  char* base = (char*)g->dev;
  int64_t* tables = (int64_t*)base;
  g->probs = (GroupedProblem*)(base + b_tab);
  g->units = (GroupedUnit*)(base + b_tab + b_prob);
  g->counters = (uint32_t*)(base + b_tab + b_prob + b_unit);
  g->partials = (void*)(base + b_tab + b_prob + b_unit + b_ctr);
  g->ptrs_dev = (uint64_t*)(base + b_tab + b_prob + b_unit + b_ctr + b_part);
  g->units_rest = g->nlib ? (GroupedUnit*)(base + b_tab + b_prob + b_unit + b_ctr + b_part + b_ptr) : g->units;
  if (!g->nlib) {
    g->nunits_rest = g->nunits;
    g->flops_rest = g->flops;
    g->bytes_rest = g->bytes;
  }
  std::vector<int64_t> host;
  host.reserve(table_elems);
  for (int64_t i = 0; i < count; ++i) {
    size_t off[9];
    for (int k = 0; k < 9; ++k) {
      off[k] = host.size();
      auto& v = hps[(size_t)i].v[k];
      host.insert(host.end(), v.begin(), v.end());
      std::vector<int64_t>().swap(v);
    }
    set_tables(probs[(size_t)i].t, tables, off);
  }
  if (!host.empty())
    TENET_HIP_CHECK(hipMemcpy(tables, host.data(), host.size() * 8, hipMemcpyHostToDevice));
  TENET_HIP_CHECK(hipMemcpy(g->probs, probs.data(), probs.size() * sizeof(GroupedProblem), hipMemcpyHostToDevice));
  if (!units.empty())
    TENET_HIP_CHECK(hipMemcpy(g->units, units.data(), units.size() * sizeof(GroupedUnit), hipMemcpyHostToDevice));
  if (g->nlib && !units_rest.empty())
    TENET_HIP_CHECK(hipMemcpy(g->units_rest, units_rest.data(), units_rest.size() * sizeof(GroupedUnit),
                              hipMemcpyHostToDevice));
  TENET_HIP_CHECK(hipMemset(g->counters, 0, b_ctr));
  return TENET_OK;
}

extern "C" int tenet_grouped_plan(tenet_ctx ctx, int dtype, int prec, int64_t count, const tenet_tdesc* descs,
                                  tenet_gplan* plan) {
  return tenet_grouped_plan_conj(ctx, dtype, prec, count, descs, nullptr, plan);
}

extern "C" int tenet_grouped_plan_conj(tenet_ctx ctx, int dtype, int prec, int64_t count, const tenet_tdesc* descs,
                                       const int32_t* conj, tenet_gplan* plan) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, prec));
  CHECK_ARG(plan && count >= 0 && (count == 0 || descs), "bad argument");
  CHECK_ARG(prec == TENET_PREC_NATIVE || (prec == TENET_PREC_FP32 && dtype == TENET_C64),
            "grouped contractions run in the dtype's native precision");
  auto* g = new tenet_gplan_s();
  const int st = grouped_plan_build(ctx, dtype, prec, count, descs, -1, g, conj);
  if (st != TENET_OK) {
    if (g->dev) (void)tenet_free(ctx, g->dev);
    delete g;
    return st;
  }
  *plan = g;
  return TENET_OK;
}

extern "C" int tenet_grouped_plan_info(tenet_gplan plan, int64_t* units, int64_t* split_tiles, int64_t shapes[4],
                                       double* flops, double* bytes) {
  CHECK_ARG(plan, "null plan");
  if (units) *units = plan->nunits;
  if (split_tiles) *split_tiles = plan->split_tiles;
  if (shapes)
    for (int i = 0; i < 4; ++i) shapes[i] = plan->shape_count[i];
  if (flops) *flops = plan->flops;
  if (bytes) *bytes = plan->bytes;
  return TENET_OK;
}

static int grouped_launch(tenet_gplan g, hipStream_t s, const uint64_t* ptrs, bool rest = false) {
  GroupedLaunch L{};
  L.probs = g->probs;
  L.units = rest ? g->units_rest : g->units;
  L.nunits = rest ? g->nunits_rest : g->nunits;
  if (L.nunits == 0) return TENET_OK;
  L.ptrs = ptrs;
  L.partials = g->partials;
  L.counters = g->counters;
  L.grid = g->grid;
  L.dtype = g->dtype;
  L.prec = g->prec;
  StatScope st(g->ctx, s, TENET_KSTAT_GEMM, rest ? g->flops_rest : g->flops, rest ? g->bytes_rest : g->bytes);
  return launch_grouped_gemm(L, s);
}

// The plan's library groups.  A group whose A, B and C pointers (host array
// ptrs) are each evenly spaced runs as one strided-batched GEMM (the tuned
// library path); otherwise from its pointer arrays at dev (the ring slot).
static int lib_launch(tenet_gplan g, hipStream_t s, void* const* ptrs, const uint64_t* dev) {
  tenet_ctx ctx = g->ctx;
  std::lock_guard<std::mutex> lk(ctx->blas_mu);
  TRY(blas_bind(ctx, s));
  const int64_t es = (int64_t)ELEM[g->dtype];
  for (const auto& lg : g->lib) {
    const int n = (int)lg.idx.size();
    const uint64_t* a = dev + lg.arr_off;
    // element strides between consecutive problems of the group, if even
    int64_t st3[3] = {0, 0, 0};
    bool even = true;
    for (int w = 0; w < 3 && even; ++w) {
      const intptr_t p0 = (intptr_t)ptrs[3 * lg.idx[0] + w];
      const intptr_t d = n > 1 ? (intptr_t)ptrs[3 * lg.idx[1] + w] - p0 : 0;
      even = d >= 0 && d % es == 0;
      for (int j = 2; j < n && even; ++j) even = (intptr_t)ptrs[3 * lg.idx[j] + w] - p0 == (intptr_t)j * d;
      st3[w] = d / es;
    }
    rocblas_status st;
    if (even) {
      void* A0 = ptrs[3 * lg.idx[0]];
      void* B0 = ptrs[3 * lg.idx[0] + 1];
      void* C0 = ptrs[3 * lg.idx[0] + 2];
      if (g->dtype == TENET_C128) {
        const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
        st = rocblas_zgemm_strided_batched(ctx->blas, lg.opA, lg.opB, (rocblas_int)lg.M, (rocblas_int)lg.N,
                                           (rocblas_int)lg.K, &one, (const rocblas_double_complex*)A0,
                                           (rocblas_int)lg.lda, st3[0], (const rocblas_double_complex*)B0,
                                           (rocblas_int)lg.ldb, st3[1], &zero, (rocblas_double_complex*)C0,
                                           (rocblas_int)lg.ldc, st3[2], n);
      } else {
        const rocblas_float_complex one = {1.0f, 0.0f}, zero = {0.0f, 0.0f};
        st = rocblas_cgemm_strided_batched(ctx->blas, lg.opA, lg.opB, (rocblas_int)lg.M, (rocblas_int)lg.N,
                                           (rocblas_int)lg.K, &one, (const rocblas_float_complex*)A0,
                                           (rocblas_int)lg.lda, st3[0], (const rocblas_float_complex*)B0,
                                           (rocblas_int)lg.ldb, st3[1], &zero, (rocblas_float_complex*)C0,
                                           (rocblas_int)lg.ldc, st3[2], n);
      }
    } else if (g->dtype == TENET_C128) {
      const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
      st = rocblas_zgemm_batched(ctx->blas, lg.opA, lg.opB, (rocblas_int)lg.M, (rocblas_int)lg.N, (rocblas_int)lg.K,
                                 &one, (const rocblas_double_complex* const*)a, (rocblas_int)lg.lda,
                                 (const rocblas_double_complex* const*)(a + n), (rocblas_int)lg.ldb, &zero,
                                 (rocblas_double_complex* const*)(a + 2 * n), (rocblas_int)lg.ldc, n);
    } else {
      const rocblas_float_complex one = {1.0f, 0.0f}, zero = {0.0f, 0.0f};
      st = rocblas_cgemm_batched(ctx->blas, lg.opA, lg.opB, (rocblas_int)lg.M, (rocblas_int)lg.N, (rocblas_int)lg.K,
                                 &one, (const rocblas_float_complex* const*)a, (rocblas_int)lg.lda,
                                 (const rocblas_float_complex* const*)(a + n), (rocblas_int)lg.ldb, &zero,
                                 (rocblas_float_complex* const*)(a + 2 * n), (rocblas_int)lg.ldc, n);
    }
    if (st != rocblas_status_success) {
      set_error(std::string("rocBLAS batched GEMM failed: ") + rocblas_status_to_string(st));
      return TENET_ERR_HIP;
    }
  }
  return TENET_OK;
}

extern "C" int tenet_grouped_exec(tenet_gplan plan, tenet_stream stream, void* const* ptrs) {
  CHECK_ARG(plan, "null plan");
  tenet_ctx ctx = plan->ctx;
  TRY(use_device(ctx));
  if (plan->count == 0) return TENET_OK;
  CHECK_ARG(ptrs, "null pointer array");
  hipStream_t s = pick(ctx, stream);
  CHECK_ARG(!capturing(s), "tenet_grouped_exec copies host pointers: capture tenet_grouped_exec_dev instead");
  const size_t es = ELEM[plan->dtype];
  for (int64_t i = 0; i < plan->count; ++i) {
    if (plan->spanC[(size_t)i] == 0) continue;
    TRY(check_span(ctx, ptrs[3 * i + 2], (size_t)plan->spanC[(size_t)i] * es, "C"));
    if (plan->spanA[(size_t)i]) TRY(check_span(ctx, ptrs[3 * i], (size_t)plan->spanA[(size_t)i] * es, "A"));
    if (plan->spanB[(size_t)i]) TRY(check_span(ctx, ptrs[3 * i + 1], (size_t)plan->spanB[(size_t)i] * es, "B"));
  }
  // pointers travel through the next ring slot (pinned host, then device:
  // the copy is asynchronous); the host array may be reused on return
  const size_t nb = (size_t)plan->slot_elems * 8;
  if (!plan->ptrs_pinned) {
    TENET_HIP_CHECK(hipHostMalloc((void**)&plan->ptrs_pinned, nb * tenet_gplan_s::RING, hipHostMallocDefault));
    for (int i = 0; i < tenet_gplan_s::RING; ++i)
      TENET_HIP_CHECK(hipEventCreateWithFlags(&plan->ring_ev[i], hipEventDisableTiming));
  }
  const int slot = plan->ring_next;
  plan->ring_next = (slot + 1) % tenet_gplan_s::RING;
  if (plan->ring_used[slot]) TENET_HIP_CHECK(hipEventSynchronize(plan->ring_ev[slot]));
  uint64_t* host = plan->ptrs_pinned + (size_t)slot * plan->slot_elems;
  uint64_t* dev = (uint64_t*)((char*)plan->ptrs_dev + align_up(nb, 256) * slot);
  std::memcpy(host, ptrs, (size_t)plan->count * 3 * 8);
  // the library groups' A, B and C pointer arrays follow the table
  for (const auto& lg : plan->lib) {
    const size_t n = lg.idx.size();
    for (size_t j = 0; j < n; ++j) {
      const int64_t i = lg.idx[j];
      host[lg.arr_off + j] = (uint64_t)(uintptr_t)ptrs[3 * i];
      host[lg.arr_off + n + j] = (uint64_t)(uintptr_t)ptrs[3 * i + 1];
      host[lg.arr_off + 2 * n + j] = (uint64_t)(uintptr_t)ptrs[3 * i + 2];
    }
  }
  TENET_HIP_CHECK(hipMemcpyAsync(dev, host, nb, hipMemcpyHostToDevice, s));
  if (!plan->lib.empty()) TRY(lib_launch(plan, s, ptrs, dev));
  TRY(grouped_launch(plan, s, dev, true));
  TENET_HIP_CHECK(hipEventRecord(plan->ring_ev[slot], s));
  plan->ring_used[slot] = true;
  return TENET_OK;
}

extern "C" int tenet_grouped_exec_dev(tenet_gplan plan, tenet_stream stream, const uint64_t* ptrs_dev) {
  CHECK_ARG(plan, "null plan");
  tenet_ctx ctx = plan->ctx;
  TRY(use_device(ctx));
  if (plan->count == 0) return TENET_OK;
  CHECK_ARG(ptrs_dev, "null pointer array");
  hipStream_t s = pick(ctx, stream);
  TRY(check_span(ctx, ptrs_dev, (size_t)plan->count * 3 * 8, "pointer array"));
  // the launch reads the caller's table (it stays valid until the launch has run)
  return grouped_launch(plan, s, ptrs_dev);
}

extern "C" int tenet_grouped_plan_destroy(tenet_gplan plan) {
  if (!plan) return TENET_OK;
  (void)hipSetDevice(plan->ctx->device);
  (void)hipDeviceSynchronize();  // no launch still reads the arena when the pool hands it out again
  if (plan->dev) (void)tenet_free(plan->ctx, plan->dev);
  if (plan->ptrs_pinned) {
    (void)hipHostFree(plan->ptrs_pinned);
    for (int i = 0; i < tenet_gplan_s::RING; ++i) (void)hipEventDestroy(plan->ring_ev[i]);
  }
  delete plan;
  return TENET_OK;
}

extern "C" int tenet_grouped_einsum(tenet_ctx ctx, tenet_stream stream, int dtype, int prec, int64_t count,
                                    const tenet_tdesc* descs, void* const* ptrs) {
  if (count == 0) return TENET_OK;
  tenet_gplan g = nullptr;
  TRY(tenet_grouped_plan(ctx, dtype, prec, count, descs, &g));
  const int st = tenet_grouped_exec(g, stream, ptrs);
  (void)tenet_grouped_plan_destroy(g);
  return st;
}

// Autotuner: time each tile shape on a fixed synthetic batch (every problem
// (256 x 1024) . (1024 x 256), inputs from the Gaussian generator) and keep
// the measured cost per (tile, K step) for the shape choice of later plans.
extern "C" int tenet_autotune_tiles(tenet_ctx ctx, int dtype, double costs_out[3]) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  hipStream_t s = ctx->stream;
  const int64_t M = 256, N = 256, K = 1024, P = 64;
  const size_t es = ELEM[dtype];
  void* buf = nullptr;
  TENET_HIP_CHECK(hipMalloc(&buf, (size_t)P * (M * K + K * N + M * N) * es));
  int st = launch_gaussian(buf, dtype, P * (M * K + K * N), 12345, 0, s);
  std::vector<tenet_tdesc> descs;
  std::vector<void*> ptrs;
  for (int64_t p = 0; p < P; ++p) {
    descs.push_back(mkdesc({M, K}, {1, 2}));
    descs.push_back(mkdesc({K, N}, {2, 3}));
    descs.push_back(mkdesc({M, N}, {1, 3}));
    char* b = (char*)buf;
    ptrs.push_back(b + (size_t)p * M * K * es);
    ptrs.push_back(b + ((size_t)P * M * K + (size_t)p * K * N) * es);
    ptrs.push_back(b + ((size_t)P * (M * K + K * N) + (size_t)p * M * N) * es);
  }
  double cost[3] = {0, 0, 0};
  for (int v = 0; v < 3 && st == TENET_OK; ++v) {
    tenet_gplan_s g;
    st = grouped_plan_build(ctx, dtype, TENET_PREC_NATIVE, P, descs.data(), v, &g);
    if (st == TENET_OK) {
      (void)hipMemcpy(g.ptrs_dev, ptrs.data(), ptrs.size() * 8, hipMemcpyHostToDevice);
      float best = 1e30f;
      for (int rep = 0; rep < 4 && st == TENET_OK; ++rep) {
        (void)hipEventRecord(ctx->ev0, s);
        st = grouped_launch(&g, s, g.ptrs_dev);
        (void)hipEventRecord(ctx->ev1, s);
        (void)hipEventSynchronize(ctx->ev1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
        if (rep > 0) best = std::min(best, ms);
      }
      const int64_t e = TILE_EDGE[v];
      cost[v] = (double)best / ((double)P * ((M + e - 1) / e) * ((N + e - 1) / e) * (K / 16));
    }
    (void)hipStreamSynchronize(s);
    if (g.dev) (void)tenet_free(ctx, g.dev);
  }
  (void)hipFree(buf);
  TRY(st);
  const int d = dtype == TENET_C128 ? 0 : 1;
  for (int v = 0; v < 3; ++v) ctx->tile_cost[d][v] = cost[v] / cost[0];
  if (costs_out)
    for (int v = 0; v < 3; ++v) costs_out[v] = ctx->tile_cost[d][v];
  return TENET_OK;
}

// A grouped SVD bound to its operands: shapes, the problem table and every
// problem's Jacobi workspace live in one pool block uploaded at build time,
// so a run is kernel launches only (graph-capturable).
struct tenet_gsvd_s {
  tenet_ctx ctx = nullptr;
  int dtype = 0;
  int64_t count = 0;
  std::vector<SvdShape> shapes;
  std::vector<SvdProblem> probs;  // host copy of the device table
  void* dev = nullptr;
  SvdProblem* dprobs = nullptr;
  int32_t* flags_base = nullptr;
  int64_t max_rounds = 0, max_pairs = 0, max_ncp = 0, max_elems = 0, max_out = 0;
  // per round (0: the fixed-pair round, 1 + r: cross round r) the
  // (problem, pair) units that have work in it, largest problems first
  std::vector<int32_t> units;  // pairs of int32
  std::vector<int64_t> uoff, ucnt;
  int2* units_dev = nullptr;
  // QR-preconditioned problems (pre set): their batched-QR table, sizes
  std::vector<QrBatchProb> qr_probs;
  QrBatchProb* qr_dev = nullptr;
  int64_t pre_max_m = 0, pre_max_n = 0, pre_max_elems = 0;
  // operands, for the single-matrix fallback of problems the sweeps leave unconverged
  std::vector<int64_t> ms, ns, maxdims;
  std::vector<const void*> As;
  std::vector<void*> Us, Vhs;
  std::vector<double*> Ss;
  double threshold = 0.0;
  int64_t* k_dev = nullptr;
  double* disc_dev = nullptr;
};

static int gsvd_build(tenet_ctx ctx, int dtype, int64_t count, const int64_t* ms, const int64_t* ns,
                      const void* const* As, void* const* Us, double* const* Ss, void* const* Vhs,
                      const int64_t* maxdims, double threshold, int64_t* k_dev, double* disc_dev, int s_mode,
                      tenet_gsvd g, hipStream_t s) {
  g->ctx = ctx;
  g->dtype = dtype;
  g->count = count;
  g->shapes.resize((size_t)count);
  g->threshold = threshold;
  g->k_dev = k_dev;
  g->disc_dev = disc_dev;
  size_t total = align_up((size_t)count * sizeof(SvdProblem), 256) + align_up((size_t)count * 32, 256) +
                 align_up((size_t)count * 8, 256);
  const size_t es = ELEM[dtype];
  const size_t s_es = s_mode == 0 ? 8 : (s_mode == 1 ? 16 : 8);
  // Without U, an m >= n problem with 64 <= n and m <= 512 is QR-preconditioned:
  // A[:, cperm] = Q R with the columns by descending norm, then the Jacobi on
  // X = R^H (n x n): V^H comes from X, Q is never formed and no rotation is
  // accumulated (Drmac and Veselic, SIAM J. Matrix Anal. Appl. 29 (2008))
  std::vector<char> pre((size_t)count, 0);
  for (int64_t i = 0; i < count; ++i) {
    pre[(size_t)i] = ctx->svd_precond && !Us[i] && ms[i] >= ns[i] && ns[i] >= 64 && ms[i] <= 512;
    if (pre[(size_t)i]) {
      g->pre_max_m = std::max(g->pre_max_m, ms[i]);
      g->pre_max_n = std::max(g->pre_max_n, ns[i]);
      g->pre_max_elems = std::max(g->pre_max_elems, ms[i] * ns[i]);
    }
  }
  for (int64_t i = 0; i < count; ++i) {
    CHECK_ARG(ms[i] > 0 && ns[i] > 0 && std::min(ms[i], ns[i]) <= 4096, "bad matrix shape");
    const int64_t m = ms[i], n = ns[i], k = std::min(m, n);
    CHECK_ARG(As[i] && Vhs[i], "null A or Vh pointer");
    TRY(check_span(ctx, As[i], (size_t)m * n * es, "A"));
    if (Us[i]) TRY(check_span(ctx, Us[i], (size_t)m * k * es, "U"));
    TRY(check_span(ctx, Vhs[i], (size_t)k * n * es, "Vh"));
    if (Ss && Ss[i]) TRY(check_span(ctx, Ss[i], (size_t)k * s_es, "S"));
    // Without U and with m <= n, the Jacobi runs on X = A^H: V^H comes from X
    // itself and the rotations need not be accumulated
    g->shapes[(size_t)i] = pre[(size_t)i] ? svd_shape(n, n, false, true) : svd_shape(m, n, false, !Us[i] && m <= n);
    const SvdShape& sh = g->shapes[(size_t)i];
    total += svd_ws_bytes(sh);
    if (pre[(size_t)i])
      total += align_up((size_t)m * n * 16, 256) + align_up((size_t)n * 16, 256) +
               align_up(qr_batched_t_bytes(m, n), 256) + align_up(qr_batched_yp_bytes(m, g->pre_max_n), 256) +
               align_up(qr_batched_tb_bytes(), 256) +
               align_up((size_t)n * 4, 256) + align_up(sizeof(QrBatchProb), 256);
    g->max_rounds = std::max(g->max_rounds, sh.nblk - 1);
    g->max_pairs = std::max(g->max_pairs, sh.npairs);
    g->max_ncp = std::max(g->max_ncp, sh.ncp);
    g->max_elems = std::max(g->max_elems, sh.nr * sh.ncp + sh.ncp * sh.ncp);
    g->max_out = std::max(g->max_out, (ms[i] + ns[i]) * sh.ncols);
    g->ms.push_back(m);
    g->ns.push_back(n);
    g->maxdims.push_back(maxdims ? maxdims[i] : 0);
    g->As.push_back(As[i]);
    g->Us.push_back(Us[i]);
    g->Vhs.push_back(Vhs[i]);
    g->Ss.push_back(Ss ? Ss[i] : nullptr);
  }
  // round unit lists: a problem of nblk blocks has npairs units in the
  // fixed-pair round and, when nblk > 2, in cross rounds 0 .. nblk - 2 (the
  // rounds launch only these units, not a pairs x problems rectangle)
  {
    std::vector<int64_t> order((size_t)count);
    for (int64_t i = 0; i < count; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
      return g->shapes[(size_t)x].nr * g->shapes[(size_t)x].ncp > g->shapes[(size_t)y].nr * g->shapes[(size_t)y].ncp;
    });
    g->uoff.assign((size_t)g->max_rounds + 1, 0);
    g->ucnt.assign((size_t)g->max_rounds + 1, 0);
    for (int64_t r = -1; r < g->max_rounds; ++r) {
      g->uoff[(size_t)(r + 1)] = (int64_t)g->units.size() / 2;
      for (int64_t i : order) {
        const SvdShape& sh = g->shapes[(size_t)i];
        if (r >= 0 && (sh.nblk <= 2 || r >= sh.nblk - 1)) continue;
        for (int64_t p = 0; p < sh.npairs; ++p) {
          g->units.push_back((int32_t)i);
          g->units.push_back((int32_t)p);
        }
      }
      g->ucnt[(size_t)(r + 1)] = (int64_t)g->units.size() / 2 - g->uoff[(size_t)(r + 1)];
    }
    total += align_up(std::max<size_t>(g->units.size(), 2) * 4, 256);
  }
  TRY(tenet_malloc(ctx, total, &g->dev));
  Bump bm(g->dev);
  g->dprobs = (SvdProblem*)bm.take((size_t)count * sizeof(SvdProblem));
  g->units_dev = (int2*)bm.take(std::max<size_t>(g->units.size(), 2) * 4);
  g->probs.resize((size_t)count);
  for (int64_t i = 0; i < count; ++i) {
    const SvdShape& sh = g->shapes[(size_t)i];
    SvdProblem& P = g->probs[(size_t)i];
    P.X = (double*)bm.take((size_t)sh.nr * sh.ncp * 16);
    P.V = (double*)bm.take((size_t)sh.ncp * sh.ncp * 16);
    P.Gpart = (double*)bm.take((size_t)sh.npairs * sh.q * (2 * sh.b) * (2 * sh.b) * 16);
    P.sigma = (double*)bm.take((size_t)sh.ncp * 8);
    P.perm = (int32_t*)bm.take((size_t)sh.ncp * 4);
    P.nr = (int32_t)sh.nr;
    P.ncols = (int32_t)sh.ncols;
    P.ncp = (int32_t)sh.ncp;
    P.b = (int32_t)sh.b;
    P.nblk = (int32_t)sh.nblk;
    P.q = 1;
    P.m = (int32_t)ms[i];
    P.n = (int32_t)ns[i];
    P.transposed = sh.transposed ? 1 : 0;
    P.accum = (Us[i] || !sh.transposed) ? 1 : 0;
    P.s_mode = s_mode;
    P.A = As[i];
    P.lda = ms[i];
    P.U = Us[i];
    P.ldu = ms[i];
    P.Vh = Vhs[i];
    P.ldvh = sh.ncols;
    P.S = Ss ? Ss[i] : nullptr;
    P.maxdim = maxdims ? maxdims[i] : 0;
    P.threshold = threshold;
    P.k_dev = k_dev ? k_dev + i : nullptr;
    P.disc_dev = disc_dev ? disc_dev + i : nullptr;
    P.tol = svd_tol(sh.nr);
    if (pre[(size_t)i]) {
      const int64_t m = ms[i], n = ns[i];
      P.pre = 1;
      P.accum = 0;
      P.ldvh = n;
      P.W = (double*)bm.take((size_t)m * n * 16);
      P.cperm = (int32_t*)bm.take((size_t)n * 4);
      QrBatchProb q{};
      q.W = P.W;
      q.tau = (double*)bm.take((size_t)n * 16);
      q.T = (double*)bm.take(qr_batched_t_bytes(m, n));
      q.Yp = (double*)bm.take(qr_batched_yp_bytes(m, g->pre_max_n));
      q.TB = (double*)bm.take(qr_batched_tb_bytes());
      q.m = (int32_t)m;
      q.n = (int32_t)n;
      g->qr_probs.push_back(q);
    }
  }
  if (!g->qr_probs.empty()) g->qr_dev = (QrBatchProb*)bm.take(g->qr_probs.size() * sizeof(QrBatchProb));
  // per problem 8 int32: rotated, converged, sweeps, spare, then the sweep's
  // largest squared relative off-diagonal (8 bytes at flag + 4)
  g->flags_base = (int32_t*)bm.take((size_t)count * 32);
  double* fro_base = (double*)bm.take((size_t)count * 8);
  for (int64_t i = 0; i < count; ++i) {
    g->probs[(size_t)i].flag = g->flags_base + 8 * i;
    g->probs[(size_t)i].fro2 = fro_base + i;
  }
  // ordered on the stream after earlier work (the pool block may have been
  // freed by a caller while that work still used it)
  TRY(upload_async(ctx, s, g->dprobs, g->probs.data(), g->probs.size() * sizeof(SvdProblem)));
  TRY(upload_async(ctx, s, g->units_dev, g->units.data(), g->units.size() * 4));
  TRY(upload_async(ctx, s, g->qr_dev, g->qr_probs.data(), g->qr_probs.size() * sizeof(QrBatchProb)));
  return TENET_OK;
}

// Pinned landing slot for a grouped run's convergence flags: two sweeps'
// copies in flight, each followed by an event.  Taken from the context's free
// list and handed back (after both copies have landed) when the run ends.
struct FlagSlotGuard {
  tenet_ctx ctx;
  tenet_ctx_s::FlagSlot* f = nullptr;
  explicit FlagSlotGuard(tenet_ctx c) : ctx(c) {}
  int take(size_t bytes) {
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      for (size_t i = 0; i < ctx->flag_slots.size(); ++i)
        if (ctx->flag_slots[i]->bytes >= bytes) {
          f = ctx->flag_slots[i];
          ctx->flag_slots.erase(ctx->flag_slots.begin() + (std::ptrdiff_t)i);
          return TENET_OK;
        }
    }
    auto* n = new tenet_ctx_s::FlagSlot();
    n->bytes = align_up(bytes, 4096);
    if (hipHostMalloc((void**)&n->host, n->bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      delete n;
      set_error("pinned host allocation failed");
      return TENET_ERR_OOM;
    }
    for (hipEvent_t& e : n->ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(n->host);
        for (hipEvent_t x : n->ev)
          if (x) (void)hipEventDestroy(x);
        delete n;
        set_error("event creation failed");
        return TENET_ERR_HIP;
      }
    f = n;
    return TENET_OK;
  }
  ~FlagSlotGuard() {
    if (!f) return;
    // no copy may still be landing in the slot when another run takes it
    for (hipEvent_t e : f->ev) (void)hipEventSynchronize(e);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->flag_slots.push_back(f);
  }
};

// One grouped run.  Eagerly: sweeps until every problem converges, then
// problems left unconverged go through the single-matrix path.  The flags of
// sweep s are copied to pinned memory behind it and read while the fixed-pair
// round of sweep s + 1 is already queued, so the stream does not idle on the
// host's convergence test; after the last sweep needed that one round is
// launched in vain (it returns at once for converged problems).  Returns
// without waiting for the final kernels.  Under capture: a fixed
// CAPTURED_SWEEPS sweeps and the context's device status bit for problems
// they did not finish.
static int gsvd_run(tenet_gsvd g, hipStream_t s) {
  tenet_ctx ctx = g->ctx;
  const int64_t count = g->count;
  const bool sync_ok = !capturing(s);
  SvdBatch B{};
  B.arr = g->dprobs;
  if (!g->qr_probs.empty()) {
    TRY(launch_svd_pre_pack(B, count, (int)g->pre_max_n, g->pre_max_elems, g->dtype, s));
    TRY(launch_qr_batched(g->qr_dev, (int64_t)g->qr_probs.size(), (int)g->pre_max_m, (int)g->pre_max_n,
                          (int)g->pre_max_n, s));
  }
  TRY(launch_svd_init(B, count, g->max_elems, (int)g->max_ncp, g->dtype, s));
  const size_t fbytes = (size_t)count * 32;
  FlagSlotGuard slot(ctx);
  if (sync_ok) TRY(slot.take(2 * fbytes));
  // converged flags as last read back (one sweep behind the launches)
  std::vector<int32_t> hflags((size_t)count * 8, 0);
  // nominal work of one sweep of problem i: the Gram of every pair over its
  // rows and the rotation of its X and V rows, in the fixed-pair round and,
  // when nblk > 2, in each cross round r < nblk - 1
  auto unit_cost = [&](int64_t i, double* f, double* b) {
    const SvdShape& sh = g->shapes[(size_t)i];
    const double n2 = 2.0 * sh.b, vr = g->probs[(size_t)i].accum ? (double)sh.ncp : 0.0;
    *f = 8.0 * sh.npairs * n2 * n2 * (2.0 * sh.nr + vr);
    *b = 16.0 * sh.npairs * n2 * (3.0 * sh.nr + 2.0 * vr);
  };
  // The work of the fixed-pair round (f0, b0) and of each cross round over
  // the problems not known to have converged
  double f0 = 0.0, b0 = 0.0;
  std::vector<double> flr((size_t)g->max_rounds, 0.0), byr((size_t)g->max_rounds, 0.0);
  auto tally = [&]() {
    f0 = b0 = 0.0;
    std::fill(flr.begin(), flr.end(), 0.0);
    std::fill(byr.begin(), byr.end(), 0.0);
    for (int64_t i = 0; i < count; ++i) {
      if (hflags[(size_t)(8 * i + 1)]) continue;
      const SvdShape& sh = g->shapes[(size_t)i];
      double f, b;
      unit_cost(i, &f, &b);
      f0 += f;
      b0 += b;
      if (sh.nblk > 2)
        for (int64_t r = 0; r < sh.nblk - 1; ++r) {
          flr[(size_t)r] += f;
          byr[(size_t)r] += b;
        }
    }
  };
  auto round0 = [&]() {
    SvdBatch Bu = B;
    Bu.units = g->units_dev + g->uoff[0];
    return svd_round_timed(ctx, s, f0, b0, Bu, count, (int)g->ucnt[0], 0, 2, 0);
  };
  // the cross rounds, the sweep's end and (eagerly) its flags copy
  auto rest = [&](int sw) -> int {
    SvdBatch Bu = B;
    for (int r = 0; r < g->max_rounds; ++r) {
      if (g->ucnt[(size_t)r + 1] == 0) continue;
      Bu.units = g->units_dev + g->uoff[(size_t)r + 1];
      TRY(svd_round_timed(ctx, s, flr[(size_t)r], byr[(size_t)r], Bu, count, (int)g->ucnt[(size_t)r + 1], r, 2, 1));
    }
    TRY(launch_svd_sweep_end(B, count, s));
    if (sync_ok) {
      TENET_HIP_CHECK(hipMemcpyAsync(slot.f->host + (size_t)(sw & 1) * count * 8, g->flags_base, fbytes,
                                     hipMemcpyDeviceToHost, s));
      TENET_HIP_CHECK(hipEventRecord(slot.f->ev[sw & 1], s));
    }
    return TENET_OK;
  };
  // Read the flags of sweep sw into hflags: 1 when every problem has
  // converged, else 0 (or an error).  *dfl / *dby: the fixed-pair round's
  // share of the problems that converged in sweep sw.
  auto read_flags = [&](int sw, double* dfl, double* dby) -> int {
    TENET_HIP_CHECK(hipEventSynchronize(slot.f->ev[sw & 1]));
    const int32_t* nf = slot.f->host + (size_t)(sw & 1) * count * 8;
    *dfl = *dby = 0.0;
    bool a = true;
    for (int64_t i = 0; i < count; ++i) {
      const bool was = hflags[(size_t)(8 * i + 1)] != 0, now = nf[8 * i + 1] != 0;
      if (!was && now) {
        double f, b;
        unit_cost(i, &f, &b);
        *dfl += f;
        *dby += b;
      }
      a = a && now;
    }
    std::memcpy(hflags.data(), nf, fbytes);
    return a ? 1 : 0;
  };
  bool all = false;
  const int max_sweeps = sync_ok ? 40 : CAPTURED_SWEEPS;
  tally();
  TRY(round0());
  TRY(rest(0));
  int sw = 1;
  for (; sw < max_sweeps; ++sw) {
    // the next sweep's fixed-pair round keeps the stream busy while the host
    // reads the previous sweep's flags
    TRY(round0());
    if (sync_ok) {
      double dfl, dby;
      const int a = read_flags(sw - 1, &dfl, &dby);
      if (a < 0) return a;
      // that round returned at once for problems that converged in sweep
      // sw - 1, though it counted them
      stats_retract(ctx, TENET_KSTAT_SVD_ROUND, s, dfl, dby);
      if (a == 1) {
        all = true;
        break;
      }
      tally();
    }
    TRY(rest(sw));
  }
  if (sync_ok && !all) {
    // the sweep limit ran out: the last sweep's flags decide
    double dfl, dby;
    const int a = read_flags(sw - 1, &dfl, &dby);
    if (a < 0) return a;
    all = a == 1;
  }
  TRY(launch_svd_finalize(B, count, (int)g->max_ncp, g->max_out, g->dtype, s));
  if (!sync_ok) return launch_svd_status(B, count, ctx->dev_status, s);
  if (!all) {
    // Problems the plain sweeps left unconverged (strongly graded spectra) go
    // one at a time through the single-matrix path, which preconditions
    // complex128 by a QR, from their untouched inputs (stream-ordered after
    // the finalisation above).
    for (int64_t i = 0; i < count; ++i) {
      if (hflags[(size_t)(8 * i + 1)]) continue;
      CHECK_ARG(g->probs[(size_t)i].s_mode == 0, "grouped SVD: unconverged problem with a complex S output");
      const int64_t m = g->ms[(size_t)i], n = g->ns[(size_t)i];
      void* w = nullptr;
      // the single-matrix path always forms U: scratch for a problem without one
      const size_t ub = g->Us[(size_t)i] ? 0 : align_up((size_t)m * std::min(m, n) * ELEM[g->dtype], 256);
      TRY(ws_get(ctx, s, svd_any_ws_bytes(ctx, g->dtype, m, n) + ub, &w));
      void* u = g->Us[(size_t)i] ? g->Us[(size_t)i] : (void*)((char*)w + svd_any_ws_bytes(ctx, g->dtype, m, n));
      // k_dev, disc_dev and S are optional, as on the grouped path above
      const int st = svd_any(ctx, s, g->dtype, m, n, g->As[(size_t)i], m, u, m, g->Ss[(size_t)i], g->Vhs[(size_t)i],
                             std::min(m, n), g->maxdims[(size_t)i], g->threshold, g->k_dev ? g->k_dev + i : nullptr,
                             g->disc_dev ? g->disc_dev + i : nullptr, w);
      if (st != TENET_OK) {
        if (st == TENET_ERR_NOCONV) set_error("grouped Jacobi SVD: some problems did not converge");
        return st;
      }
    }
  }
  return TENET_OK;
}

extern "C" int tenet_grouped_svd_plan(tenet_ctx ctx, int dtype, int64_t count, const int64_t* ms, const int64_t* ns,
                                      const void* const* As, void* const* Us, void* const* Ss, void* const* Vhs,
                                      const int64_t* maxdims, double threshold, int64_t* k_dev, double* disc_dev,
                                      int s_mode, tenet_gsvd* plan) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(plan && count > 0 && ms && ns && As && Us && Vhs, "bad argument");
  CHECK_ARG(s_mode == TENET_S_REAL || s_mode == TENET_S_DTYPE, "bad S mode");
  *plan = nullptr;
  if (capturing(ctx->stream)) {
    set_error("tenet_grouped_svd_plan: build the plan before the capture (it uploads its tables)");
    return TENET_ERR_UNSUPPORTED;
  }
  auto* g = new tenet_gsvd_s();
  const int sm = s_mode == TENET_S_REAL ? 0 : (dtype == TENET_C128 ? 1 : 2);
  const int st = gsvd_build(ctx, dtype, count, ms, ns, As, Us, reinterpret_cast<double* const*>(Ss), Vhs, maxdims,
                            threshold, k_dev, disc_dev, sm, g, ctx->stream);
  if (st == TENET_OK) TENET_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (st != TENET_OK) {
    if (g->dev) (void)tenet_free(ctx, g->dev);
    delete g;
    return st;
  }
  *plan = g;
  return TENET_OK;
}

extern "C" int tenet_grouped_svd_exec(tenet_gsvd plan, tenet_stream stream) {
  CHECK_ARG(plan, "null plan");
  TRY(use_device(plan->ctx));
  return gsvd_run(plan, pick(plan->ctx, stream));
}

extern "C" int tenet_grouped_svd_destroy(tenet_gsvd plan) {
  if (!plan) return TENET_OK;
  if (plan->dev) {
    (void)hipSetDevice(plan->ctx->device);
    (void)hipDeviceSynchronize();
    (void)tenet_free(plan->ctx, plan->dev);
  }
  delete plan;
  return TENET_OK;
}

extern "C" int tenet_grouped_svd(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, const int64_t* ms,
                                 const int64_t* ns, const void* const* As, void* const* Us, double* const* Ss,
                                 void* const* Vhs, const int64_t* maxdims, double threshold, int64_t* k_dev,
                                 double* disc_dev) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(count >= 0, "bad count");
  if (count == 0) return TENET_OK;
  CHECK_ARG(ms && ns && As && Us && Vhs, "null argument");
  hipStream_t s = pick(ctx, stream);
  if (capturing(s)) {
    set_error("tenet_grouped_svd cannot be captured (it allocates its plan): use tenet_grouped_svd_plan / exec");
    return TENET_ERR_UNSUPPORTED;
  }
  tenet_gsvd_s g;
  int st = gsvd_build(ctx, dtype, count, ms, ns, As, Us, Ss, Vhs, maxdims, threshold, k_dev, disc_dev, 0, &g, s);
  if (st == TENET_OK) st = gsvd_run(&g, s);
  // the run's last kernels may still be in flight: the block returns to the
  // pool once they have finished
  if (g.dev) (void)tenet_free_async(ctx, g.dev, (tenet_stream)s);
  return st;
}

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------
extern "C" int tenet_fill_gaussian(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, void* dst,
                                   uint64_t seed, uint64_t offset) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(count >= 0, "negative count");
  if (count == 0) return TENET_OK;
  TRY(check_span(ctx, dst, (size_t)count * ELEM[dtype], "dst"));
  return launch_gaussian(dst, dtype, count, seed, offset, pick(ctx, stream));
}

extern "C" int tenet_scale(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, void* x,
                           const double alpha[2]) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(count >= 0 && alpha, "bad argument");
  if (count == 0) return TENET_OK;
  TRY(check_span(ctx, x, (size_t)count * ELEM[dtype], "x"));
  return launch_scale(x, dtype, count, alpha[0], alpha[1], pick(ctx, stream));
}

extern "C" int tenet_sample_select(tenet_ctx ctx, tenet_stream stream, int dtype, const void* T, int d, int chir,
                                   int64_t nsamp, int n, int site, uint64_t seed, int32_t* bits, void* Vn) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(T && bits && d >= 1 && d <= 8 && chir >= 1 && nsamp >= 0 && n >= 1 && site >= 0 && site < n,
            "bad argument");
  if (nsamp == 0) return TENET_OK;
  TRY(check_span(ctx, T, (size_t)d * chir * nsamp * ELEM[dtype], "T"));
  TRY(check_span(ctx, bits, (size_t)nsamp * n * 4, "bits"));
  if (Vn) TRY(check_span(ctx, Vn, (size_t)chir * nsamp * ELEM[dtype], "Vn"));
  return launch_sample_select(T, dtype, d, chir, nsamp, n, site, seed, bits, Vn, pick(ctx, stream));
}

extern "C" int tenet_sigma_scales(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t k, const double* S,
                                  void* sqrt_out, void* inv_sqrt_out) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(k >= 0 && S && sqrt_out && inv_sqrt_out, "bad argument");
  if (k == 0) return TENET_OK;
  TRY(check_span(ctx, sqrt_out, (size_t)k * ELEM[dtype], "sqrt_out"));
  TRY(check_span(ctx, inv_sqrt_out, (size_t)k * ELEM[dtype], "inv_sqrt_out"));
  return launch_sigma_scales(S, k, sqrt_out, inv_sqrt_out, dtype, pick(ctx, stream));
}

extern "C" int tenet_grouped_gate(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, int d,
                                  const int64_t* chil, const int64_t* chir, const void* const* theta,
                                  const void* const* gates, const void* const* lams, void* const* theta_out,
                                  void* const* x_out) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(count >= 0, "bad count");
  if (count == 0) return TENET_OK;
  CHECK_ARG(chil && chir && theta && gates && theta_out, "null argument");
  CHECK_ARG(d >= 2 && d <= 4, "gate: physical dimension outside 2..4");
  const size_t es = ELEM[dtype];
  hipStream_t s = pick(ctx, stream);
  GateBatch b{};
  b.d = d;
  int64_t max_elems = 0;
  for (int64_t i0 = 0; i0 < count; i0 += GATE_BATCH) {
    b.count = (int32_t)std::min<int64_t>(GATE_BATCH, count - i0);
    max_elems = 0;
    for (int j = 0; j < b.count; ++j) {
      const int64_t i = i0 + j;
      CHECK_ARG(chil[i] > 0 && chir[i] > 0, "gate: bad bond dimension");
      const void* lam = lams ? lams[i] : nullptr;
      void* xo = x_out ? x_out[i] : nullptr;
      CHECK_ARG((lam == nullptr) == (xo == nullptr), "gate: Lambda and X come together");
      const size_t nb = (size_t)chil[i] * d * d * chir[i] * es;
      TRY(check_span(ctx, theta[i], nb, "theta"));
      TRY(check_span(ctx, gates[i], (size_t)d * d * d * d * es, "gate"));
      TRY(check_span(ctx, theta_out[i], nb, "theta_out"));
      const uintptr_t t0 = (uintptr_t)theta[i], o0 = (uintptr_t)theta_out[i];
      CHECK_ARG(o0 + nb <= t0 || t0 + nb <= o0, "gate: theta_out overlaps theta");
      if (xo) {
        TRY(check_span(ctx, xo, nb, "X"));
        TRY(check_span(ctx, lam, (size_t)chil[i] * es, "Lambda"));
        const uintptr_t x0 = (uintptr_t)xo;
        CHECK_ARG(x0 == t0 || x0 + nb <= t0 || t0 + nb <= x0, "gate: X partially overlaps theta");
        CHECK_ARG(x0 + nb <= o0 || o0 + nb <= x0, "gate: X overlaps theta_out");
      }
      b.p[j] = GateProb{theta[i], gates[i], lam, theta_out[i], xo, chil[i], chir[i]};
      max_elems = std::max(max_elems, chil[i] * chir[i]);
    }
    TRY(launch_gate_apply(b, max_elems, dtype, s));
  }
  return TENET_OK;
}

extern "C" int tenet_nrm2sq(tenet_ctx ctx, tenet_stream stream, int dtype, int64_t count, const void* x,
                            double* out) {
  TRY(use_device(ctx));
  TRY(valid_dtype_prec(dtype, 0));
  CHECK_ARG(count >= 0 && out, "bad argument");
  hipStream_t s = pick(ctx, stream);
  if (count == 0) return launch_fill_zero(out, sizeof(double), s);
  TRY(check_span(ctx, x, (size_t)count * ELEM[dtype], "x"));
  void* w = nullptr;
  TRY(ws_get(ctx, s, 1024 * sizeof(double) + 256, &w));
  return launch_nrm2(x, dtype, count, (double*)w, out, s);
}

// ---------------------------------------------------------------------------
// graphs and timers
// ---------------------------------------------------------------------------
extern "C" int tenet_graph_begin(tenet_ctx ctx, tenet_stream stream) {
  TRY(use_device(ctx));
  TENET_HIP_CHECK(hipStreamBeginCapture(pick(ctx, stream), hipStreamCaptureModeRelaxed));
  return TENET_OK;
}
extern "C" int tenet_graph_end(tenet_ctx ctx, tenet_stream stream, tenet_graph* graph) {
  TRY(use_device(ctx));
  CHECK_ARG(graph, "bad argument");
  auto* g = new tenet_graph_s();
  g->device = ctx->device;
  g->stream = pick(ctx, stream);
  hipError_t e = hipStreamEndCapture(g->stream, &g->graph);
  if (e != hipSuccess) {
    delete g;
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return TENET_ERR_HIP;
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    delete g;
    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return TENET_ERR_HIP;
  }
  *graph = g;
  return TENET_OK;
}
extern "C" int tenet_graph_launch(tenet_graph graph, tenet_stream stream) {
  CHECK_ARG(graph, "bad argument");
  TENET_HIP_CHECK(hipSetDevice(graph->device));
  TENET_HIP_CHECK(hipGraphLaunch(graph->exec, stream ? (hipStream_t)stream : graph->stream));
  return TENET_OK;
}
extern "C" int tenet_graph_status(tenet_ctx ctx, int* status) {
  TRY(use_device(ctx));
  CHECK_ARG(status, "null status");
  int32_t v = 0;
  TENET_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  TENET_HIP_CHECK(hipMemcpy(&v, ctx->dev_status, sizeof v, hipMemcpyDeviceToHost));
  if (v) {
    const int32_t zero = 0;
    TENET_HIP_CHECK(hipMemcpy(ctx->dev_status, &zero, sizeof zero, hipMemcpyHostToDevice));
  }
  *status = v;
  return TENET_OK;
}
extern "C" int tenet_graph_destroy(tenet_graph graph) {
  if (!graph) return TENET_OK;
  (void)hipSetDevice(graph->device);
  (void)hipGraphExecDestroy(graph->exec);
  (void)hipGraphDestroy(graph->graph);
  delete graph;
  return TENET_OK;
}

extern "C" int tenet_stats_enable(tenet_ctx ctx, int family) {
  TRY(use_device(ctx));
  CHECK_ARG(family >= TENET_KSTAT_OFF && family <= TENET_KSTAT_AMPLITUDE, "unknown kernel family");
  TRY(stats_drain(ctx));
  ctx->stat_family = family;
  ctx->stat_launches = ctx->stat_sampled = 0;
  ctx->stat_ms = ctx->stat_flops = ctx->stat_bytes = 0.0;
  return TENET_OK;
}

extern "C" int tenet_stats_sample_every(tenet_ctx ctx, int every) {
  TRY(use_device(ctx));
  CHECK_ARG(every >= 1, "sampling stride must be >= 1");
  TRY(stats_drain(ctx));
  ctx->stat_every = every;
  return TENET_OK;
}

extern "C" int tenet_stats_read(tenet_ctx ctx, int64_t* launches, double* total_ms, double* flops, double* bytes) {
  TRY(use_device(ctx));
  TRY(stats_drain(ctx));
  if (launches) *launches = ctx->stat_launches;
  if (total_ms)
    *total_ms = ctx->stat_sampled ? ctx->stat_ms * (double)ctx->stat_launches / (double)ctx->stat_sampled : 0.0;
  if (flops) *flops = ctx->stat_flops;
  if (bytes) *bytes = ctx->stat_bytes;
  ctx->stat_launches = ctx->stat_sampled = 0;
  ctx->stat_ms = ctx->stat_flops = ctx->stat_bytes = 0.0;
  return TENET_OK;
}

extern "C" int tenet_timer_start(tenet_ctx ctx, tenet_stream stream) {
  TRY(use_device(ctx));
  TENET_HIP_CHECK(hipEventRecord(ctx->ev0, pick(ctx, stream)));
  return TENET_OK;
}
extern "C" int tenet_timer_stop(tenet_ctx ctx, tenet_stream stream, float* ms) {
  TRY(use_device(ctx));
  CHECK_ARG(ms, "null output");
  TENET_HIP_CHECK(hipEventRecord(ctx->ev1, pick(ctx, stream)));
  TENET_HIP_CHECK(hipEventSynchronize(ctx->ev1));
  TENET_HIP_CHECK(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return TENET_OK;
}
