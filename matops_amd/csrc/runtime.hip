// Per-thread runtime state behind the C ABI: last-error text, a growable
// device workspace and a stream per host thread and device.
//
// The reference calls its kernels concurrently from several goroutines on
// disjoint blocks of C (mmult.go:241-255), so nothing here is shared between
// host threads: each thread owns its stream and workspace.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <hip/hip_ext.h>

#include "mops_common.hpp"

namespace mops {

namespace {
thread_local char g_err[512] = {0};

struct ThreadState {
  Workspace ws[16];
  hipStream_t stream[16] = {nullptr};
  hipStream_t aux[16] = {nullptr};
  hipStream_t bulk[16] = {nullptr};
  bool bulk_tried[16] = {};
  hipEvent_t ev[16][6] = {};
  int* status[16] = {};                 // sticky status words (device ints)
  int* flags[16] = {};                  // signal flag pools
  long flags_next[16] = {};             // next free flag of the pool
  int flags_ready[16] = {};             // ready value of the last launch
  hipEvent_t ws_done[16] = {};          // end of the last call's work
  hipStream_t ws_done_stream[16] = {};  // the stream it was recorded on
  bool ws_done_set[16] = {};
};
thread_local ThreadState g_ts;
thread_local int* g_status_override = nullptr;

__global__ void k_clear_status(int* p) {
  if (threadIdx.x == 0) *p = 0;
}

int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return 0;
  return d < 0 || d >= 16 ? 0 : d;
}
}  // namespace

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

const char* last_error() { return g_err; }

int ws_reserve(size_t bytes) {
  const int d = current_device();
  Workspace& w = g_ts.ws[d];
  bytes = align_up(bytes, 256) + 256;
  if (w.base && w.cap >= bytes) {
    w.used = 0;
    return ST_OK;
  }
  if (w.base) {
    // the old buffer may still be read by queued work on this thread's streams
    (void)hipDeviceSynchronize();
    (void)hipFree(w.base);
    w.base = nullptr;
    w.cap = 0;
  }
  size_t cap = bytes < (size_t(64) << 20) ? (size_t(64) << 20) : bytes;
  void* p = nullptr;
  if (hipMalloc(&p, cap) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device workspace of %zu bytes could not be allocated", cap);
    return ST_NOMEM;
  }
  w.base = static_cast<char*>(p);
  w.cap = cap;
  w.used = 0;
  w.device = d;
  return ST_OK;
}

WsUse::WsUse(hipStream_t s) : s_(s) {
  const int d = current_device();
  if (g_ts.ws_done_set[d] && g_ts.ws_done_stream[d] != s &&
      hipStreamWaitEvent(s, g_ts.ws_done[d], 0) != hipSuccess) {
    // fall back to waiting on the host: never overlap two calls' workspace
    (void)hipGetLastError();
    (void)hipEventSynchronize(g_ts.ws_done[d]);
  }
}

WsUse::~WsUse() {
  const int d = current_device();
  if (!g_ts.ws_done[d] &&
      hipEventCreateWithFlags(&g_ts.ws_done[d], hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    g_ts.ws_done[d] = nullptr;
    g_ts.ws_done_set[d] = false;
    return;
  }
  if (hipEventRecord(g_ts.ws_done[d], s_) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(s_);
    g_ts.ws_done_set[d] = false;
    return;
  }
  g_ts.ws_done_stream[d] = s_;
  g_ts.ws_done_set[d] = true;
}

void* ws_carve(size_t bytes) {
  Workspace& w = g_ts.ws[current_device()];
  size_t off = align_up(w.used, 256);
  bytes = align_up(bytes, 256);
  if (!w.base || off + bytes > w.cap) return nullptr;
  w.used = off + bytes;
  return w.base + off;
}

void ws_reset() { g_ts.ws[current_device()].used = 0; }

size_t ws_mark() { return g_ts.ws[current_device()].used; }

void ws_release(size_t mark) {
  Workspace& w = g_ts.ws[current_device()];
  if (mark <= w.used) w.used = mark;
}

int* status_word() {
  const int d = current_device();
  if (!g_ts.status[d]) {
    void* p = nullptr;
    if (hipMalloc(&p, 256) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    // synchronous on purpose: the word is zero before any kernel can see it
    if (hipMemset(p, 0, 256) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(p);
      return nullptr;
    }
    g_ts.status[d] = static_cast<int*>(p);
  }
  return g_ts.status[d];
}

constexpr long FLAG_POOL = 1l << 20;

bool signal_flags(long count, int** flags, int* ready) {
  if (count <= 0 || count > FLAG_POOL) return false;
  const int d = current_device();
  bool clear = false;
  if (!g_ts.flags[d]) {
    void* p = nullptr;
    if (hipMalloc(&p, FLAG_POOL * sizeof(int)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    g_ts.flags[d] = static_cast<int*>(p);
    clear = true;
  } else if (g_ts.flags_ready[d] == 0x7fffffff) {
    // every ready value has been used: earlier ones could read as set again
    if (hipDeviceSynchronize() != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    clear = true;
  }
  if (clear) {
    // synchronous on purpose: the pool is zero before any kernel can see it
    if (hipMemset(g_ts.flags[d], 0, FLAG_POOL * sizeof(int)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    g_ts.flags_ready[d] = 0;
    g_ts.flags_next[d] = 0;
  }
  if (g_ts.flags_next[d] + count > FLAG_POOL) g_ts.flags_next[d] = 0;
  *flags = g_ts.flags[d] + g_ts.flags_next[d];
  g_ts.flags_next[d] += (count + 63) / 64 * 64;  // 256-byte aligned ranges
  *ready = ++g_ts.flags_ready[d];
  return true;
}

unsigned wait_spin_limit(unsigned dflt) {
  // -1: not read yet, -2: not set
  static std::atomic<long long> g_limit{-1};
  long long v = g_limit.load(std::memory_order_relaxed);
  if (v == -1) {
    const char* e = getenv("MOPS_SPIN_LIMIT");
    v = -2;
    if (e && *e) {
      char* end = nullptr;
      const unsigned long long x = strtoull(e, &end, 10);
      if (end && *end == 0 && x <= 0xffffffffull) v = (long long)x;
    }
    g_limit.store(v, std::memory_order_relaxed);
  }
  return v >= 0 ? (unsigned)v : dflt;
}

int* active_status() { return g_status_override ? g_status_override : status_word(); }

StatusOverride::StatusOverride(int* p) : prev_(g_status_override) { g_status_override = p; }
StatusOverride::~StatusOverride() { g_status_override = prev_; }

int take_status(hipStream_t s) {
  int* p = g_status_override ? g_status_override : g_ts.status[current_device()];
  if (!p) return ST_OK;  // never allocated: no kernel could have raised a bit
  int v = 0;
  MOPS_CHECK_HIP(hipMemcpyAsync(&v, p, sizeof(int), hipMemcpyDeviceToHost, s));
  MOPS_CHECK_HIP(hipStreamSynchronize(s));
  if (v == 0) return ST_OK;
  hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(64), 0, s, p);
  MOPS_CHECK_HIP(hipGetLastError());
  MOPS_CHECK_HIP(hipStreamSynchronize(s));
  set_error("an internal wait timed out (%s%s); the results of the calls since the last "
            "status check are not valid",
            (v & STATUS_GEMM_WAIT) ? "split-k GEMM partial tiles" : "",
            (v & STATUS_TRSM_WAIT) ? ((v & STATUS_GEMM_WAIT) ? ", chained triangular solve"
                                                             : "chained triangular solve")
                                   : "");
  return ST_HIP;
}

void thread_release() {
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (int d = 0; d < 16; ++d) {
    if (!g_ts.ws[d].base && !g_ts.stream[d] && !g_ts.aux[d] && !g_ts.bulk[d] && !g_ts.status[d] &&
        !g_ts.flags[d])
      continue;
    if (hipSetDevice(d) != hipSuccess) continue;
    if (g_ts.stream[d]) (void)hipStreamSynchronize(g_ts.stream[d]);
    // the last call's work (on whatever stream) is done with the workspace
    if (g_ts.ws_done_set[d]) (void)hipEventSynchronize(g_ts.ws_done[d]);
    if (g_ts.ws[d].base) (void)hipFree(g_ts.ws[d].base);
    if (g_ts.status[d]) (void)hipFree(g_ts.status[d]);
    g_ts.status[d] = nullptr;
    if (g_ts.flags[d]) (void)hipFree(g_ts.flags[d]);
    g_ts.flags[d] = nullptr;
    g_ts.flags_next[d] = 0;
    g_ts.flags_ready[d] = 0;
    if (g_ts.stream[d]) (void)hipStreamDestroy(g_ts.stream[d]);
    if (g_ts.aux[d]) {
      (void)hipStreamSynchronize(g_ts.aux[d]);
      (void)hipStreamDestroy(g_ts.aux[d]);
    }
    if (g_ts.bulk[d]) {
      (void)hipStreamSynchronize(g_ts.bulk[d]);
      (void)hipStreamDestroy(g_ts.bulk[d]);
    }
    for (int e = 0; e < 6; ++e)
      if (g_ts.ev[d][e]) (void)hipEventDestroy(g_ts.ev[d][e]);
    if (g_ts.ws_done[d]) (void)hipEventDestroy(g_ts.ws_done[d]);
    g_ts.ws_done[d] = nullptr;
    g_ts.ws_done_set[d] = false;
    g_ts.ws_done_stream[d] = nullptr;
    g_ts.ws[d] = Workspace();
    g_ts.stream[d] = nullptr;
    g_ts.aux[d] = nullptr;
    g_ts.bulk[d] = nullptr;
    g_ts.bulk_tried[d] = false;
    for (int e = 0; e < 6; ++e) g_ts.ev[d][e] = nullptr;
  }
  (void)hipSetDevice(cur);
}

namespace {
struct TraceRec {
  hipEvent_t a, b;
  int cls;
  double flops;
};
struct TraceState {
  bool on = false;
  int open_cls = -1;
  hipEvent_t open_ev = nullptr;
  TraceRec* recs = nullptr;
  int n = 0, cap = 0;
};
thread_local TraceState g_tr;
}  // namespace

bool trace_on() { return g_tr.on; }

void trace_mark_begin(hipStream_t s, int cls) {
  if (!g_tr.on) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  g_tr.open_ev = e;
  g_tr.open_cls = cls;
}

void trace_mark_end(hipStream_t s, int cls, double flops) {
  if (!g_tr.on || g_tr.open_cls != cls || !g_tr.open_ev) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  if (g_tr.n == g_tr.cap) {
    int nc = g_tr.cap ? 2 * g_tr.cap : 256;
    TraceRec* nr = new TraceRec[nc];
    for (int i = 0; i < g_tr.n; ++i) nr[i] = g_tr.recs[i];
    delete[] g_tr.recs;
    g_tr.recs = nr;
    g_tr.cap = nc;
  }
  g_tr.recs[g_tr.n++] = TraceRec{g_tr.open_ev, e, cls, flops};
  g_tr.open_ev = nullptr;
  g_tr.open_cls = -1;
}

void trace_enable(int on) {
  for (int i = 0; i < g_tr.n; ++i) {
    (void)hipEventDestroy(g_tr.recs[i].a);
    (void)hipEventDestroy(g_tr.recs[i].b);
  }
  g_tr.n = 0;
  g_tr.on = on != 0;
}

int trace_collect(int cls, int* count, double* ms, double* flops) {
  int c = 0;
  double t = 0.0, f = 0.0;
  for (int i = 0; i < g_tr.n; ++i) {
    if (g_tr.recs[i].cls != cls) continue;
    float e = 0.f;
    if (hipEventSynchronize(g_tr.recs[i].b) != hipSuccess) return ST_HIP;
    if (hipEventElapsedTime(&e, g_tr.recs[i].a, g_tr.recs[i].b) != hipSuccess) return ST_HIP;
    ++c;
    t += e;
    f += g_tr.recs[i].flops;
  }
  *count = c;
  *ms = t;
  *flops = f;
  return ST_OK;
}

hipStream_t aux_stream() {
  const int d = current_device();
  if (!g_ts.aux[d]) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    g_ts.aux[d] = s;
  }
  return g_ts.aux[d];
}

hipStream_t bulk_stream() {
  const int d = current_device();
  if (!g_ts.bulk_tried[d]) {
    g_ts.bulk_tried[d] = true;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess)
      return nullptr;
    if (ncu <= 2 * BULK_RESERVED_CUS) return nullptr;
    uint32_t mask[32];
    const int words = (ncu + 31) / 32;
    if (words > 32) return nullptr;
    for (int i = 0; i < words; ++i) mask[i] = 0xffffffffu;
    if (ncu % 32) mask[words - 1] = (1u << (ncu % 32)) - 1;
    for (int c = 0; c < BULK_RESERVED_CUS; ++c) mask[c / 32] &= ~(1u << (c % 32));
    hipStream_t st = nullptr;
    if (hipExtStreamCreateWithCUMask(&st, (uint32_t)words, mask) != hipSuccess) return nullptr;
    g_ts.bulk[d] = st;
  }
  return g_ts.bulk[d];
}

hipEvent_t sync_event(int i) {
  const int d = current_device();
  if (i < 0 || i >= 6) return nullptr;
  if (!g_ts.ev[d][i]) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    g_ts.ev[d][i] = e;
  }
  return g_ts.ev[d][i];
}

hipStream_t host_stream() {
  const int d = current_device();
  if (!g_ts.stream[d]) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    g_ts.stream[d] = s;
  }
  return g_ts.stream[d];
}

}  // namespace mops
