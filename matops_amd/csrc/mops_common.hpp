// Shared definitions for the gfx950 dense FP64 path: flag bits, error state,
// per-thread device workspace and small device helpers.
//
// Flag bits keep the values of the reference's cgo boundary
// (calgo/cmops.h:13-24, calgo/cmops.go:19-31) so a caller can pass the
// reference's Flags value through unchanged.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mops {

enum : int {
  F_NOTRANS = 0,
  F_TRANSA = 0x1,
  F_TRANSB = 0x2,
  F_LOWER = 0x4,
  F_UPPER = 0x8,
  F_LEFT = 0x10,
  F_RIGHT = 0x20,
  F_UNIT = 0x40,
  F_TRANS = 0x80,
  F_NOSCALE = 0x800,
};

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Status codes returned through the C ABI.
enum : int {
  ST_OK = 0,
  ST_BADARG = -1,     // negative: -(index of the bad argument) when known
  ST_HIP = -1000,     // HIP runtime failure, message in mops_last_error()
  ST_NOMEM = -1001,   // device allocation failed
  ST_NOTSPD = -2000,  // Cholesky met a NaN square root (info holds the index)
};

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();
const char* last_error();

#define MOPS_CHECK_HIP(call)                                                  \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::mops::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,   \
                        hipGetErrorString(e_));                               \
      return ::mops::ST_HIP;                                                  \
    }                                                                         \
  } while (0)

#define MOPS_TRY(expr)        \
  do {                        \
    int st_ = (expr);         \
    if (st_ != 0) return st_; \
  } while (0)

// Device workspace owned by the calling host thread, one per device. It only
// grows; requests are served from a bump region that is reset by the
// top-level entry point, so nested helpers can carve sub-buffers.
struct Workspace {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  int device = -1;
};

// Returns the calling thread's workspace for the current device, grown to at
// least `bytes` of free space beyond what is already carved. Growing discards
// previously carved buffers, so callers reserve once up front (reserve()) and
// then carve().
int ws_reserve(size_t bytes);
void* ws_carve(size_t bytes);  // 256-byte aligned; nullptr when out of space
void ws_reset();
// Scoped carving: buffers carved after ws_mark() are given back by
// ws_release(mark). Safe for work queued on one stream (it runs in order).
size_t ws_mark();
void ws_release(size_t mark);
// Frees the calling thread's workspace and streams on every device.
void thread_release();

// Held by every entry point for as long as it queues work that uses the
// workspace on stream s. The workspace is reused by the next call of the
// same thread, which may name another stream: the constructor makes s wait
// for the work the previous call queued (when that was on another stream),
// the destructor marks the end of this call's work on s.
class WsUse {
 public:
  explicit WsUse(hipStream_t s);
  ~WsUse();
  WsUse(const WsUse&) = delete;
  WsUse& operator=(const WsUse&) = delete;

 private:
  hipStream_t s_;
};

// Sticky status word of the calling thread on the current device (device
// int, zero until a kernel raises a bit; nullptr if it cannot be allocated).
// Kernels whose bounded waits have no other way to report a timeout OR
// their bit into it; take_status() reads and clears it.
enum : int { STATUS_GEMM_WAIT = 1, STATUS_TRSM_WAIT = 2 };
int* status_word();
// Synchronizes s, then reads and clears the status word: ST_OK when no bit
// was raised, else ST_HIP with the text naming the wait that timed out.
int take_status(hipStream_t s);
// The status word kernels of the calling thread write to: status_word(),
// or the pointer installed by a StatusOverride in scope (entry points that
// may not allocate pass one inside their own scratch).
int* active_status();
// Flags of a kernel whose workgroups signal each other (split-k GEMM,
// chained triangular solve): count ints of the calling thread's flag pool
// on the current device, and *ready, the value that marks a flag as set in
// this launch. Every launch gets a value no earlier launch of the thread
// used, so whatever earlier launches left in the pool (also one whose wait
// gave up) never reads as set, and no launch has to clear its flags first.
// Launches take consecutive ranges of the pool, so launches in flight on
// other streams hold other flags. False when count exceeds the pool (1 Mi
// flags) or the pool cannot be allocated.
bool signal_flags(long count, int** flags, int* ready);
// Status word of the no-allocation GEMM on the current device (inside its
// scratch array), nullptr until that entry point first ran there.
int* noalloc_status_word();
class StatusOverride {
 public:
  explicit StatusOverride(int* p);
  ~StatusOverride();
  StatusOverride(const StatusOverride&) = delete;
  StatusOverride& operator=(const StatusOverride&) = delete;

 private:
  int* prev_;
};

// Spin limit of the bounded waits: `dflt`, unless the environment variable
// MOPS_SPIN_LIMIT (read once per process) gives another value. A limit of 0
// makes every bounded wait give up at once, which lets tests of the failure
// path run without any kernel waiting at all.
unsigned wait_spin_limit(unsigned dflt);

__device__ __forceinline__ void raise_status(int* p, int bit) {
  if (p) __hip_atomic_fetch_or(p, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-thread stream used by the host-pointer entry points.
hipStream_t host_stream();
// Per-thread second stream (default priority) for look-ahead work, and six
// reusable synchronisation events (index 0..5) of the calling thread.
hipStream_t aux_stream();
hipEvent_t sync_event(int i);
// Per-thread stream (default priority) whose kernels stay off the first
// BULK_RESERVED_CUS compute units, for bulk updates that run beside a
// latency-bound chain of small kernels on another stream: the chain always
// finds free CUs. nullptr when the device cannot make one.
constexpr int BULK_RESERVED_CUS = 32;
hipStream_t bulk_stream();

// Optional timing of selected launches with HIP events recorded on the
// launching stream (per host thread). Classes name what is timed.
enum : int { TR_SYRK_UPDATE = 0, TR_GEMM_UPDATE = 1, TR_PANEL = 2, TR_NCLASS = 3 };
bool trace_on();
// Records an event pair around a launch sequence: call trace_mark_begin
// before and trace_mark_end (with the algorithmic flops) after.
void trace_mark_begin(hipStream_t s, int cls);
void trace_mark_end(hipStream_t s, int cls, double flops);
void trace_enable(int on);
int trace_collect(int cls, int* count, double* ms, double* flops);

static inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
static inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace mops
