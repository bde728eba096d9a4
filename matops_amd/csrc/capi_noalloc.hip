// C ABI: the GEMM entry point that allocates nothing (mops_d_dgemm_noalloc).
//
// Same arguments and results as mops_d_dgemm. The only device memory that
// entry point takes beyond its operands is the scratch of the k split
// (partial tiles and chunk flags, carved from the calling thread's
// workspace, which may grow with hipMalloc) and the thread's status word
// (hipMalloc'd on first use). Here both live in one file-scope device array
// per GPU, `scratch`, so a call makes no allocation of any kind; callers
// under stream capture, in tight serving loops or in harnesses that forbid
// hipMalloc can use it.
//
// Layout of `scratch` (NOALLOC_SCRATCH_BYTES = 34 MiB): bytes [0, 256) hold
// the status word the split-k waits raise (read and cleared by
// mops_d_status), the rest the split scratch. The split plan bounds what a
// product can need: (S - 1) x tiles < 1024 partial 64x64 FP64 tiles (32 MiB)
// for the 64x64 instance, fewer for the skinny one, plus a few KiB of flags
// (gemm_plan.hpp); sizes beyond the array are rejected all the same.
//
// All calls on one GPU share its array: calls that overlap on one GPU
// (several streams or host threads at once) race on the partial tiles and
// flags and must be serialized by the caller. The flags a launch polls are
// cleared by a small kernel before it (no memset node, so a captured graph
// replays correctly).
#include <atomic>
#include <mutex>

#include "../../include/mops.h"
#include "mops_kernels.hpp"

using namespace mops;

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(34) << 20;
constexpr size_t NOALLOC_STATUS_BYTES = 256;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on each GPU, found once per device with that
// device current (each GPU holds its own copy of the array). The array's
// name in host code is only a stand-in for the symbol, never an address.
std::atomic<char*> g_scratch_ptr[NOALLOC_MAX_DEVICES];
std::once_flag g_scratch_once[NOALLOC_MAX_DEVICES];
hipError_t g_scratch_err[NOALLOC_MAX_DEVICES];

__global__ void k_clear_words(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Returns the current device's scratch address (nullptr with the error text
// set on failure). The first call on a device also clears the status word,
// on `s`, ahead of any kernel that may raise it.
char* scratch_ptr(hipStream_t s) {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) {
    (void)hipGetLastError();
    set_error("noalloc: no current HIP device");
    return nullptr;
  }
  if (d < 0 || d >= NOALLOC_MAX_DEVICES) {
    set_error("noalloc: device %d beyond the %d supported", d, NOALLOC_MAX_DEVICES);
    return nullptr;
  }
  std::call_once(g_scratch_once[d], [&] {
    void* p = nullptr;
    g_scratch_err[d] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_scratch_err[d] != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, s, static_cast<int*>(p),
                       (int)(NOALLOC_STATUS_BYTES / sizeof(int)));
    g_scratch_err[d] = hipGetLastError();
    if (g_scratch_err[d] == hipSuccess) g_scratch_ptr[d].store(static_cast<char*>(p));
  });
  char* p = g_scratch_ptr[d].load();
  if (!p)
    set_error("noalloc: the address of the scratch array is not available: %s",
              hipGetErrorString(g_scratch_err[d]));
  return p;
}

}  // namespace

namespace mops {
// Status word of the no-allocation entry point on the current device, or
// nullptr when that entry point has not run there (read by mops_d_status).
int* noalloc_status_word() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= NOALLOC_MAX_DEVICES) {
    (void)hipGetLastError();
    return nullptr;
  }
  return reinterpret_cast<int*>(g_scratch_ptr[d].load());
}
}  // namespace mops

extern "C" {

int mops_d_dgemm_noalloc(int flags, int M, int N, int P, double alpha, const double* A, int lda,
                         const double* B, int ldb, double beta, double* C, int ldc,
                         void* stream) {
  clear_error();
  const bool ta = flags & F_TRANSA, tb = flags & F_TRANSB;
  if (M < 0) return -2;
  if (N < 0) return -3;
  if (P < 0) return -4;
  const int arows = ta ? P : M, brows = tb ? N : P;
  if (lda < (arows > 1 ? arows : 1)) return -7;
  if (ldb < (brows > 1 ? brows : 1)) return -9;
  if (ldc < (M > 1 ? M : 1)) return -12;
  const size_t need = gemm_split_ws_bytes(M, N, P, sizeof(double));
  if (need > NOALLOC_SCRATCH_BYTES - NOALLOC_STATUS_BYTES) {
    set_error("mops_d_dgemm_noalloc: the k split of a %d x %d x %d product needs %zu bytes of "
              "scratch, more than the %zu this entry point holds",
              M, N, P, need, NOALLOC_SCRATCH_BYTES - NOALLOC_STATUS_BYTES);
    return ST_BADARG;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = scratch_ptr(s);
  if (!base) return ST_HIP;
  StatusOverride so(reinterpret_cast<int*>(base));
  SplitScratchOverride sc(base + NOALLOC_STATUS_BYTES, NOALLOC_SCRATCH_BYTES - NOALLOC_STATUS_BYTES);
  return gemm_d(s, ta, tb, M, N, P, alpha, A, lda, B, ldb, beta, C, ldc);
}

}  // extern "C"
