// Work memory of the allocation-free CG solve (bpx_cg_solve_noalloc).
//
// The solve's device work buffers (p, Ap, the CG state, the r.r and p.Ap
// partials and the r.r / alpha histories) live in one file-scope device
// array of fixed size instead of memory the library allocates, so a caller
// that must not allocate device memory can still run the CG. Each GPU the
// library runs on holds its own copy of the array (about 33 MiB) until the
// process ends. Host code never uses the array's name as an address: the
// device address comes from hipGetSymbolAddress, once per device, with that
// device current, and every kernel argument is derived from it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr size_t kPad = 32;  // region starts on 256-byte boundaries
constexpr size_t pad(size_t n) { return (n + kPad - 1) / kPad * kPad; }

constexpr size_t kOffP = 0;
constexpr size_t kOffAp = kOffP + pad(kNoAllocMaxN);
constexpr size_t kOffHistRR = kOffAp + pad(kNoAllocMaxN);
constexpr size_t kOffHistAlpha = kOffHistRR + pad((size_t)kNoAllocMaxIter + 1);
constexpr size_t kOffPartRR = kOffHistAlpha + pad((size_t)kNoAllocMaxIter + 1);
constexpr size_t kOffPartPAp = kOffPartRR + pad(kRedBlocks);
constexpr size_t kOffState = kOffPartPAp + pad(kRedBlocks);
constexpr size_t kStateDoubles = (sizeof(CgState) + sizeof(double) - 1) / sizeof(double);
constexpr size_t kScratchDoubles = kOffState + pad(kStateDoubles);

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_init_mutex;
double* g_scratch[kMaxDevices] = {};
std::mutex g_solve_mutex[kMaxDevices];

}  // namespace

int noalloc_scratch_bytes(size_t* bytes) {
  *bytes = kScratchDoubles * sizeof(double);
  return BPX_OK;
}

// The device whose array a solve on stream s uses, its array's address
// (looked up once per device, with that device current) and the device's
// lock, which serialises the solves of this process on one device.
int noalloc_work(hipStream_t s, int* device, NoAllocWork* w) {
  int dev = -1;
  if (s) {
    BPX_HIP(hipStreamGetDevice(s, &dev));
  } else {
    BPX_HIP(hipGetDevice(&dev));
  }
  if (dev < 0 || dev >= kMaxDevices) return set_error(BPX_ERR_UNSUPPORTED, "cg (no-alloc): device index above 63");
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (!g_scratch[dev]) {
      int cur = -1;
      BPX_HIP(hipGetDevice(&cur));
      if (cur != dev) BPX_HIP(hipSetDevice(dev));
      void* ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (cur != dev) (void)hipSetDevice(cur);
      if (e != hipSuccess) return hip_error(e, "cg (no-alloc): hipGetSymbolAddress(scratch)");
      if (!ptr) return set_error(BPX_ERR_HIP, "cg (no-alloc): hipGetSymbolAddress(scratch) returned NULL");
      g_scratch[dev] = static_cast<double*>(ptr);
    }
    base = g_scratch[dev];
  }
  *device = dev;
  w->p = base + kOffP;
  w->Ap = base + kOffAp;
  w->hist_rr = base + kOffHistRR;
  w->hist_alpha = base + kOffHistAlpha;
  w->part_rr = base + kOffPartRR;
  w->part_pAp = base + kOffPartPAp;
  w->state = reinterpret_cast<CgState*>(base + kOffState);
  w->state_doubles = (int)kStateDoubles;
  w->lock = &g_solve_mutex[dev];
  return BPX_OK;
}

}  // namespace bpx
