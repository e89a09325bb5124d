// Allocation-free inclusive scan: dmgs_inclusive_scan_u32 with its scratch in
// a file-scope device array instead of a caller-provided workspace.
//
// For callers that cannot allocate device memory at call time (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
// the call takes only the input and output.  The scan's scratch is the
// per-block sums (one word per 4096 inputs plus 64), held in `scratch` below;
// each GPU the library runs on holds its own copy until the process ends.
// Host code never passes the array's name to a kernel (on the host it is only
// a stand-in for the symbol): the device address comes from
// hipGetSymbolAddress, looked up once per device with that device current.
//
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU (two streams, two host threads) race and must be serialised by the
// caller.  The block sums are written in full by each call before they are
// read, so nothing has to be cleared between calls.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/dmgs.h"
#include "dmgs_common.h"
#include "dmgs_kernels.h"

namespace dmgs {
void set_last_error(const char* msg);
}

namespace {

constexpr size_t kNoAllocScanMaxN = size_t(1) << 30;
constexpr size_t kNoAllocScanWords = (kNoAllocScanMaxN + 4095) / 4096 + 64;  // 262 208 words, 1 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kNoAllocScanWords];

std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};

int nfail(int code, const char* msg) {
  dmgs::set_last_error(msg);
  return code;
}

// Device address of `scratch` on the current device (looked up on the first
// call on that device).
int scratch_on_current_device(hipStream_t stream, uint32_t** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nfail(DMGS_EHIP, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return nfail(DMGS_EINVAL, "device index outside the scratch table (64 devices)");
  if (stream) {
    hipDevice_t sdev = -1;
    e = hipStreamGetDevice(stream, &sdev);
    if (e != hipSuccess) return nfail(DMGS_EHIP, hipGetErrorString(e));
    if (sdev != dev) return nfail(DMGS_EINVAL, "the stream belongs to another device than the current one");
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char buf[256];
      snprintf(buf, sizeof(buf), "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
               hipGetErrorString(e));
      return nfail(DMGS_EHIP, buf);
    }
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return DMGS_OK;
}

}  // namespace

extern "C" {

int dmgs_inclusive_scan_u32_noalloc(const uint32_t* in, uint32_t* out, int64_t n, dmgs_stream_t stream) {
  if (n < 0) return nfail(DMGS_EINVAL, "invalid n");
  if ((uint64_t)n > kNoAllocScanMaxN) {
    char buf[160];
    snprintf(buf, sizeof(buf), "n = %lld exceeds the allocation-free scan's maximum of %zu elements",
             (long long)n, kNoAllocScanMaxN);
    return nfail(DMGS_EINVAL, buf);
  }
  if (n == 0) return DMGS_OK;
  if (!in || !out) return nfail(DMGS_EINVAL, "null pointer");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  uint32_t* ws = nullptr;
  int rc = scratch_on_current_device(st, &ws);
  if (rc) return rc;
  if (dmgs::scan_scratch_words((size_t)n) > kNoAllocScanWords) return nfail(DMGS_EINVAL, "internal: scan scratch size");
  dmgs::launch_inclusive_scan(in, out, (size_t)n, ws, st);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return nfail(DMGS_EHIP, hipGetErrorString(e));
  return DMGS_OK;
}

}  // extern "C"
