// smdRadixSortNoAlloc: the key/value sort of radixSort (include/
// gpusmoldyn_legacy.h) for callers that cannot allocate device memory (code
// under stream capture, serving loops, harnesses that forbid hipMalloc).
//
// radixSort keeps its digit counters in a buffer the library allocates on
// first use and grows.  This entry point uses a fixed file-scope device array
// instead, sized for kNoAllocMaxKeys keys (about 4 MiB per GPU, held until the
// process ends), and calls no allocating routine.  Its address on each GPU is
// looked up once (hipGetSymbolAddress, with that GPU current) and kept per
// device; host code passes that address, never the array's name.
//
// Every call on one GPU shares that GPU's array, and the sort rewrites the
// counters it reads on every call (nothing needs clearing between calls).
// Calls that overlap on one GPU (from several host threads, or on streams that
// run concurrently) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/gpusmoldyn.h"
#include "../../include/gpusmoldyn_legacy.h"
#include "smd_kernels.h"
#include "smd_setup.h"

namespace {

constexpr uint32_t kNoAllocMaxKeys = 1u << 24;
constexpr int kNoAllocRadixBits = 8;
constexpr size_t kCountOffset = 64;  // word 0: key count; counters start 256 B in
// 8-bit digits, 4096-key tiles above 6M keys: 256 x (4096 tiles + 1) per-tile
// counters and 256 digit totals for 2^24 keys
constexpr size_t kScratchWords = kCountOffset + 256u * (4096u + 1u) + 256u;
constexpr int kMaxDevices = 64;

std::mutex gMu;
uint32_t* gScratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

// The key count the sort kernels read, stored by the GPU itself (no host
// buffer to keep alive, nothing to wait for).
__global__ void k_store_count(uint32_t* __restrict__ p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}

// The device address of `scratch` on the current device (looked up once per
// device), or nullptr with the error set.
uint32_t* scratch_on_current_device(int& status) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    smd::set_error(e != hipSuccess ? std::string("HIP error '") + hipGetErrorString(e) + "' in hipGetDevice"
                                   : std::string("device id out of range"));
    status = SMD_ERR_HIP;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(gMu);
  if (!gScratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      smd::set_error(std::string("HIP error '") + hipGetErrorString(e) + "' in hipGetSymbolAddress(scratch)");
      status = SMD_ERR_HIP;
      return nullptr;
    }
    gScratch[dev] = (uint32_t*)p;
  }
  return gScratch[dev];
}

}  // namespace

extern "C" int smdRadixSortNoAlloc(unsigned int* keys, unsigned int* values, unsigned int* tempKeys,
                                   unsigned int* tempValues, unsigned int numElements, unsigned int keyBits) {
  if (numElements == 0) return SMD_OK;
  if (!keys || !values || !tempKeys || !tempValues) {
    smd::set_error("smdRadixSortNoAlloc: null array");
    return SMD_ERR_INVALID;
  }
  if (keyBits == 0 || keyBits > 32) {
    smd::set_error("smdRadixSortNoAlloc: keyBits must be in 1..32");
    return SMD_ERR_INVALID;
  }
  if (numElements > kNoAllocMaxKeys) {
    smd::set_error("smdRadixSortNoAlloc: " + std::to_string(numElements) + " keys exceed the maximum of " +
                   std::to_string(kNoAllocMaxKeys) + " (fixed scratch array); use radixSort");
    return SMD_ERR_CAPACITY;
  }
  if (kCountOffset + smd::sort3k_counts_words(numElements, 0, kNoAllocRadixBits) > kScratchWords) {
    smd::set_error("smdRadixSortNoAlloc: scratch array too small for this key count");
    return SMD_ERR_CAPACITY;
  }
  int status = SMD_OK;
  uint32_t* base = scratch_on_current_device(status);
  if (!base) return status;
  uint32_t* dN = base;
  smd::SortBuffers b;
  b.keys = keys;
  b.vals = values;
  b.keysAlt = tempKeys;
  b.valsAlt = tempValues;
  b.counts = base + kCountOffset;
  b.status = nullptr;
  b.algo = 1;  // count / scan / scatter: its counters are exactly sort3k_counts_words
  b.gatherSrc = nullptr;
  b.gatherDst = nullptr;
  b.gatherDeadKey = 0;
  b.sizeHint = 0;
  b.packIndex = 0;
  // same stream as radixSort (the null stream), asynchronous like it
  hipLaunchKernelGGL(k_store_count, dim3(1), dim3(64), 0, nullptr, dN, (uint32_t)numElements);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = smd::launch_sort(b, dN, numElements, keyBits, kNoAllocRadixBits, false, nullptr);
  if (e == hipSuccess && b.keys != keys) {
    e = hipMemcpyAsync(keys, b.keys, numElements * sizeof(uint32_t), hipMemcpyDeviceToDevice, nullptr);
    if (e == hipSuccess)
      e = hipMemcpyAsync(values, b.vals, numElements * sizeof(uint32_t), hipMemcpyDeviceToDevice, nullptr);
  }
  if (e != hipSuccess) {
    smd::set_error(std::string("HIP error '") + hipGetErrorString(e) + "' in smdRadixSortNoAlloc");
    return SMD_ERR_HIP;
  }
  return SMD_OK;
}
