// Scratch of the allocation-free WCSPH step (sfl_step_no_alloc).
//
// A WCSPH step needs, beyond the particle state and its sorted copy (both
// held by the world since its particles were added), a control block (flags,
// striped report reductions, the cell box of the sort) and the sort's dense
// cell table (per-cell counts, the exclusive scan of the counts, the scan's
// tile sums and total). The allocation-free step places those in `scratch`,
// one file-scope device array: each GPU the library runs on holds its own
// copy from the first such call until the process ends, and every call on a
// GPU shares it. Calls that overlap on one GPU race; each call returns only
// after its last kernel has finished, so calls made one after another from
// any number of worlds are safe.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "world.hpp"

namespace sfl {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Zeroes `n16` 16-byte words (grid-stride; any grid covers the range).
__global__ __launch_bounds__(256) void k_fill_zero(uint4* __restrict__ p, size_t n16) {
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    p[i] = z;
}

// The device address of `scratch` on `device`, which must be the current
// device. hipGetSymbolAddress runs once per device, on the first call there;
// the host-side name of the array is never used as an address.
hipError_t no_alloc_scratch(int device, unsigned char** out) {
  static std::mutex lock;
  static std::vector<unsigned char*> addresses;
  std::lock_guard<std::mutex> hold(lock);
  if (device < 0) return hipErrorInvalidDevice;
  if ((size_t)device >= addresses.size()) addresses.resize((size_t)device + 1, nullptr);
  if (!addresses[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    addresses[device] = static_cast<unsigned char*>(ptr);
  }
  *out = addresses[device];
  return hipSuccess;
}

// Clears `bytes` (a multiple of 16, 16-byte aligned) at `p` with a kernel:
// no memset node, so the clear also holds where a caller captures the work.
void launch_fill_zero(hipStream_t s, void* p, size_t bytes) {
  size_t n16 = bytes / 16;
  if (!n16) return;
  size_t blocks = (n16 + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_fill_zero, dim3((uint32_t)blocks), dim3(256), 0, s, static_cast<uint4*>(p), n16);
}

}  // namespace sfl
