// Test helper: fills the LDS of every CU with a 32-bit pattern, so that a test
// can show that a kernel launched next reads no LDS it has not written first
// (a workgroup starts with whatever the previous one on its CU left there).
// Built into tests/liblds_fill.so by __graft_entry__.build().
#include <hip/hip_runtime.h>

namespace {

constexpr int kLdsBytes = 160 * 1024;  // the whole LDS of a gfx950 CU
constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) ldsFill(unsigned pattern) {
  extern __shared__ unsigned lds[];
  for (int i = threadIdx.x; i < kLdsBytes / 4; i += kThreads) lds[i] = pattern;
  __syncthreads();
}

}  // namespace

// Launches `blocks` workgroups of kThreads, each holding the whole LDS of its
// CU, and waits for them; returns the HIP error code (0 on success).
extern "C" int lds_fill(unsigned pattern, int blocks) {
  hipError_t e = hipFuncSetAttribute((const void*)ldsFill, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
  if (e != hipSuccess) return (int)e;
  hipLaunchKernelGGL(ldsFill, dim3(blocks), dim3(kThreads), kLdsBytes, 0, pattern);
  e = hipGetLastError();
  if (e != hipSuccess) return (int)e;
  return (int)hipDeviceSynchronize();
}
