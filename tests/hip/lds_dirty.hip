// Test helper: fill the LDS of every CU with 0xFF bytes (a NaN as FP64 and FP32), so that a later
// kernel that reads LDS it has not written in the same launch sees NaNs rather than whatever the
// previous workgroup on that CU left there (usually finite values, which hide such reads when
// they are masked by a 0 weight). LDS is not cleared between launches.
#include <hip/hip_runtime.h>

__global__ void k_lds_dirty(unsigned words) {
  extern __shared__ unsigned dyn[];
  volatile unsigned* s = dyn;  // volatile: the stores must happen although nothing reads them
  for (unsigned i = threadIdx.x; i < words; i += blockDim.x) s[i] = 0xFFFFFFFFu;
  __syncthreads();
}

// Returns 0 on success, else the HIP error code.
extern "C" int lds_dirty(void) {
  int dev = 0, bytes = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return (int)e;
  // a few workgroups per CU, each taking the whole LDS a workgroup may have
  hipLaunchKernelGGL(k_lds_dirty, dim3(8 * cus), dim3(256), (size_t)bytes, 0, (unsigned)(bytes / 4));
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return (int)e;
}
