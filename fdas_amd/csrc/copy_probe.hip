// Copy kernel of known size, used to calibrate the rocprofv3 FETCH_SIZE and
// WRITE_SIZE counters (bytes per dispatch) before they are read as HBM traffic
// of the FDAS kernels.  256 MiB read + 256 MiB written per dispatch, 3 dispatches.
#include <hip/hip_runtime.h>
#include <cstdio>
__global__ void copyk(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x, s = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) b[i] = a[i];
}
int main() {
  const size_t bytes = 256ull << 20, n = bytes / 16;   // 256 MiB read + 256 MiB write
  float4 *a, *b;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
  hipMemset(a, 0, bytes); hipMemset(b, 0, bytes); hipDeviceSynchronize();
  for (int r = 0; r < 3; ++r) copyk<<<8192, 256>>>(a, b, n);
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  printf("copy of %zu bytes x3 done\n", bytes);
  return 0;
}
