// HBM traffic calibration: three launches of a plain copy kernel of known size (256 MiB
// read + 256 MiB written per launch).  rocprofv3's FETCH_SIZE / WRITE_SIZE counters read on
// this kernel give the factor that turns those counters into bytes for the walk kernels.
#include <hip/hip_runtime.h>
#include <cstdio>

__global__ void copy_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

int main()
{
    const size_t bytes = 256ull << 20, n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    for (int it = 0; it < 3; ++it) hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 2;
    printf("copy_probe: 3 launches of %zu bytes read + %zu bytes written\n", bytes, bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
