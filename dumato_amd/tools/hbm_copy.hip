// Calibration kernel for rocprofv3 HBM counters: copies N int32 from one
// buffer to another (4N bytes read, 4N bytes written) several times.  The
// FETCH_SIZE / WRITE_SIZE a counter run reports for this kernel, divided by
// 4N, is the factor that turns the same counters on the enumeration kernel
// into bytes.
//   hbm_copy [N] [reps]     default N = 256 Mi ints (1 GiB), reps = 3
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_i32(const int *__restrict__ a, int *__restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char **argv) {
    size_t n = argc > 1 ? strtoull(argv[1], nullptr, 10) : (size_t)256 << 20;
    int reps = argc > 2 ? atoi(argv[2]) : 3;
    int *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 4) != hipSuccess || hipMalloc(&b, n * 4) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    hipMemset(a, 1, n * 4);
    hipMemset(b, 0, n * 4);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (int r = 0; r < reps; r++) {
        hipEventRecord(e0);
        hipLaunchKernelGGL(copy_i32, dim3(8192), dim3(256), 0, 0, a, b, n);
        hipEventRecord(e1);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        printf("copy %zu bytes read + %zu written: %.3f ms, %.1f GB/s\n", n * 4, n * 4, ms,
               2.0 * n * 4 / ms / 1e6);
    }
    hipFree(a);
    hipFree(b);
    return 0;
}
