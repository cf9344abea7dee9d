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

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            return 1;                                                             \
        }                                                                         \
    } while (0)

int main(int argc, char **argv) {
    size_t n = argc > 1 ? strtoull(argv[1], nullptr, 10) : (size_t)256 << 20;
    int reps = argc > 2 ? atoi(argv[2]) : 3;
    int *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, n * 4));
    CHECK(hipMalloc(&b, n * 4));
    CHECK(hipMemset(a, 1, n * 4));
    CHECK(hipMemset(b, 0, n * 4));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int r = 0; r < reps; r++) {
        CHECK(hipEventRecord(e0));
        hipLaunchKernelGGL(copy_i32, dim3(8192), dim3(256), 0, 0, a, b, n);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        printf("copy %zu bytes read + %zu written: %.3f ms, %.1f GB/s\n", n * 4, n * 4, ms,
               2.0 * n * 4 / ms / 1e6);
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
