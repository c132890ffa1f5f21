// Streaming read-2/write-2 copy of a known size: the calibration kernel for
// HBM traffic counters (FETCH_SIZE / WRITE_SIZE per dispatch) and the measured
// copy ceiling the sweep's bandwidth is compared with. Usage:
//   hbm_copy_probe [MiB per array, default 1780] [repetitions, default 20]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy2(const f32x4 *__restrict__ a, const f32x4 *__restrict__ b,
                                             f32x4 *__restrict__ c, f32x4 *__restrict__ d, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        c[e] = a[e];
        d[e] = b[e];
    }
}

#define CHECK(x)                                                             \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                              \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                        \
        }                                                                    \
    } while (0)

int main(int argc, char **argv) {
    const long mib = argc > 1 ? atol(argv[1]) : 1780;
    const int reps = argc > 2 ? atoi(argv[2]) : 20;
    const long bytes = mib << 20, n = bytes / 16;
    f32x4 *a, *b, *c, *d;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMalloc(&c, bytes));
    CHECK(hipMalloc(&d, bytes));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int blocks = 32768;
    for (int it = 0; it < 3; ++it) hipLaunchKernelGGL(copy2, dim3(blocks), dim3(256), 0, 0, a, b, c, d, n);
    CHECK(hipEventRecord(e0));
    for (int it = 0; it < reps; ++it) hipLaunchKernelGGL(copy2, dim3(blocks), dim3(256), 0, 0, a, b, c, d, n);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    ms /= reps;
    printf("{\"bytes_read_per_launch\": %ld, \"bytes_written_per_launch\": %ld, \"avg_ms\": %.5f, \"GBs\": %.1f}\n",
           2 * bytes, 2 * bytes, ms, 4.0 * bytes / (ms * 1e-3) / 1e9);
    return 0;
}
