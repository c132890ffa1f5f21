// Streaming read-2/write-2 copy of a known size: the calibration kernel for
// HBM traffic counters (FETCH_SIZE / WRITE_SIZE per dispatch) and the measured
// copy ceiling the sweep's bandwidth is compared with. Usage:
//   hbm_copy_probe [MiB per array, default 1780] [repetitions, default 20] [all]
// With "all" it also times, over grids of 1024..65536 blocks, the same copy
// with non-temporal loads and stores, a read-only pass over the two inputs
// and a write-only pass over the two outputs: the ceilings of each traffic
// mix on this device; then the copy with four elements per thread in flight
// in blocks of 1024 threads, plain and non-temporal.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

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

__global__ void __launch_bounds__(256) copy2_nt(const f32x4 *__restrict__ a, const f32x4 *__restrict__ b,
                                                f32x4 *__restrict__ c, f32x4 *__restrict__ d, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const f32x4 x = __builtin_nontemporal_load(&a[e]), y = __builtin_nontemporal_load(&b[e]);
        __builtin_nontemporal_store(x, &c[e]);
        __builtin_nontemporal_store(y, &d[e]);
    }
}

// the copy with four 16-byte elements per thread and array in flight before
// any store, for blocks of up to 1024 threads: does more in flight per thread
// move the ceiling?
template <bool NT>
__global__ void __launch_bounds__(1024) copy2_x4(const f32x4 *__restrict__ a, const f32x4 *__restrict__ b,
                                                 f32x4 *__restrict__ c, f32x4 *__restrict__ d, long n) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long e0 = blockIdx.x * (long)blockDim.x + threadIdx.x; e0 < n; e0 += 4 * stride) {
        f32x4 x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long e = e0 + u * stride;
            if (e < n) {
                x[u] = NT ? __builtin_nontemporal_load(&a[e]) : a[e];
                y[u] = NT ? __builtin_nontemporal_load(&b[e]) : b[e];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long e = e0 + u * stride;
            if (e < n) {
                if (NT) {
                    __builtin_nontemporal_store(x[u], &c[e]);
                    __builtin_nontemporal_store(y[u], &d[e]);
                } else {
                    c[e] = x[u];
                    d[e] = y[u];
                }
            }
        }
    }
}

// reads both inputs; the sum is stored only if it takes an impossible value
__global__ void __launch_bounds__(256) read2(const f32x4 *__restrict__ a, const f32x4 *__restrict__ b,
                                             float *__restrict__ out, long n) {
    float s = 0.f;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const f32x4 x = a[e], y = b[e];
        s += x.x + y.y;
    }
    if (s == 12345.f) out[0] = s;
}

__global__ void __launch_bounds__(256) write2(f32x4 *__restrict__ c, f32x4 *__restrict__ d, long n) {
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        c[e] = z;
        d[e] = z;
    }
}

int main(int argc, char **argv) {
    const long mib = argc > 1 ? atol(argv[1]) : 1780;
    const int reps = argc > 2 ? atoi(argv[2]) : 20;
    const bool all = argc > 3 && strcmp(argv[3], "all") == 0;
    const long bytes = mib << 20, n = bytes / 16;
    f32x4 *a, *b, *c, *d;
    float *o;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMalloc(&c, bytes));
    CHECK(hipMalloc(&d, bytes));
    CHECK(hipMalloc(&o, 64));
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
    if (!all) return 0;
    const char *names[] = {"copy", "copy_nontemporal", "read_only", "write_only"};
    const int grids[] = {1024, 4096, 16384, 65536};
    for (int kind = 0; kind < 4; ++kind)
        for (int g : grids) {
            auto launch = [&]() {
                if (kind == 0) hipLaunchKernelGGL(copy2, dim3(g), dim3(256), 0, 0, a, b, c, d, n);
                if (kind == 1) hipLaunchKernelGGL(copy2_nt, dim3(g), dim3(256), 0, 0, a, b, c, d, n);
                if (kind == 2) hipLaunchKernelGGL(read2, dim3(g), dim3(256), 0, 0, a, b, o, n);
                if (kind == 3) hipLaunchKernelGGL(write2, dim3(g), dim3(256), 0, 0, c, d, n);
            };
            for (int it = 0; it < 3; ++it) launch();
            CHECK(hipEventRecord(e0));
            for (int it = 0; it < reps; ++it) launch();
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            const double moved = (kind <= 1 ? 4.0 : 2.0) * bytes;
            printf("{\"kernel\": \"%s\", \"grid\": %d, \"avg_ms\": %.5f, \"GBs\": %.1f}\n", names[kind], g, ms,
                   moved / (ms * 1e-3) / 1e9);
        }
    // four elements in flight per thread, blocks of 1024 threads
    for (int nt = 0; nt < 2; ++nt)
        for (int g : {8192, 32768, 65536}) {
            auto launch = [&]() {
                if (nt) hipLaunchKernelGGL(copy2_x4<true>, dim3(g), dim3(1024), 0, 0, a, b, c, d, n);
                else hipLaunchKernelGGL(copy2_x4<false>, dim3(g), dim3(1024), 0, 0, a, b, c, d, n);
            };
            for (int it = 0; it < 3; ++it) launch();
            CHECK(hipEventRecord(e0));
            for (int it = 0; it < reps; ++it) launch();
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            printf("{\"kernel\": \"%s\", \"grid\": %d, \"block\": 1024, \"avg_ms\": %.5f, \"GBs\": %.1f}\n",
                   nt ? "copy_x4_nontemporal" : "copy_x4", g, ms, 4.0 * bytes / (ms * 1e-3) / 1e9);
        }
    return 0;
}
