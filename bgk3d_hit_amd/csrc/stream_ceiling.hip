// stream_ceiling.hip -- achievable HBM bandwidth for the access pattern of a
// D3Q19 step on one GPU, the practical ceiling under the 8 TB/s peak.
//
// Plain kernels, no arithmetic: a float4 copy of one contiguous array (the
// best case of the memory system), a copy of 19 arrays into 19 others with
// one site per thread (what every one-pass stream-collide must move: 19
// loads + 19 stores per site, 38 concurrent streams), the same copy in the
// library's interleaved layout ([z][y][q][x] per copy) with its 128x4 tiles
// and XCD-aware tile order, and a read of 19 arrays.
// Usage: stream_ceiling [n] (lattice n^3, FP32, default 512).  Prints one
// JSON object per kernel.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            fprintf(stderr, "stream_ceiling: %s at line %d\n", hipGetErrorString(e_), \
                    __LINE__);                                                      \
            exit(1);                                                                \
        }                                                                           \
    } while (0)

struct In19 {
    const float *p[19];
};
struct Out19 {
    float *p[19];
};

__global__ void k_fill(float *p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 1.0f;
}

__global__ void k_copy_float4(const float4 *__restrict__ a, float4 *__restrict__ b, long long n4) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) b[i] = a[i];
}

__global__ void k_copy19(In19 a, Out19 b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[19];
#pragma unroll
    for (int q = 0; q < 19; q++) v[q] = a.p[q][i];
#pragma unroll
    for (int q = 0; q < 19; q++) b.p[q][i] = v[q];
}

// tiles of 128 x 4 sites of one plane, each XCD (workgroups are dispatched
// round-robin over 8) given one contiguous range of tiles
__device__ __forceinline__ int xcd_order(int b, int g) {
    const int q = g >> 3, r = g & 7, xcd = b & 7, slot = b >> 3;
    return (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}
__global__ void k_copy19_interleaved(const float *__restrict__ a, float *__restrict__ b, int n) {
    const int t = xcd_order((int)blockIdx.x, (int)gridDim.x);
    const int tiles_x = n / 128, tiles_y = n / 4;
    const int tx = t % tiles_x, r = t / tiles_x, ty = r % tiles_y, z = r / tiles_y;
    const long long x = tx * 128 + threadIdx.x, y = ty * 4 + threadIdx.y;
    const long long o = x + y * 19LL * n + (long long)z * 19LL * n * n;
    float v[19];
#pragma unroll
    for (int q = 0; q < 19; q++) v[q] = a[o + (long long)q * n];
#pragma unroll
    for (int q = 0; q < 19; q++) b[o + (long long)q * n] = v[q];
}

__global__ void k_read19(In19 a, float *sink, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < 19; q++) s += a.p[q][i];
    if (s == -1.0f) sink[0] = s;  // never true: keeps the loads
}

template <typename F>
static float time_ms(F launch, int reps) {
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    launch();  // warm-up
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) launch();
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipGetLastError());
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return ms / reps;
}

int main(int argc, char **argv) {
    const long long l = argc > 1 ? atoll(argv[1]) : 512;
    if (l < 128 || l > 1024 || l % 128 != 0) {
        fprintf(stderr, "stream_ceiling: n must be a multiple of 128 in [128, 1024]\n");
        return 2;
    }
    const long long n = l * l * l;
    const int reps = 20;
    float *base = nullptr;
    CHECK(hipMalloc(&base, sizeof(float) * n * 38));
    const int bs = 256;
    k_fill<<<(unsigned)((n * 38 + bs - 1) / bs), bs>>>(base, n * 38);
    CHECK(hipDeviceSynchronize());

    // one contiguous array of 19 n floats copied into the next 19 n
    const long long n4 = n * 19 / 4;
    const float4 *a4 = (const float4 *)base;
    float4 *b4 = (float4 *)(base + n * 19);
    float ms = time_ms([&] { k_copy_float4<<<(unsigned)((n4 + bs - 1) / bs), bs>>>(a4, b4, n4); },
                       reps);
    printf("{\"kernel\": \"copy_float4\", \"bytes\": %lld, \"ms\": %.4f, \"gbs\": %.1f}\n",
           2 * n4 * 16, ms, 2.0 * n4 * 16 / (ms * 1e6));

    In19 a;
    Out19 b;
    for (int q = 0; q < 19; q++) {
        a.p[q] = base + q * n;
        b.p[q] = base + (19 + q) * n;
    }
    for (int t : {128, 256}) {
        ms = time_ms([&] { k_copy19<<<(unsigned)((n + t - 1) / t), t>>>(a, b, n); }, reps);
        printf("{\"kernel\": \"copy19\", \"threads\": %d, \"lattice\": %lld, \"bytes\": %lld, "
               "\"ms\": %.4f, \"gbs\": %.1f}\n",
               t, l, 38 * n * 4, ms, 38.0 * n * 4 / (ms * 1e6));
    }
    {
        const unsigned nb = (unsigned)((l / 128) * (l / 4) * l);
        ms = time_ms([&] { k_copy19_interleaved<<<nb, dim3(128, 4)>>>(base, base + 19 * n, (int)l); },
                     reps);
        printf("{\"kernel\": \"copy19_interleaved\", \"lattice\": %lld, \"bytes\": %lld, "
               "\"ms\": %.4f, \"gbs\": %.1f}\n",
               l, 38 * n * 4, ms, 38.0 * n * 4 / (ms * 1e6));
    }
    ms = time_ms([&] { k_read19<<<(unsigned)((n + bs - 1) / bs), bs>>>(a, base + 37 * n, n); },
                 reps);
    printf("{\"kernel\": \"read19\", \"lattice\": %lld, \"bytes\": %lld, \"ms\": %.4f, "
           "\"gbs\": %.1f}\n",
           l, 19 * n * 4, ms, 19.0 * n * 4 / (ms * 1e6));
    CHECK(hipFree(base));
    return 0;
}
