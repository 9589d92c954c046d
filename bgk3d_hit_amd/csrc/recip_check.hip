// recip_check.hip -- bin/recip_check: recip_rn<float> (d3q19.h), the FP32
// reciprocal of the collision's density, against the IEEE quotient 1.0f / x
// for every one of the 2^32 float bit patterns (those in [2^-125, 2^126)
// take the estimate + Newton-Raphson path, the others the quotient).  Bit
// equality is required, NaN payloads included.  Prints the count of
// mismatches and the first one; exit status 0 only when there are none.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "d3q19.h"

__global__ __launch_bounds__(256) void k_recip_check(unsigned long long *count,
                                                     unsigned int *first) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    unsigned long long bad = 0;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         i < (1ULL << 32); i += stride) {
        const uint32_t bits = (uint32_t)i;
        const float x = __builtin_bit_cast(float, bits);
        const float q = 1.0f / x;
        const float r = bgk::recip_rn<float>(x);
        if (__builtin_bit_cast(uint32_t, r) != __builtin_bit_cast(uint32_t, q)) {
            bad++;
            atomicMin(first, bits);
        }
    }
    if (bad) atomicAdd(count, bad);
}

int main() {
    unsigned long long *count = nullptr;
    unsigned int *first = nullptr;
    if (hipMalloc(&count, sizeof *count) != hipSuccess ||
        hipMalloc(&first, sizeof *first) != hipSuccess) {
        fprintf(stderr, "recip_check: hipMalloc failed\n");
        return 2;
    }
    const unsigned long long zero = 0;
    const unsigned int none = 0xffffffffu;
    if (hipMemcpy(count, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(first, &none, sizeof none, hipMemcpyHostToDevice) != hipSuccess) {
        fprintf(stderr, "recip_check: hipMemcpy failed\n");
        return 2;
    }
    hipLaunchKernelGGL(k_recip_check, dim3(16384), dim3(256), 0, 0, count, first);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        fprintf(stderr, "recip_check: kernel failed\n");
        return 2;
    }
    unsigned long long h = 0;
    unsigned int f = 0;
    if (hipMemcpy(&h, count, sizeof h, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&f, first, sizeof f, hipMemcpyDeviceToHost) != hipSuccess) {
        fprintf(stderr, "recip_check: hipMemcpy failed\n");
        return 2;
    }
    printf("recip_check: 4294967296 float bit patterns, mismatches %llu", h);
    if (h) printf(", first 0x%08x", f);
    printf("\n");
    (void)hipFree(count);
    (void)hipFree(first);
    return h ? 1 : 0;
}
