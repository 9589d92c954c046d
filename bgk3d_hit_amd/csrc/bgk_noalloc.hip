// bgk_noalloc.hip -- bgk_diagnostics_noalloc: the diagno + dissipation pass
// of bgk_diagnostics on populations the caller owns, with no allocation.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The
// reduction partials and the combined sums live in `scratch`, a file-scope
// device array of fixed size; every call on a device shares that device's
// copy of it, so calls that overlap on one GPU race.  The per-block
// partials are written slot by slot and the combine reads exactly the slots
// the pass wrote, so nothing in `scratch` has to start at zero and no
// memset is issued (the call is graph-capturable as it stands).
#include <mutex>

#include "bgk_internal.h"

namespace bgk {

// reduction slots: two sets (diagno, dissipation) of 8 doubles per block,
// then the 16 combined sums
constexpr long long NOALLOC_PARTS = BGK_NOALLOC_MAX_PARTS;
constexpr long long NOALLOC_DOUBLES = 2 * 8 * NOALLOC_PARTS + 16;

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// the device address of `scratch` on each device (the array's name is a
// host-side stand-in, not an address a kernel may use)
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

static int scratch_of_current_device(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(BGK_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(BGK_EHIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BGK_OK;
}

// bgk_diagnostics' host-side finish (diagno's divisions, SRC/diagno.F90:117-121,
// and the two dissipation sums) written to device memory
__global__ void k_noalloc_finish(const double *sums, int nx, int ny, int nz, double *out) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 5; k++) out[k] = ((sums[k] / (double)nx) / (double)ny) / (double)nz;
    out[5] = sums[8];
    out[6] = sums[9];
}

}  // namespace bgk

using namespace bgk;

extern "C" int bgk_diagnostics_noalloc(const void *pop, int lx, int ly, int lz, int storage,
                                       int flags, double svisc, double *out, void *stream) {
    if (!pop || !out) return fail(BGK_EINVAL, "null population or output pointer");
    if (lx < 1 || ly < 1 || lz < 1) return fail(BGK_EINVAL, "lattice sizes must be positive");
    if (storage_bytes(storage) == 0) return fail(BGK_EINVAL, "unknown storage kind");
    if (flags & ~BGK_FLAG_NOSHIFT)
        return fail(BGK_EINVAL, "bgk_diagnostics_noalloc takes BGK_FLAG_NOSHIFT only");
    if ((long long)lx * ly * lz >= (1LL << 31))
        return fail(BGK_EINVAL, "more than 2^31 sites per population");
    Lattice L;
    lattice_geometry(L, lx, ly, lz, storage, BGK_LAYOUT_PERIODIC, flags, 0);
    const int parts = reduction_parts(L);
    if (parts > NOALLOC_PARTS)
        return fail(BGK_EINVAL, "lattice needs " + std::to_string(parts) +
                                    " reduction slots, more than the " +
                                    std::to_string(NOALLOC_PARTS) +
                                    " of the allocation-free scratch (use bgk_diagnostics)");
    double *sp = nullptr;
    int r = scratch_of_current_device(&sp);
    if (r) return r;
    const size_t es = storage_bytes(storage);
    for (int q = 0; q < 19; q++) L.pop[0][q] = (char *)pop + (size_t)q * lx * es;
    hipStream_t s = (hipStream_t)stream;
    double *partials = sp, *sums = sp + 2 * 8 * NOALLOC_PARTS;
    BGK_HIP(launch_diag_diss(L, 0, nullptr, svisc, partials, sums, s));
    hipLaunchKernelGGL(k_noalloc_finish, dim3(1), dim3(64), 0, s, sums, lx, ly, lz, out);
    BGK_HIP(hipGetLastError());
    return BGK_OK;
}
