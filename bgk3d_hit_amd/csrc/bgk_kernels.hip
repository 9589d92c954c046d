// bgk_kernels.hip -- gfx950 kernels of the BGK3D_HIT hot path.
//
//   k_collide          fused pull stream + BGK collide (SRC/col_MC.F90:104-361),
//                      one thread per site, x-contiguous tiles, periodic wrap
//                      in the index math (or slab ghost planes in z)
//   k_collide2         two time steps per pass (launchers here, kernel in
//                      bgk_collide.h); on the slab for the bulk planes
//   k_collide3         three time steps per pass; on the slab for the bulk planes
//   k_diagno           at-site moment sums (SRC/diagno.F90:73-115)
//   k_dissipation      pulled Pi^neq / kinetic energy sums
//                      (SRC/dissipation.F90:70-253)
//   k_pack_z/unpack_z  slab halo messages of the 5 crossing populations per
//                      z face (SRC/bcond.comm.step3.F90:79-143)
//   k_pack_faces       stale-halo images for the reference-compatible
//                      dissipation (periodic layout)
//   k_init_equilibrium population initialisation (SRC/init.F90:97-137)
//
// All reductions accumulate in FP64 with a fixed grid and a fixed
// combination order, so they are deterministic run to run.
#include <algorithm>
#include <type_traits>

#include "bgk_collide.h"

namespace bgk {

template <typename S, typename C>
static hipError_t collide_t(const Lattice &L, int src, const Box &box, double omega,
                            hipStream_t s) {
    if (box.sites() <= 0) return hipSuccess;
    int ntiles;
    Geo g = make_geo<S>(L, box, &ntiles);
    Pops<S> a = pops_of<S>(L, src), b = pops_of<S>(L, 1 - src);
    const C w = (C)omega, cte1 = (C)(L.shifted ? 1.0 : 0.0);
    if (offsets_fit_32(L, sizeof(S)))
        hipLaunchKernelGGL((k_collide<S, C, true>), dim3(ntiles), dim3(Tile<S>::X, Tile<S>::Y), 0,
                           s, a, b, g, w, cte1);
    else
        hipLaunchKernelGGL((k_collide<S, C, false>), dim3(ntiles), dim3(Tile<S>::X, Tile<S>::Y), 0,
                           s, a, b, g, w, cte1);
    return hipGetLastError();
}

template <typename S, typename C, bool O32, bool SH>
static hipError_t collide2_sh(const Lattice &L, int src, double omega, void *faces,
                              hipStream_t s) {
    using T = Tile2<S, C, O32>;
    const int zc = two_step_zc(L, T::X, T::Y, T::ZC);
    if (zc <= 0) return hipErrorInvalidValue;  // caller checks two_step_ok first
    int nblocks;
    Geo2 g = make_geo2(L, T::X, T::Y, zc, &nblocks);
    g.faces = faces;
    const CollideOp<C, false, SH> op{(C)omega, (C)(L.shifted ? 1.0 : 0.0), nullptr, nullptr, 0};
    if (faces)
        hipLaunchKernelGGL((k_collide2<S, C, O32, T::X, T::Y, T::MINB, false, true, SH>),
                           dim3(nblocks), dim3(collide2_threads(T::X, T::Y)), 0, s,
                           pops_of<S>(L, src), pops_of<S>(L, 1 - src), g, op);
    else
        hipLaunchKernelGGL((k_collide2<S, C, O32, T::X, T::Y, T::MINB, false, false, SH>),
                           dim3(nblocks), dim3(collide2_threads(T::X, T::Y)), 0, s,
                           pops_of<S>(L, src), pops_of<S>(L, 1 - src), g, op);
    return hipGetLastError();
}
template <typename S, typename C, bool O32>
static hipError_t collide2_o(const Lattice &L, int src, double omega, void *faces,
                             hipStream_t s) {
    return L.shifted ? collide2_sh<S, C, O32, true>(L, src, omega, faces, s)
                     : collide2_sh<S, C, O32, false>(L, src, omega, faces, s);
}
// two steps from copy `src` into copy 1-src (periodic lattice, or the bulk
// planes of a slab)
template <typename S, typename C>
static hipError_t collide2_t(const Lattice &L, int src, double omega, void *faces,
                             hipStream_t s) {
    return offsets_fit_32(L, sizeof(S)) ? collide2_o<S, C, true>(L, src, omega, faces, s)
                                        : collide2_o<S, C, false>(L, src, omega, faces, s);
}

// three steps from copy `src` into copy 1-src (periodic lattice, or the
// bulk planes 3 .. nz-4 of a slab; 32-bit offsets)
template <typename S, typename C, bool SH>
static hipError_t collide3_sh(const Lattice &L, int src, double omega, void *faces1,
                              void *faces2, hipStream_t s) {
    using T = Tile3<S, C>;
    const int zc = three_step_zc(L, T::X, T::Y, T::ZC);
    if (zc <= 0) return hipErrorInvalidValue;  // caller checks three_step_ok first
    int nblocks;
    Geo2 g = make_geo2(L, T::X, T::Y, zc, &nblocks);
    // periodic: whole tile columns first where that takes fewer plane
    // iterations (the slab keeps chunks: its boundary path on the second
    // stream needs CUs that whole-column workgroups hold for hundreds of
    // plane iterations; 1024x1024x256 measured 73.6-76.3k against
    // 75.4-77.0k MLUPS with whole columns)
    if (g.periodic_z) {
        const Split3 sp = three_step_split(L, T::X, T::Y, T::ZC);
        g.full_cols = sp.full_cols;
        g.zc = sp.zc;
        nblocks = sp.nblocks;
    }
    const C w = (C)omega, cte1 = (C)(L.shifted ? 1.0 : 0.0);
    if (faces1 || faces2) {
        if (!g.periodic_z || !faces1 || !faces2) return hipErrorInvalidValue;
        g.faces = faces1;
        g.faces2 = faces2;
        hipLaunchKernelGGL((k_collide3<S, C, T::X, T::Y, SH, false, true>), dim3(nblocks),
                           dim3(collide3_threads(T::X, T::Y)), 0, s, pops_of<S>(L, src),
                           pops_of<S>(L, 1 - src), g, w, cte1);
    } else if (g.periodic_z) {
        hipLaunchKernelGGL((k_collide3<S, C, T::X, T::Y, SH>), dim3(nblocks),
                           dim3(collide3_threads(T::X, T::Y)), 0, s, pops_of<S>(L, src),
                           pops_of<S>(L, 1 - src), g, w, cte1);
    } else {  // slab: the bulk third-step planes 3 .. nz-4
        g.zlo = 3;
        g.zhi = L.nz - 3;
        nblocks = g.tiles_x * g.tiles_y * ((g.zhi - g.zlo + zc - 1) / zc);
        hipLaunchKernelGGL((k_collide3<S, C, T::X, T::Y, SH, true>), dim3(nblocks),
                           dim3(collide3_threads(T::X, T::Y)), 0, s, pops_of<S>(L, src),
                           pops_of<S>(L, 1 - src), g, w, cte1);
    }
    return hipGetLastError();
}
template <typename S, typename C>
static hipError_t collide3_t(const Lattice &L, int src, double omega, void *faces1,
                             void *faces2, hipStream_t s) {
    return L.shifted ? collide3_sh<S, C, true>(L, src, omega, faces1, faces2, s)
                     : collide3_sh<S, C, false>(L, src, omega, faces1, faces2, s);
}

// --------------------------------------------------------------- reductions
constexpr int RED_THREADS = 256;
constexpr int FIN_THREADS = 1024;              // threads of the final combine
constexpr long long RED_TARGET_PARTS = 16384;  // partial sums per reduction

__global__ __launch_bounds__(FIN_THREADS) void k_finalize(const double *partials, int nblocks,
                                                          double *out) {
    double v[RED_NV];
#pragma unroll
    for (int k = 0; k < RED_NV; k++) v[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += FIN_THREADS) {
#pragma unroll
        for (int k = 0; k < RED_NV; k++) v[k] += partials[(size_t)b * RED_NV + k];
    }
    block_sum_store<RED_NV>(v, out);
}

// Reduction geometry: one block per (x,y) tile of the storage kind's collide
// tile shape and per chunk of zb consecutive z planes; the partial of a block
// goes to the slot of its tile, so the combination order is fixed.
struct RedGeo {
    int nx, ny, nz, nxy, zoff, periodic_z;
    int tiles_x, tiles_y, zb, nparts;
    long long ystride, zstride;  // population pitches (Lattice)
};

template <typename S>
static RedGeo make_red_geo(const Lattice &L) {
    RedGeo g;
    g.nx = L.nx;
    g.ny = L.ny;
    g.nz = L.nz;
    g.nxy = L.nx * L.ny;
    g.zoff = L.zoff;
    g.periodic_z = (L.layout == BGK_LAYOUT_PERIODIC);
    g.tiles_x = (L.nx + Tile<S>::X - 1) / Tile<S>::X;
    g.tiles_y = (L.ny + Tile<S>::Y - 1) / Tile<S>::Y;
    const long long txy = (long long)g.tiles_x * g.tiles_y;
    g.zb = (int)std::max<long long>(1, (txy * L.nz + RED_TARGET_PARTS - 1) / RED_TARGET_PARTS);
    g.nparts = (int)(txy * ((L.nz + g.zb - 1) / g.zb));
    g.ystride = L.ystride;
    g.zstride = L.zstride;
    return g;
}

int reduction_parts(const Lattice &L) {
    switch (L.storage) {
    case BGK_STORE_F64: return make_red_geo<double>(L).nparts;
    case BGK_STORE_F32:
    case BGK_STORE_F32_MIXED: return make_red_geo<float>(L).nparts;
    case BGK_STORE_F16: return make_red_geo<half_t>(L).nparts;
    default: return make_red_geo<bf16_t>(L).nparts;
    }
}

// Reduction body over the block's sites (one tile, one z chunk).
// WHAT & 1: at-site moments into vd = {sum rho, sum ux, sum uy, sum uz,
// sum |u|^2/2, uz of the last site (l,m,n)} (SRC/diagno.F90:96-111);
// WHAT & 2: pulled moments + Pi^neq into vs = {sum 2 svisc P, sum |u|^2/2}
// (SRC/dissipation.F90:76-249).  GC = 0: pulled sources wrap periodically
// (slab: the ghost planes as stored), GC = 1: sources outside the box read
// the recorded face images gx/gy/gz (reference stale-halo semantics,
// periodic layout).  Each sum runs over the same sites in the same order
// whichever parts are taken, so a fused reduction equals the separate ones
// bit for bit.
template <typename S, typename C, int WHAT, int GC>
__device__ __forceinline__ void reduce_body(const Pops<S> &P, const S *gcache, const RedGeo &g,
                                            C cte1, C svisc, double *pd, double *ps) {
    const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int tx = t % g.tiles_x;
    const int r = t / g.tiles_x;
    const int ty = r % g.tiles_y;
    const int zc = r / g.tiles_y;
    const int x = tx * Tile<S>::X + (int)threadIdx.x;
    const int y = ty * Tile<S>::Y + (int)threadIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz, nxy = g.nxy;
    double vd[RED_NV], vs[RED_NV];
#pragma unroll
    for (int k = 0; k < RED_NV; k++) vd[k] = vs[k] = 0.0;
    if (x < nx && y < ny) {
        const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
        const int ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
        const int X[3] = {xm, x, xp};
        const int Yc[3] = {ym, y, yp};
        const long long Y[3] = {ym * g.ystride, y * g.ystride, yp * g.ystride};
        const int z1 = min(nz, (zc + 1) * g.zb);
        for (int z = zc * g.zb; z < z1; z++) {
            const int zs = z + g.zoff;
            if (WHAT & 1) {
                const size_t site = (size_t)x + (size_t)Y[1] + (size_t)g.zstride * zs;
                C f[19];
#pragma unroll
                for (int q = 0; q < 19; q++) f[q] = Storage<S>::template load<C>(P.p[q][site]);
                C rho, ux, uy, uz;
                moments_site<C>(f, cte1, rho, ux, uy, uz);
                const C sk = (C)0.5 * (ux * ux + uy * uy + uz * uz);
                vd[0] += (double)rho;
                vd[1] += (double)ux;
                vd[2] += (double)uy;
                vd[3] += (double)uz;
                vd[4] += (double)sk;
                if (x == nx - 1 && y == ny - 1 && z == nz - 1) vd[5] = (double)uz;
            }
            if (WHAT & 2) {
                int zm, zp;
                if (g.periodic_z) {
                    zm = (z == 0) ? nz - 1 : z - 1;
                    zp = (z == nz - 1) ? 0 : z + 1;
                } else {
                    zm = zs - 1;
                    zp = zs + 1;
                }
                const int Zc[3] = {zm, zs, zp};
                const long long Z[3] = {zm * g.zstride, zs * g.zstride, zp * g.zstride};
                C f[19];
#pragma unroll
                for (int q = 0; q < 19; q++) {
                    S raw = P.p[q][X[1 - cx(q)] + Y[1 - cy(q)] + Z[1 - cz(q)]];
                    if (GC) {
                        // reference ghosts: a source outside the box reads the
                        // face image recorded when this copy was last exchanged
                        const int sx = x - cx(q), sy = y - cy(q), sz = z - cz(q);
                        const bool ox = (sx < 0 || sx >= nx), oy = (sy < 0 || sy >= ny);
                        const bool oz = (sz < 0 || sz >= nz);
                        if (ox | oy | oz) {
                            const int wx = X[1 - cx(q)], wy = Yc[1 - cy(q)];
                            const int wz = Zc[1 - cz(q)];
                            const S *gx = gcache;
                            const S *gy = gcache + (size_t)19 * ny * nz;
                            const S *gz = gy + (size_t)19 * nx * nz;
                            if (ox)
                                raw = gx[(size_t)q * ny * nz + wy + (size_t)ny * wz];
                            else if (oy)
                                raw = gy[(size_t)q * nx * nz + wx + (size_t)nx * wz];
                            else
                                raw = gz[(size_t)q * nxy + wx + (size_t)nx * wy];
                        }
                    }
                    f[q] = Storage<S>::template load<C>(raw);
                }
                C dsite, tsite;
                dissipation_site<C>(f, cte1, svisc, dsite, tsite);
                vs[0] += (double)dsite;
                vs[1] += (double)tsite;
            }
        }
    }
    if (WHAT & 1) block_sum_store<RED_NV>(vd, pd + (size_t)t * RED_NV);
    if (WHAT == 3) __syncthreads();  // block_sum_store's LDS is reused
    if (WHAT & 2) block_sum_store<RED_NV>(vs, ps + (size_t)t * RED_NV);
}

// waves per SIMD the reduction kernels are compiled for (the second
// argument of __launch_bounds__): four for FP32 storage and compute reading
// the state itself (at most 128 VGPRs, two 8-wave workgroups per CU: 512^3
// diagno+dissipation 2.41 against 2.67 ms), one (the compiler's choice)
// otherwise (FP16/BF16 gained nothing, FP64 and the face-image variants
// would spill)
template <typename S, typename C, int GC>
constexpr int red_waves() {
    return (std::is_same<S, float>::value && std::is_same<C, float>::value && GC == 0) ? 4 : 1;
}

template <typename S, typename C>
__global__ __launch_bounds__(Tile<S>::X * Tile<S>::Y, (red_waves<S, C, 0>())) void k_diagno(Pops<S> P, RedGeo g, C cte1,
                                                                    double *partials) {
    reduce_body<S, C, 1, 0>(P, nullptr, g, cte1, (C)0, partials, nullptr);
}
template <typename S, typename C, int GC>
__global__ __launch_bounds__(Tile<S>::X * Tile<S>::Y, (red_waves<S, C, GC>())) void k_dissipation(Pops<S> P,
                                                                         const S *gcache,
                                                                         RedGeo g, C cte1,
                                                                         C svisc,
                                                                         double *partials) {
    reduce_body<S, C, 2, GC>(P, gcache, g, cte1, svisc, nullptr, partials);
}
// both in one pass over the populations (bgk_diagnostics)
template <typename S, typename C, int GC>
__global__ __launch_bounds__(Tile<S>::X * Tile<S>::Y, (red_waves<S, C, GC>())) void k_diag_diss(Pops<S> P,
                                                                       const S *gcache,
                                                                       RedGeo g, C cte1,
                                                                       C svisc, double *pd,
                                                                       double *ps) {
    reduce_body<S, C, 3, GC>(P, gcache, g, cte1, svisc, pd, ps);
}

template <typename S, typename C>
static hipError_t diagno_t(const Lattice &L, int cur, double *partials, double *out,
                           hipStream_t s) {
    const RedGeo g = make_red_geo<S>(L);
    hipLaunchKernelGGL((k_diagno<S, C>), dim3(g.nparts), dim3(Tile<S>::X, Tile<S>::Y), 0, s,
                       pops_of<S>(L, cur), g, (C)(L.shifted ? 1.0 : 0.0), partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s, partials, g.nparts, out);
    return hipGetLastError();
}

template <typename S, typename C>
static hipError_t dissipation_t(const Lattice &L, int cur, const void *gcache, double svisc,
                                double *partials, double *out, hipStream_t s) {
    const RedGeo g = make_red_geo<S>(L);
    const C cte1 = (C)(L.shifted ? 1.0 : 0.0);
    if (gcache)
        hipLaunchKernelGGL((k_dissipation<S, C, 1>), dim3(g.nparts), dim3(Tile<S>::X, Tile<S>::Y),
                           0, s, pops_of<S>(L, cur), (const S *)gcache, g, cte1, (C)svisc,
                           partials);
    else
        hipLaunchKernelGGL((k_dissipation<S, C, 0>), dim3(g.nparts), dim3(Tile<S>::X, Tile<S>::Y),
                           0, s, pops_of<S>(L, cur), (const S *)nullptr, g, cte1, (C)svisc,
                           partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s, partials, g.nparts, out);
    return hipGetLastError();
}

template <typename S, typename C>
static hipError_t diag_diss_t(const Lattice &L, int cur, const void *gcache, double svisc,
                              double *partials, double *out, hipStream_t s) {
    const RedGeo g = make_red_geo<S>(L);
    const C cte1 = (C)(L.shifted ? 1.0 : 0.0);
    double *pd = partials, *ps = partials + (size_t)g.nparts * RED_NV;
    if (gcache)
        hipLaunchKernelGGL((k_diag_diss<S, C, 1>), dim3(g.nparts), dim3(Tile<S>::X, Tile<S>::Y),
                           0, s, pops_of<S>(L, cur), (const S *)gcache, g, cte1, (C)svisc, pd, ps);
    else
        hipLaunchKernelGGL((k_diag_diss<S, C, 0>), dim3(g.nparts), dim3(Tile<S>::X, Tile<S>::Y),
                           0, s, pops_of<S>(L, cur), (const S *)nullptr, g, cte1, (C)svisc, pd,
                           ps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s, pd, g.nparts, out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s, ps, g.nparts, out + RED_NV);
    return hipGetLastError();
}

// Two time steps from copy `src` into copy 1-src and the diagnostics of the
// result in one pass (k_collide3 with RED; periodic lattice): partials
// [2][nblocks][RED_NV], combined into out[0..8) (diagno's sums, as
// k_diag_diss's first set) and out[8..16) (dissipation's).
template <typename S, typename C, bool SH>
static hipError_t collide3_diag_sh(const Lattice &L, int src, double omega, double svisc,
                                   double *partials, double *out, hipStream_t s) {
    using T = Tile3<S, C>;
    const int zc = three_step_zc(L, T::X, T::Y, T::ZC);
    if (zc <= 0 || L.layout != BGK_LAYOUT_PERIODIC) return hipErrorInvalidValue;
    int nblocks;
    Geo2 g = make_geo2(L, T::X, T::Y, zc, &nblocks);
    const Split3 sp = three_step_split(L, T::X, T::Y, T::ZC);  // as collide3_sh
    g.full_cols = sp.full_cols;
    g.zc = sp.zc;
    nblocks = sp.nblocks;
    g.red = partials;
    g.svisc = svisc;
    const C w = (C)omega, cte1 = (C)(L.shifted ? 1.0 : 0.0);
    hipLaunchKernelGGL((k_collide3<S, C, T::X, T::Y, SH, false, false, true>), dim3(nblocks),
                       dim3(collide3_threads(T::X, T::Y)), 0, s, pops_of<S>(L, src),
                       pops_of<S>(L, 1 - src), g, w, cte1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s, partials, nblocks, out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s,
                       partials + (size_t)nblocks * RED_NV, nblocks, out + RED_NV);
    return hipGetLastError();
}
template <typename S, typename C>
static hipError_t collide3_diag_t(const Lattice &L, int src, double omega, double svisc,
                                  double *partials, double *out, hipStream_t s) {
    return L.shifted ? collide3_diag_sh<S, C, true>(L, src, omega, svisc, partials, out, s)
                     : collide3_diag_sh<S, C, false>(L, src, omega, svisc, partials, out, s);
}
template <typename S, typename C>
static int collide3_diag_parts_t(const Lattice &L) {
    using T = Tile3<S, C>;
    const int zc = three_step_zc(L, T::X, T::Y, T::ZC);
    if (zc <= 0) return 0;
    return three_step_split(L, T::X, T::Y, T::ZC).nblocks;  // collide3_diag_sh's grid
}

// ------------------------------------------------------------- checksum
// Order-independent 64-bit hash of the raw stored bits of the interior
// populations: sum over (q, site) of splitmix64(bits ^ key(q, site)) mod 2^64.
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
template <typename S>
__device__ __forceinline__ unsigned long long raw_bits(S v) {
    if constexpr (sizeof(S) == 8) return __builtin_bit_cast(unsigned long long, v);
    else if constexpr (sizeof(S) == 4) return __builtin_bit_cast(unsigned int, v);
    else return __builtin_bit_cast(unsigned short, v);
}

template <typename S>
__global__ __launch_bounds__(RED_THREADS) void k_checksum(Pops<S> P, long long nsite,
                                                          SiteMap m,
                                                          unsigned long long *partials) {
    unsigned long long h = 0;
    for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < nsite;
         i += (long long)gridDim.x * RED_THREADS) {
        const long long o = site_elem(m, i);
#pragma unroll
        for (int q = 0; q < 19; q++)
            h += mix64(raw_bits(P.p[q][o]) ^ ((unsigned long long)q << 58) ^
                       ((unsigned long long)i * 0xD1B54A32D192ED03ull));
    }
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
    __shared__ unsigned long long sm[RED_THREADS / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < RED_THREADS / 64; w++) t += sm[w];
        partials[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_checksum_final(const unsigned long long *partials,
                                                                int nblocks,
                                                                unsigned long long *out) {
    unsigned long long h = 0;
    for (int b = threadIdx.x; b < nblocks; b += RED_THREADS) h += partials[b];
    for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
    __shared__ unsigned long long sm[RED_THREADS / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = h;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < RED_THREADS / 64; w++) t += sm[w];
        *out = t;
    }
}

template <typename S>
static hipError_t checksum_t(const Lattice &L, int cur, void *partials, int nblocks, void *out,
                             hipStream_t s) {
    hipLaunchKernelGGL((k_checksum<S>), dim3(nblocks), dim3(RED_THREADS), 0, s, pops_of<S>(L, cur),
                       L.nsite, site_map(L), (unsigned long long *)partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_checksum_final, dim3(1), dim3(RED_THREADS), 0, s,
                       (const unsigned long long *)partials, nblocks, (unsigned long long *)out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- halo path
// crossing populations (0-based) per z face, SRC/bcond.comm.step3.F90:82-140
__constant__ int c_zplus[5] = {3, 5, 6, 12, 17};   // 04 06 07 13 18: leave through +z
__constant__ int c_zminus[5] = {1, 8, 10, 14, 15}; // 02 09 11 15 16: leave through -z

// message layout: [face][5][nx*ny], face 0 = the +z message (stored plane
// `top`: the top interior plane), face 1 = the -z message (stored plane
// `bot`: the bottom interior plane)
template <typename S>
__global__ __launch_bounds__(256) void k_pack_z(Pops<S> P, int nx, int nxy, long long ystride,
                                                long long zstride, int top, int bot, S *send) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nxy) return;
    const int q5 = blockIdx.y;  // 0..4
    const size_t o = (size_t)(p % nx) + (size_t)(p / nx) * ystride;
    send[(size_t)(0 * 5 + q5) * nxy + p] = P.p[c_zplus[q5]][(size_t)top * zstride + o];
    send[(size_t)(1 * 5 + q5) * nxy + p] = P.p[c_zminus[q5]][(size_t)bot * zstride + o];
}

// the +z message fills the bottom ghost plane `lo`, the -z message the top
// ghost plane `hi` (SRC/bcond.comm.step3.F90:97-106, 133-142)
template <typename S>
__global__ __launch_bounds__(256) void k_unpack_z(Pops<S> P, int nx, int nxy, long long ystride,
                                                  long long zstride, int lo, int hi,
                                                  const S *recv) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nxy) return;
    const int q5 = blockIdx.y;
    const size_t o = (size_t)(p % nx) + (size_t)(p / nx) * ystride;
    P.p[c_zplus[q5]][(size_t)lo * zstride + o] = recv[(size_t)(0 * 5 + q5) * nxy + p];
    P.p[c_zminus[q5]][(size_t)hi * zstride + o] = recv[(size_t)(1 * 5 + q5) * nxy + p];
}

template <typename S>
static Pops<S> pops_at(void *const *base) {
    Pops<S> P;
    for (int q = 0; q < 19; q++) P.p[q] = (S *)base[q];
    return P;
}
template <typename S>
static hipError_t pack_planes_t(const Lattice &L, void *const *pops, int top, int bot,
                                void *send, hipStream_t s) {
    const int nxy = L.nx * L.ny;
    hipLaunchKernelGGL((k_pack_z<S>), dim3((nxy + 255) / 256, 5), dim3(256), 0, s,
                       pops_at<S>(pops), L.nx, nxy, L.ystride, L.zstride, top, bot, (S *)send);
    return hipGetLastError();
}
template <typename S>
static hipError_t unpack_planes_t(const Lattice &L, void *const *pops, int lo, int hi,
                                  const void *recv, hipStream_t s) {
    const int nxy = L.nx * L.ny;
    hipLaunchKernelGGL((k_unpack_z<S>), dim3((nxy + 255) / 256, 5), dim3(256), 0, s,
                       pops_at<S>(pops), L.nx, nxy, L.ystride, L.zstride, lo, hi,
                       (const S *)recv);
    return hipGetLastError();
}
template <typename S>
static hipError_t pack_z_t(const Lattice &L, int cur, void *send, hipStream_t s) {
    return pack_planes_t<S>(L, L.pop[cur], L.nz, 1, send, s);
}
template <typename S>
static hipError_t unpack_z_t(const Lattice &L, int cur, const void *recv, hipStream_t s) {
    return unpack_planes_t<S>(L, L.pop[cur], 0, L.nz + 1, recv, s);
}

// one step over the slab box `box` from populations `src` into `dst` (both
// in the lattice's layout), where interior plane z is stored at plane
// z + src_off (z + dst_off): the slab's two-step boundary path, which takes
// single steps through a small plane buffer
template <typename S, typename C>
static hipError_t collide_planes_t(const Lattice &L, void *const *src, int src_off,
                                   void *const *dst, int dst_off, const Box &box, double omega,
                                   hipStream_t s) {
    if (box.sites() <= 0) return hipSuccess;
    int ntiles;
    Geo g = make_geo<S>(L, box, &ntiles);
    g.zoff = 0;
    g.periodic_z = 0;
    Pops<S> a, b;
    for (int q = 0; q < 19; q++) {
        a.p[q] = (S *)src[q] + (ptrdiff_t)src_off * (ptrdiff_t)L.zstride;
        b.p[q] = (S *)dst[q] + (ptrdiff_t)dst_off * (ptrdiff_t)L.zstride;
    }
    // the shifted pointers may lie below an allocation; the kernel adds the
    // (signed) plane offset to them and only the in-plane part is an offset
    if (offsets_fit_32(L, sizeof(S)))
        hipLaunchKernelGGL((k_collide<S, C, true>), dim3(ntiles), dim3(Tile<S>::X, Tile<S>::Y), 0,
                           s, a, b, g, (C)omega, (C)(L.shifted ? 1.0 : 0.0));
    else
        hipLaunchKernelGGL((k_collide<S, C, false>), dim3(ntiles), dim3(Tile<S>::X, Tile<S>::Y), 0,
                           s, a, b, g, (C)omega, (C)(L.shifted ? 1.0 : 0.0));
    return hipGetLastError();
}

// face images of the periodic layout: gx[q] = plane x = (c_x>0 ? nx-1 : 0)
// of population q for the 10 populations with c_x != 0, likewise gy, gz.
template <typename S>
__global__ __launch_bounds__(256) void k_pack_faces(Pops<S> P, int nx, int ny, int nz,
                                                    long long ys, long long zs, S *g) {
    const int axis = blockIdx.y;
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    const int nxy = nx * ny;
    S *gx = g;
    S *gy = g + (size_t)19 * ny * nz;
    S *gz = gy + (size_t)19 * nx * nz;
    if (axis == 0) {
        if (p >= (long long)ny * nz) return;
        const int y = (int)(p % ny), z = (int)(p / ny);
#pragma unroll
        for (int q = 0; q < 18; q++) {
            if (cx(q) == 0) continue;
            const int x = cx(q) > 0 ? nx - 1 : 0;
            gx[(size_t)q * ny * nz + p] = P.p[q][x + ys * y + zs * z];
        }
    } else if (axis == 1) {
        if (p >= (long long)nx * nz) return;
        const int x = (int)(p % nx), z = (int)(p / nx);
#pragma unroll
        for (int q = 0; q < 18; q++) {
            if (cy(q) == 0) continue;
            const int y = cy(q) > 0 ? ny - 1 : 0;
            gy[(size_t)q * nx * nz + p] = P.p[q][x + ys * y + zs * z];
        }
    } else {
        if (p >= (long long)nxy) return;
#pragma unroll
        for (int q = 0; q < 18; q++) {
            if (cz(q) == 0) continue;
            const int z = cz(q) > 0 ? nz - 1 : 0;
            gz[(size_t)q * nxy + p] = P.p[q][p % nx + ys * (p / nx) + zs * z];
        }
    }
}

template <typename S>
static hipError_t pack_faces_t(const Lattice &L, int cur, void *g, hipStream_t s) {
    long long mx = (long long)L.ny * L.nz;
    if ((long long)L.nx * L.nz > mx) mx = (long long)L.nx * L.nz;
    if ((long long)L.nx * L.ny > mx) mx = (long long)L.nx * L.ny;
    hipLaunchKernelGGL((k_pack_faces<S>), dim3((unsigned)((mx + 255) / 256), 3), dim3(256), 0, s,
                       pops_of<S>(L, cur), L.nx, L.ny, L.nz, L.ystride, L.zstride, (S *)g);
    return hipGetLastError();
}

// ------------------------------------------------------- transfers and init
// Host transfers go through a staging buffer in chunks of `count` sites
// starting at interior site i0.
template <typename S, typename C>
__global__ __launch_bounds__(256) void k_plane_from_f64(S *dst, SiteMap m, long long i0,
                                                        const double *src, long long count) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[site_elem(m, i0 + i)] = Storage<S>::template store<C>((C)src[i]);
}
template <typename S, typename C>
__global__ __launch_bounds__(256) void k_plane_to_f64(const S *src, SiteMap m, long long i0,
                                                      double *dst, long long count) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = (double)Storage<S>::template load<C>(src[site_elem(m, i0 + i)]);
}

template <typename S, typename C>
static hipError_t from_f64_t(const Lattice &L, int copy, int q, const double *src, long long i0,
                             long long count, hipStream_t s) {
    hipLaunchKernelGGL((k_plane_from_f64<S, C>), dim3((unsigned)((count + 255) / 256)),
                       dim3(256), 0, s, (S *)L.pop[copy][q], site_map(L), i0, src, count);
    return hipGetLastError();
}
template <typename S, typename C>
static hipError_t to_f64_t(const Lattice &L, int copy, int q, double *dst, long long i0,
                           long long count, hipStream_t s) {
    hipLaunchKernelGGL((k_plane_to_f64<S, C>), dim3((unsigned)((count + 255) / 256)),
                       dim3(256), 0, s, (const S *)L.pop[copy][q], site_map(L), i0, dst, count);
    return hipGetLastError();
}

// SRC/init.F90:97-137 with crho = 1: populations of the equilibrium of u.
// u holds 3 planes (ux, uy, uz) of `ld` values for sites i0 .. i0+count-1.
template <typename S, typename C>
__global__ __launch_bounds__(256) void k_init_equilibrium(Pops<S> P, const double *u,
                                                          long long ld, long long count,
                                                          SiteMap m, long long i0, C cte1) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const C rf = 3, qf = 1.5, tre = 3, crho = 1, zero = 0;
    const C p0 = K<C>::p0, p1 = K<C>::p1, p2 = K<C>::p2;
    const C xj = (C)u[i], yj = (C)u[ld + i], zj = (C)u[2 * ld + i];
    const C cvsq = xj * xj + yj * yj + zj * zj;
    C c[19];
    c[0] = rf * (xj - yj) + qf * (tre * (xj - yj) * (xj - yj) - cvsq);
    c[1] = rf * (xj - zj) + qf * (tre * (xj - zj) * (xj - zj) - cvsq);
    c[2] = rf * (xj + yj) + qf * (tre * (xj + yj) * (xj + yj) - cvsq);
    c[3] = rf * (xj + zj) + qf * (tre * (xj + zj) * (xj + zj) - cvsq);
    c[4] = rf * (xj) + qf * (tre * (xj) * (xj)-cvsq);
    c[5] = rf * (zj) + qf * (tre * (zj) * (zj)-cvsq);
    c[6] = rf * (yj + zj) + qf * (tre * (yj + zj) * (yj + zj) - cvsq);
    c[7] = rf * (yj) + qf * (tre * (yj) * (yj)-cvsq);
    c[8] = rf * (yj - zj) + qf * (tre * (yj - zj) * (yj - zj) - cvsq);
    c[9] = rf * (-xj - yj) + qf * (tre * (xj + yj) * (xj + yj) - cvsq);
    c[10] = rf * (-xj - zj) + qf * (tre * (xj + zj) * (xj + zj) - cvsq);
    c[11] = rf * (-xj + yj) + qf * (tre * (xj - yj) * (xj - yj) - cvsq);
    c[12] = rf * (-xj + zj) + qf * (tre * (xj - zj) * (xj - zj) - cvsq);
    c[13] = rf * (-xj) + qf * (tre * (xj) * (xj)-cvsq);
    c[14] = rf * (-zj) + qf * (tre * (zj) * (zj)-cvsq);
    c[15] = rf * (-yj - zj) + qf * (tre * (yj + zj) * (yj + zj) - cvsq);
    c[16] = rf * (-yj) + qf * (tre * (yj) * (yj)-cvsq);
    c[17] = rf * (-yj + zj) + qf * (tre * (yj - zj) * (yj - zj) - cvsq);
    c[18] = rf * (zero) + qf * (tre * (zero) * (zero)-cvsq);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const C w = (q == 18) ? p0 : ((cx(q) * cx(q) + cy(q) * cy(q) + cz(q) * cz(q)) == 1 ? p1 : p2);
        P.p[q][site_elem(m, i0 + i)] = Storage<S>::template store<C>(crho * w * (cte1 + c[q]));
    }
}

template <typename S, typename C>
static hipError_t init_eq_t(const Lattice &L, int copy, const double *u, long long ld,
                            long long i0, long long count, hipStream_t s) {
    // init.F90 uses cte1 = 0 when shifted (SRC/init.F90:65-69)
    hipLaunchKernelGGL((k_init_equilibrium<S, C>), dim3((unsigned)((count + 255) / 256)),
                       dim3(256), 0, s, pops_of<S>(L, copy), u, ld, count, site_map(L), i0,
                       (C)(L.shifted ? 0.0 : 1.0));
    return hipGetLastError();
}

// ------------------------------------------------------------ box gather
// populations of the box [x0,x1) x [y0,y1) x [z0,z1) as doubles,
// out[q][z][y][x] (x fastest), for host-side probes and slices
template <typename S, typename C>
__global__ __launch_bounds__(256) void k_gather_box(Pops<S> P, Box b, long long ys, long long zs,
                                                    int zoff, double *out) {
    const long long bx = b.x1 - b.x0, by = b.y1 - b.y0, bz = b.z1 - b.z0;
    const long long n = bx * by * bz;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int x = b.x0 + (int)(i % bx), y = b.y0 + (int)((i / bx) % by),
              z = b.z0 + (int)(i / (bx * by));
    const size_t site = (size_t)x + (size_t)ys * y + (size_t)zs * (z + zoff);
#pragma unroll
    for (int q = 0; q < 19; q++)
        out[(size_t)q * n + i] = (double)Storage<S>::template load<C>(P.p[q][site]);
}

template <typename S, typename C>
static hipError_t gather_box_t(const Lattice &L, int cur, const Box &b, double *out,
                               hipStream_t s) {
    const long long n = b.sites();
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_gather_box<S, C>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       pops_of<S>(L, cur), b, L.ystride, L.zstride, L.zoff, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ kind dispatch
#define BGK_DISPATCH(storage, FN, ...)                                   \
    switch (storage) {                                                   \
    case BGK_STORE_F64: return FN<double, double>(__VA_ARGS__);          \
    case BGK_STORE_F32: return FN<float, float>(__VA_ARGS__);            \
    case BGK_STORE_F32_MIXED: return FN<float, double>(__VA_ARGS__);     \
    case BGK_STORE_F16: return FN<half_t, float>(__VA_ARGS__);           \
    case BGK_STORE_BF16: return FN<bf16_t, float>(__VA_ARGS__);          \
    default: return hipErrorInvalidValue;                                \
    }
#define BGK_DISPATCH_S(storage, FN, ...)                                 \
    switch (storage) {                                                   \
    case BGK_STORE_F64: return FN<double>(__VA_ARGS__);                  \
    case BGK_STORE_F32:                                                  \
    case BGK_STORE_F32_MIXED: return FN<float>(__VA_ARGS__);             \
    case BGK_STORE_F16: return FN<half_t>(__VA_ARGS__);                  \
    case BGK_STORE_BF16: return FN<bf16_t>(__VA_ARGS__);                 \
    default: return hipErrorInvalidValue;                                \
    }

hipError_t launch_collide(const Lattice &L, int src, const Box &box, double omega,
                          hipStream_t s) {
    BGK_DISPATCH(L.storage, collide_t, L, src, box, omega, s);
}
template <typename S, typename C>
static bool two_step_ok_t(const Lattice &L) {
    if (!L.two_step) return false;
    // slab: boundary planes 0..2 and nz-3..nz-1 must be distinct
    if (L.layout == BGK_LAYOUT_SLAB_Z && L.nz < 6) return false;
    return offsets_fit_32(L, sizeof(S))
               ? two_step_zc(L, Tile2<S, C, true>::X, Tile2<S, C, true>::Y,
                             Tile2<S, C, true>::ZC) > 0
               : two_step_zc(L, Tile2<S, C, false>::X, Tile2<S, C, false>::Y,
                             Tile2<S, C, false>::ZC) > 0;
}
bool two_step_ok(const Lattice &L) {
    switch (L.storage) {
    case BGK_STORE_F64: return two_step_ok_t<double, double>(L);
    case BGK_STORE_F32: return two_step_ok_t<float, float>(L);
    case BGK_STORE_F32_MIXED: return two_step_ok_t<float, double>(L);
    case BGK_STORE_F16: return two_step_ok_t<half_t, float>(L);
    case BGK_STORE_BF16: return two_step_ok_t<bf16_t, float>(L);
    default: return false;
    }
}
template <typename S, typename C>
static bool three_step_ok_t(const Lattice &L) {
    using T = Tile3<S, C>;
    return offsets_fit_32(L, sizeof(S)) && three_step_zc(L, T::X, T::Y, T::ZC) > 0;
}
bool three_step_ok(const Lattice &L) {
    if (!L.two_step || !L.three_step) return false;
    switch (L.storage) {
    case BGK_STORE_F64: return three_step_ok_t<double, double>(L);
    case BGK_STORE_F32: return three_step_ok_t<float, float>(L);
    case BGK_STORE_F32_MIXED: return three_step_ok_t<float, double>(L);
    case BGK_STORE_F16: return three_step_ok_t<half_t, float>(L);
    case BGK_STORE_BF16: return three_step_ok_t<bf16_t, float>(L);
    default: return false;
    }
}
hipError_t launch_collide3(const Lattice &L, int src, double omega, void *faces1, void *faces2,
                           hipStream_t s) {
    BGK_DISPATCH(L.storage, collide3_t, L, src, omega, faces1, faces2, s);
}
hipError_t launch_collide3_diag(const Lattice &L, int src, double omega, double svisc,
                                double *partials, double *out, hipStream_t s) {
    BGK_DISPATCH(L.storage, collide3_diag_t, L, src, omega, svisc, partials, out, s);
}
int collide3_diag_parts(const Lattice &L) {
    switch (L.storage) {
    case BGK_STORE_F64: return collide3_diag_parts_t<double, double>(L);
    case BGK_STORE_F32: return collide3_diag_parts_t<float, float>(L);
    case BGK_STORE_F32_MIXED: return collide3_diag_parts_t<float, double>(L);
    case BGK_STORE_F16: return collide3_diag_parts_t<half_t, float>(L);
    case BGK_STORE_BF16: return collide3_diag_parts_t<bf16_t, float>(L);
    default: return 0;
    }
}
hipError_t launch_collide2(const Lattice &L, int src, double omega, void *faces,
                           hipStream_t s) {
    BGK_DISPATCH(L.storage, collide2_t, L, src, omega, faces, s);
}
hipError_t launch_diagno(const Lattice &L, int cur, double *partials, double *out,
                         hipStream_t s) {
    BGK_DISPATCH(L.storage, diagno_t, L, cur, partials, out, s);
}
hipError_t launch_dissipation(const Lattice &L, int cur, const void *gcache, double svisc,
                              double *partials, double *out, hipStream_t s) {
    BGK_DISPATCH(L.storage, dissipation_t, L, cur, gcache, svisc, partials, out, s);
}
hipError_t launch_diag_diss(const Lattice &L, int cur, const void *gcache, double svisc,
                            double *partials, double *out, hipStream_t s) {
    BGK_DISPATCH(L.storage, diag_diss_t, L, cur, gcache, svisc, partials, out, s);
}
hipError_t launch_pack_planes(const Lattice &L, void *const *pops, int top, int bot, void *send,
                              hipStream_t s) {
    BGK_DISPATCH_S(L.storage, pack_planes_t, L, pops, top, bot, send, s);
}
hipError_t launch_unpack_planes(const Lattice &L, void *const *pops, int lo, int hi,
                                const void *recv, hipStream_t s) {
    BGK_DISPATCH_S(L.storage, unpack_planes_t, L, pops, lo, hi, recv, s);
}
hipError_t launch_collide_planes(const Lattice &L, void *const *src, int src_off,
                                 void *const *dst, int dst_off, const Box &box, double omega,
                                 hipStream_t s) {
    BGK_DISPATCH(L.storage, collide_planes_t, L, src, src_off, dst, dst_off, box, omega, s);
}
hipError_t launch_pack_z(const Lattice &L, int cur, void *send, hipStream_t s) {
    BGK_DISPATCH_S(L.storage, pack_z_t, L, cur, send, s);
}
hipError_t launch_unpack_z(const Lattice &L, int cur, const void *recv, hipStream_t s) {
    BGK_DISPATCH_S(L.storage, unpack_z_t, L, cur, recv, s);
}
hipError_t launch_pack_faces(const Lattice &L, int cur, void *gcache, hipStream_t s) {
    BGK_DISPATCH_S(L.storage, pack_faces_t, L, cur, gcache, s);
}
hipError_t launch_plane_from_f64(const Lattice &L, int copy, int q, const double *src,
                                 long long i0, long long count, hipStream_t s) {
    BGK_DISPATCH(L.storage, from_f64_t, L, copy, q, src, i0, count, s);
}
hipError_t launch_plane_to_f64(const Lattice &L, int copy, int q, double *dst, long long i0,
                               long long count, hipStream_t s) {
    BGK_DISPATCH(L.storage, to_f64_t, L, copy, q, dst, i0, count, s);
}
hipError_t launch_init_equilibrium(const Lattice &L, int copy, const double *u, long long ld,
                                   long long i0, long long count, hipStream_t s) {
    BGK_DISPATCH(L.storage, init_eq_t, L, copy, u, ld, i0, count, s);
}
hipError_t launch_checksum(const Lattice &L, int cur, void *partials, int nblocks, void *out,
                           hipStream_t s) {
    BGK_DISPATCH_S(L.storage, checksum_t, L, cur, partials, nblocks, out, s);
}
hipError_t launch_gather_box(const Lattice &L, int cur, const Box &b, double *out,
                             hipStream_t s) {
    BGK_DISPATCH(L.storage, gather_box_t, L, cur, b, out, s);
}
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s) {
    return hipMemsetAsync(p, 0, bytes, s);
}

}  // namespace bgk
