// bgk_collide.h -- the fused pull stream + BGK collide kernels
// (SRC/col_MC.F90:104-361): k_collide, one time step per pass, and the
// two-step pass (two_step_pass / k_collide2), which advances two time steps
// per pass over HBM.  Shared by the time-step path (bgk_kernels.hip) and the
// adjoint (bgk_adjoint.hip: taping forward, and the same two-step pass body
// with the adjoint step as its per-site operation).
#pragma once

#include <type_traits>

#include "bgk_internal.h"
#include "d3q19.h"

namespace bgk {

template <typename S>
struct Pops {
    S *p[19];
};

// Reduction partials: 8 doubles per workgroup slot.
constexpr int RED_NV = 8;

// Reduce NV doubles over the block (any size up to 1024 threads, all threads
// taking part) and write them to out[0..NV).
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double *out) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] += __shfl_xor(v[k], off);
    }
    __shared__ double sm[16][NV];
    const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);
    const int nwaves = (int)(blockDim.x * blockDim.y) / 64;
    const int lane = tid & 63, wave = tid >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) sm[wave][k] = v[k];
    }
    __syncthreads();
    if (tid < NV) {
        double s = 0.0;
        for (int w = 0; w < nwaves; w++) s += sm[w][tid];
        out[tid] = s;
    }
}

template <typename S>
static Pops<S> pops_of(const Lattice &L, int copy) {
    Pops<S> P;
    for (int q = 0; q < 19; q++) P.p[q] = (S *)L.pop[copy][q];
    return P;
}

// Tile (workgroup) shape per storage kind, from a sweep of 32x8 .. 256x1
// tiles at 512^3 and 1024^3: a wider x extent leaves fewer partially used
// 128-B lines at the tile edges for the 10 x-shifted populations.
template <typename S>
struct Tile {
    static constexpr int X = 128, Y = 4;  // float
};
template <>
struct Tile<double> {
    static constexpr int X = 128, Y = 2;
};
template <>
struct Tile<half_t> {
    static constexpr int X = 256, Y = 1;
};
template <>
struct Tile<bf16_t> {
    static constexpr int X = 256, Y = 1;
};

// Launch geometry of a box of sites in TX x TY tiles (one z plane deep).
// Population q of stored site (x, y, zs) is element x + y ystride + zs zstride
// of its population pointer.
struct Geo {
    int nx, ny, nz, nxy, zoff, periodic_z;
    int tiles_x, tiles_y;
    int x0, x1, y0, y1, z0;
    long long ystride, zstride;
};

template <typename S>
static Geo make_geo(const Lattice &L, const Box &b, int *ntiles) {
    Geo g;
    g.nx = L.nx;
    g.ny = L.ny;
    g.nz = L.nz;
    g.nxy = L.nx * L.ny;
    g.zoff = L.zoff;
    g.periodic_z = (L.layout == BGK_LAYOUT_PERIODIC);
    g.tiles_x = (b.x1 - b.x0 + Tile<S>::X - 1) / Tile<S>::X;
    g.tiles_y = (b.y1 - b.y0 + Tile<S>::Y - 1) / Tile<S>::Y;
    g.x0 = b.x0;
    g.x1 = b.x1;
    g.y0 = b.y0;
    g.y1 = b.y1;
    g.z0 = b.z0;
    g.ystride = L.ystride;
    g.zstride = L.zstride;
    *ntiles = g.tiles_x * g.tiles_y * (b.z1 - b.z0);
    return g;
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Remap so that each
// XCD works on one contiguous range of tiles: x-adjacent tiles then share an
// L2 and the cache line that straddles their boundary is fetched once.
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int q = G >> 3, r = G & 7;
    const int xcd = b & 7, slot = b >> 3;
    return (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Block order of the two-step pass, where each workgroup fills a CU: the
// workgroups that run at the same time (one round: gsz per XCD on 8 XCDs)
// take 8*gsz consecutive tiles, XCD k the k-th group of gsz.  The groups of
// one round are neighbours in y on different XCDs and fetch the halo rows
// they share at the same time; the next round continues with the next 8*gsz
// tiles.  Measured against xcd_remap (each XCD one contiguous eighth of the
// launch): 2.5 % faster in FP32, 1 % in FP16 at 512^3.  The workgroups past
// the last full round (all of them on a lattice of less than one round, as
// 128^3) are spread as by xcd_remap, each XCD one contiguous range of the
// remaining tiles: in launch order x- and y-adjacent tiles would sit on
// different XCDs and each fetch the rows they share (128^3: 67 floats per
// site fetched per three-step pass, 19 needed).
__device__ __forceinline__ int xcd_round_remap(int b, int G, int gsz) {
    const int full = (G / (8 * gsz)) * 8 * gsz;
    if (b >= full) return full + xcd_remap(b - full, G - full);
    const int k = b & 7, slot = b >> 3;
    return ((slot / gsz) * 8 + k) * gsz + slot % gsz;
}

// Element access by byte offset from a per-population base pointer.  The
// kernels add a site's z plane to the base (uniform across the workgroup, so
// the sum stays in scalar registers) and pass the in-plane part as the
// offset.  With O32 (the in-plane offsets fit 32 bits) the offset is a
// 32-bit unsigned value: each load or store takes one vector offset register
// (global_load saddr form) instead of a 64-bit address per population.
template <bool O32>
using off_t_ = typename std::conditional<O32, uint32_t, size_t>::type;

// The base is a global-address-space pointer (plane_ptr); the element moves
// as an unsigned integer of its size.
template <int N>
struct uint_of;
template <>
struct uint_of<2> {
    using type = uint16_t;
};
template <>
struct uint_of<4> {
    using type = uint32_t;
};
template <>
struct uint_of<8> {
    using type = uint64_t;
};
template <typename S, typename O>
__device__ __forceinline__ S ld_at(__attribute__((address_space(1))) S *base, O off) {
    using R = typename uint_of<sizeof(S)>::type;
    const R r = *(__attribute__((address_space(1))) const R *)(
        (__attribute__((address_space(1))) const char *)base + off);
    return __builtin_bit_cast(S, r);
}
template <typename S, typename O>
__device__ __forceinline__ void st_at(__attribute__((address_space(1))) S *base, O off, S v) {
    using R = typename uint_of<sizeof(S)>::type;
    *(__attribute__((address_space(1))) R *)((__attribute__((address_space(1))) char *)base +
                                             off) = __builtin_bit_cast(R, v);
}

// p + e for an element offset e that is uniform across the workgroup (a z
// plane), kept in scalar registers: the asm makes the sum opaque, so the
// compiler cannot fold a lane's in-plane offset into it (which would give
// every access a 64-bit vector address).
template <typename S>
using gptr_t = __attribute__((address_space(1))) S *;
template <typename S>
__device__ __forceinline__ gptr_t<S> plane_ptr(S *p, ptrdiff_t e) {
    gptr_t<S> r = (gptr_t<S>)(p + e);
    asm volatile("" : "+s"(r));
    return r;
}

// Raw buffer access (buffer_load/store with a 128-bit resource in scalar
// registers, a 32-bit vector offset and a 32-bit scalar offset): used where
// the 19 populations lie at equal byte distances from population 0 (the
// interleaved layout), so one resource per z plane serves all of them and
// the per-population distance is a loop-invariant scalar offset.  Stride 0,
// 2^32 - 1 records, 32-bit data format (word 3 of the CDNA descriptor).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, -1, 0x00020000);
}
template <typename S>
__device__ __forceinline__ S buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(S) == 2)
        return __builtin_bit_cast(S, (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
    else if constexpr (sizeof(S) == 4)
        return __builtin_bit_cast(S, (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    else
        return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// Cache policy of the population stores: the sc1 bit (bit 4 of the
// auxiliary operand).  A pass never reads back what it stores, and with sc1
// the stores displace less of the L2's read working set (the halo rows and
// planes that neighbouring workgroups read again): measured on the
// three-step pass at 512^3 FP32, read traffic 16.2 -> 14.0 GB per launch
// and 1-3% less time at FP32 and FP64; values are unchanged.  nt (bit 1)
// cut reads further (13.2 GB) but cost 4% time.
constexpr int POP_STORE_POLICY = 16;
template <typename S>
__device__ __forceinline__ void buf_st(S v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                       uint32_t soff) {
    constexpr int A = POP_STORE_POLICY;
    if constexpr (sizeof(S) == 2)
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, voff, soff, A);
    else if constexpr (sizeof(S) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, A);
    else
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(unsigned int __attribute__((ext_vector_type(2))), v), r, voff,
            soff, A);
}

// whether the in-plane byte offsets (x + y ystride, plus the population
// offsets of the slab plane buffer's pointers) fit 32 bits
inline bool offsets_fit_32(const Lattice &L, size_t elem_bytes) {
    return !L.wide_offsets && (unsigned long long)L.zstride * elem_bytes <= (1ULL << 32);
}

// ------------------------------------------------------------------ collide
// TAPE: also write the pulled-moment velocity u = j/rho of every site to
// tape[0..2][site] (the adjoint's record of the forward pass).
template <typename S, typename C, bool O32, bool TAPE = false>
__global__ __launch_bounds__(Tile<S>::X * Tile<S>::Y) void k_collide(Pops<S> src, Pops<S> dst, Geo g, C omega,
                                                 C cte1, C *__restrict__ tape = nullptr,
                                                 long long tape_ld = 0) {
    using O = off_t_<O32>;
    constexpr O es = sizeof(S);
    const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int tx = t % g.tiles_x;
    const int r = t / g.tiles_x;
    const int ty = r % g.tiles_y;
    const int tz = r / g.tiles_y;
    const int x = g.x0 + tx * Tile<S>::X + (int)threadIdx.x;
    const int y = g.y0 + ty * Tile<S>::Y + (int)threadIdx.y;
    const int z = g.z0 + tz;
    if (x >= g.x1 || y >= g.y1) return;

    const int xm = (x == 0) ? g.nx - 1 : x - 1;
    const int xp = (x == g.nx - 1) ? 0 : x + 1;
    const int ym = (y == 0) ? g.ny - 1 : y - 1;
    const int yp = (y == g.ny - 1) ? 0 : y + 1;
    const int zs = z + g.zoff;
    int zm, zp;
    if (g.periodic_z) {
        zm = (z == 0) ? g.nz - 1 : z - 1;
        zp = (z == g.nz - 1) ? 0 : z + 1;
    } else {  // slab: ghost planes 0 and nz+1 hold the neighbours' images
        zm = zs - 1;
        zp = zs + 1;
    }
    // in-plane byte offsets of the x and y neighbours; element offsets of the
    // z planes (uniform)
    const O X[3] = {(O)xm * es, (O)x * es, (O)xp * es};
    const O Y[3] = {(O)ym * (O)g.ystride * es, (O)y * (O)g.ystride * es,
                    (O)yp * (O)g.ystride * es};
    const ptrdiff_t Z[3] = {(ptrdiff_t)zm * g.zstride, (ptrdiff_t)zs * g.zstride,
                            (ptrdiff_t)zp * g.zstride};

    C f[19];
#pragma unroll
    for (int q = 0; q < 19; q++) {
        // pull: population q arrives from x - c_q
        const O off = X[1 - cx(q)] + Y[1 - cy(q)];
        f[q] = Storage<S>::template load<C>(ld_at(plane_ptr(src.p[q], Z[1 - cz(q)]), off));
    }
    C b[19];
    C ux, uy, uz;
    collide_site<C>(f, b, omega, cte1, ux, uy, uz);
    const O c = X[1] + Y[1];
    if (TAPE) {  // compact interior site index
        const size_t i = (size_t)x + (size_t)g.nx * y + (size_t)g.nxy * z;
        tape[i] = ux;
        tape[tape_ld + i] = uy;
        tape[2 * tape_ld + i] = uz;
    }
#pragma unroll
    for (int q = 0; q < 19; q++)
        st_at(plane_ptr(dst.p[q], Z[1]), c, Storage<S>::template store<C>(b[q]));
}

// --------------------------------------------------------- two-step collide
// Two time steps in one pass over HBM (temporal blocking).  A workgroup owns
// a TX x TY column of sites over a chunk of ZC planes and marches it in z.
// At plane z it computes the first step S1 on the tile grown by one site on
// every x/y side, reading the source copy S0, keeps S1 in LDS, and computes
// the second step S2 of plane z-1 from there into the destination copy.  S1
// goes through the storage type exactly as a store and reload would, so the
// result equals two single steps bit for bit.  x and y wrap periodically; z
// wraps on the periodic lattice, and on a slab the pass covers the bulk
// planes whose first step needs no ghost plane.
//
// Only the S1 values that a later neighbour pull reads are kept in LDS, each
// for as long as it is needed (for the stream-collide a population with
// c_z = d at plane p is read by the second step at plane p + d):
//   c_z = -1 with an x/y component (4): plane z, read within the same iteration
//   c_z =  0 with an x/y component (8): plane z-1, one slot
//   c_z = +1 with an x/y component (4): planes z-1 and z-2, two slots
// The populations (0,0,+-1) and the rest population move only in z; they stay
// in registers of the thread that owns the column.
//
// Grown-tile shape (X x Y interior sites, ZC planes per workgroup, MINB
// workgroups per CU), from a sweep at 512^3: 128 x 5 (960 threads) for FP32
// compute with 32-bit offsets; 64 x 5 (512 threads, twice the registers per
// thread) for FP64 compute or 64-bit offsets.
template <typename S, typename C, bool O32>
struct Tile2 {
    static constexpr bool wide = sizeof(C) == 4 && O32;
    static constexpr int X = wide ? 128 : 64, Y = 5, ZC = 64, MINB = 1;
};

// Planes per workgroup of the two-step pass.  A workgroup fills a CU, so the
// pass takes ceil(workgroups / CUs) rounds of (zc + 2) plane iterations; zc
// in [8, ZC] minimises that (ties: the longer chunk, fewer redundant first-
// step planes).  Measured at 128^3-512^3 the model picks the fastest chunk
// of the sweep (e.g. 16 planes at 128^3 FP32: 208 workgroups, one round;
// 13 planes would take 260, two rounds, and lose the gain).  0 (take single
// steps) when the launch would not reach half of the CUs.
constexpr int TWO_STEP_MIN_ZC = 8;
// nz: the number of second-step planes of the pass (the slab's bulk passes
// cover nz - 4 of its planes).  L.two_step == 2 fuses at any size.
inline int two_step_zc(const Lattice &L, int tx, int ty, int zc_max, int nz) {
    if (nz <= 0) return 0;
    const long long txy = (long long)((L.nx + tx - 1) / tx) * ((L.ny + ty - 1) / ty);
    const long long ncu = L.ncu > 0 ? L.ncu : 1;
    const int lo = nz < TWO_STEP_MIN_ZC ? nz : TWO_STEP_MIN_ZC;
    const int hi = nz < zc_max ? nz : zc_max;
    int best = lo;
    long long best_cost = -1;
    for (int zc = lo; zc <= hi; zc++) {
        const long long nb = txy * ((nz + zc - 1) / zc);
        const long long cost = ((nb + ncu - 1) / ncu) * (zc + 2);
        if (best_cost < 0 || cost <= best_cost) {
            best_cost = cost;
            best = zc;
        }
    }
    const long long nb = txy * ((nz + best - 1) / best);
    if (L.two_step != 2 && 2 * nb < ncu) return 0;
    return best;
}
inline int two_step_zc(const Lattice &L, int tx, int ty, int zc_max) {
    return two_step_zc(L, tx, ty, zc_max, L.layout == BGK_LAYOUT_PERIODIC ? L.nz : L.nz - 4);
}

// A pass reads population q of site x from x - SGN c_q: SGN = +1 is the
// stream-collide (pull), SGN = -1 the adjoint step (which reads x + c_q).
// Source plane offset of population q:
template <int SGN>
__host__ __device__ constexpr int src_dz(int q) {
    return -SGN * cz(q);
}
// population group: 0 read from the newest plane (source dz = +1), 1 from the
// plane before (dz = 0), 2 from two planes before (dz = -1), all with an x/y
// component; 3 no x/y component
template <int SGN>
__host__ __device__ constexpr int pgroup(int q) {
    return (cx(q) == 0 && cy(q) == 0) ? 3
                                      : (src_dz<SGN>(q) > 0 ? 0 : (src_dz<SGN>(q) == 0 ? 1 : 2));
}
// index of q within its group
template <int SGN>
__host__ __device__ constexpr int pslot(int q) {
    int k = 0;
    for (int i = 0; i < q; i++) k += (pgroup<SGN>(i) == pgroup<SGN>(q)) ? 1 : 0;
    return k;
}
// first LDS plane of a group (group 2: plus 4 per slot)
__host__ __device__ constexpr int pbase(int grp) { return grp == 0 ? 0 : (grp == 1 ? 4 : 12); }

// Geometry of a two-step pass: second-step planes [zlo, zhi) in chunks of
// zc; periodic in z (wrap in the index math) or a slab whose bulk pass reads
// only planes zlo-2 .. zhi+1, all inside the slab; interior plane z is stored
// at z + zoff.
struct Geo2 {
    int nx, ny, nz, nxy;
    int tiles_x, tiles_y, zc;
    int zlo, zhi, zoff, periodic_z;
    long long ystride, zstride;  // element pitches of rows and planes (as Geo)
    unsigned pstride;            // BUF: bytes from population q to q+1
    int xcd_group;               // workgroups per XCD in one round (CUs / 8)
    int full_cols;  // k_collide3: the first full_cols workgroups each take a
                    // whole tile column (all of [zlo, zhi)); the rest of the
                    // columns go in chunks of zc planes (0: all in chunks)
    void *faces;   // FACES: face images of the first step (k_pack_faces layout)
    void *faces2;  // FACES, three-step pass: face images of the second step
    double *red;   // RED: [2][gridDim.x][RED_NV] partial sums (diagno, dissipation)
    double svisc;  // RED: the dissipation's viscosity
};

static Geo2 make_geo2(const Lattice &L, int tx, int ty, int zc, int *nblocks) {
    Geo2 g;
    g.nx = L.nx;
    g.ny = L.ny;
    g.nz = L.nz;
    g.nxy = L.nx * L.ny;
    g.tiles_x = (L.nx + tx - 1) / tx;
    g.tiles_y = (L.ny + ty - 1) / ty;
    g.zc = zc;
    g.periodic_z = L.layout == BGK_LAYOUT_PERIODIC;
    g.zoff = L.zoff;
    g.zlo = g.periodic_z ? 0 : 2;
    g.zhi = g.periodic_z ? L.nz : L.nz - 2;
    g.ystride = L.ystride;
    g.zstride = L.zstride;
    g.pstride = (unsigned)((const char *)L.pop[0][1] - (const char *)L.pop[0][0]);
    g.xcd_group = L.ncu >= 8 ? L.ncu / 8 : 1;
    g.full_cols = 0;
    g.faces = nullptr;
    g.faces2 = nullptr;
    g.red = nullptr;
    g.svisc = 0.0;
    *nblocks = g.tiles_x * g.tiles_y * ((g.zhi - g.zlo + g.zc - 1) / g.zc);
    return g;
}

// threads: one per site of the TX-wide x range of the TY+2 grown rows, plus
// one wave for the two grown columns
constexpr int collide2_threads(int tx, int ty) { return tx * (ty + 2) + 64; }

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// waits for every global load in flight, which would drain the next plane's
// loads issued before it; the two-step pass shares data between threads
// through LDS alone (it reads only the source copy and writes only the
// destination copy).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Face images of the periodic layout (the layout of k_pack_faces): the value
// s[q] of site (gx, gy, p) goes to the image of each face that population q
// leaves through, gx[q][gy + ny p], gy[q][gx + nx p], gz[q][gx + nx gy].
// BUF: through one buffer resource (32-bit offsets: the images hold
// 19 (ny nz + nx nz + nx ny) values), a vector offset per face and the
// population's distance as a scalar offset, so that no per-population
// address stays live in vector registers across the pass's loop.
template <typename S, bool BUF = false>
__device__ __forceinline__ void write_faces(void *faces, const Geo2 &g, const S (&s)[19], int gx,
                                            int gy, int p) {
    if (gx != 0 && gx != g.nx - 1 && gy != 0 && gy != g.ny - 1 && p != 0 && p != g.nz - 1)
        return;  // not on a face (most sites)
    if constexpr (BUF) {
        // one branch per face (uniform over most waves), then the five
        // populations that leave through it
        constexpr uint32_t es = sizeof(S);
        const __amdgpu_buffer_rsrc_t R = raw_rsrc(faces);
        const uint32_t nyz = (uint32_t)g.ny * g.nz, nxz = (uint32_t)g.nx * g.nz;
        const uint32_t oy = 19 * nyz, oz = oy + 19 * nxz;  // elements before gy, gz
        auto face = [&](int axis, int sgn, uint32_t voff) {
            asm volatile("" : "+v"(voff));
#pragma unroll
            for (int q = 0; q < 18; q++) {
                const int c = axis == 0 ? cx(q) : (axis == 1 ? cy(q) : cz(q));
                if (c != sgn) continue;
                const uint32_t e0 = axis == 0 ? (uint32_t)q * nyz
                                              : (axis == 1 ? oy + (uint32_t)q * nxz
                                                           : oz + (uint32_t)q * g.nxy);
                buf_st<S>(s[q], R, voff, e0 * es);
            }
        };
        const uint32_t vx = ((uint32_t)gy + (uint32_t)g.ny * p) * es;
        const uint32_t vy = ((uint32_t)gx + (uint32_t)g.nx * p) * es;
        const uint32_t vz = ((uint32_t)gx + (uint32_t)g.nx * gy) * es;
        if (gx == g.nx - 1) face(0, 1, vx);
        if (gx == 0) face(0, -1, vx);
        if (gy == g.ny - 1) face(1, 1, vy);
        if (gy == 0) face(1, -1, vy);
        if (p == g.nz - 1) face(2, 1, vz);
        if (p == 0) face(2, -1, vz);
        return;
    }
    S *fx = (S *)faces;
    S *fy = fx + (size_t)19 * g.ny * g.nz;
    S *fz = fy + (size_t)19 * g.nx * g.nz;
#pragma unroll
    for (int q = 0; q < 18; q++) {
        if (cx(q) != 0 && gx == (cx(q) > 0 ? g.nx - 1 : 0))
            fx[(size_t)q * g.ny * g.nz + gy + (size_t)g.ny * p] = s[q];
        if (cy(q) != 0 && gy == (cy(q) > 0 ? g.ny - 1 : 0))
            fy[(size_t)q * g.nx * g.nz + gx + (size_t)g.nx * p] = s[q];
        if (cz(q) != 0 && p == (cz(q) > 0 ? g.nz - 1 : 0))
            fz[(size_t)q * g.nxy + gx + (size_t)g.nx * gy] = s[q];
    }
}

// Per-site operations of the two-step pass.  `first` runs at every site of
// the grown tile (element index i; own = the site belongs to this workgroup's
// chunk and tile), `second` at the workgroup's own sites.  Extra holds inputs
// loaded with a plane's populations (prefetched one plane ahead).

// stream-collide; TAPE: record u = j/rho of both steps (slices tape0, tape1);
// SH: shifted storage only (collide_site<C, true>, cte1 = 1); KEEP = false:
// the pass does not store its final state (a taped re-run's last pass, whose
// state nothing reads; only its tape is kept)
template <typename C, bool TAPE, bool SH = false, bool KEEP = true>
struct CollideOp {
    static constexpr int SGN = 1;
    static constexpr bool KEEP_STATE = KEEP;
    // two load register sets (see two_step_pass); the taping variant keeps
    // one, its tape pointers would not fit beside the second set
    static constexpr bool PINGPONG = !TAPE;
    struct Extra {};
    C omega, cte1;
    C *tape0, *tape1;
    size_t ld;  // sites per tape slice component
    __device__ __forceinline__ Extra extra_first(size_t) const { return {}; }
    __device__ __forceinline__ Extra extra_second(size_t) const { return {}; }
    __device__ __forceinline__ void first(const C (&f)[19], C (&b)[19], const Extra &, size_t i,
                                          bool own) const {
        C ux, uy, uz;
        collide_site<C, SH>(f, b, omega, cte1, ux, uy, uz);
        if (TAPE && own) {
            tape0[i] = ux;
            tape0[ld + i] = uy;
            tape0[2 * ld + i] = uz;
        }
    }
    __device__ __forceinline__ void second(const C (&f)[19], C (&b)[19], const Extra &,
                                           size_t i) const {
        C ux, uy, uz;
        collide_site<C, SH>(f, b, omega, cte1, ux, uy, uz);
        if (TAPE) {
            tape1[i] = ux;
            tape1[ld + i] = uy;
            tape1[2 * ld + i] = uz;
        }
    }
};

// The body of a two-step pass from `src` into `dst` (see the comment above
// Tile2).  Every thread runs the same number of iterations and barriers.
// FACES (periodic lattice): also write the face planes of the first step's
// state to g.faces, in the layout of k_pack_faces, for the dissipation that
// reads the stale halo images of the reference (BGK_FLAG_REFCOMPAT_HALOS).
// BUF (with O32): populations at equal distances g.pstride (the lattice's
// interleaved copies); accesses go through one buffer resource per plane.
// ZM: how the pass finds its z range -- 0 from g at run time (periodic or
// slab), 1 periodic at compile time.  The collide pass keeps 0 (its code as
// measured fastest), the adjoint takes 1 (4 % faster there).
template <typename S, typename C, bool O32, int TX, int TY, class Op, bool FACES = false,
          bool BUF = false, int ZM = 0>
__device__ __forceinline__ void two_step_pass(const Pops<S> &src, const Pops<S> &dst,
                                              const Geo2 &g, const Op &op,
                                              S (*lds)[TY + 2][TX + 2]) {
    static_assert(!BUF || O32, "buffer access takes 32-bit offsets");
    using O = off_t_<O32>;
    using X_t = typename Op::Extra;
    constexpr int SGN = Op::SGN;
    constexpr O es = sizeof(S);
    constexpr int EX = TX + 2, EY = TY + 2;
    constexpr int NMAIN = TX * EY;  // threads of the x range [0, TX) of every row
    // the z-only populations: QN read from the newest plane, QO from two
    // planes before, QR the rest population
    constexpr int QN = SGN > 0 ? 14 : 5, QO = SGN > 0 ? 5 : 14, QR = 18;

    const int t = xcd_round_remap((int)blockIdx.x, (int)gridDim.x, g.xcd_group);
    const int tx = t % g.tiles_x;
    const int r = t / g.tiles_x;
    const int ty = r % g.tiles_y;
    const int za = (ZM == 1 ? 0 : g.zlo) + (r / g.tiles_y) * g.zc;
    const int zb = min(ZM == 1 ? g.nz : g.zhi, za + g.zc);

    // site of the grown tile handled by this thread: (ex, ey) in [0,EX) x [0,EY);
    // the last wave takes the columns ex = 0 and ex = EX-1
    const int tid = (int)threadIdx.x;
    int ex, ey;
    bool active;
    if (tid < NMAIN) {
        ex = tid % TX + 1;
        ey = tid / TX;
        active = true;
    } else {
        const int e = tid - NMAIN;
        active = e < 2 * EY;
        ex = (e < EY) ? 0 : EX - 1;
        ey = e % EY;
    }
    const int lx = tx * TX + ex - 1, ly = ty * TY + ey - 1;  // may leave the box
    const int gx = ((lx % g.nx) + g.nx) % g.nx;
    const int gy = ((ly % g.ny) + g.ny) % g.ny;
    // the thread that computes S1 of an interior column also computes its S2
    const bool owner = tid < NMAIN && ey >= 1 && ey <= TY && lx < g.nx && ly < g.ny;

    const int xm = (gx == 0) ? g.nx - 1 : gx - 1;
    const int xp = (gx == g.nx - 1) ? 0 : gx + 1;
    const int ym = (gy == 0) ? g.ny - 1 : gy - 1;
    const int yp = (gy == g.ny - 1) ? 0 : gy + 1;
    // in-plane byte offsets of the x and y neighbours (z planes are added to
    // the population pointers, uniform across the workgroup)
    const O X[3] = {(O)xm * es, (O)gx * es, (O)xp * es};
    const O Y[3] = {(O)ym * (O)g.ystride * es, (O)gy * (O)g.ystride * es,
                    (O)yp * (O)g.ystride * es};
    const ptrdiff_t plane = (ptrdiff_t)g.zstride;

    // z in [zlo-1, zhi]; a slab's planes need no wrap
    auto wrapz = [&](int z) {
        if constexpr (ZM == 1) return z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z);
        return !g.periodic_z ? z : (z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z));
    };
    // compact index of the column's site at stored plane p + zoff (auxiliary
    // arrays of the per-site operation: tape, adjoint inputs)
    auto site_of = [&](int p) {
        return (size_t)gx + (size_t)g.nx * gy + (size_t)g.nxy * (p + (ZM == 1 ? 0 : g.zoff));
    };
    auto load_plane = [&](int z, S(&raw)[19], X_t &xr) {
        const int p = wrapz(z);
        int zm, zp;
        ptrdiff_t Z[3];
        if constexpr (ZM == 1) {
            zm = (p == 0) ? g.nz - 1 : p - 1;
            zp = (p == g.nz - 1) ? 0 : p + 1;
            Z[0] = (ptrdiff_t)zm * plane;
            Z[1] = (ptrdiff_t)p * plane;
            Z[2] = (ptrdiff_t)zp * plane;
        } else {
            zm = !g.periodic_z ? p - 1 : ((p == 0) ? g.nz - 1 : p - 1);
            zp = !g.periodic_z ? p + 1 : ((p == g.nz - 1) ? 0 : p + 1);
            Z[0] = (ptrdiff_t)(zm + g.zoff) * plane;
            Z[1] = (ptrdiff_t)(p + g.zoff) * plane;
            Z[2] = (ptrdiff_t)(zp + g.zoff) * plane;
        }
        if constexpr (BUF) {
            const S *b0 = src.p[0];
            const __amdgpu_buffer_rsrc_t R[3] = {raw_rsrc(b0 + Z[0]), raw_rsrc(b0 + Z[1]),
                                                 raw_rsrc(b0 + Z[2])};
#pragma unroll
            for (int q = 0; q < 19; q++)
                raw[q] = buf_ld<S>(R[1 - SGN * cz(q)], X[1 - SGN * cx(q)] + Y[1 - SGN * cy(q)],
                                   (uint32_t)q * g.pstride);
            xr = op.extra_first(site_of(p));
            return;
        }
#pragma unroll
        for (int q = 0; q < 19; q++) {
            // opaque per plane, so that the compiler does not keep the 64-bit
            // extension of the offset live across the loop (a 32-bit offset
            // register is what the scalar-base load form takes)
            O off = X[1 - SGN * cx(q)] + Y[1 - SGN * cy(q)];
            asm volatile("" : "+v"(off));
            raw[q] = ld_at(plane_ptr(src.p[q], Z[1 - SGN * cz(q)]), off);
        }
        xr = op.extra_first(site_of(p));
    };

    S rO_1 = S{}, rO_2 = S{}, rR_1 = S{};  // S1 of this column at planes z-1, z-2
    const int nit = zb - za + 2;           // iterations z = za-1 .. zb
    // iteration `it`: S1 of plane z from cur (loaded one iteration earlier),
    // loads of plane z+1 into nxt, S2 of plane z-1
    auto iteration = [&](int it, S(&cur)[19], X_t &xcur, S(&nxt)[19], X_t &xnxt) {
        const int z = za - 1 + it;
        const int pb = pbase(2) + 4 * (it & 1);  // slot of S1(z-2), then of S1(z)
        S s1[19];
        // every thread loads (idle lanes of the last wave at valid addresses),
        // so the registers in flight need no merge at the end of `if (active)`
        C f[19];
#pragma unroll
        for (int q = 0; q < 19; q++) {
            f[q] = Storage<S>::template load<C>(cur[q]);
            // a copy of its own: when cur and nxt are one register set, the
            // next loads land in cur's registers, which carry them around the
            // loop without a copy (a copy at the loop end would wait for them)
            asm volatile("" : "+v"(f[q]));
        }
        const X_t x1 = xcur;
        if (it + 1 < nit) load_plane(z + 1, nxt, xnxt);  // in flight during this iteration
        const bool second = owner && it >= 2;
        const int p2 = wrapz(z - 1);
        X_t x2{};
        if (second) x2 = op.extra_second(site_of(p2));
        if (active) {
            C b[19];
            op.first(f, b, x1, site_of(wrapz(z)), owner && z >= za && z < zb);
#pragma unroll
            for (int q = 0; q < 19; q++) s1[q] = Storage<S>::template store<C>(b[q]);
#pragma unroll
            for (int q = 0; q < 19; q++)
                if (pgroup<SGN>(q) == 0) lds[pbase(0) + pslot<SGN>(q)][ey][ex] = s1[q];
            if (FACES && owner && z >= za && z < zb)  // S1 on the faces it leaves through
                write_faces<S>(g.faces, g, s1, gx, gy, wrapz(z));
        }
        lds_barrier();
        if (second) {  // second step at plane z-1
            C f2[19];
#pragma unroll
            for (int q = 0; q < 19; q++) {
                S v;
                const int sy = ey - SGN * cy(q), sx = ex - SGN * cx(q);
                if (pgroup<SGN>(q) == 0)
                    v = lds[pbase(0) + pslot<SGN>(q)][sy][sx];  // S1(z)
                else if (pgroup<SGN>(q) == 1)
                    v = lds[pbase(1) + pslot<SGN>(q)][sy][sx];  // S1(z-1)
                else if (pgroup<SGN>(q) == 2)
                    v = lds[pb + pslot<SGN>(q)][sy][sx];  // S1(z-2)
                else if (q == QO)
                    v = rO_2;
                else if (q == QN)
                    v = s1[q];
                else
                    v = rR_1;
                f2[q] = Storage<S>::template load<C>(v);
            }
            C b[19];
            op.second(f2, b, x2, site_of(p2));
            const ptrdiff_t zo = (ptrdiff_t)(p2 + (ZM == 1 ? 0 : g.zoff)) * plane;
            if constexpr (!Op::KEEP_STATE) {
                // no store of the final state
            } else if constexpr (BUF) {
                const __amdgpu_buffer_rsrc_t R = raw_rsrc(dst.p[0] + zo);
#pragma unroll
                for (int q = 0; q < 19; q++)
                    buf_st<S>(Storage<S>::template store<C>(b[q]), R, X[1] + Y[1],
                              (uint32_t)q * g.pstride);
            } else {
                O c = X[1] + Y[1];
                asm volatile("" : "+v"(c));
#pragma unroll
                for (int q = 0; q < 19; q++)
                    st_at(plane_ptr(dst.p[q], zo), c, Storage<S>::template store<C>(b[q]));
            }
        }
        lds_barrier();
        if (active) {
#pragma unroll
            for (int q = 0; q < 19; q++) {
                if (pgroup<SGN>(q) == 1) lds[pbase(1) + pslot<SGN>(q)][ey][ex] = s1[q];
                if (pgroup<SGN>(q) == 2) lds[pb + pslot<SGN>(q)][ey][ex] = s1[q];
            }
        }
        rO_2 = rO_1;
        rO_1 = s1[QO];
        rR_1 = s1[QR];
    };

    S ra[19];
    X_t xa;
    load_plane(za - 1, ra, xa);
    if constexpr (O32 && !FACES && Op::PINGPONG) {
        // two register sets taking turns: the loads of plane z+1 land in the
        // set plane z-1 used, so the compute copy of a plane costs no moves
        // (not for the 64-bit-offset, face-writing and taping variants,
        // where the second set would spill)
        S rb[19];
        X_t xb;
        int it = 0;
        for (; it + 1 < nit; it += 2) {
            iteration(it, ra, xa, rb, xb);
            iteration(it + 1, rb, xb, ra, xa);
        }
        if (it < nit) iteration(it, ra, xa, rb, xb);
    } else {
        for (int it = 0; it < nit; it++) iteration(it, ra, xa, ra, xa);
    }
}

template <typename S, typename C, bool O32, int TX, int TY, int MINB, bool TAPE,
          bool FACES = false, bool SH = false, bool KEEP = true>
__global__ __launch_bounds__(collide2_threads(TX, TY), MINB) void k_collide2(
    Pops<S> src, Pops<S> dst, Geo2 g, CollideOp<C, TAPE, SH, KEEP> op) {
    __shared__ S lds[20][TY + 2][TX + 2];
    two_step_pass<S, C, O32, TX, TY, CollideOp<C, TAPE, SH, KEEP>, FACES, O32>(src, dst, g, op,
                                                                              lds);
}

// ------------------------------------------------------ three steps per pass
// The two-step pass one level deeper: a workgroup sweeps z once and advances
// its tile three time steps, the first two steps' states kept on chip.  Per
// plane iteration z: S1(z) at every site of the (TX+4) x (TY+4) grown tile,
// S2(z-1) at the (TX+2) x (TY+2) middle sites (pulled from S1 of planes z,
// z-1, z-2), S3(z-2) at the TX x TY own sites (pulled from S2 likewise).  The
// x/y-moving populations of S1 and of S2 are shared through LDS in the
// group/slot layout of the two-step pass (20 planes each); the z-only ones
// stay in the column's registers.  Every intermediate value is rounded to the
// storage type exactly as a store and reload would round it, so a pass equals
// three single steps bit for bit.  Periodic lattice (wrap in z), or a slab's
// bulk, whose pass reads only stored planes zlo-3 .. zhi+2, all inside the
// slab; 32-bit offsets through one buffer resource per z plane (populations
// g.pstride bytes apart).
//
// Cost against the two-step pass: a chunk of zc planes takes zc + 4 plane
// iterations (two-step: zc + 2) and three collisions per site become
// (TX+4)(TY+4)/(TX TY) + (TX+2)(TY+2)/(TX TY) + 1 (64 x 10: 3.73, i.e. 1.24
// per step, against 1.21 per step of the two-step pass's 128 x 5 tiles),
// while the populations cross HBM once per three steps instead of two.
// FP64 compute: 32 x 10 own tiles (36 x 14 grown, 146 KB of LDS in FP64),
// 512 threads, so that a thread may hold 256 VGPRs (collisions per step:
// 1.28).  Measured at 512^3 FP64, ms per pass: 32 x 10 11.73, 32 x 8 12.58,
// 32 x 6 13.70, 64 x 3 14.01 (64 x 4 would spill).
template <typename S, typename C>
struct Tile3 {
    static constexpr bool f64 = sizeof(C) == 8;
    static constexpr int X = f64 ? 32 : 64, Y = 10, ZC = 64;
};

// threads: one per site of the TX-wide x range of the TY+4 grown rows, plus
// one wave for the four grown columns (4 (TY+4) <= 64)
constexpr int collide3_threads(int tx, int ty) { return tx * (ty + 4) + 64; }

// Where a three-step pass issues the loads of the next plane: populations
// [0, A) after the first stage's collision, [A, B) after the second's,
// [B, 19) after the third's (by the compute kind's size).  Measured at 512^3,
// ms per pass against all 19 loads at the top of the iteration: FP32 5.09-5.33
// -> 4.56-4.68 (split 10/9; 12/7 4.64, 8/11 4.61-4.63, 10/5/4 4.62-4.66,
// 7/6/6 4.75-4.88), FP16 4.43 -> 4.12-4.24, BF16 4.48-4.53 -> 4.03-4.15,
// FP64 10.94 -> 9.80-9.82 (7/6/6; 10/9 10.18-10.51).  RED (the steps +
// diagnostics pass, whose second stage also stores S2 and sums its moments
// and whose third stage stores nothing): [10, 19) after the third stage,
// 5.25-5.36 -> 5.01-5.08 ms per pass in FP32 (all 19 after the first
// stage: 5.38-5.50).
template <int CSIZE, bool RED = false>
struct Load3 {
    static constexpr int A = 10, B = 19;
};
template <int CSIZE>
struct Load3<CSIZE, true> {
    static constexpr int A = 10, B = 10;
};
template <bool RED>
struct Load3<8, RED> {
    static constexpr int A = 7, B = 13;
};
template <>
struct Load3<8, true> {
    static constexpr int A = 7, B = 13;
};

// Planes per workgroup of the three-step pass: as two_step_zc, with zc + 4
// plane iterations per chunk.  0 when the launch would not reach half of the
// CUs (unless L.two_step == 2) or the lattice is too thin in z (periodic:
// fewer than 4 planes; slab: fewer than 12, whose pass covers planes
// 3 .. nz-4).
inline int three_step_zc(const Lattice &L, int tx, int ty, int zc_max) {
    const bool slab = L.layout != BGK_LAYOUT_PERIODIC;
    if (L.nz < (slab ? SLAB3_MIN_NZ : 4)) return 0;
    const int nz = slab ? L.nz - 6 : L.nz;
    const long long txy = (long long)((L.nx + tx - 1) / tx) * ((L.ny + ty - 1) / ty);
    const long long ncu = L.ncu > 0 ? L.ncu : 1;
    const int lo = nz < TWO_STEP_MIN_ZC ? nz : TWO_STEP_MIN_ZC;
    const int hi = nz < zc_max ? nz : zc_max;
    int best = lo;
    long long best_cost = -1;
    for (int zc = lo; zc <= hi; zc++) {
        const long long nb = txy * ((nz + zc - 1) / zc);
        const long long cost = ((nb + ncu - 1) / ncu) * (zc + 4);
        if (best_cost < 0 || cost <= best_cost) {
            best_cost = cost;
            best = zc;
        }
    }
    const long long nb = txy * ((nz + best - 1) / best);
    if (L.two_step != 2 && 2 * nb < ncu) return 0;
    return best;
}

// Whole tile columns first in a three-step pass (Geo2::full_cols):
// a chunk of zc planes costs zc + 4 plane iterations (the first two steps'
// halo planes), so on a lattice with c >= ncu tile columns, k rounds of
// whole columns (k ncu workgroups of nz + 4 iterations) and the other
// c - k ncu columns in chunks can take fewer plane iterations per CU than
// all of them in chunks; the chunks follow the model of three_step_zc.
// 512^3, 64x10 tiles, 256 CUs: 256 whole columns and 160 columns in
// 64-plane chunks, 516 + 5 x 68 = 856 iterations against 13 x 68 = 884.
struct Split3 {
    int full_cols, zc, nblocks;
};
inline Split3 three_step_split(const Lattice &L, int tx, int ty, int zc_max) {
    const int zc0 = three_step_zc(L, tx, ty, zc_max);
    const long long c = (long long)((L.nx + tx - 1) / tx) * ((L.ny + ty - 1) / ty);
    // third-step planes: all of a periodic lattice, a slab's bulk 3 .. nz-4
    const int nz = L.layout == BGK_LAYOUT_PERIODIC ? L.nz : L.nz - 6;
    Split3 best{0, zc0, zc0 > 0 ? (int)(c * ((nz + zc0 - 1) / zc0)) : 0};
    const long long ncu = L.ncu > 0 ? L.ncu : 1;
    if (zc0 <= 0 || c < ncu) return best;
    const int lo = nz < TWO_STEP_MIN_ZC ? nz : TWO_STEP_MIN_ZC;
    const int hi = nz < zc_max ? nz : zc_max;
    auto chunked = [&](long long cols, int *zc_out) {
        long long bc = -1;
        for (int zc = lo; zc <= hi; zc++) {
            const long long n = cols * ((nz + zc - 1) / zc);
            const long long cost = ((n + ncu - 1) / ncu) * (zc + 4);
            if (bc < 0 || cost <= bc) {
                bc = cost;
                *zc_out = zc;
            }
        }
        return bc;
    };
    int zr = zc0;
    long long best_cost = chunked(c, &zr);
    for (long long k = 1; k * ncu <= c; k++) {
        const long long rem = c - k * ncu;
        long long cost = k * (nz + 4);
        zr = zc0;
        if (rem > 0) cost += chunked(rem, &zr);
        if (cost < best_cost) {
            best_cost = cost;
            best = Split3{(int)(k * ncu), zr, (int)(k * ncu + rem * ((nz + zr - 1) / zr))};
        }
    }
    return best;
}

// Tile column and z range [za, zb) of this workgroup in a three-step pass
// over the third-step planes [zlo, zhi): the first g.full_cols workgroups
// (whole rounds) take whole columns, the rest of the columns go in chunks
// of g.zc planes; each part in the round block order (xcd_round_remap).
struct Tile3Pos {
    int tx, ty, za, zb;
};
__device__ __forceinline__ Tile3Pos tile3_pos(const Geo2 &g, int zlo, int zhi) {
    const int b = (int)blockIdx.x, nb = (int)gridDim.x;
    int col;
    Tile3Pos p;
    if (b < g.full_cols) {
        col = xcd_round_remap(b, g.full_cols, g.xcd_group);
        p.za = zlo;
        p.zb = zhi;
    } else {
        const int rc = g.tiles_x * g.tiles_y - g.full_cols;
        const int t = xcd_round_remap(b - g.full_cols, nb - g.full_cols, g.xcd_group);
        col = g.full_cols + t % rc;
        p.za = zlo + (t / rc) * g.zc;
        p.zb = min(zhi, p.za + g.zc);
    }
    p.tx = col % g.tiles_x;
    p.ty = col / g.tiles_x;
    return p;
}

// The grown-tile row that main wave w of k_collide3 computes.  With
// TX = 64 a wave is one row, and waves w, w+4, ... issue on one SIMD, so a
// stage takes the time of its busiest SIMD's waves.  The second stage's
// rows 1 and EY-2 and the edge wave (wave EY) compute only part of their
// sites' populations (s2kind); rows EY-2 and EY-4 swap waves so that row
// EY-2 shares the edge wave's SIMD: 64x10 tiles then give the second
// stage's busiest SIMD two full collisions and two partial ones, against
// three full and the edge wave's partial one in row order; the third
// stage keeps at most three collisions per SIMD.
template <int TX, int EY>
__device__ constexpr int k3_row(int w) {
    if constexpr (TX == 64 && EY >= 8)
        return w == EY - 4 ? EY - 2 : (w == EY - 2 ? EY - 4 : w);
    else
        return w;
}

// SLAB: the slab's bulk (z offset g.zoff, no wrap), compiled apart so that
// the periodic pass carries none of it.  FACES (periodic lattice): also write
// the face images of the first step's state to g.faces and of the second
// step's to g.faces2 (the stale halo images the reference's dissipation reads,
// BGK_FLAG_REFCOMPAT_HALOS); that variant carries one load register set, the
// face writes' registers would not fit beside the second.
// RED (periodic lattice): two time steps and the diagnostics of their
// result in one pass (the reference's loop body at a diagnostic step,
// SRC/bgk3d_HIT.F90:127-146): the owners store S2 and sum its at-site
// moments (SRC/diagno.F90:96-111), and the third stage, instead of a
// collision, sums the dissipation of S2's pulled populations
// (SRC/dissipation.F90:97-249); per workgroup one slot of diagno sums
// {rho, ux, uy, uz, |u|^2/2, uz at (nx-1, ny-1, nz-1)} and one of
// {2 svisc P, |u|^2/2} in g.red, in FP64 as the reduction kernels sum.  One
// load register set, as FACES.
template <typename S, typename C, int TX, int TY, bool SH, bool SLAB = false, bool FACES = false,
          bool RED = false>
__global__ __launch_bounds__(collide3_threads(TX, TY), 1) void k_collide3(Pops<S> src,
                                                                          Pops<S> dst, Geo2 g,
                                                                          C omega, C cte1) {
    static_assert(!(RED && (SLAB || FACES)), "RED runs on the periodic lattice alone");
    // the second stage trimmed to what the third reads (s2kind, k3_row)
    constexpr bool TRIM2 = (sizeof(S) == 2 || sizeof(C) == 8) && !FACES && !RED;
    constexpr int EX = TX + 4, EY = TY + 4;
    constexpr int NMAIN = TX * EY;
    static_assert(4 * EY <= 64, "the edge wave covers four grown columns");
    constexpr uint32_t es = sizeof(S);
    constexpr int QN = 14, QO = 5, QR = 18;  // z-only: newest plane, two before, rest
    __shared__ S l1[20][EY][EX];             // S1, site (ey, ex)
    __shared__ S l2[20][EY - 2][EX - 2];     // S2, site (ey, ex) at [ey-1][ex-1]

    const Tile3Pos pos = tile3_pos(g, SLAB ? g.zlo : 0, SLAB ? g.zhi : g.nz);
    const int tx = pos.tx, ty = pos.ty, za = pos.za, zb = pos.zb;

    const int tid = (int)threadIdx.x;
    int ex, ey;
    bool active;
    if (tid < NMAIN) {
        ex = tid % TX + 2;
        ey = TRIM2 ? k3_row<TX, EY>(tid / TX) : tid / TX;
        active = true;
    } else {
        const int e = tid - NMAIN;
        const int col = e / EY;  // grown columns 0, 1, EX-2, EX-1
        active = e < 4 * EY;
        ex = col < 2 ? col : EX - 4 + col;
        ey = e % EY;
    }
    const int lx = tx * TX + ex - 2, ly = ty * TY + ey - 2;  // may leave the box
    const int gx = ((lx % g.nx) + g.nx) % g.nx;
    const int gy = ((ly % g.ny) + g.ny) % g.ny;
    const bool mid = active && ex >= 1 && ex <= EX - 2 && ey >= 1 && ey <= EY - 2;
    const bool owner = tid < NMAIN && ey >= 2 && ey <= EY - 3 && lx < g.nx && ly < g.ny;
    // What the third stage reads of a second-stage site: all of it at the
    // own tile's sites; at row 1 only the populations with c_y = +1, at row
    // EY-2 those with c_y = -1, at the edge columns 1 and EX-2 (the edge
    // wave) those with c_x != 0.  The others are left uncomputed (the
    // compiler drops their equilibrium and relaxation); the kind is uniform
    // over the second stage's lanes of each wave.
    // Taken for 16-bit storage (the collisions bound the pass; 107 and 124
    // VGPRs) and FP64 compute (242 VGPRs of its 256; 1.5-4 % faster).  FP32
    // storage spilled registers with it (and trimming the edge wave alone
    // measured slower).
    const int s2kind =
        !TRIM2 ? 0 : (tid >= NMAIN ? 3 : (ey == 1 ? 1 : (ey == EY - 2 ? 2 : 0)));

    const int xm = (gx == 0) ? g.nx - 1 : gx - 1;
    const int xp = (gx == g.nx - 1) ? 0 : gx + 1;
    const int ym = (gy == 0) ? g.ny - 1 : gy - 1;
    const int yp = (gy == g.ny - 1) ? 0 : gy + 1;
    const uint32_t X[3] = {(uint32_t)xm * es, (uint32_t)gx * es, (uint32_t)xp * es};
    const uint32_t ys = (uint32_t)g.ystride * es;
    const uint32_t Y[3] = {(uint32_t)ym * ys, (uint32_t)gy * ys, (uint32_t)yp * ys};
    const ptrdiff_t plane = (ptrdiff_t)g.zstride;

    // z in [zlo-2, zhi+1]; a slab's planes need no wrap
    const int zoff = SLAB ? g.zoff : 0;
    auto wrapz = [&](int z) {
        return SLAB ? z : (z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z));
    };
    // the loads of populations [QA, QB) of plane z
    auto load_part = [&](int z, S(&raw)[19], auto qa, auto qb) {
        constexpr int QA = decltype(qa)::value, QB = decltype(qb)::value;
        const int p = wrapz(z);
        const int zm = SLAB ? p - 1 : ((p == 0) ? g.nz - 1 : p - 1);
        const int zp = SLAB ? p + 1 : ((p == g.nz - 1) ? 0 : p + 1);
        const S *b0 = src.p[0];
        const __amdgpu_buffer_rsrc_t R[3] = {raw_rsrc(b0 + (ptrdiff_t)(zm + zoff) * plane),
                                             raw_rsrc(b0 + (ptrdiff_t)(p + zoff) * plane),
                                             raw_rsrc(b0 + (ptrdiff_t)(zp + zoff) * plane)};
#pragma unroll
        for (int q = QA; q < QB; q++)
            raw[q] = buf_ld<S>(R[1 - cz(q)], X[1 - cx(q)] + Y[1 - cy(q)], (uint32_t)q * g.pstride);
    };
    // The next plane's loads go out in parts after the stages' collisions
    // (Load3): issued all at once at the top of the iteration, the 285
    // loads of a workgroup queue behind each other and the youngest waves
    // of each SIMD wait thousands of cycles to issue theirs before they can
    // start the first stage.
    using LP = Load3<sizeof(C), RED>;
    std::integral_constant<int, 0> q0;
    std::integral_constant<int, LP::A> qA;
    std::integral_constant<int, LP::B> qB;
    std::integral_constant<int, 19> q19;
    // pulled value of population q for the site (sy, sx) of a state whose
    // x/y-moving populations are in `lds` (newest plane group 0, slot pb of
    // the plane two before); the z-only ones from the column's registers
    auto pull = [&](auto &lds, int q, int sy, int sx, int pb, S newest, S older, S rest) {
        const int yy = sy - cy(q), xx = sx - cx(q);
        if (pgroup<1>(q) == 0) return lds[pbase(0) + pslot<1>(q)][yy][xx];
        if (pgroup<1>(q) == 1) return lds[pbase(1) + pslot<1>(q)][yy][xx];
        if (pgroup<1>(q) == 2) return lds[pb + pslot<1>(q)][yy][xx];
        return q == QO ? older : (q == QN ? newest : rest);
    };

    S r1O_1 = S{}, r1O_2 = S{}, r1R_1 = S{};  // S1 of this column at planes z-1, z-2
    S r2O_1 = S{}, r2O_2 = S{}, r2R_1 = S{};  // S2 of this column, one plane behind
    const int nit = zb - za + 4;              // iterations z = za-2 .. zb+1
    // the owner's site of plane p to the destination copy
    auto store_site = [&](const S(&v)[19], int p) {
        const ptrdiff_t zo = (ptrdiff_t)(p + zoff) * plane;
        if constexpr (sizeof(S) == 2) {
            // global stores from per-population scalar bases: a value in
            // the high half of a packed pair stores with the d16_hi form
            // (the offset opaque per plane, so that it stays the 32-bit
            // offset of the scalar-base store form, see two_step_pass)
            uint32_t c = X[1] + Y[1];
            asm volatile("" : "+v"(c));
#pragma unroll
            for (int q = 0; q < 19; q++) st_at(plane_ptr(dst.p[q], zo), c, v[q]);
        } else {
            const __amdgpu_buffer_rsrc_t R = raw_rsrc(dst.p[0] + zo);
#pragma unroll
            for (int q = 0; q < 19; q++)
                buf_st<S>(v[q], R, X[1] + Y[1], (uint32_t)q * g.pstride);
        }
    };
    // RED: this thread's sums (diagno's five, dissipation's two) and uz of
    // the last site
    double rd[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, rs[2] = {0.0, 0.0};
    C ulast = (C)0;
    auto iteration = [&](int it, S(&cur)[19], S(&nxt)[19]) {
        const int z = za - 2 + it;
        const int pb = pbase(2) + 4 * (it & 1);  // slot of S(z-2), then of S(z)
        S s1[19], s2[19];
        C f[19];
#pragma unroll
        for (int q = 0; q < 19; q++) {
            f[q] = Storage<S>::template load<C>(cur[q]);
            asm volatile("" : "+v"(f[q]));  // see two_step_pass
        }
        const bool more = it + 1 < nit;
        if (active) {  // S1 at plane z
            C b[19], ux, uy, uz;
            collide_site<C, SH>(f, b, omega, cte1, ux, uy, uz);
            store_all<S, C>(b, s1);
#pragma unroll
            for (int q = 0; q < 19; q++)
                if (pgroup<1>(q) == 0) l1[pbase(0) + pslot<1>(q)][ey][ex] = s1[q];
        }
        if (more) load_part(z + 1, nxt, q0, qA);
        lds_barrier();
        const bool do2 = mid && it >= 2;
        if (do2) {  // S2 at plane z-1
            C f2[19];
#pragma unroll
            for (int q = 0; q < 19; q++)
                f2[q] = Storage<S>::template load<C>(pull(l1, q, ey, ex, pb, s1[QN], r1O_2, r1R_1));
            C ux, uy, uz;
            const Moments<C> mo = collide_moments<C, SH>(f2, cte1, ux, uy, uz);
            auto second = [&](auto sel) {
                constexpr uint32_t M = decltype(sel)::value;
                C b[19];
                collide_relax<M>(f2, mo, omega, b);
                store_sel<M, S, C>(b, s2);
#pragma unroll
                for (int q = 0; q < 19; q++)
                    if (((M >> q) & 1u) && pgroup<1>(q) == 0)
                        l2[pbase(0) + pslot<1>(q)][ey - 1][ex - 1] = s2[q];
            };
            if (s2kind == 0)
                second(std::integral_constant<uint32_t, ALL_POPS>{});
            else if (s2kind == 1)
                second(std::integral_constant<uint32_t, pops_cy(1)>{});
            else if (s2kind == 2)
                second(std::integral_constant<uint32_t, pops_cy(-1)>{});
            else
                second(std::integral_constant<uint32_t, ALL_POPS & ~pops_cx(0)>{});
            if (FACES && owner && z - 1 >= za && z - 1 < zb)
                write_faces<S, true>(g.faces2, g, s2, gx, gy, wrapz(z - 1));
            if constexpr (RED) {
                if (owner && z - 1 >= za && z - 1 < zb) {
                    store_site(s2, z - 1);
                    C f2s[19], rho, ux, uy, uz;
#pragma unroll
                    for (int q = 0; q < 19; q++) f2s[q] = Storage<S>::template load<C>(s2[q]);
                    moments_site<C>(f2s, cte1, rho, ux, uy, uz);
                    const C sk = (C)0.5 * (ux * ux + uy * uy + uz * uz);
                    rd[0] += (double)rho;
                    rd[1] += (double)ux;
                    rd[2] += (double)uy;
                    rd[3] += (double)uz;
                    rd[4] += (double)sk;
                    if (gx == g.nx - 1 && gy == g.ny - 1 && z - 1 == g.nz - 1) ulast = uz;
                }
            }
        }
        if (more) load_part(z + 1, nxt, qA, qB);
        lds_barrier();
        if (active) {
#pragma unroll
            for (int q = 0; q < 19; q++) {
                if (pgroup<1>(q) == 1) l1[pbase(1) + pslot<1>(q)][ey][ex] = s1[q];
                if (pgroup<1>(q) == 2) l1[pb + pslot<1>(q)][ey][ex] = s1[q];
            }
            // S1 on the faces it leaves through (the collision's temporaries
            // are dead here)
            if (FACES && owner && z >= za && z < zb)
                write_faces<S, true>(g.faces, g, s1, gx, gy, wrapz(z));
        }
        if (owner && it >= 4) {  // S3 at plane z-2
            C f3[19];
#pragma unroll
            for (int q = 0; q < 19; q++)
                f3[q] = Storage<S>::template load<C>(
                    pull(l2, q, ey - 1, ex - 1, pb, s2[QN], r2O_2, r2R_1));
            if constexpr (RED) {  // dissipation of S2 at plane z-2
                C dsite, tsite;
                dissipation_site<C>(f3, cte1, (C)g.svisc, dsite, tsite);
                rs[0] += (double)dsite;
                rs[1] += (double)tsite;
            } else {
                C b[19], ux, uy, uz;
                collide_site<C, SH>(f3, b, omega, cte1, ux, uy, uz);
                S s3[19];
                store_all<S, C>(b, s3);
                store_site(s3, wrapz(z - 2));
            }
        }
        if (more) load_part(z + 1, nxt, qB, q19);
        lds_barrier();
        if (do2) {
            auto late = [&](auto sel) {
                constexpr uint32_t M = decltype(sel)::value;
#pragma unroll
                for (int q = 0; q < 19; q++) {
                    if (!((M >> q) & 1u)) continue;
                    if (pgroup<1>(q) == 1) l2[pbase(1) + pslot<1>(q)][ey - 1][ex - 1] = s2[q];
                    if (pgroup<1>(q) == 2) l2[pb + pslot<1>(q)][ey - 1][ex - 1] = s2[q];
                }
            };
            if (s2kind == 0) {
                late(std::integral_constant<uint32_t, ALL_POPS>{});
                r2O_2 = r2O_1;  // the column's z-only populations (own sites)
                r2O_1 = s2[QO];
                r2R_1 = s2[QR];
            } else if (s2kind == 1) {
                late(std::integral_constant<uint32_t, pops_cy(1)>{});
            } else if (s2kind == 2) {
                late(std::integral_constant<uint32_t, pops_cy(-1)>{});
            } else {
                late(std::integral_constant<uint32_t, ALL_POPS & ~pops_cx(0)>{});
            }
        }
        r1O_2 = r1O_1;
        r1O_1 = s1[QO];
        r1R_1 = s1[QR];
    };

    S ra[19];
    load_part(za - 2, ra, q0, q19);
    // One load register set where the loads of the next plane land in the
    // registers the first stage has consumed: the FACES and RED forms (a
    // second set would spill) and FP16 storage, measured faster with it
    // (512^3, three rounds: 92.4-93.0 against 86.6-89.3 k MLUPS in the
    // bench; FP32, BF16 and FP64 within the spread either way).
    if constexpr (FACES || RED || std::is_same<S, half_t>::value) {
        for (int it = 0; it < nit; it++) iteration(it, ra, ra);
    } else {
        S rb[19];
        int it = 0;
        for (; it + 1 < nit; it += 2) {  // two register sets taking turns
            iteration(it, ra, rb);
            iteration(it + 1, rb, ra);
        }
        if (it < nit) iteration(it, ra, rb);
    }
    if constexpr (RED) {
        double v[RED_NV] = {rd[0], rd[1], rd[2], rd[3], rd[4], (double)ulast, 0.0, 0.0};
        block_sum_store<RED_NV>(v, g.red + (size_t)blockIdx.x * RED_NV);
        __syncthreads();  // block_sum_store's LDS is reused
        double w[RED_NV] = {rs[0], rs[1], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        block_sum_store<RED_NV>(w, g.red + ((size_t)gridDim.x + blockIdx.x) * RED_NV);
    }
}

// ------------------------------------ three steps per pass, per-site operation
// The three-step pass over a per-site operation (as two_step_pass): the
// adjoint's taping forward and its backward sweep.  k_collide3 keeps its own
// copy of this body: compiled through a generic form its code changed (same
// instruction counts, other register allocation and order), which cost 3.5 %
// in BF16 with this form and 1.5 % in FP16 with LDS passed by pointer
// (measured on one box against the kept kernel), so the two are kept apart;
// both are tested against single steps bit for bit.

// first, second and third run at S1, S2 and S3; own = the site is one of the
// workgroup's own sites of its chunk.  TAPE: record u = j/rho of the three
// steps at the own sites (slices tape0..tape2).
template <typename C, bool TAPE, bool SH, bool KEEP = true>
struct CollideOp3 {
    static constexpr int SGN = 1;
    static constexpr bool KEEP_STATE = KEEP;  // as CollideOp
    static constexpr bool PINGPONG = !TAPE;
    struct Extra {};
    C omega, cte1;
    C *tape0, *tape1, *tape2;
    size_t ld;  // sites per tape slice component
    __device__ __forceinline__ Extra extra_first(size_t) const { return {}; }
    __device__ __forceinline__ Extra extra_second(size_t) const { return {}; }
    __device__ __forceinline__ Extra extra_third(size_t) const { return {}; }
    __device__ __forceinline__ void tape(C *t, size_t i, C ux, C uy, C uz) const {
        t[i] = ux;
        t[ld + i] = uy;
        t[2 * ld + i] = uz;
    }
    __device__ __forceinline__ void first(const C (&f)[19], C (&b)[19], const Extra &, size_t i,
                                          bool own) const {
        C ux, uy, uz;
        collide_site<C, SH>(f, b, omega, cte1, ux, uy, uz);
        if (TAPE && own) tape(tape0, i, ux, uy, uz);
    }
    __device__ __forceinline__ void second(const C (&f)[19], C (&b)[19], const Extra &, size_t i,
                                           bool own) const {
        C ux, uy, uz;
        collide_site<C, SH>(f, b, omega, cte1, ux, uy, uz);
        if (TAPE && own) tape(tape1, i, ux, uy, uz);
    }
    __device__ __forceinline__ void third(const C (&f)[19], C (&b)[19], const Extra &,
                                          size_t i) const {
        C ux, uy, uz;
        collide_site<C, SH>(f, b, omega, cte1, ux, uy, uz);
        if (TAPE) tape(tape2, i, ux, uy, uz);
    }
};

// Periodic lattice, 32-bit offsets through one buffer resource per z plane
// (populations g.pstride bytes apart).  Op::SGN = -1 mirrors the pull (the
// adjoint reads x + c_q).  Every thread runs the same iterations and barriers.
template <typename S, typename C, int TX, int TY, class Op>
__device__ __forceinline__ void three_step_pass(const Pops<S> &src, const Pops<S> &dst,
                                                const Geo2 &g, const Op &op) {
    using X_t = typename Op::Extra;
    constexpr int SGN = Op::SGN;
    constexpr int EX = TX + 4, EY = TY + 4;
    constexpr int NMAIN = TX * EY;
    static_assert(4 * EY <= 64, "the edge wave covers four grown columns");
    constexpr uint32_t es = sizeof(S);
    // the z-only populations: QN read from the newest plane, QO from two
    // planes before, QR the rest population
    constexpr int QN = SGN > 0 ? 14 : 5, QO = SGN > 0 ? 5 : 14, QR = 18;
    __shared__ S l1[20][EY][EX];          // S1, site (ey, ex)
    __shared__ S l2[20][EY - 2][EX - 2];  // S2, site (ey, ex) at [ey-1][ex-1]

    const Tile3Pos pos = tile3_pos(g, 0, g.nz);
    const int tx = pos.tx, ty = pos.ty, za = pos.za, zb = pos.zb;

    const int tid = (int)threadIdx.x;
    int ex, ey;
    bool active;
    if (tid < NMAIN) {
        ex = tid % TX + 2;
        ey = tid / TX;
        active = true;
    } else {
        const int e = tid - NMAIN;
        const int col = e / EY;  // grown columns 0, 1, EX-2, EX-1
        active = e < 4 * EY;
        ex = col < 2 ? col : EX - 4 + col;
        ey = e % EY;
    }
    const int lx = tx * TX + ex - 2, ly = ty * TY + ey - 2;  // may leave the box
    const int gx = ((lx % g.nx) + g.nx) % g.nx;
    const int gy = ((ly % g.ny) + g.ny) % g.ny;
    const bool mid = active && ex >= 1 && ex <= EX - 2 && ey >= 1 && ey <= EY - 2;
    const bool owner = tid < NMAIN && ey >= 2 && ey <= EY - 3 && lx < g.nx && ly < g.ny;

    const int xm = (gx == 0) ? g.nx - 1 : gx - 1;
    const int xp = (gx == g.nx - 1) ? 0 : gx + 1;
    const int ym = (gy == 0) ? g.ny - 1 : gy - 1;
    const int yp = (gy == g.ny - 1) ? 0 : gy + 1;
    const uint32_t X[3] = {(uint32_t)xm * es, (uint32_t)gx * es, (uint32_t)xp * es};
    const uint32_t ys = (uint32_t)g.ystride * es;
    const uint32_t Y[3] = {(uint32_t)ym * ys, (uint32_t)gy * ys, (uint32_t)yp * ys};
    const ptrdiff_t plane = (ptrdiff_t)g.zstride;

    auto wrapz = [&](int z) { return z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z); };
    // compact index of the column's site at plane p (tape, adjoint inputs)
    auto site_of = [&](int p) {
        return (size_t)gx + (size_t)g.nx * gy + (size_t)g.nxy * p;
    };
    // the loads of populations [QA, QB) of plane z (and, with the first part,
    // the per-site operation's inputs of the first stage); the next plane's
    // go out in parts after the stages' collisions, as in k_collide3 (Load3)
    auto load_part = [&](int z, S(&raw)[19], X_t &xr, auto qa, auto qb) {
        constexpr int QA = decltype(qa)::value, QB = decltype(qb)::value;
        const int p = wrapz(z);
        const int zm = (p == 0) ? g.nz - 1 : p - 1;
        const int zp = (p == g.nz - 1) ? 0 : p + 1;
        const S *b0 = src.p[0];
        const __amdgpu_buffer_rsrc_t R[3] = {raw_rsrc(b0 + (ptrdiff_t)zm * plane),
                                             raw_rsrc(b0 + (ptrdiff_t)p * plane),
                                             raw_rsrc(b0 + (ptrdiff_t)zp * plane)};
#pragma unroll
        for (int q = QA; q < QB; q++)
            raw[q] = buf_ld<S>(R[1 - SGN * cz(q)], X[1 - SGN * cx(q)] + Y[1 - SGN * cy(q)],
                               (uint32_t)q * g.pstride);
        if constexpr (QA == 0) xr = op.extra_first(site_of(p));
    };
    using LP = Load3<sizeof(C)>;
    std::integral_constant<int, 0> q0;
    std::integral_constant<int, LP::A> qA;
    std::integral_constant<int, LP::B> qB;
    std::integral_constant<int, 19> q19;
    // pulled value of population q for the site (sy, sx) of a state whose
    // x/y-moving populations are in `lds` (newest plane group 0, slot pb of
    // the plane two before); the z-only ones from the column's registers
    auto pull = [&](auto &lds, int q, int sy, int sx, int pb, S newest, S older, S rest) {
        const int yy = sy - SGN * cy(q), xx = sx - SGN * cx(q);
        if (pgroup<SGN>(q) == 0) return lds[pbase(0) + pslot<SGN>(q)][yy][xx];
        if (pgroup<SGN>(q) == 1) return lds[pbase(1) + pslot<SGN>(q)][yy][xx];
        if (pgroup<SGN>(q) == 2) return lds[pb + pslot<SGN>(q)][yy][xx];
        return q == QO ? older : (q == QN ? newest : rest);
    };

    S r1O_1 = S{}, r1O_2 = S{}, r1R_1 = S{};  // S1 of this column at planes z-1, z-2
    S r2O_1 = S{}, r2O_2 = S{}, r2R_1 = S{};  // S2 of this column, one plane behind
    const int nit = zb - za + 4;              // iterations z = za-2 .. zb+1
    auto iteration = [&](int it, S(&cur)[19], X_t &xcur, S(&nxt)[19], X_t &xnxt) {
        const int z = za - 2 + it;
        const int pb = pbase(2) + 4 * (it & 1);  // slot of S(z-2), then of S(z)
        S s1[19], s2[19];
        C f[19];
#pragma unroll
        for (int q = 0; q < 19; q++) {
            f[q] = Storage<S>::template load<C>(cur[q]);
            asm volatile("" : "+v"(f[q]));  // see two_step_pass
        }
        const X_t x1 = xcur;
        const bool more = it + 1 < nit;
        const bool do2 = mid && it >= 2, do3 = owner && it >= 4;
        X_t x2{}, x3{};
        if (do2) x2 = op.extra_second(site_of(wrapz(z - 1)));
        if (do3) x3 = op.extra_third(site_of(wrapz(z - 2)));
        if (active) {  // S1 at plane z
            C b[19];
            op.first(f, b, x1, site_of(wrapz(z)), owner && z >= za && z < zb);
#pragma unroll
            for (int q = 0; q < 19; q++) s1[q] = Storage<S>::template store<C>(b[q]);
#pragma unroll
            for (int q = 0; q < 19; q++)
                if (pgroup<SGN>(q) == 0) l1[pbase(0) + pslot<SGN>(q)][ey][ex] = s1[q];
        }
        if (more) load_part(z + 1, nxt, xnxt, q0, qA);
        lds_barrier();
        if (do2) {  // S2 at plane z-1
            C f2[19], b[19];
#pragma unroll
            for (int q = 0; q < 19; q++)
                f2[q] = Storage<S>::template load<C>(pull(l1, q, ey, ex, pb, s1[QN], r1O_2, r1R_1));
            op.second(f2, b, x2, site_of(wrapz(z - 1)), owner && z - 1 >= za && z - 1 < zb);
#pragma unroll
            for (int q = 0; q < 19; q++) s2[q] = Storage<S>::template store<C>(b[q]);
#pragma unroll
            for (int q = 0; q < 19; q++)
                if (pgroup<SGN>(q) == 0) l2[pbase(0) + pslot<SGN>(q)][ey - 1][ex - 1] = s2[q];
        }
        if (more) load_part(z + 1, nxt, xnxt, qA, qB);
        lds_barrier();
        if (active) {
#pragma unroll
            for (int q = 0; q < 19; q++) {
                if (pgroup<SGN>(q) == 1) l1[pbase(1) + pslot<SGN>(q)][ey][ex] = s1[q];
                if (pgroup<SGN>(q) == 2) l1[pb + pslot<SGN>(q)][ey][ex] = s1[q];
            }
        }
        if (do3) {  // S3 at plane z-2
            C f3[19], b[19];
#pragma unroll
            for (int q = 0; q < 19; q++)
                f3[q] = Storage<S>::template load<C>(
                    pull(l2, q, ey - 1, ex - 1, pb, s2[QN], r2O_2, r2R_1));
            op.third(f3, b, x3, site_of(wrapz(z - 2)));
            if constexpr (Op::KEEP_STATE) {
                const __amdgpu_buffer_rsrc_t R =
                    raw_rsrc(dst.p[0] + (ptrdiff_t)wrapz(z - 2) * plane);
#pragma unroll
                for (int q = 0; q < 19; q++)
                    buf_st<S>(Storage<S>::template store<C>(b[q]), R, X[1] + Y[1],
                              (uint32_t)q * g.pstride);
            }
        }
        if (more) load_part(z + 1, nxt, xnxt, qB, q19);
        lds_barrier();
        if (do2) {
#pragma unroll
            for (int q = 0; q < 19; q++) {
                if (pgroup<SGN>(q) == 1) l2[pbase(1) + pslot<SGN>(q)][ey - 1][ex - 1] = s2[q];
                if (pgroup<SGN>(q) == 2) l2[pb + pslot<SGN>(q)][ey - 1][ex - 1] = s2[q];
            }
            r2O_2 = r2O_1;
            r2O_1 = s2[QO];
            r2R_1 = s2[QR];
        }
        r1O_2 = r1O_1;
        r1O_1 = s1[QO];
        r1R_1 = s1[QR];
    };

    S ra[19];
    X_t xa;
    load_part(za - 2, ra, xa, q0, q19);
    if constexpr (Op::PINGPONG) {
        S rb[19];
        X_t xb;
        int it = 0;
        for (; it + 1 < nit; it += 2) {  // two register sets taking turns
            iteration(it, ra, xa, rb, xb);
            iteration(it + 1, rb, xb, ra, xa);
        }
        if (it < nit) iteration(it, ra, xa, rb, xb);
    } else {
        for (int it = 0; it < nit; it++) iteration(it, ra, xa, ra, xa);
    }
}

}  // namespace bgk
