// bgk_collide.h -- the fused pull stream + BGK collide kernel, shared by the
// time-step path (bgk_kernels.hip) and the adjoint's taping forward pass
// (bgk_adjoint.hip).  SRC/col_MC.F90:104-361.
#pragma once

#include "bgk_internal.h"
#include "d3q19.h"

namespace bgk {

template <typename S>
struct Pops {
    S *p[19];
};

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
struct Geo {
    int nx, ny, nz, nxy, zoff, periodic_z;
    int tiles_x, tiles_y;
    int x0, x1, y0, y1, z0;
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

// ------------------------------------------------------------------ collide
// TAPE: also write the pulled-moment velocity u = j/rho of every site to
// tape[0..2][site] (the adjoint's record of the forward pass).
template <typename S, typename C, bool TAPE = false>
__global__ __launch_bounds__(Tile<S>::X * Tile<S>::Y) void k_collide(Pops<S> src, Pops<S> dst, Geo g, C omega,
                                                 C cte1, C *__restrict__ tape = nullptr,
                                                 long long tape_ld = 0) {
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
    const int X[3] = {xm, x, xp};
    const int Y[3] = {ym * g.nx, y * g.nx, yp * g.nx};
    const int Z[3] = {zm * g.nxy, zs * g.nxy, zp * g.nxy};

    C f[19];
#pragma unroll
    for (int q = 0; q < 19; q++) {
        // pull: population q arrives from x - c_q
        const int idx = X[1 - cx(q)] + Y[1 - cy(q)] + Z[1 - cz(q)];
        f[q] = Storage<S>::template load<C>(src.p[q][idx]);
    }
    C b[19];
    C ux, uy, uz;
    collide_site<C>(f, b, omega, cte1, ux, uy, uz);
    const int c = x + Y[1] + Z[1];
    if (TAPE) {
        tape[c] = ux;
        tape[tape_ld + c] = uy;
        tape[2 * tape_ld + c] = uz;
    }
#pragma unroll
    for (int q = 0; q < 18; q++) dst.p[q][c] = Storage<S>::template store<C>(b[q]);
    src.p[18][c] = Storage<S>::template store<C>(b[18]);  // rest: in place
}

}  // namespace bgk
