// bgk_internal.h -- device context and kernel launchers shared by the C-ABI
// (bgk_capi.hip) and the kernels (bgk_kernels.hip, bgk_adjoint.hip,
// bgk_hit.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/bgk3d_hit.h"

namespace bgk {

// a box of interior sites [x0,x1) x [y0,y1) x [z0,z1), 0-based
struct Box {
    int x0, x1, y0, y1, z0, z1;
    long long sites() const { return (long long)(x1 - x0) * (y1 - y0) * (z1 - z0); }
};

// Everything a kernel launch needs to find the populations.
//
// Layout in HBM: each of the two copies (the reference's a/b arrays) is one
// block ordered [z][y][q][x]: stored plane z, row y, then the 19 populations
// of that row side by side, x fastest.  Population q of stored site
// (x, y, zs) is element x + y*ystride + zs*zstride of pop[copy][q], with
// pop[copy][q] = copy base + q*nx, ystride = 19*nx, zstride = 19*nx*ny.  A
// row of all 19 populations is one contiguous 19*nx*s byte run, so the 38
// load and store streams of a step stay within a few DRAM pages (measured
// +10 % on the collide kernels against 19 separate arrays per copy).
//   periodic layout: interior plane z is stored plane z
//   slab layout    : interior plane z is stored plane z+1, ghosts 0 and nz+1
struct Lattice {
    int nx, ny, nz;  // interior sites
    int zoff;        // 0 periodic, 1 slab
    int nzs;         // stored planes: nz or nz+2
    int layout, storage;
    int shifted;
    long long nsite;       // nx*ny*nz
    long long plane_elems; // nx*ny*nzs stored sites
    long long ystride;     // elements between rows y and y+1 of a population
    long long zstride;     // elements between stored planes z and z+1
    long long copy_elems;  // elements of one copy (19*nx*ny*nzs)
    void *pop[2][19];      // population base pointers of copy 0/1
    int two_step;          // bgk_step may fuse two time steps per pass: 0 no,
                           // 1 where it pays, 2 at any size
    int three_step;        // bgk_step may fuse three time steps per pass
                           // (where two_step allows it)
    int ncu;               // compute units of the device
    int wide_offsets;      // BGK_FLAG_WIDE_OFFSETS: never take 32-bit offsets
};

// Element offset of interior site i (x fastest, then y, then z) of a
// population, for kernels that walk the sites in order.
struct SiteMap {
    int nx, ny, zoff;
    long long ystride, zstride;
};
inline SiteMap site_map(const Lattice &L) {
    return SiteMap{L.nx, L.ny, L.zoff, L.ystride, L.zstride};
}
__host__ __device__ inline long long site_elem(const SiteMap &m, long long i) {
    const long long nxy = (long long)m.nx * m.ny;
    const long long z = i / nxy, r = i - z * nxy;
    const long long y = r / m.nx, x = r - y * m.nx;
    return x + y * m.ystride + (z + m.zoff) * m.zstride;
}

struct Timing {
    bool on = false;
    std::vector<hipEvent_t> pool;  // reusable events
    size_t used = 0;
    struct Rec {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    static constexpr int NCLS = 16;  // room for the TCLS_ classes
    double ms[NCLS] = {};
    long long launches[NCLS] = {};
};

}  // namespace bgk

struct bgk_ctx {
    bgk::Lattice lat;
    int flags = 0;
    double omega = 1.0;
    int border = 10;
    int cur = 0;  // which copy holds the current state
    int device = 0;
    long long steps = 0;
    size_t bytes = 0;

    hipStream_t s0 = nullptr;  // main stream (interior work)
    hipStream_t s1 = nullptr;  // halo / boundary stream (slab overlap)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    void *pop_alloc = nullptr;  // one allocation holding the two copies

    // slab halo message: [face 0 = z+ (plane nz), 1 = z- (plane 1)][5][nx*ny]
    void *halo_send = nullptr;
    size_t halo_bytes = 0;
    // slab fused boundary paths, one allocation in the lattice's layout:
    // slab_s1_pop at planes 0..11 (two-step: first-step planes -1..2 and
    // nz-3..nz at 0..7; three-step: -1..4 and nz-5..nz at 0..11),
    // slab_s2_pop at planes 12..21 (three-step second step: -1..3 and
    // nz-4..nz at 0..9); 8 planes and no slab_s2_pop where three-step passes
    // cannot run (BGK_FLAG_NO_THREE_STEP, fewer than 12 planes)
    void *slab_s1 = nullptr;
    void *slab_s1_pop[19] = {};  // plane buffer of the fused passes' first step
    void *slab_s2_pop[19] = {};  // of the three-step pass's second step

    // stale-halo images (reference ghost semantics), per copy:
    // gx[19][ny*nz], gy[19][nx*nz], gz[19][nx*ny]
    void *gcache[2] = {nullptr, nullptr};
    size_t gcache_bytes = 0;

    // reductions
    double *red_partials = nullptr;  // [max(red_blocks, reduction_parts)][8]
    double *red_out = nullptr;       // [8] device
    double *red_host = nullptr;      // [8] pinned host
    int red_blocks = 0;

    // staging for host <-> device population transfers (one plane of doubles)
    double *staging = nullptr;

    bgk::Timing timing;
};

namespace bgk {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define BGK_HIP(call)                                 \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

size_t storage_bytes(int storage);

// sizes, pitches and pass options of a lattice (bgk_create's arguments);
// the population pointers are left null for the caller to set
void lattice_geometry(Lattice &L, int lx, int ly, int lz, int storage, int layout, int flags,
                      int device);

// timing helpers (no-ops unless BGK_FLAG_TIMING)
// TCLS_PASS2 / TCLS_PASS3: collide launches of two / three time steps
// (counted with the collide launches in bgk_timing); TCLS_TAPE / TCLS_ADJOINT:
// the adjoint's taped forward launches and its adjoint-step launches
// TCLS_PASS2D: passes of two time steps and the diagnostics of their result
// (counted with the collide launches in bgk_timing)
enum { TCLS_COLLIDE = 0, TCLS_REDUCE = 1, TCLS_HALO = 2, TCLS_STEP = 3, TCLS_PASS2 = 4,
       TCLS_PASS3 = 5, TCLS_TAPE = 6, TCLS_ADJOINT = 7, TCLS_PASS2D = 8, TCLS_N = 9 };
static_assert(TCLS_N <= Timing::NCLS, "timing classes");
hipError_t timing_begin(bgk_ctx *c, int cls, hipStream_t s, hipEvent_t *out);
hipError_t timing_end(bgk_ctx *c, int cls, hipStream_t s, hipEvent_t begin);

// ---- launchers (bgk_kernels.hip) ----
// fused pull-stream + collide from copy `src` to copy 1-src over `box`
hipError_t launch_collide(const Lattice &L, int src, const Box &box, double omega,
                          hipStream_t s);
// two time steps in one pass from copy `src` into copy 1-src; only when
// two_step_ok(L).  faces (periodic layout, may be null):
// receives the face images of the intermediate step (k_pack_faces layout).
bool two_step_ok(const Lattice &L);
hipError_t launch_collide2(const Lattice &L, int src, double omega, void *faces,
                           hipStream_t s);
// three time steps in one pass from copy `src` into copy 1-src (periodic
// layout; slab: the bulk planes 3 .. nz-4); only when three_step_ok(L).
// faces1 / faces2 (periodic layout, both or neither): receive the face
// images of the first / second step (k_pack_faces layout).
bool three_step_ok(const Lattice &L);
// thinnest slab the three-step pass takes (its bulk covers planes 3..nz-4)
constexpr int SLAB3_MIN_NZ = 12;
hipError_t launch_collide3(const Lattice &L, int src, double omega, void *faces1, void *faces2,
                           hipStream_t s);
// two time steps from copy `src` into copy 1-src and the diagnostics of
// the result in one pass (periodic layout, where three_step_ok(L) holds):
// partials holds 2 * 8 * collide3_diag_parts(L) doubles, out 16 (out[0..7]
// as launch_diagno's, out[8..15] as launch_dissipation's, consistent mode)
int collide3_diag_parts(const Lattice &L);
hipError_t launch_collide3_diag(const Lattice &L, int src, double omega, double svisc,
                                double *partials, double *out, hipStream_t s);
// diagno and dissipation in one pass: partials holds 2 * 8 * parts doubles,
// out 16 (out[0..7] as launch_diagno's, out[8..15] as launch_dissipation's)
hipError_t launch_diag_diss(const Lattice &L, int cur, const void *gcache, double svisc,
                            double *partials, double *out, hipStream_t s);
// reductions: at-site (diagno) and pulled (dissipation)
//   mode 0: pulled gather with periodic wrap (or slab ghosts as stored)
//   mode 1: pulled gather reading face images from gcache (periodic layout)
//   partials: reduction_parts(L) slots of 8 doubles
int reduction_parts(const Lattice &L);
hipError_t launch_diagno(const Lattice &L, int cur, double *partials, double *out,
                         hipStream_t s);
hipError_t launch_dissipation(const Lattice &L, int cur, const void *gcache, double svisc,
                              double *partials, double *out, hipStream_t s);
// slab halo path
//   pack: +z message from stored plane top, -z message from stored plane bot;
//   unpack: +z message into stored plane lo, -z message into stored plane hi
hipError_t launch_pack_planes(const Lattice &L, void *const *pops, int top, int bot, void *send,
                              hipStream_t s);
hipError_t launch_unpack_planes(const Lattice &L, void *const *pops, int lo, int hi,
                                const void *recv, hipStream_t s);
// one step over a slab box between arbitrary population arrays: interior
// plane z read from stored plane z + src_off of src, written to z + dst_off
hipError_t launch_collide_planes(const Lattice &L, void *const *src, int src_off,
                                 void *const *dst, int dst_off, const Box &box, double omega,
                                 hipStream_t s);
hipError_t launch_pack_z(const Lattice &L, int cur, void *send, hipStream_t s);
hipError_t launch_unpack_z(const Lattice &L, int cur, const void *recv, hipStream_t s);
// stale-halo images of the periodic layout
hipError_t launch_pack_faces(const Lattice &L, int cur, void *gcache, hipStream_t s);
// population transfer: interior sites i0 .. i0+count-1 of one population
// from/to a device array of doubles
hipError_t launch_plane_from_f64(const Lattice &L, int copy, int q, const double *src,
                                 long long i0, long long count, hipStream_t s);
hipError_t launch_plane_to_f64(const Lattice &L, int copy, int q, double *dst, long long i0,
                               long long count, hipStream_t s);
// equilibrium of a device velocity field u = 3 planes of `ld` doubles
// describing sites i0 .. i0+count-1
hipError_t launch_init_equilibrium(const Lattice &L, int copy, const double *u, long long ld,
                                   long long i0, long long count, hipStream_t s);
// order-independent hash of the stored interior populations of copy `cur`
// (partials: nblocks 64-bit words, out: one 64-bit word, both on the device)
hipError_t launch_checksum(const Lattice &L, int cur, void *partials, int nblocks, void *out,
                           hipStream_t s);
// populations of box b of copy `cur` into out[19][box sites] (doubles, device)
hipError_t launch_gather_box(const Lattice &L, int cur, const Box &b, double *out,
                             hipStream_t s);
// zero the ghost planes of the slab layout of one copy
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);

// ---- HIT synthesis (bgk_hit.cpp / bgk_hit.hip) ----
struct HitSpectrum {
    int Kx, Ky, Kz;                 // mode box half widths
    std::vector<double> coef;       // [3][2Kz+1][2Ky+1][Kx+1] complex (re,im) pairs
    std::vector<double> twx, twy, twz;  // cos,sin tables: 2*lx, 2*ly, 2*lz
};
int hit_spectrum(int lx, int ly, int lz, double k0, double urms, uint64_t seed, HitSpectrum &sp);
void hit_synthesize_host(int lx, int ly, int lz, const HitSpectrum &sp, double *u);
int hit_synthesize_device(bgk_ctx *c, const HitSpectrum &sp, double *u_dev);

// ---- adjoint (bgk_adjoint.hip) ----
int adjoint_ke(bgk_ctx *c, const double *u0_dev, int nsteps, int ckpt_every, int levels,
               double *J, double *grad_dev, double stats[4]);
int forward_ke(bgk_ctx *c, const double *u0_dev, int nsteps, double *J);

// shared helpers used by the adjoint
int reduce_ke(bgk_ctx *c, int copy, double *J);
// nsteps of bgk_step on the periodic layout, queued on `stream` (null: the
// context's main stream s0)
int step_periodic(bgk_ctx *c, int nsteps, bool count = true, hipStream_t stream = nullptr);

}  // namespace bgk
