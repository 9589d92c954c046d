// bgk_capi.hip -- the extern "C" entry points of include/bgk3d_hit.h.
//
// The context replaces the reference's module globals (SRC/storage.F90) and
// its OpenACC data region (SRC/bgk3d_HIT.F90:111-115): both population copies
// stay resident in HBM for the context's lifetime and the a/b pointer swap
// (SRC/col_MC.F90:365-421) is the index `cur`.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "bgk_internal.h"

namespace bgk {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char *what) {
    g_err = std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) +
            ") in " + what;
    return e == hipErrorOutOfMemory ? BGK_ENOMEM : BGK_EHIP;
}

size_t storage_bytes(int storage) {
    switch (storage) {
    case BGK_STORE_F64: return 8;
    case BGK_STORE_F32:
    case BGK_STORE_F32_MIXED: return 4;
    case BGK_STORE_F16:
    case BGK_STORE_BF16: return 2;
    default: return 0;
    }
}

// ---------------------------------------------------------------- timing
static hipError_t next_event(bgk_ctx *c, hipEvent_t *ev) {
    Timing &T = c->timing;
    if (T.used == T.pool.size()) {
        hipEvent_t e;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        T.pool.push_back(e);
    }
    *ev = T.pool[T.used++];
    return hipSuccess;
}

static hipError_t harvest(bgk_ctx *c) {
    Timing &T = c->timing;
    if (T.recs.empty()) return hipSuccess;
    hipError_t r = hipEventSynchronize(T.recs.back().b);
    if (r != hipSuccess) return r;
    for (auto &rec : T.recs) {
        r = hipEventSynchronize(rec.b);
        if (r != hipSuccess) return r;
        float ms = 0.f;
        r = hipEventElapsedTime(&ms, rec.a, rec.b);
        if (r != hipSuccess) return r;
        T.ms[rec.cls] += ms;
        T.launches[rec.cls] += 1;
    }
    T.recs.clear();
    T.used = 0;
    return hipSuccess;
}

static void timing_clear(Timing &T) {
    for (int k = 0; k < TCLS_N; k++) {
        T.ms[k] = 0;
        T.launches[k] = 0;
    }
}

hipError_t timing_begin(bgk_ctx *c, int cls, hipStream_t s, hipEvent_t *out) {
    *out = nullptr;
    if (!c->timing.on) return hipSuccess;
    if (c->timing.recs.size() > 8192) {
        hipError_t r = harvest(c);
        if (r != hipSuccess) return r;
    }
    hipError_t r = next_event(c, out);
    if (r != hipSuccess) return r;
    (void)cls;
    return hipEventRecord(*out, s);
}

hipError_t timing_end(bgk_ctx *c, int cls, hipStream_t s, hipEvent_t begin) {
    if (!c->timing.on || !begin) return hipSuccess;
    hipEvent_t e;
    hipError_t r = next_event(c, &e);
    if (r != hipSuccess) return r;
    r = hipEventRecord(e, s);
    if (r != hipSuccess) return r;
    c->timing.recs.push_back({cls, begin, e});
    return hipSuccess;
}

// ------------------------------------------------------------- region boxes
static Box full_box(const Lattice &L) { return Box{0, L.nx, 0, L.ny, 0, L.nz}; }

// bulk / border split of col_MC_masked: the reference mask
// (SRC/init.F90:44-62, tested at SRC/col_MC_masked.F90:114): a site is bulk
// (mask 0) when it lies at least `border` sites from every face of the
// rank's box, border (mask 1) otherwise; with 2*border >= a side there is
// no bulk.  The same on both layouts (a slab's ghost planes are outside the
// box): with border >= 1 the bulk reads no ghost plane.
static void mask_boxes(const bgk_ctx *c, int opt, std::vector<Box> &out) {
    const Lattice &L = c->lat;
    out.clear();
    const int b = c->border;
    const bool has_bulk = (2 * b < L.nx) && (2 * b < L.ny) && (2 * b < L.nz);
    if (!has_bulk) {
        if (opt == 1) out.push_back(full_box(L));
        return;
    }
    if (opt == 0) {
        out.push_back(Box{b, L.nx - b, b, L.ny - b, b, L.nz - b});
        return;
    }
    if (b == 0) return;
    out.push_back(Box{0, L.nx, 0, L.ny, 0, b});
    out.push_back(Box{0, L.nx, 0, L.ny, L.nz - b, L.nz});
    out.push_back(Box{0, L.nx, 0, b, b, L.nz - b});
    out.push_back(Box{0, L.nx, L.ny - b, L.ny, b, L.nz - b});
    out.push_back(Box{0, b, b, L.ny - b, b, L.nz - b});
    out.push_back(Box{L.nx - b, L.nx, b, L.ny - b, b, L.nz - b});
}

// the slab's overlapped single step (step_slab): the bulk planes 1..nz-2,
// which read no ghost plane, and the two boundary planes that do
static void slab_plane_boxes(const Lattice &L, std::vector<Box> &bulk, std::vector<Box> &bnd) {
    bulk.clear();
    bnd.clear();
    if (L.nz > 2) bulk.push_back(Box{0, L.nx, 0, L.ny, 1, L.nz - 1});
    bnd.push_back(Box{0, L.nx, 0, L.ny, 0, 1});
    if (L.nz > 1) bnd.push_back(Box{0, L.nx, 0, L.ny, L.nz - 1, L.nz});
}

static int collide_boxes(bgk_ctx *c, const std::vector<Box> &boxes, hipStream_t s) {
    for (const Box &b : boxes) {
        hipEvent_t ev;
        BGK_HIP(timing_begin(c, TCLS_COLLIDE, s, &ev));
        BGK_HIP(launch_collide(c->lat, c->cur, b, c->omega, s));
        BGK_HIP(timing_end(c, TCLS_COLLIDE, s, ev));
    }
    return BGK_OK;
}

// slab halo exchange of copy `cur` on stream s: pack the 5 crossing
// populations of both z faces into the message buffer, deliver, unpack into
// the ghost planes.  In a periodic self-exchange this rank is its own up and
// down neighbour, so the delivered message is the packed buffer itself; a
// multi-GPU build sends halo_send to the neighbours' halo_recv here
// (ncclSend / ncclRecv) and unpacks from halo_recv.
static int halo_exchange(bgk_ctx *c, hipStream_t s) {
    hipEvent_t ev;
    BGK_HIP(timing_begin(c, TCLS_HALO, s, &ev));
    BGK_HIP(launch_pack_z(c->lat, c->cur, c->halo_send, s));
    BGK_HIP(launch_unpack_z(c->lat, c->cur, c->halo_send, s));
    BGK_HIP(timing_end(c, TCLS_HALO, s, ev));
    return BGK_OK;
}

static int record_faces(bgk_ctx *c, hipStream_t s) {
    hipEvent_t ev;
    BGK_HIP(timing_begin(c, TCLS_HALO, s, &ev));
    BGK_HIP(launch_pack_faces(c->lat, c->cur, c->gcache[c->cur], s));
    BGK_HIP(timing_end(c, TCLS_HALO, s, ev));
    return BGK_OK;
}

// With reference halo images, each copy's image slot holds the faces the
// reference's ghosts of that array would hold: after time T, slot[cur] the
// faces of S_{T-2} and slot[1-cur] those of S_{T-1}.  A single step records
// S_T into slot[cur] before it; a two-step pass from S_T (copy a) to S_{T+2}
// (copy 1-a) records S_T into slot[1-a] and writes the faces of its
// intermediate S_{T+1} into slot[a] from inside the pass; a three-step pass
// to S_{T+3} writes the faces of S_{T+1} into slot[1-a] and those of S_{T+2}
// into slot[a] from inside the pass.  The images are read only through
// slot[cur] (the dissipation) and a later single step keeps slot[1-cur]; so
// after a call of n steps only S_{n-2} and S_{n-1} must be recorded, and a
// pass writes images only if it ends at step n-1 or later (the last pass of
// the call, and the one before it when a single step ends the call).  The
// passes before go without them (512^3 three-step pass with images: FP32
// 5.04 against 4.48 ms, FP16 4.85 against 3.88 ms).
int step_periodic(bgk_ctx *c, int nsteps, bool count, hipStream_t stream) {
    hipStream_t s = stream ? stream : c->s0;
    std::vector<Box> all{full_box(c->lat)};
    const bool refimg = (c->flags & BGK_FLAG_REFCOMPAT_HALOS) != 0;
    const int n0 = nsteps;
    if (three_step_ok(c->lat)) {
        // a remainder of one step: two two-step passes rather than a
        // three-step pass and a single step (512^3 FP32: 8.0 against 8.9 ms)
        int n3 = nsteps / 3;
        if (nsteps % 3 == 1 && n3 > 0 && two_step_ok(c->lat)) n3--;
        for (; n3 > 0; n3--, nsteps -= 3) {
            const int a = c->cur;
            const bool images = refimg && nsteps - 3 <= 1;
            hipEvent_t ev;
            BGK_HIP(timing_begin(c, TCLS_PASS3, s, &ev));
            BGK_HIP(launch_collide3(c->lat, a, c->omega, images ? c->gcache[1 - a] : nullptr,
                                    images ? c->gcache[a] : nullptr, s));
            BGK_HIP(timing_end(c, TCLS_PASS3, s, ev));
            c->cur ^= 1;
        }
    }
    if (two_step_ok(c->lat)) {
        for (; nsteps >= 2; nsteps -= 2) {
            const int a = c->cur;
            const bool images = refimg && nsteps - 2 <= 1;
            hipEvent_t ev;
            if (images) {
                BGK_HIP(timing_begin(c, TCLS_HALO, s, &ev));
                BGK_HIP(launch_pack_faces(c->lat, a, c->gcache[1 - a], s));
                BGK_HIP(timing_end(c, TCLS_HALO, s, ev));
            }
            BGK_HIP(timing_begin(c, TCLS_PASS2, s, &ev));
            BGK_HIP(launch_collide2(c->lat, a, c->omega, images ? c->gcache[a] : nullptr,
                                    s));
            BGK_HIP(timing_end(c, TCLS_PASS2, s, ev));
            c->cur ^= 1;
        }
    }
    for (int t = 0; t < nsteps; t++) {
        if (refimg && nsteps - t <= 2) {
            int r = record_faces(c, s);
            if (r) return r;
        }
        int r = collide_boxes(c, all, s);
        if (r) return r;
        c->cur ^= 1;
    }
    if (count) c->steps += n0;
    return BGK_OK;
}

// Slab time step with overlap (the structure of SRC/bcond.comm.step9.F90:
// pack, exchange, bulk collide overlapping the exchange, unpack, border
// collide).  Stream s0 runs the bulk planes 1..nz-2, stream s1 the halo path
// and the two boundary planes.  Cross-stream dependencies per step t:
//   s1 boundary(t) reads planes 1 and nz-2 written by s0 bulk(t-1)
//   s0 bulk(t) reads planes 0 and nz-1 written by s1 boundary(t-1) and
//   overwrites the buffer s1 boundary(t-1) read.
static int step_slab(bgk_ctx *c, int nsteps) {
    std::vector<Box> bulk, border;
    slab_plane_boxes(c->lat, bulk, border);
    // both streams start after all previously queued work of the context
    BGK_HIP(hipEventRecord(c->ev0, c->s0));
    BGK_HIP(hipStreamWaitEvent(c->s1, c->ev0, 0));
    BGK_HIP(hipEventRecord(c->ev1, c->s1));
    for (int t = 0; t < nsteps; t++) {
        // s1: halo path does not depend on s0's previous step
        int r = halo_exchange(c, c->s1);
        if (r) return r;
        BGK_HIP(hipStreamWaitEvent(c->s1, c->ev0, 0));  // bulk(t-1) done
        // s0: bulk(t) after boundary(t-1)
        BGK_HIP(hipStreamWaitEvent(c->s0, c->ev1, 0));
        r = collide_boxes(c, bulk, c->s0);
        if (r) return r;
        r = collide_boxes(c, border, c->s1);
        if (r) return r;
        BGK_HIP(hipEventRecord(c->ev0, c->s0));
        BGK_HIP(hipEventRecord(c->ev1, c->s1));
        c->cur ^= 1;
        c->steps++;
    }
    // join: later work on s0 sees the boundary planes
    BGK_HIP(hipStreamWaitEvent(c->s0, c->ev1, 0));
    return BGK_OK;
}

// Ghost planes as single steps leave them (what the reference-compatible
// dissipation reads, SRC/dissipation.F90:76-94): after step t the current
// copy's ghosts hold the images of S_{t-2}, the other copy's those of
// S_{t-1}.  A fused pass from S_T to S_{T+n} has both messages at hand (the
// exchanges of its first-step and (n-1)-th-step planes, or of S_T itself)
// and unpacks them into the two copies' ghosts as well.
//
// Two steps per pass on the slab.  Stream s0: the two-step pass over the
// bulk planes 2..nz-3 (it reads planes 0..nz-1 of the source copy only).
// Stream s1: the halo exchange of the source copy (the single step's
// message), the first step of planes 0..2 and nz-3..nz-1 into the plane
// buffer slab_s1 (interior plane z at buffer plane z+1, resp. z+7-nz), the
// exchange of the buffer's ghost planes 0 and 7 (again the 5 crossing
// populations per face), and the second step of planes 0, 1, nz-2, nz-1 into
// the destination copy.  Cross-stream dependencies per pair k:
//   s1 first step(k) reads planes 2, 3, nz-4, nz-3 written by s0 bulk(k-1);
//   s0 bulk(k) reads planes 0, 1, nz-2, nz-1 written by s1 boundary(k-1) and
//   overwrites the copy that s1 boundary(k-1) read.
static int step_slab2(bgk_ctx *c, int npairs) {
    const Lattice &L = c->lat;
    const int nz = L.nz;
    const Box lo3{0, L.nx, 0, L.ny, 0, 3}, hi3{0, L.nx, 0, L.ny, nz - 3, nz};
    const Box lo2{0, L.nx, 0, L.ny, 0, 2}, hi2{0, L.nx, 0, L.ny, nz - 2, nz};
    BGK_HIP(hipEventRecord(c->ev0, c->s0));
    BGK_HIP(hipStreamWaitEvent(c->s1, c->ev0, 0));
    BGK_HIP(hipEventRecord(c->ev1, c->s1));
    for (int k = 0; k < npairs; k++) {
        const int a = c->cur, b = 1 - c->cur;
        int r = halo_exchange(c, c->s1);  // does not depend on s0's previous pass
        if (r) return r;
        // the destination's ghosts as single steps leave them: S_T's images
        BGK_HIP(launch_unpack_planes(L, L.pop[b], 0, nz + 1, c->halo_send, c->s1));
        BGK_HIP(hipStreamWaitEvent(c->s1, c->ev0, 0));  // bulk(k-1) done
        BGK_HIP(hipStreamWaitEvent(c->s0, c->ev1, 0));  // boundary(k-1) done
        hipEvent_t ev;
        BGK_HIP(timing_begin(c, TCLS_PASS2, c->s0, &ev));
        BGK_HIP(launch_collide2(L, a, c->omega, nullptr, c->s0));
        BGK_HIP(timing_end(c, TCLS_PASS2, c->s0, ev));
        BGK_HIP(timing_begin(c, TCLS_COLLIDE, c->s1, &ev));
        BGK_HIP(launch_collide_planes(L, L.pop[a], L.zoff, c->slab_s1_pop, 1, lo3, c->omega,
                                      c->s1));
        BGK_HIP(launch_collide_planes(L, L.pop[a], L.zoff, c->slab_s1_pop, 7 - nz, hi3,
                                      c->omega, c->s1));
        BGK_HIP(timing_end(c, TCLS_COLLIDE, c->s1, ev));
        BGK_HIP(timing_begin(c, TCLS_HALO, c->s1, &ev));
        BGK_HIP(launch_pack_planes(L, c->slab_s1_pop, 6, 1, c->halo_send, c->s1));
        BGK_HIP(launch_unpack_planes(L, c->slab_s1_pop, 0, 7, c->halo_send, c->s1));
        // and the source's (read above) as single steps leave them: S_{T+1}'s
        BGK_HIP(launch_unpack_planes(L, L.pop[a], 0, nz + 1, c->halo_send, c->s1));
        BGK_HIP(timing_end(c, TCLS_HALO, c->s1, ev));
        BGK_HIP(timing_begin(c, TCLS_COLLIDE, c->s1, &ev));
        BGK_HIP(launch_collide_planes(L, c->slab_s1_pop, 1, L.pop[b], L.zoff, lo2, c->omega,
                                      c->s1));
        BGK_HIP(launch_collide_planes(L, c->slab_s1_pop, 7 - nz, L.pop[b], L.zoff, hi2,
                                      c->omega, c->s1));
        BGK_HIP(timing_end(c, TCLS_COLLIDE, c->s1, ev));
        BGK_HIP(hipEventRecord(c->ev0, c->s0));
        BGK_HIP(hipEventRecord(c->ev1, c->s1));
        c->cur ^= 1;
        c->steps += 2;
    }
    // join: later work on s0 sees the boundary planes
    BGK_HIP(hipStreamWaitEvent(c->s0, c->ev1, 0));
    return BGK_OK;
}

// Three steps per pass on the slab, the two-step structure one level deeper.
// Stream s0: the three-step pass over the bulk planes 3..nz-4 (it reads
// planes 0..nz-1 of the source copy only).  Stream s1: the halo exchange of
// the source copy; the first step of planes 0..4 and nz-5..nz-1 into plane
// buffer B1 (slab_s1_pop: interior plane z at buffer plane z+1, resp.
// z+11-nz), the exchange of B1's ghost planes 0 and 11; the second step of
// planes 0..3 and nz-4..nz-1 into B2 (slab_s2_pop: z+1, resp. z+9-nz), the
// exchange of B2's ghost planes 0 and 9; the third step of planes 0..2 and
// nz-3..nz-1 into the destination copy.  One exchange of the 5 crossing
// populations per face and step, as single steps.  Cross-stream
// dependencies per pass k as in step_slab2 (the boundary path reads planes
// 0..5 and nz-6..nz-1, the bulk writes planes 3..nz-4).
static int step_slab3(bgk_ctx *c, int ntriples) {
    const Lattice &L = c->lat;
    const int nz = L.nz;
    auto zbox = [&](int z0, int z1) { return Box{0, L.nx, 0, L.ny, z0, z1}; };
    void *const *B1 = c->slab_s1_pop;
    void *const *B2 = c->slab_s2_pop;
    BGK_HIP(hipEventRecord(c->ev0, c->s0));
    BGK_HIP(hipStreamWaitEvent(c->s1, c->ev0, 0));
    BGK_HIP(hipEventRecord(c->ev1, c->s1));
    for (int k = 0; k < ntriples; k++) {
        const int a = c->cur, b = 1 - c->cur;
        int r = halo_exchange(c, c->s1);  // does not depend on s0's previous pass
        if (r) return r;
        BGK_HIP(hipStreamWaitEvent(c->s1, c->ev0, 0));  // bulk(k-1) done
        BGK_HIP(hipStreamWaitEvent(c->s0, c->ev1, 0));  // boundary(k-1) done
        hipEvent_t ev;
        BGK_HIP(timing_begin(c, TCLS_PASS3, c->s0, &ev));
        BGK_HIP(launch_collide3(L, a, c->omega, nullptr, nullptr, c->s0));
        BGK_HIP(timing_end(c, TCLS_PASS3, c->s0, ev));
        // first step
        BGK_HIP(timing_begin(c, TCLS_COLLIDE, c->s1, &ev));
        BGK_HIP(launch_collide_planes(L, L.pop[a], L.zoff, B1, 1, zbox(0, 5), c->omega, c->s1));
        BGK_HIP(launch_collide_planes(L, L.pop[a], L.zoff, B1, 11 - nz, zbox(nz - 5, nz),
                                      c->omega, c->s1));
        BGK_HIP(timing_end(c, TCLS_COLLIDE, c->s1, ev));
        BGK_HIP(timing_begin(c, TCLS_HALO, c->s1, &ev));
        BGK_HIP(launch_pack_planes(L, B1, 10, 1, c->halo_send, c->s1));
        BGK_HIP(launch_unpack_planes(L, B1, 0, 11, c->halo_send, c->s1));
        // the destination's ghosts as single steps leave them: S_{T+1}'s images
        BGK_HIP(launch_unpack_planes(L, L.pop[b], 0, nz + 1, c->halo_send, c->s1));
        BGK_HIP(timing_end(c, TCLS_HALO, c->s1, ev));
        // second step
        BGK_HIP(timing_begin(c, TCLS_COLLIDE, c->s1, &ev));
        BGK_HIP(launch_collide_planes(L, B1, 1, B2, 1, zbox(0, 4), c->omega, c->s1));
        BGK_HIP(launch_collide_planes(L, B1, 11 - nz, B2, 9 - nz, zbox(nz - 4, nz), c->omega,
                                      c->s1));
        BGK_HIP(timing_end(c, TCLS_COLLIDE, c->s1, ev));
        BGK_HIP(timing_begin(c, TCLS_HALO, c->s1, &ev));
        BGK_HIP(launch_pack_planes(L, B2, 8, 1, c->halo_send, c->s1));
        BGK_HIP(launch_unpack_planes(L, B2, 0, 9, c->halo_send, c->s1));
        // and the source's (read above) as single steps leave them: S_{T+2}'s
        BGK_HIP(launch_unpack_planes(L, L.pop[a], 0, nz + 1, c->halo_send, c->s1));
        BGK_HIP(timing_end(c, TCLS_HALO, c->s1, ev));
        // third step into the destination copy
        BGK_HIP(timing_begin(c, TCLS_COLLIDE, c->s1, &ev));
        BGK_HIP(launch_collide_planes(L, B2, 1, L.pop[b], L.zoff, zbox(0, 3), c->omega, c->s1));
        BGK_HIP(launch_collide_planes(L, B2, 9 - nz, L.pop[b], L.zoff, zbox(nz - 3, nz),
                                      c->omega, c->s1));
        BGK_HIP(timing_end(c, TCLS_COLLIDE, c->s1, ev));
        BGK_HIP(hipEventRecord(c->ev0, c->s0));
        BGK_HIP(hipEventRecord(c->ev1, c->s1));
        c->cur ^= 1;
        c->steps += 3;
    }
    // join: later work on s0 sees the boundary planes
    BGK_HIP(hipStreamWaitEvent(c->s0, c->ev1, 0));
    return BGK_OK;
}

int reduce_ke(bgk_ctx *c, int copy, double *J) {
    BGK_HIP(launch_diagno(c->lat, copy, c->red_partials, c->red_out, c->s0));
    BGK_HIP(hipMemcpyAsync(c->red_host, c->red_out, 8 * sizeof(double), hipMemcpyDeviceToHost,
                           c->s0));
    BGK_HIP(hipStreamSynchronize(c->s0));
    *J = c->red_host[4];
    if (!std::isfinite(*J)) return fail(BGK_ENONFINITE, "kinetic energy is not finite");
    return BGK_OK;
}

static int check_ctx(bgk_ctx *c) {
    if (!c) return fail(BGK_EINVAL, "null context");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    return BGK_OK;
}

static const long long CHUNK = 1LL << 23;  // sites per host transfer chunk

void lattice_geometry(Lattice &L, int lx, int ly, int lz, int storage, int layout, int flags,
                      int device) {
    L.ncu = 0;
    if (hipDeviceGetAttribute(&L.ncu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        L.ncu < 1)
        L.ncu = 256;
    const int nzs = layout == BGK_LAYOUT_SLAB_Z ? lz + 2 : lz;
    L.nx = lx;
    L.ny = ly;
    L.nz = lz;
    L.layout = layout;
    L.storage = storage;
    L.zoff = layout == BGK_LAYOUT_SLAB_Z ? 1 : 0;
    L.nzs = nzs;
    L.shifted = (flags & BGK_FLAG_NOSHIFT) ? 0 : 1;
    L.nsite = (long long)lx * ly * lz;
    L.plane_elems = (long long)lx * ly * nzs;
    L.ystride = 19LL * lx;
    L.zstride = 19LL * lx * ly;
    L.copy_elems = 19LL * L.plane_elems;
    L.wide_offsets = (flags & BGK_FLAG_WIDE_OFFSETS) ? 1 : 0;
    L.two_step = !(flags & BGK_FLAG_SINGLE_STEP)
                     ? ((flags & BGK_FLAG_TWO_STEP_ANY_SIZE) ? 2 : 1)
                     : 0;
    L.three_step = (flags & BGK_FLAG_NO_THREE_STEP) ? 0 : 1;
    for (int k = 0; k < 2; k++)
        for (int q = 0; q < 19; q++) L.pop[k][q] = nullptr;
}

}  // namespace bgk

using namespace bgk;

extern "C" {

int bgk_version(void) { return 100; }

int bgk_last_error(char *buf, int len) {
    if (!buf || len <= 0) return BGK_EINVAL;
    snprintf(buf, (size_t)len, "%s", g_err.c_str());
    return BGK_OK;
}

int bgk_device_count(int *count) {
    if (!count) return fail(BGK_EINVAL, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return BGK_OK;
}

int bgk_create(int lx, int ly, int lz, int storage, int layout, int flags, double omega,
               bgk_ctx **out) {
    if (!out) return fail(BGK_EINVAL, "null output pointer");
    *out = nullptr;
    if (lx < 1 || ly < 1 || lz < 1)
        return fail(BGK_EINVAL, "lattice sizes must be positive");
    if (storage_bytes(storage) == 0) return fail(BGK_EINVAL, "unknown storage kind");
    if (layout != BGK_LAYOUT_PERIODIC && layout != BGK_LAYOUT_SLAB_Z)
        return fail(BGK_EINVAL, "unknown layout");
    if (layout == BGK_LAYOUT_SLAB_Z && lz < 2)
        return fail(BGK_EINVAL, "slab layout needs at least 2 planes");
    if (layout == BGK_LAYOUT_SLAB_Z && (flags & BGK_FLAG_REFCOMPAT_HALOS))
        return fail(BGK_EINVAL, "refcompat halo images apply to the periodic layout");
    if (!(omega >= 0.0 && omega < 2.0))
        return fail(BGK_EINVAL, "omega must lie in [0, 2)");
    const int nzs = layout == BGK_LAYOUT_SLAB_Z ? lz + 2 : lz;
    const long long plane = (long long)lx * ly * nzs;
    if (plane >= (1LL << 31))
        return fail(BGK_EINVAL, "more than 2^31 stored sites per population");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(BGK_ENODEV, "no GPU visible to the HIP runtime");

    bgk_ctx *c = new bgk_ctx();
    (void)hipGetDevice(&c->device);
    Lattice &L = c->lat;
    lattice_geometry(L, lx, ly, lz, storage, layout, flags, c->device);
    c->flags = flags;
    c->omega = omega;
    c->timing.on = (flags & BGK_FLAG_TIMING) != 0;

    const size_t es = storage_bytes(storage);
    const size_t cbytes = ((size_t)L.copy_elems * es + 255) & ~(size_t)255;
    int rc = BGK_OK;
    auto bail = [&](int code) {
        bgk_destroy(c);
        return code;
    };
    // two copies in the interleaved [z][y][q][x] layout (see Lattice)
    hipError_t e = hipMalloc(&c->pop_alloc, cbytes * 2);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc(populations)"));
    c->bytes += cbytes * 2;
    char *base = (char *)c->pop_alloc;
    for (int q = 0; q < 19; q++) {
        L.pop[0][q] = base + (size_t)q * lx * es;
        L.pop[1][q] = base + cbytes + (size_t)q * lx * es;
    }

    if ((e = hipStreamCreateWithFlags(&c->s0, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamCreate"));
    if ((e = hipStreamCreateWithFlags(&c->s1, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamCreate"));
    if ((e = hipEventCreateWithFlags(&c->ev0, hipEventDisableTiming)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));
    if ((e = hipEventCreateWithFlags(&c->ev1, hipEventDisableTiming)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));
    if ((e = hipMemsetAsync(c->pop_alloc, 0, cbytes * 2, c->s0)) != hipSuccess)
        return bail(hip_fail(e, "hipMemsetAsync(populations)"));

    if (layout == BGK_LAYOUT_SLAB_Z) {
        c->halo_bytes = (size_t)2 * 5 * lx * ly * es;
        if ((e = hipMalloc(&c->halo_send, c->halo_bytes)) != hipSuccess)
            return bail(hip_fail(e, "hipMalloc(halo)"));
        c->bytes += c->halo_bytes;
        if (L.two_step) {  // plane buffers of the fused boundary paths in the
                           // lattice's layout (same pitches): 22 planes where
                           // three-step passes may run, else 8
            const bool three = L.three_step && L.nz >= SLAB3_MIN_NZ;
            const size_t pbytes = (size_t)L.zstride * es;
            const size_t pb = ((three ? 22 : 8) * pbytes + 255) & ~(size_t)255;
            if ((e = hipMalloc(&c->slab_s1, pb)) != hipSuccess)
                return bail(hip_fail(e, "hipMalloc(slab plane buffer)"));
            c->bytes += pb;
            for (int q = 0; q < 19; q++) {
                c->slab_s1_pop[q] = (char *)c->slab_s1 + (size_t)q * lx * es;
                c->slab_s2_pop[q] =
                    three ? (char *)c->slab_s1 + 12 * pbytes + (size_t)q * lx * es : nullptr;
            }
            if ((e = hipMemsetAsync(c->slab_s1, 0, pb, c->s0)) != hipSuccess)
                return bail(hip_fail(e, "hipMemsetAsync(slab plane buffer)"));
        }
    }
    if (flags & BGK_FLAG_REFCOMPAT_HALOS) {
        c->gcache_bytes =
            (size_t)19 * ((size_t)ly * lz + (size_t)lx * lz + (size_t)lx * ly) * es;
        for (int k = 0; k < 2; k++) {
            if ((e = hipMalloc(&c->gcache[k], c->gcache_bytes)) != hipSuccess)
                return bail(hip_fail(e, "hipMalloc(halo images)"));
            if ((e = hipMemsetAsync(c->gcache[k], 0, c->gcache_bytes, c->s0)) != hipSuccess)
                return bail(hip_fail(e, "hipMemsetAsync(halo images)"));
        }
        c->bytes += 2 * c->gcache_bytes;
    }
    long long nb = (L.nsite + 255) / 256;
    c->red_blocks = (int)std::min<long long>(std::max<long long>(nb, 1), 2048);  // checksum
    const int nparts = std::max(c->red_blocks, reduction_parts(L));
    // two sets of 8 partial sums per block (diagno and dissipation in one pass)
    if ((e = hipMalloc(&c->red_partials, sizeof(double) * 16 * nparts)) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(reduction)"));
    if ((e = hipMalloc(&c->red_out, sizeof(double) * 16)) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(reduction)"));
    if ((e = hipHostMalloc(&c->red_host, sizeof(double) * 16, hipHostMallocDefault)) !=
        hipSuccess)
        return bail(hip_fail(e, "hipHostMalloc(reduction)"));
    const long long chunk = std::min(L.nsite, CHUNK);
    if ((e = hipMalloc(&c->staging, sizeof(double) * 3 * chunk)) != hipSuccess)
        return bail(hip_fail(e, "hipMalloc(staging)"));
    c->bytes += sizeof(double) * (3 * chunk + 8 * (size_t)nparts + 8);
    if ((e = hipStreamSynchronize(c->s0)) != hipSuccess)
        return bail(hip_fail(e, "hipStreamSynchronize"));
    (void)rc;
    *out = c;
    return BGK_OK;
}

int bgk_destroy(bgk_ctx *c) {
    if (!c) return BGK_OK;
    (void)hipSetDevice(c->device);
    if (c->s0) (void)hipStreamSynchronize(c->s0);
    if (c->s1) (void)hipStreamSynchronize(c->s1);
    for (hipEvent_t e : c->timing.pool) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->s0) (void)hipStreamDestroy(c->s0);
    if (c->s1) (void)hipStreamDestroy(c->s1);
    (void)hipFree(c->pop_alloc);
    (void)hipFree(c->halo_send);
    (void)hipFree(c->slab_s1);
    (void)hipFree(c->gcache[0]);
    (void)hipFree(c->gcache[1]);
    (void)hipFree(c->red_partials);
    (void)hipFree(c->red_out);
    if (c->red_host) (void)hipHostFree(c->red_host);
    (void)hipFree(c->staging);
    delete c;
    return BGK_OK;
}

int bgk_set_omega(bgk_ctx *c, double omega) {
    if (!c) return fail(BGK_EINVAL, "null context");
    if (!(omega >= 0.0 && omega < 2.0)) return fail(BGK_EINVAL, "omega must lie in [0, 2)");
    c->omega = omega;
    return BGK_OK;
}

int bgk_set_border(bgk_ctx *c, int border) {
    if (!c) return fail(BGK_EINVAL, "null context");
    if (border < 0) return fail(BGK_EINVAL, "border must be >= 0");
    c->border = border;
    return BGK_OK;
}

int bgk_info(bgk_ctx *c, int64_t info[8]) {
    if (!c || !info) return fail(BGK_EINVAL, "null argument");
    info[0] = c->lat.nx;
    info[1] = c->lat.ny;
    info[2] = c->lat.nz;
    info[3] = c->lat.storage;
    info[4] = c->lat.layout;
    info[5] = c->flags;
    info[6] = (int64_t)c->bytes;
    info[7] = c->steps;
    return BGK_OK;
}

int bgk_set_populations(bgk_ctx *c, const double *f) {
    int r = check_ctx(c);
    if (r) return r;
    if (!f) return fail(BGK_EINVAL, "null population array");
    const Lattice &L = c->lat;
    const long long chunk = std::min(L.nsite, CHUNK);
    for (int q = 0; q < 19; q++)
        for (long long i0 = 0; i0 < L.nsite; i0 += chunk) {
            const long long n = std::min(chunk, L.nsite - i0);
            BGK_HIP(hipMemcpyAsync(c->staging, f + (size_t)q * L.nsite + i0, sizeof(double) * n,
                                   hipMemcpyHostToDevice, c->s0));
            BGK_HIP(launch_plane_from_f64(L, c->cur, q, c->staging, i0, n, c->s0));
        }
    BGK_HIP(hipStreamSynchronize(c->s0));
    return BGK_OK;
}

int bgk_get_populations(bgk_ctx *c, double *f) {
    int r = check_ctx(c);
    if (r) return r;
    if (!f) return fail(BGK_EINVAL, "null population array");
    const Lattice &L = c->lat;
    const long long chunk = std::min(L.nsite, CHUNK);
    for (int q = 0; q < 19; q++)
        for (long long i0 = 0; i0 < L.nsite; i0 += chunk) {
            const long long n = std::min(chunk, L.nsite - i0);
            BGK_HIP(launch_plane_to_f64(L, c->cur, q, c->staging, i0, n, c->s0));
            BGK_HIP(hipMemcpyAsync(f + (size_t)q * L.nsite + i0, c->staging, sizeof(double) * n,
                                   hipMemcpyDeviceToHost, c->s0));
            BGK_HIP(hipStreamSynchronize(c->s0));
        }
    return BGK_OK;
}

int bgk_init_equilibrium(bgk_ctx *c, const double *u) {
    int r = check_ctx(c);
    if (r) return r;
    if (!u) return fail(BGK_EINVAL, "null velocity array");
    const Lattice &L = c->lat;
    const long long chunk = std::min(L.nsite, CHUNK);
    for (long long i0 = 0; i0 < L.nsite; i0 += chunk) {
        const long long n = std::min(chunk, L.nsite - i0);
        for (int d = 0; d < 3; d++)
            BGK_HIP(hipMemcpyAsync(c->staging + (size_t)d * n, u + (size_t)d * L.nsite + i0,
                                   sizeof(double) * n, hipMemcpyHostToDevice, c->s0));
        BGK_HIP(launch_init_equilibrium(L, c->cur, c->staging, n, i0, n, c->s0));
        BGK_HIP(hipStreamSynchronize(c->s0));
    }
    return BGK_OK;
}

int bgk_hit_velocity(int lx, int ly, int lz, double k0, double urms, uint64_t seed, double *u) {
    if (!u) return fail(BGK_EINVAL, "null velocity array");
    HitSpectrum sp;
    int r = hit_spectrum(lx, ly, lz, k0, urms, seed, sp);
    if (r) return r;
    hit_synthesize_host(lx, ly, lz, sp, u);
    return BGK_OK;
}

int bgk_init_hit(bgk_ctx *c, double k0, double urms, uint64_t seed) {
    int r = check_ctx(c);
    if (r) return r;
    const Lattice &L = c->lat;
    HitSpectrum sp;
    r = hit_spectrum(L.nx, L.ny, L.nz, k0, urms, seed, sp);
    if (r) return r;
    double *u = nullptr;
    hipError_t e = hipMalloc(&u, sizeof(double) * 3 * L.nsite);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(velocity field)");
    r = hit_synthesize_device(c, sp, u);
    if (r == BGK_OK) {
        e = launch_init_equilibrium(L, c->cur, u, L.nsite, 0, L.nsite, c->s0);
        if (e == hipSuccess) e = hipStreamSynchronize(c->s0);
        if (e != hipSuccess) r = hip_fail(e, "init_equilibrium");
    }
    (void)hipFree(u);
    return r;
}

int bgk_boundaries(bgk_ctx *c) {
    int r = check_ctx(c);
    if (r) return r;
    if (c->lat.layout == BGK_LAYOUT_SLAB_Z) return halo_exchange(c, c->s0);
    if (c->flags & BGK_FLAG_REFCOMPAT_HALOS) return record_faces(c, c->s0);
    return BGK_OK;
}

int bgk_col_MC(bgk_ctx *c) {
    int r = check_ctx(c);
    if (r) return r;
    std::vector<Box> all{full_box(c->lat)};
    r = collide_boxes(c, all, c->s0);
    if (r) return r;
    c->cur ^= 1;
    c->steps++;
    return BGK_OK;
}

int bgk_col_MC_masked(bgk_ctx *c, int opt) {
    int r = check_ctx(c);
    if (r) return r;
    if (opt != 0 && opt != 1) return fail(BGK_EINVAL, "opt must be 0 (bulk) or 1 (border)");
    std::vector<Box> boxes;
    mask_boxes(c, opt, boxes);
    r = collide_boxes(c, boxes, c->s0);
    if (r) return r;
    if (opt == 1) {  // SRC/col_MC_masked.F90:369: swap after the border update
        c->cur ^= 1;
        c->steps++;
    }
    return BGK_OK;
}

int bgk_step(bgk_ctx *c, int nsteps) {
    int r = check_ctx(c);
    if (r) return r;
    if (nsteps < 0) return fail(BGK_EINVAL, "nsteps must be >= 0");
    hipEvent_t ev;
    BGK_HIP(timing_begin(c, TCLS_STEP, c->s0, &ev));
    if (c->lat.layout == BGK_LAYOUT_SLAB_Z) {
        r = BGK_OK;
        if (three_step_ok(c->lat)) {  // remainders as on the periodic lattice
            int n3 = nsteps / 3;
            if (nsteps % 3 == 1 && n3 > 0 && two_step_ok(c->lat)) n3--;
            r = step_slab3(c, n3);
            nsteps -= 3 * n3;
        }
        if (!r && two_step_ok(c->lat)) {
            r = step_slab2(c, nsteps / 2);
            nsteps %= 2;
        }
        if (!r) r = step_slab(c, nsteps);
    } else {
        r = step_periodic(c, nsteps);
    }
    if (r) return r;
    BGK_HIP(timing_end(c, TCLS_STEP, c->s0, ev));
    return BGK_OK;
}

// diagno's outputs from the 6 reduced sums v (SRC/diagno.F90:110-121)
static int diagno_finish(const Lattice &L, int mode, const double *v, double out[5]) {
    // SRC/diagno.F90:110 leaves ztot = ytot + zj(l,m,n) after the loop
    const double ztot = (mode == BGK_DIAG_REFCOMPAT) ? v[2] + v[5] : v[3];
    const double tot[5] = {v[0], v[1], v[2], ztot, v[4]};
    for (int k = 0; k < 5; k++)  // SRC/diagno.F90:117-121 division order
        out[k] = ((tot[k] / (double)L.nx) / (double)L.ny) / (double)L.nz;
    for (int k = 0; k < 5; k++)
        if (!std::isfinite(out[k]))
            return fail(BGK_ENONFINITE, "diagno: non-finite moments (populations diverged)");
    return BGK_OK;
}

// where the dissipation's pulled gather finds the sources outside the box:
// *gc = face images (refcompat, periodic) or null (index wrap / slab ghosts,
// refreshed first in consistent mode)
static int dissipation_sources(bgk_ctx *c, int mode, const void **gc) {
    const Lattice &L = c->lat;
    *gc = nullptr;
    if (L.layout == BGK_LAYOUT_SLAB_Z) {
        // slab ghosts as the last exchange left them are the reference's
        // view; consistent mode refreshes them first
        if (mode == BGK_DIAG_CONSISTENT) return halo_exchange(c, c->s0);
    } else if (mode == BGK_DIAG_REFCOMPAT) {
        if (!(c->flags & BGK_FLAG_REFCOMPAT_HALOS))
            return fail(BGK_ESTATE,
                        "refcompat dissipation needs a context created with "
                        "BGK_FLAG_REFCOMPAT_HALOS");
        *gc = c->gcache[c->cur];
    }
    return BGK_OK;
}

static int dissipation_finish(const double *v, double out[2]) {
    out[0] = v[0];
    out[1] = v[1];
    if (!std::isfinite(out[0]) || !std::isfinite(out[1]))
        return fail(BGK_ENONFINITE, "dissipation: non-finite sums (populations diverged)");
    return BGK_OK;
}

int bgk_diagno(bgk_ctx *c, int mode, double out[5]) {
    int r = check_ctx(c);
    if (r) return r;
    if (!out) return fail(BGK_EINVAL, "null output");
    if (mode != BGK_DIAG_CONSISTENT && mode != BGK_DIAG_REFCOMPAT)
        return fail(BGK_EINVAL, "unknown diagnostic mode");
    const Lattice &L = c->lat;
    hipEvent_t ev;
    BGK_HIP(timing_begin(c, TCLS_REDUCE, c->s0, &ev));
    BGK_HIP(launch_diagno(L, c->cur, c->red_partials, c->red_out, c->s0));
    BGK_HIP(timing_end(c, TCLS_REDUCE, c->s0, ev));
    BGK_HIP(hipMemcpyAsync(c->red_host, c->red_out, 8 * sizeof(double), hipMemcpyDeviceToHost,
                           c->s0));
    BGK_HIP(hipStreamSynchronize(c->s0));
    return diagno_finish(L, mode, c->red_host, out);
}

int bgk_dissipation(bgk_ctx *c, double svisc, int mode, double out[2]) {
    int r = check_ctx(c);
    if (r) return r;
    if (!out) return fail(BGK_EINVAL, "null output");
    if (mode != BGK_DIAG_CONSISTENT && mode != BGK_DIAG_REFCOMPAT)
        return fail(BGK_EINVAL, "unknown diagnostic mode");
    const Lattice &L = c->lat;
    const void *gc = nullptr;
    if ((r = dissipation_sources(c, mode, &gc))) return r;
    hipEvent_t ev;
    BGK_HIP(timing_begin(c, TCLS_REDUCE, c->s0, &ev));
    BGK_HIP(launch_dissipation(L, c->cur, gc, svisc, c->red_partials, c->red_out, c->s0));
    BGK_HIP(timing_end(c, TCLS_REDUCE, c->s0, ev));
    BGK_HIP(hipMemcpyAsync(c->red_host, c->red_out, 8 * sizeof(double), hipMemcpyDeviceToHost,
                           c->s0));
    BGK_HIP(hipStreamSynchronize(c->s0));
    return dissipation_finish(c->red_host, out);
}

int bgk_diagnostics(bgk_ctx *c, double svisc, int mode, double out[7]) {
    int r = check_ctx(c);
    if (r) return r;
    if (!out) return fail(BGK_EINVAL, "null output");
    if (mode != BGK_DIAG_CONSISTENT && mode != BGK_DIAG_REFCOMPAT)
        return fail(BGK_EINVAL, "unknown diagnostic mode");
    const Lattice &L = c->lat;
    const void *gc = nullptr;
    if ((r = dissipation_sources(c, mode, &gc))) return r;
    hipEvent_t ev;
    BGK_HIP(timing_begin(c, TCLS_REDUCE, c->s0, &ev));
    BGK_HIP(launch_diag_diss(L, c->cur, gc, svisc, c->red_partials, c->red_out, c->s0));
    BGK_HIP(timing_end(c, TCLS_REDUCE, c->s0, ev));
    BGK_HIP(hipMemcpyAsync(c->red_host, c->red_out, 16 * sizeof(double), hipMemcpyDeviceToHost,
                           c->s0));
    BGK_HIP(hipStreamSynchronize(c->s0));
    if ((r = diagno_finish(L, mode, c->red_host, out))) return r;
    return dissipation_finish(c->red_host + 8, out + 5);
}

// The reference's loop body at a diagnostic step (SRC/bgk3d_HIT.F90:127-146:
// the time steps, then diagno and dissipation of the state they leave).  On
// the periodic lattice, where three-step passes run and no halo images are
// kept, and for step counts of the form 3k + 2 or 3k + 1 (k >= 1), the last
// two steps and the reductions share one pass (launch_collide3_diag); 3k
// steps keep the separate reduction pass (a fused pass there would turn the
// last three-step pass into a two-step and a fused pass, more work than the
// reduction it saves).
int bgk_step_diagnostics(bgk_ctx *c, int nsteps, double svisc, double out[7]) {
    int r = check_ctx(c);
    if (r) return r;
    if (!out) return fail(BGK_EINVAL, "null output");
    if (nsteps < 0) return fail(BGK_EINVAL, "nsteps must be >= 0");
    const Lattice &L = c->lat;
    const long long slots = std::max(c->red_blocks, reduction_parts(L));
    const bool fused = L.layout == BGK_LAYOUT_PERIODIC &&
                       !(c->flags & BGK_FLAG_REFCOMPAT_HALOS) && nsteps % 3 != 0 &&
                       (nsteps % 3 == 2 || nsteps >= 4) && three_step_ok(L) &&
                       collide3_diag_parts(L) > 0 && collide3_diag_parts(L) <= slots;
    if (!fused) {
        if ((r = bgk_step(c, nsteps))) return r;
        return bgk_diagnostics(c, svisc, BGK_DIAG_CONSISTENT, out);
    }
    hipEvent_t ev, ev2;
    BGK_HIP(timing_begin(c, TCLS_STEP, c->s0, &ev));
    if ((r = step_periodic(c, nsteps - 2))) return r;
    BGK_HIP(timing_begin(c, TCLS_PASS2D, c->s0, &ev2));
    BGK_HIP(launch_collide3_diag(L, c->cur, c->omega, svisc, c->red_partials, c->red_out,
                                 c->s0));
    BGK_HIP(timing_end(c, TCLS_PASS2D, c->s0, ev2));
    c->cur ^= 1;
    c->steps += 2;
    BGK_HIP(timing_end(c, TCLS_STEP, c->s0, ev));
    BGK_HIP(hipMemcpyAsync(c->red_host, c->red_out, 16 * sizeof(double), hipMemcpyDeviceToHost,
                           c->s0));
    BGK_HIP(hipStreamSynchronize(c->s0));
    if ((r = diagno_finish(L, BGK_DIAG_CONSISTENT, c->red_host, out))) return r;
    return dissipation_finish(c->red_host + 8, out + 5);
}

int bgk_get_box(bgk_ctx *c, int x0, int x1, int y0, int y1, int z0, int z1, double *out) {
    int r = check_ctx(c);
    if (r) return r;
    const Lattice &L = c->lat;
    if (!out) return fail(BGK_EINVAL, "null output");
    if (x0 < 0 || y0 < 0 || z0 < 0 || x1 > L.nx || y1 > L.ny || z1 > L.nz || x0 >= x1 ||
        y0 >= y1 || z0 >= z1)
        return fail(BGK_EINVAL, "box outside the lattice or empty");
    const Box b{x0, x1, y0, y1, z0, z1};
    const long long n = b.sites();
    const long long chunk = std::min(L.nsite, CHUNK);
    // the staging buffer holds 3*chunk doubles; larger boxes go z-plane by z-plane
    if (19 * n <= 3 * chunk) {
        BGK_HIP(launch_gather_box(L, c->cur, b, c->staging, c->s0));
        BGK_HIP(hipMemcpyAsync(out, c->staging, sizeof(double) * 19 * n, hipMemcpyDeviceToHost,
                               c->s0));
        BGK_HIP(hipStreamSynchronize(c->s0));
        return BGK_OK;
    }
    const long long plane = (long long)(x1 - x0) * (y1 - y0);
    if (19 * plane > 3 * chunk) return fail(BGK_EINVAL, "box plane too large for get_box");
    for (int z = z0; z < z1; z++) {
        const Box bz{x0, x1, y0, y1, z, z + 1};
        BGK_HIP(launch_gather_box(L, c->cur, bz, c->staging, c->s0));
        for (int q = 0; q < 19; q++)
            BGK_HIP(hipMemcpyAsync(out + (size_t)q * n + (size_t)(z - z0) * plane,
                                   c->staging + (size_t)q * plane, sizeof(double) * plane,
                                   hipMemcpyDeviceToHost, c->s0));
        BGK_HIP(hipStreamSynchronize(c->s0));
    }
    return BGK_OK;
}

int bgk_checksum(bgk_ctx *c, uint64_t *out) {
    int r = check_ctx(c);
    if (r) return r;
    if (!out) return fail(BGK_EINVAL, "null output");
    // red_partials holds 8 doubles per block: room for one 64-bit word per block
    BGK_HIP(launch_checksum(c->lat, c->cur, c->red_partials, c->red_blocks, c->red_out, c->s0));
    BGK_HIP(hipMemcpyAsync(c->red_host, c->red_out, sizeof(uint64_t), hipMemcpyDeviceToHost,
                           c->s0));
    BGK_HIP(hipStreamSynchronize(c->s0));
    memcpy(out, c->red_host, sizeof(uint64_t));
    return BGK_OK;
}

int bgk_synchronize(bgk_ctx *c) {
    int r = check_ctx(c);
    if (r) return r;
    BGK_HIP(hipStreamSynchronize(c->s1));
    BGK_HIP(hipStreamSynchronize(c->s0));
    return BGK_OK;
}

int bgk_timing(bgk_ctx *c, double t[8], int reset) {
    int r = check_ctx(c);
    if (r) return r;
    if (!t) return fail(BGK_EINVAL, "null output");
    BGK_HIP(harvest(c));
    Timing &T = c->timing;
    t[0] = T.ms[TCLS_COLLIDE] + T.ms[TCLS_PASS2] + T.ms[TCLS_PASS3] + T.ms[TCLS_PASS2D];
    t[1] = (double)(T.launches[TCLS_COLLIDE] + T.launches[TCLS_PASS2] +
                    T.launches[TCLS_PASS3] + T.launches[TCLS_PASS2D]);
    t[2] = T.ms[TCLS_REDUCE];
    t[3] = (double)T.launches[TCLS_REDUCE];
    t[4] = T.ms[TCLS_HALO];
    t[5] = (double)T.launches[TCLS_HALO];
    t[6] = T.ms[TCLS_STEP];
    t[7] = (double)T.launches[TCLS_STEP];
    if (reset) timing_clear(T);
    return BGK_OK;
}

int bgk_timing_passes(bgk_ctx *c, double t[6], int reset) {
    int r = check_ctx(c);
    if (r) return r;
    if (!t) return fail(BGK_EINVAL, "null output");
    BGK_HIP(harvest(c));
    Timing &T = c->timing;
    const int cls[3] = {TCLS_COLLIDE, TCLS_PASS2, TCLS_PASS3};
    for (int k = 0; k < 3; k++) {
        t[2 * k] = T.ms[cls[k]];
        t[2 * k + 1] = (double)T.launches[cls[k]];
    }
    if (reset) timing_clear(T);
    return BGK_OK;
}

int bgk_timing_fused(bgk_ctx *c, double t[2], int reset) {
    int r = check_ctx(c);
    if (r) return r;
    if (!t) return fail(BGK_EINVAL, "null output");
    BGK_HIP(harvest(c));
    Timing &T = c->timing;
    t[0] = T.ms[TCLS_PASS2D];
    t[1] = (double)T.launches[TCLS_PASS2D];
    if (reset) timing_clear(T);
    return BGK_OK;
}

int bgk_timing_adjoint(bgk_ctx *c, double t[4], int reset) {
    int r = check_ctx(c);
    if (r) return r;
    if (!t) return fail(BGK_EINVAL, "null output");
    BGK_HIP(harvest(c));
    Timing &T = c->timing;
    t[0] = T.ms[TCLS_TAPE];
    t[1] = (double)T.launches[TCLS_TAPE];
    t[2] = T.ms[TCLS_ADJOINT];
    t[3] = (double)T.launches[TCLS_ADJOINT];
    if (reset) timing_clear(T);
    return BGK_OK;
}

int bgk_adjoint_ke(bgk_ctx *c, const double *u0, int nsteps, int ckpt_every, int levels,
                   double *J, double *grad, double stats[4]) {
    int r = check_ctx(c);
    if (r) return r;
    if (!u0 || !grad || !J) return fail(BGK_EINVAL, "null argument");
    const size_t nb = sizeof(double) * 3 * c->lat.nsite;
    double *du = nullptr, *dg = nullptr;
    hipError_t e = hipMalloc(&du, nb);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(u0)");
    e = hipMalloc(&dg, nb);
    if (e != hipSuccess) {
        (void)hipFree(du);
        return hip_fail(e, "hipMalloc(grad)");
    }
    e = hipMemcpy(du, u0, nb, hipMemcpyHostToDevice);
    if (e != hipSuccess) r = hip_fail(e, "hipMemcpy(u0)");
    if (!r) r = adjoint_ke(c, du, nsteps, ckpt_every, levels, J, dg, stats);
    if (!r) {
        e = hipMemcpy(grad, dg, nb, hipMemcpyDeviceToHost);
        if (e != hipSuccess) r = hip_fail(e, "hipMemcpy(grad)");
    }
    (void)hipFree(du);
    (void)hipFree(dg);
    return r;
}

int bgk_adjoint_ke_device(bgk_ctx *c, const double *u0_dev, int nsteps, int ckpt_every,
                          int levels, double *J, double *grad_dev, double stats[4]) {
    int r = check_ctx(c);
    if (r) return r;
    if (!u0_dev || !grad_dev || !J) return fail(BGK_EINVAL, "null argument");
    return adjoint_ke(c, u0_dev, nsteps, ckpt_every, levels, J, grad_dev, stats);
}

int bgk_forward_ke(bgk_ctx *c, const double *u0, int nsteps, double *J) {
    int r = check_ctx(c);
    if (r) return r;
    if (!u0 || !J) return fail(BGK_EINVAL, "null argument");
    const size_t nb = sizeof(double) * 3 * c->lat.nsite;
    double *du = nullptr;
    hipError_t e = hipMalloc(&du, nb);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(u0)");
    e = hipMemcpy(du, u0, nb, hipMemcpyHostToDevice);
    if (e != hipSuccess) r = hip_fail(e, "hipMemcpy(u0)");
    if (!r) r = forward_ke(c, du, nsteps, J);
    (void)hipFree(du);
    return r;
}

}  // extern "C"
