// bgk_adjoint.hip -- discrete adjoint of the fused stream-collide (not in the
// reference): gradient of J = sum_x |u(x)|^2/2 at the final step with respect
// to the initial velocity field.
//
// Forward step (periodic, pull):  g_k(x) = f_k(x - c_k),
//   f'_i(x) = C_i(g(x)) = g_i - omega (g_i - e_i(g)),
//   e_i = w_i [rho + 3 c_i.j + 9/2 (c_i.j)^2/rho - 3/2 |j|^2/rho] - w_i (shifted),
//   rho = sum g + 1, j = sum c g, u = j / rho.
// Adjoint step: lambda^t_k(x - c_k) = mu_k(x) with
//   mu_k = (1 - omega) lambda_k + omega [A + (B - 3 S u) . c_k],
// stored as M^t_k(x) = mu_k(x), so that lambda^t_i(x) = M^t_i(x + c_i): each
// step reads M^{t+1} with shifted loads and writes M^t at the site (the
// forward's own access pattern: aligned stores, no push),
//   A = sum_i w_i lambda_i (1 - 9/2 (c_i.u)^2 + 3/2 |u|^2),
//   B = sum_i w_i lambda_i (3 + 9 c_i.u) c_i,   S = sum_i w_i lambda_i,
// i.e. J^T applied per site as a rank-4 moment correction plus the
// (1 - omega) diagonal.  The Jacobian depends on
// the site's pulled-moment velocity u only, so the forward pass records u
// (3 values per site and step, the "tape") instead of whole states.
// Final condition: lambda^T_i = (u.c_i - |u|^2) / rho (at-site moments).
// Initial map: f^0_i = w_i (3 c_i.u0 + 9/2 (c_i.u0)^2 - 3/2 |u0|^2), so
//   dJ/du0 = sum_i lambda^0_i w_i (3 c_i + 9 (c_i.u0) c_i - 3 u0).
//
// Checkpointing: levels = 0 tapes every step of the horizon (no
// recompute); levels = 1 keeps the state every ckpt_every steps and
// re-runs each segment with the tape; levels = 2 adds coarse checkpoints
// every ckpt_every*M steps (M ~ sqrt(number of segments)) and recomputes the
// fine checkpoints of one coarse segment at a time.
#include <math.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "bgk_collide.h"

namespace bgk {

// lambda^T from at-site moments of the final state, stored in the shifted
// form M^T_i(x + c_i) = lambda^T_i(x)
template <typename S, typename C>
__global__ __launch_bounds__(256) void k_lambda_final(Pops<S> P, SiteMap m, C *M, int nx, int ny,
                                                      int nz) {
    const long long nsite = (long long)nx * ny * nz;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= nsite) return;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long long)nx * ny));
    const long long o = site_elem(m, i);
    C f[19];
#pragma unroll
    for (int q = 0; q < 19; q++) f[q] = Storage<S>::template load<C>(P.p[q][o]);
    C rho, ux, uy, uz;
    moments_site<C>(f, (C)1, rho, ux, uy, uz);
    const C usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        int tx = x + cx(q), ty = y + cy(q), tz = z + cz(q);
        tx = tx < 0 ? tx + nx : (tx >= nx ? tx - nx : tx);
        ty = ty < 0 ? ty + ny : (ty >= ny ? ty - ny : ty);
        tz = tz < 0 ? tz + nz : (tz >= nz ? tz - nz : tz);
        M[(size_t)q * nx + tx + (size_t)m.ystride * ty + (size_t)m.zstride * tz] =
            (ux * (C)cx(q) + uy * (C)cy(q) + uz * (C)cz(q) - usq) / rho;
    }
}

// Index of the population with velocity s e_k (axis), and of the one with
// components sk on axis k and sl on axis l (diagonal), in the velocity table.
__host__ __device__ constexpr int comp_c(int q, int k) {
    return k == 0 ? cx(q) : (k == 1 ? cy(q) : cz(q));
}
__host__ __device__ constexpr int axis_q(int k, int s) {
    int r = -1;
    for (int q = 0; q < 19; q++)
        if (comp_c(q, k) == s && comp_c(q, (k + 1) % 3) == 0 && comp_c(q, (k + 2) % 3) == 0)
            r = q;
    return r;
}
__host__ __device__ constexpr int diag_q(int k, int l, int sk, int sl) {
    int r = -1;
    for (int q = 0; q < 19; q++)
        if (comp_c(q, k) == sk && comp_c(q, l) == sl && comp_c(q, 3 - k - l) == 0) r = q;
    return r;
}
__host__ __device__ constexpr int rest_q() {
    int r = -1;
    for (int q = 0; q < 19; q++)
        if (cx(q) == 0 && cy(q) == 0 && cz(q) == 0) r = q;
    return r;
}

// mu_k = (1 - omega) lambda_k + omega [A + G . c_k] at one site, through the
// weighted moments of lambda: m0 = sum w lambda, m1 = sum w lambda c,
// P = sum w lambda c c^T, so that
//   A = sum_i w_i lambda_i (1 - 9/2 (c_i.u)^2 + 3/2 |u|^2)
//     = m0 (1 + 3/2 |u|^2) - 9/2 u.P u,
//   G = sum_i w_i lambda_i (3 + 9 c_i.u) c_i - 3 S u = 3 (m1 - m0 u) + 9 P u.
// The moments come from unweighted sums over the axis populations and the
// four diagonals of each coordinate plane, each weight applied once (the
// adjoint has no counterpart in the reference, so its arithmetic is free to
// take this form, with fused multiply-adds): ~130 VALU operations per site
// against ~330 for the per-population sums.
// fused multiply-add in the compute kind
__device__ __forceinline__ float fma_k(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_k(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename C>
__device__ __forceinline__ void adjoint_site(const C (&lam)[19], C ux, C uy, C uz, C omega,
                                             C (&mu)[19]) {
    const C w0 = (C)(1.0 / 3.0), w1 = (C)(1.0 / 18.0), w2 = (C)(1.0 / 36.0);
    const C u[3] = {ux, uy, uz};
    // axis populations: c_k lambda and c_k^2 lambda summed along each axis
    C a1[3], a2[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const C p = lam[axis_q(k, 1)], m = lam[axis_q(k, -1)];
        a1[k] = p - m;
        a2[k] = p + m;
    }
    // diagonals of the planes (x,y), (x,z), (y,z): total, c_k c_l sum, and
    // the first moments along k and l
    C T[3], Xc[3], Dk[3], Dl[3];
#pragma unroll
    for (int pl = 0; pl < 3; pl++) {
        const int k = pl == 2 ? 1 : 0, l = pl == 0 ? 1 : 2;
        const C pp = lam[diag_q(k, l, 1, 1)], pm = lam[diag_q(k, l, 1, -1)];
        const C mp = lam[diag_q(k, l, -1, 1)], mm = lam[diag_q(k, l, -1, -1)];
        const C s1 = pp + mm, s2 = pm + mp, t1 = pp - mm, t2 = pm - mp;
        T[pl] = s1 + s2;
        Xc[pl] = s1 - s2;
        Dk[pl] = t1 + t2;
        Dl[pl] = t1 - t2;
    }
    const C d1[3] = {Dk[0] + Dk[1], Dl[0] + Dk[2], Dl[1] + Dl[2]};
    const C d2[3] = {T[0] + T[1], T[0] + T[2], T[1] + T[2]};
    C m1[3], Pd[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        m1[k] = fma_k(a1[k], w1, d1[k] * w2);
        Pd[k] = fma_k(a2[k], w1, d2[k] * w2);
    }
    const C Pxy = Xc[0] * w2, Pxz = Xc[1] * w2, Pyz = Xc[2] * w2;
    const C m0 = fma_k(lam[rest_q()], w0,
                               fma_k(a2[0] + a2[1] + a2[2], w1, (T[0] + T[1] + T[2]) * w2));
    // P u, u.P u, |u|^2
    const C Pu[3] = {fma_k(Pd[0], u[0], fma_k(Pxy, u[1], Pxz * u[2])),
                     fma_k(Pxy, u[0], fma_k(Pd[1], u[1], Pyz * u[2])),
                     fma_k(Pxz, u[0], fma_k(Pyz, u[1], Pd[2] * u[2]))};
    const C uPu = fma_k(u[0], Pu[0], fma_k(u[1], Pu[1], u[2] * Pu[2]));
    const C usq = fma_k(u[0], u[0], fma_k(u[1], u[1], u[2] * u[2]));
    const C A = fma_k((C)-4.5, uPu, m0 * fma_k((C)1.5, usq, (C)1));
    C oG[3];
#pragma unroll
    for (int k = 0; k < 3; k++)
        oG[k] = omega * fma_k((C)9, Pu[k], (C)3 * fma_k(-m0, u[k], m1[k]));
    const C oA = omega * A, om1 = (C)1 - omega;
    // omega (A + G.c_q) for each population: the axis values oA +- oG_k, the
    // diagonals from them
    C ax[3][2];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        ax[k][0] = oA + oG[k];
        ax[k][1] = oA - oG[k];
    }
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int n = (cx(q) != 0) + (cy(q) != 0) + (cz(q) != 0);
        C t;
        if (n == 0) {
            t = oA;
        } else {
            const int k = cx(q) != 0 ? 0 : (cy(q) != 0 ? 1 : 2);
            t = ax[k][comp_c(q, k) > 0 ? 0 : 1];
            if (n == 2) {
                const int l = (k == 0 && cy(q) != 0) ? 1 : 2;
                t = comp_c(q, l) > 0 ? t + oG[l] : t - oG[l];
            }
        }
        mu[q] = fma_k(om1, lam[q], t);
    }
}

// one adjoint step: lambda^{t+1}_i(x) = M^{t+1}_i(x + c_i) (shifted loads),
// tape u^t(x); writes M^t_k(x) = mu_k(x)
template <typename C>
__global__ __launch_bounds__(256) void k_adjoint_step(const C *__restrict__ Min,
                                                      C *__restrict__ Mout,
                                                      const C *__restrict__ tape, int nx, int ny,
                                                      int nz, long long ys, long long zs,
                                                      int tiles_x, int tiles_y, C omega) {
    const int t = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int tx = t % tiles_x;
    const int r = t / tiles_x;
    const int ty = r % tiles_y;
    const int z = r / tiles_y;
    const int x = tx * 64 + (int)threadIdx.x;
    const int y = ty * 4 + (int)threadIdx.y;
    if (x >= nx || y >= ny) return;
    const long long N = (long long)nx * ny * nz;
    const int nxy = nx * ny;
    const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
    const int ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const int zm = (z == 0) ? nz - 1 : z - 1, zp = (z == nz - 1) ? 0 : z + 1;
    const long long X[3] = {xm, x, xp};
    const long long Y[3] = {ym * ys, y * ys, yp * ys};
    const long long Z[3] = {zm * zs, z * zs, zp * zs};
    C lam[19];
#pragma unroll
    for (int q = 0; q < 19; q++)
        lam[q] = Min[(size_t)q * nx + X[1 + cx(q)] + Y[1 + cy(q)] + Z[1 + cz(q)]];
    const size_t c = (size_t)(x + Y[1] + Z[1]);                   // lambda element
    const size_t i = (size_t)x + (size_t)nx * y + (size_t)nxy * z;  // tape site
    C mu[19];
    adjoint_site<C>(lam, tape[i], tape[N + i], tape[2 * N + i], omega, mu);
#pragma unroll
    for (int q = 0; q < 19; q++) Mout[(size_t)q * nx + c] = mu[q];
}

// Two adjoint steps in one pass (the two-step pass of bgk_collide.h with the
// pull mirrored: lambda is read at x + c_q): M^{t+1} -> M^t (u^t from tape0)
// kept in LDS -> M^{t-1} (u^{t-1} from tape1).  Bit-identical to two
// k_adjoint_step launches.
template <typename C>
struct AdjointOp {
    static constexpr int SGN = -1;
    static constexpr bool PINGPONG = true;
    static constexpr bool KEEP_STATE = true;
    struct Extra {
        C u[3];
    };
    C omega;
    const C *tape0, *tape1;
    size_t ld;
    __device__ __forceinline__ Extra extra_first(size_t i) const {
        return Extra{{tape0[i], tape0[ld + i], tape0[2 * ld + i]}};
    }
    __device__ __forceinline__ Extra extra_second(size_t i) const {
        return Extra{{tape1[i], tape1[ld + i], tape1[2 * ld + i]}};
    }
    __device__ __forceinline__ void first(const C (&f)[19], C (&b)[19], const Extra &e, size_t,
                                          bool) const {
        adjoint_site<C>(f, e.u[0], e.u[1], e.u[2], omega, b);
    }
    __device__ __forceinline__ void second(const C (&f)[19], C (&b)[19], const Extra &e,
                                           size_t) const {
        adjoint_site<C>(f, e.u[0], e.u[1], e.u[2], omega, b);
    }
};

template <typename C, bool O32, int TX, int TY, int MINB>
__global__ __launch_bounds__(collide2_threads(TX, TY), MINB) void k_adjoint2(Pops<C> src,
                                                                             Pops<C> dst, Geo2 g,
                                                                             AdjointOp<C> op) {
    __shared__ C lds[20][TY + 2][TX + 2];
    two_step_pass<C, C, O32, TX, TY, AdjointOp<C>, false, O32, 1>(src, dst, g, op, lds);
}

// Three adjoint steps in one pass (three_step_pass with the pull mirrored):
// M^{t+1} -> M^t (u^t from tape0) -> M^{t-1} (tape1), both kept in LDS ->
// M^{t-2} (tape2).  Bit-identical to three k_adjoint_step launches.
template <typename C>
struct AdjointOp3 {
    static constexpr int SGN = -1;
    static constexpr bool PINGPONG = true;
    static constexpr bool KEEP_STATE = true;
    struct Extra {
        C u[3];
    };
    C omega;
    const C *tape0, *tape1, *tape2;
    size_t ld;
    __device__ __forceinline__ Extra at(const C *t, size_t i) const {
        return Extra{{t[i], t[ld + i], t[2 * ld + i]}};
    }
    __device__ __forceinline__ Extra extra_first(size_t i) const { return at(tape0, i); }
    __device__ __forceinline__ Extra extra_second(size_t i) const { return at(tape1, i); }
    __device__ __forceinline__ Extra extra_third(size_t i) const { return at(tape2, i); }
    __device__ __forceinline__ void first(const C (&f)[19], C (&b)[19], const Extra &e, size_t,
                                          bool) const {
        adjoint_site<C>(f, e.u[0], e.u[1], e.u[2], omega, b);
    }
    __device__ __forceinline__ void second(const C (&f)[19], C (&b)[19], const Extra &e, size_t,
                                           bool) const {
        adjoint_site<C>(f, e.u[0], e.u[1], e.u[2], omega, b);
    }
    __device__ __forceinline__ void third(const C (&f)[19], C (&b)[19], const Extra &e,
                                          size_t) const {
        adjoint_site<C>(f, e.u[0], e.u[1], e.u[2], omega, b);
    }
};

template <typename C, int TX, int TY>
__global__ __launch_bounds__(collide3_threads(TX, TY), 1) void k_adjoint3(Pops<C> src,
                                                                          Pops<C> dst, Geo2 g,
                                                                          AdjointOp3<C> op) {
    three_step_pass<C, C, TX, TY, AdjointOp3<C>>(src, dst, g, op);
}

// three taped forward steps in one pass (u of each step to its tape slice);
// KEEP = false: the final state is not stored (see CollideOp)
template <typename S, typename C, int TX, int TY, bool KEEP = true>
__global__ __launch_bounds__(collide3_threads(TX, TY), 1) void k_tape3(
    Pops<S> src, Pops<S> dst, Geo2 g, CollideOp3<C, true, true, KEEP> op) {
    three_step_pass<S, C, TX, TY, CollideOp3<C, true, true, KEEP>>(src, dst, g, op);
}

// dJ/du0 = sum_i lambda^0_i w_i (3 c_i + 9 (c_i.u0) c_i - 3 u0),
// lambda^0_i(x) = M^0_i(x + c_i)
template <typename C>
__global__ __launch_bounds__(256) void k_grad_u0(const C *M, const double *u0, double *grad,
                                                 int nx, int ny, int nz, long long ys,
                                                 long long zs) {
    const long long N = (long long)nx * ny * nz;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long long)nx * ny));
    const double ux = u0[i], uy = u0[N + i], uz = u0[2 * N + i];
    double gx = 0, gy = 0, gz = 0;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        int sx = x + cx(q), sy = y + cy(q), sz = z + cz(q);
        sx = sx < 0 ? sx + nx : (sx >= nx ? sx - nx : sx);
        sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
        sz = sz < 0 ? sz + nz : (sz >= nz ? sz - nz : sz);
        const double L =
            (double)M[(size_t)q * nx + sx + (size_t)ys * sy + (size_t)zs * sz] * wq(q);
        const double cu = cx(q) * ux + cy(q) * uy + cz(q) * uz;
        gx += L * (3.0 * cx(q) + 9.0 * cu * cx(q) - 3.0 * ux);
        gy += L * (3.0 * cy(q) + 9.0 * cu * cy(q) - 3.0 * uy);
        gz += L * (3.0 * cz(q) + 9.0 * cu * cz(q) - 3.0 * uz);
    }
    grad[i] = gx;
    grad[N + i] = gy;
    grad[2 * N + i] = gz;
}

// ------------------------------------------------------------ host driver
namespace {

struct Adj {
    bgk_ctx *c;
    size_t state_bytes; // one copy of the populations (all 19, interleaved)
    size_t csize;       // sizeof compute kind
    size_t es;          // sizeof storage kind
    void *own[2];       // the lattice's own two copies
    std::vector<void *> allocs;
    size_t bytes = 0;
    long long recomputed = 0;

    Adj(bgk_ctx *ctx, size_t storage_bytes) : c(ctx), es(storage_bytes) {
        own[0] = c->lat.pop[0][0];
        own[1] = c->lat.pop[1][0];
        state_bytes = ((size_t)c->lat.copy_elems * es + 255) & ~(size_t)255;
    }
    ~Adj() {
        // the lattice returns to its own copies (whose contents are whatever
        // the adjoint left there) before the checkpoint buffers are freed
        // (hipFree waits for the device)
        set_copy(0, own[0]);
        set_copy(1, own[1]);
        for (void *p : allocs) (void)hipFree(p);
    }
    // point copy k of the lattice at the block `base` (the lattice's
    // interleaved layout): checkpoints change places with the lattice's
    // copies instead of being copied
    void set_copy(int k, const void *base) {
        Lattice &L = c->lat;
        for (int q = 0; q < 19; q++) L.pop[k][q] = (char *)base + (size_t)q * L.nx * es;
    }
    int alloc(void **p, size_t n, const char *what) {
        hipError_t e = hipMalloc(p, n);
        if (e != hipSuccess) return hip_fail(e, what);
        allocs.push_back(*p);
        bytes += n;
        return BGK_OK;
    }
    // copy the current state (copy c->cur, whose block starts at population
    // 0's pointer) to / from a slot, on stream s
    int save(void *slot, hipStream_t s) {
        const Lattice &L = c->lat;
        BGK_HIP(hipMemcpyAsync(slot, L.pop[c->cur][0], state_bytes, hipMemcpyDeviceToDevice, s));
        return BGK_OK;
    }
    int restore(const void *slot, hipStream_t s) {
        const Lattice &L = c->lat;
        BGK_HIP(hipMemcpyAsync(L.pop[c->cur][0], slot, state_bytes, hipMemcpyDeviceToDevice, s));
        return BGK_OK;
    }
};

template <typename S, typename C>
hipError_t launch_tape_step(bgk_ctx *c, C *tape, hipStream_t s) {
    const Lattice &L = c->lat;
    int ntiles;
    Geo g = make_geo<S>(L, Box{0, L.nx, 0, L.ny, 0, L.nz}, &ntiles);
    if (offsets_fit_32(L, sizeof(S)))
        hipLaunchKernelGGL((k_collide<S, C, true, true>), dim3(ntiles),
                           dim3(Tile<S>::X, Tile<S>::Y), 0, s, pops_of<S>(L, c->cur),
                           pops_of<S>(L, 1 - c->cur), g, (C)c->omega, (C)1, tape, L.nsite);
    else
        hipLaunchKernelGGL((k_collide<S, C, false, true>), dim3(ntiles),
                           dim3(Tile<S>::X, Tile<S>::Y), 0, s, pops_of<S>(L, c->cur),
                           pops_of<S>(L, 1 - c->cur), g, (C)c->omega, (C)1, tape, L.nsite);
    c->cur ^= 1;
    return hipGetLastError();
}

// The lattice as a pass of the backward sizes its z chunks (two_step_zc,
// three_step_zc minimise rounds of one workgroup per CU times the plane
// iterations of a chunk): the re-tape and the adjoint sweep run on two
// streams at once, so each of their passes has about half of the CUs.  At
// 128^3 that takes 32-plane chunks (104 workgroups, 36 plane iterations)
// instead of 16-plane ones (208 workgroups, 20 iterations each, two rounds
// when shared).  ncu = 0: the device's count.
inline Lattice with_cus(const Lattice &L, int ncu) {
    Lattice s = L;
    if (ncu > 0) s.ncu = ncu;
    return s;
}

// two taped forward steps (slices tape0, tape1) in one pass
template <typename S, typename C, bool O32, bool KEEP>
hipError_t launch_tape2_k(bgk_ctx *c, C *tape0, C *tape1, hipStream_t s, int ncu) {
    const Lattice &L = c->lat;
    using T = Tile2<S, C, O32>;
    int nblocks;
    const Geo2 g =
        make_geo2(L, T::X, T::Y, two_step_zc(with_cus(L, ncu), T::X, T::Y, T::ZC), &nblocks);
    const CollideOp<C, true, true, KEEP> op{(C)c->omega, (C)1, tape0, tape1, (size_t)L.nsite};
    hipLaunchKernelGGL((k_collide2<S, C, O32, T::X, T::Y, T::MINB, true, false, true, KEEP>),
                       dim3(nblocks), dim3(collide2_threads(T::X, T::Y)), 0, s,
                       pops_of<S>(L, c->cur), pops_of<S>(L, 1 - c->cur), g, op);
    c->cur ^= 1;
    return hipGetLastError();
}
// keep = false: the final state is not stored (the last pass of a re-tape)
template <typename S, typename C, bool O32>
hipError_t launch_tape2_o(bgk_ctx *c, C *tape0, C *tape1, hipStream_t s, int ncu, bool keep) {
    return keep ? launch_tape2_k<S, C, O32, true>(c, tape0, tape1, s, ncu)
                : launch_tape2_k<S, C, O32, false>(c, tape0, tape1, s, ncu);
}

// Own tiles of the adjoint's three-step passes (taping and adjoint steps):
// 64x8 in the 4-byte compute kind, where the forward's passes take 64x10.
// At 128^3 the backward's two streams each size their z chunks for half of
// the CUs: 64x8 gives 32 tile columns and 4 chunks, 128 workgroups per
// stream, where 64x10 gave 26 columns and 104 workgroups (48 CUs idle),
// with the same 36 plane iterations per workgroup and 13 waves each instead
// of 15: backward 34.3-34.6 -> 29.0-30.8 ms per 500 steps, 256^3 unchanged.
// FP64 compute keeps 32x10 (32x8 measured slower in the forward).
template <typename C>
struct Tile3A {
    static constexpr bool f64 = sizeof(C) == 8;
    static constexpr int X = f64 ? 32 : 64, Y = f64 ? 10 : 8, ZC = 64;
};

// three taped forward steps (slices tape0..tape2) in one pass
template <typename S, typename C, bool KEEP>
hipError_t launch_tape3_k(bgk_ctx *c, C *tape0, C *tape1, C *tape2, hipStream_t s, int ncu) {
    const Lattice &L = c->lat;
    using T = Tile3A<C>;
    int nblocks;
    Geo2 g = make_geo2(L, T::X, T::Y, 1, &nblocks);
    const Split3 sp = three_step_split(with_cus(L, ncu), T::X, T::Y, T::ZC);
    g.full_cols = sp.full_cols;
    g.zc = sp.zc;
    nblocks = sp.nblocks;
    const CollideOp3<C, true, true, KEEP> op{(C)c->omega, (C)1, tape0, tape1, tape2,
                                             (size_t)L.nsite};
    hipLaunchKernelGGL((k_tape3<S, C, T::X, T::Y, KEEP>), dim3(nblocks),
                       dim3(collide3_threads(T::X, T::Y)), 0, s, pops_of<S>(L, c->cur),
                       pops_of<S>(L, 1 - c->cur), g, op);
    c->cur ^= 1;
    return hipGetLastError();
}
template <typename S, typename C>
hipError_t launch_tape3(bgk_ctx *c, C *tape0, C *tape1, C *tape2, hipStream_t s, int ncu,
                        bool keep) {
    return keep ? launch_tape3_k<S, C, true>(c, tape0, tape1, tape2, s, ncu)
                : launch_tape3_k<S, C, false>(c, tape0, tape1, tape2, s, ncu);
}

// Three-step passes of the adjoint's forward and backward: where
// three_step_ok holds, lambda (in the compute kind) takes 32-bit offsets and
// a pass has at least half a round of workgroups (one per CU), as the plain
// forward's passes.  At 128^3 (208 workgroups) the backward with them takes
// 38.3 ms per 500 steps against 39.3 with two-step passes (with them in the
// sweep or in the re-tape alone, 39.3); at 256^3 the backward is 10 % faster.
// Of n steps, n3 such passes (a remainder of one goes to two two-step
// passes, as in bgk_step).
template <typename C>
int adjoint_triples(const Lattice &L, int n, int ncu) {
    if (!three_step_ok(L) || !offsets_fit_32(L, sizeof(C))) return 0;
    using T = Tile3A<C>;
    const Lattice Lz = with_cus(L, ncu);
    const int zc = three_step_zc(Lz, T::X, T::Y, T::ZC);
    const long long nb = (long long)((L.nx + T::X - 1) / T::X) * ((L.ny + T::Y - 1) / T::Y) *
                         ((L.nz + zc - 1) / zc);
    if (L.two_step != 2 && 2 * nb < Lz.ncu) return 0;
    int n3 = n / 3;
    if (n % 3 == 1 && n3 > 0 && two_step_ok(L)) n3--;
    return n3;
}

// taped forward steps on stream s: slice k of `tape` records step k of the
// n steps (z chunks for ncu CUs, with_cus).  keep_last = false: the last
// fused pass does not store its state (a re-tape's end, whose state nothing
// reads; single steps always store).
template <typename S, typename C>
int run_taped(bgk_ctx *c, C *tape, int n, hipStream_t s, int ncu, bool keep_last = true) {
    const Lattice &L = c->lat;
    const size_t slice = (size_t)3 * L.nsite;
    hipEvent_t ev;
    const int n3 = adjoint_triples<C>(L, n, ncu);
    const bool two = two_step_ok(L);
    int k = 0;
    for (int i = 0; i < n3; i++, k += 3) {
        const bool keep = keep_last || k + 3 < n;
        BGK_HIP(timing_begin(c, TCLS_TAPE, s, &ev));
        BGK_HIP((launch_tape3<S, C>(c, tape + k * slice, tape + (k + 1) * slice,
                                    tape + (k + 2) * slice, s, ncu, keep)));
        BGK_HIP(timing_end(c, TCLS_TAPE, s, ev));
    }
    if (two) {
        const bool o32 = offsets_fit_32(L, sizeof(S));
        for (; k + 1 < n; k += 2) {
            const bool keep = keep_last || k + 2 < n;
            BGK_HIP(timing_begin(c, TCLS_TAPE, s, &ev));
            BGK_HIP(o32 ? (launch_tape2_o<S, C, true>(c, tape + k * slice,
                                                      tape + (k + 1) * slice, s, ncu, keep))
                        : (launch_tape2_o<S, C, false>(c, tape + k * slice,
                                                       tape + (k + 1) * slice, s, ncu, keep)));
            BGK_HIP(timing_end(c, TCLS_TAPE, s, ev));
        }
    }
    for (; k < n; k++) {
        BGK_HIP(timing_begin(c, TCLS_TAPE, s, &ev));
        BGK_HIP((launch_tape_step<S, C>(c, tape + k * slice, s)));
        BGK_HIP(timing_end(c, TCLS_TAPE, s, ev));
    }
    return BGK_OK;
}

template <typename C, bool O32>
hipError_t launch_adjoint2_o(bgk_ctx *c, const C *lin, C *lout, const C *tape0,
                             const C *tape1, int ncu) {
    const Lattice &L = c->lat;
    using T = Tile2<C, C, O32>;
    int nblocks;
    Geo2 g = make_geo2(L, T::X, T::Y, two_step_zc(with_cus(L, ncu), T::X, T::Y, T::ZC), &nblocks);
    // lambda: the lattice's interleaved layout in the compute kind
    g.pstride = (unsigned)(L.nx * sizeof(C));
    Pops<C> a, b;
    for (int q = 0; q < 19; q++) {
        a.p[q] = (C *)lin + (size_t)q * L.nx;
        b.p[q] = lout + (size_t)q * L.nx;
    }
    const AdjointOp<C> op{(C)c->omega, tape0, tape1, (size_t)L.nsite};
    hipLaunchKernelGGL((k_adjoint2<C, O32, T::X, T::Y, T::MINB>), dim3(nblocks),
                       dim3(collide2_threads(T::X, T::Y)), 0, c->s0, a, b, g, op);
    return hipGetLastError();
}

template <typename C>
hipError_t launch_adjoint3(bgk_ctx *c, const C *lin, C *lout, const C *tape0, const C *tape1,
                           const C *tape2, int ncu) {
    const Lattice &L = c->lat;
    using T = Tile3A<C>;
    int nblocks;
    Geo2 g = make_geo2(L, T::X, T::Y, 1, &nblocks);
    const Split3 sp = three_step_split(with_cus(L, ncu), T::X, T::Y, T::ZC);
    g.full_cols = sp.full_cols;
    g.zc = sp.zc;
    nblocks = sp.nblocks;
    // lambda: the lattice's interleaved layout in the compute kind
    g.pstride = (unsigned)(L.nx * sizeof(C));
    Pops<C> a, b;
    for (int q = 0; q < 19; q++) {
        a.p[q] = (C *)lin + (size_t)q * L.nx;
        b.p[q] = lout + (size_t)q * L.nx;
    }
    const AdjointOp3<C> op{(C)c->omega, tape0, tape1, tape2, (size_t)L.nsite};
    hipLaunchKernelGGL((k_adjoint3<C, T::X, T::Y>), dim3(nblocks),
                       dim3(collide3_threads(T::X, T::Y)), 0, c->s0, a, b, g, op);
    return hipGetLastError();
}

template <typename C>
hipError_t launch_adjoint_step(bgk_ctx *c, const C *lin, C *lout, const C *tape);

// adjoint steps n-1 .. 0 of a segment whose tape slice k is step k; lin
// holds M after step n and ends holding M after step 0 (z chunks for ncu
// CUs, with_cus)
template <typename C>
int adjoint_sweep(bgk_ctx *c, C *&lin, C *&lout, const C *tape, int n, int ncu) {
    const Lattice &L = c->lat;
    const size_t slice = (size_t)3 * L.nsite;
    hipEvent_t ev;
    int k = n - 1;
    for (int n3 = adjoint_triples<C>(L, n, ncu); n3 > 0; n3--, k -= 3) {
        BGK_HIP(timing_begin(c, TCLS_ADJOINT, c->s0, &ev));
        BGK_HIP(launch_adjoint3<C>(c, lin, lout, tape + k * slice, tape + (k - 1) * slice,
                                   tape + (k - 2) * slice, ncu));
        BGK_HIP(timing_end(c, TCLS_ADJOINT, c->s0, ev));
        std::swap(lin, lout);
    }
    if (two_step_ok(L)) {
        const bool o32 = offsets_fit_32(L, sizeof(C));
        for (; k >= 1; k -= 2) {
            BGK_HIP(timing_begin(c, TCLS_ADJOINT, c->s0, &ev));
            BGK_HIP(o32 ? (launch_adjoint2_o<C, true>(c, lin, lout, tape + k * slice,
                                                      tape + (k - 1) * slice, ncu))
                        : (launch_adjoint2_o<C, false>(c, lin, lout, tape + k * slice,
                                                       tape + (k - 1) * slice, ncu)));
            BGK_HIP(timing_end(c, TCLS_ADJOINT, c->s0, ev));
            std::swap(lin, lout);
        }
    }
    for (; k >= 0; k--) {
        BGK_HIP(timing_begin(c, TCLS_ADJOINT, c->s0, &ev));
        BGK_HIP(launch_adjoint_step<C>(c, lin, lout, tape + k * slice));
        BGK_HIP(timing_end(c, TCLS_ADJOINT, c->s0, ev));
        std::swap(lin, lout);
    }
    return BGK_OK;
}

template <typename C>
hipError_t launch_adjoint_step(bgk_ctx *c, const C *lin, C *lout, const C *tape) {
    const Lattice &L = c->lat;
    const int tx = (L.nx + 63) / 64, ty = (L.ny + 3) / 4;
    hipLaunchKernelGGL(k_adjoint_step<C>, dim3(tx * ty * L.nz), dim3(64, 4), 0, c->s0, lin, lout,
                       tape, L.nx, L.ny, L.nz, L.ystride, L.zstride, tx, ty, (C)c->omega);
    return hipGetLastError();
}

// time steps of the first pass step_periodic (first_pass_steps) and
// run_taped (taped_first_pass) take for n steps; they take the same passes
// for the rest of them afterwards
int first_pass_steps(const Lattice &L, int n) {
    if (n <= 0) return 0;
    if (three_step_ok(L)) {
        int n3 = n / 3;
        if (n % 3 == 1 && n3 > 0 && two_step_ok(L)) n3--;
        if (n3 > 0) return 3;
    }
    return (two_step_ok(L) && n >= 2) ? 2 : 1;
}

template <typename C>
int taped_first_pass(const Lattice &L, int n, int ncu) {
    if (n <= 0) return 0;
    if (adjoint_triples<C>(L, n, ncu) > 0) return 3;
    return (two_step_ok(L) && n >= 2) ? 2 : 1;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
}

// two events per role, used alternately by consecutive segments
struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~EventPair() {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
    int create() {
        for (hipEvent_t &x : e) BGK_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        return BGK_OK;
    }
};

template <typename S, typename C>
int adjoint_t(bgk_ctx *c, const double *u0, int T, int K, int levels, double *J, double *grad,
              double stats[4]) {
    const Lattice &L = c->lat;
    const long long N = L.nsite;
    Adj a(c, sizeof(S));
    a.csize = sizeof(C);
    if (K < 1) K = 1;
    if (levels == 0) K = std::max(T, 1);
    const int nseg = (T + K - 1) / K;
    const int M = (levels == 2) ? std::max(1, (int)ceil(sqrt((double)nseg))) : 1;
    const int K2 = K * M;                       // coarse interval
    const int ncoarse = (T + K2 - 1) / K2;
    const int tape_steps = std::min(K, std::max(T, 1));
    std::vector<void *> coarse(ncoarse), fine(M > 1 ? M : 0);
    int r;
    for (int i = 0; i < ncoarse; i++)
        if ((r = a.alloc(&coarse[i], a.state_bytes, "hipMalloc(checkpoint)"))) return r;
    for (int i = 0; i < (int)fine.size(); i++)
        if ((r = a.alloc(&fine[i], a.state_bytes, "hipMalloc(fine checkpoint)"))) return r;
    // two tape buffers (one segment each) when segments are re-taped, so that
    // the re-tape of the next segment overlaps the sweep of this one
    C *tape[2] = {nullptr, nullptr}, *lamA = nullptr, *lamB = nullptr;
    const size_t tape_bytes = sizeof(C) * 3 * (size_t)N * tape_steps;
    for (int b = 0; b < (nseg > 1 ? 2 : 1); b++)
        if ((r = a.alloc((void **)&tape[b], tape_bytes, "hipMalloc(tape)"))) return r;
    if ((r = a.alloc((void **)&lamA, sizeof(C) * 19 * (size_t)N, "hipMalloc(lambda)"))) return r;
    if ((r = a.alloc((void **)&lamB, sizeof(C) * 19 * (size_t)N, "hipMalloc(lambda)"))) return r;
    EventPair ev_tape, ev_sweep;
    if ((r = ev_tape.create()) || (r = ev_sweep.create())) return r;
    hipStream_t s1 = c->s1;

    // ---- forward (stream s0): f^0 from u0, coarse checkpoints every K2
    // steps; the tapes of the last two segments (the backward's first two)
    // are recorded on the way, the last into tape[0] and the one before into
    // tape[1], so that the backward re-tapes from its third segment on
    const int ntaped = nseg > 1 ? 2 : (nseg > 0 ? 1 : 0);
    const int taped_s0 = nseg > 0 ? (nseg - ntaped) * K : 0;  // first taped step
    const int last_s0 = nseg > 0 ? (nseg - 1) * K : 0;        // the last segment
    auto t0 = std::chrono::steady_clock::now();
    BGK_HIP(launch_init_equilibrium(L, c->cur, u0, N, 0, N, c->s0));
    // Checkpoint i is kept in place: the buffer holding S_{i K2} becomes
    // checkpoint i once the next pass has read it, and the lattice carries on
    // in a spare buffer (the checkpoint allocations), so no state is copied.
    std::vector<void *> spare(coarse.rbegin(), coarse.rend()), ck(ncoarse, nullptr);
    for (int t = 0; t < T;) {
        // up to the next checkpoint; the last segment taped (fused passes
        // where they pay)
        const int stop = t < taped_s0 ? taped_s0 : (t < last_s0 ? last_s0 : T);
        const int n = std::min(K2 - t % K2, stop - t);
        const bool taped = t >= taped_s0;
        C *tp = !taped          ? nullptr
                : t >= last_s0 ? tape[0] + (size_t)(t - last_s0) * 3 * N
                               : tape[1] + (size_t)(t - taped_s0) * 3 * N;
        auto run = [&](int m) {
            return taped ? run_taped<S, C>(c, tp, m, c->s0, L.ncu)
                         : step_periodic(c, m, false, c->s0);
        };
        int done = 0;
        if (t % K2 == 0) {
            const int k = c->cur;
            ck[t / K2] = c->lat.pop[k][0];
            done = taped ? taped_first_pass<C>(L, n, L.ncu) : first_pass_steps(L, n);
            if ((r = run(done))) return r;
            a.set_copy(k, spare.back());
            spare.pop_back();
            if (tp) tp += (size_t)done * 3 * N;
        }
        if (n > done && (r = run(n - done))) return r;
        t += n;
    }
    if ((r = reduce_ke(c, c->cur, J))) return r;  // synchronises
    hipLaunchKernelGGL((k_lambda_final<S, C>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                       c->s0, pops_of<S>(L, c->cur), site_map(L), lamA, L.nx, L.ny, L.nz);
    BGK_HIP(hipGetLastError());
    BGK_HIP(hipStreamSynchronize(c->s0));
    const double fwd_ms = ms_since(t0);

    // ---- backward.  The segments in backward order; segments 0 (the last)
    // and 1 were taped by the forward into tape[0] and tape[1], segment k
    // uses tape[k % 2].  The sweeps run on s0; everything on the lattice
    // state (fine checkpoints of a coarse segment, restores, re-taping) runs
    // on s1, so the re-tape of segment k+1 overlaps the sweep of segment k
    // (k >= 1; the sweep of segment 0 runs alone).  Cross-stream order:
    //   sweep k (k >= 2) waits for the re-tape of segment k (ev_tape);
    //   the re-tape of segment k+1 into tape[(k+1) % 2] waits for the sweep
    //   of segment k-1, which read that buffer (ev_sweep).
    t0 = std::chrono::steady_clock::now();
    struct Seg {
        int cs, j, s0, s1;
    };
    std::vector<Seg> segs;
    for (int cs = ncoarse - 1; cs >= 0; cs--) {
        const int c0 = cs * K2, c1 = std::min(T, c0 + K2);
        const int nf = (c1 - c0 + K - 1) / K;
        for (int j = nf - 1; j >= 0; j--)
            segs.push_back(Seg{cs, j, c0 + j * K, std::min(c1, c0 + j * K + K)});
    }
    // CUs a re-tape or a sweep beside one has (with_cus): half of them
    // when re-tapes run (more than two segments)
    const int ncu_shared = segs.size() > 2 ? std::max(1, L.ncu / 2) : L.ncu;
    int prepared_cs = ncoarse;  // coarse segment whose fine checkpoints are in `fine`
    auto retape = [&](int k) -> int {
        const Seg &g = segs[k];
        int rr;
        if (k >= 2) BGK_HIP(hipStreamWaitEvent(s1, ev_sweep.e[k % 2], 0));
        if (M > 1 && g.cs != prepared_cs) {
            // fine checkpoints 0 .. g.j of this coarse segment (g.j is its
            // first segment to re-tape)
            if ((rr = a.restore(ck[g.cs], s1))) return rr;
            for (int jj = 0; jj <= g.j; jj++) {
                if ((rr = a.save(fine[jj], s1))) return rr;
                if (jj == g.j) break;
                if ((rr = step_periodic(c, K, false, s1))) return rr;
                a.recomputed += K;
            }
            prepared_cs = g.cs;
        }
        // the first taped pass reads the checkpoint in place
        const int n = g.s1 - g.s0, kk = c->cur, first = taped_first_pass<C>(L, n, ncu_shared);
        void *mine = c->lat.pop[kk][0];
        a.set_copy(kk, M > 1 ? fine[g.j] : ck[g.cs]);
        // the re-tape's final state is not stored (nothing reads it)
        if ((rr = run_taped<S, C>(c, tape[k % 2], first, s1, ncu_shared, n > first))) return rr;
        a.set_copy(kk, mine);
        if (n > first &&
            (rr = run_taped<S, C>(c, tape[k % 2] + (size_t)first * 3 * N, n - first, s1,
                                  ncu_shared, false)))
            return rr;
        a.recomputed += g.s1 - g.s0;
        BGK_HIP(hipEventRecord(ev_tape.e[k % 2], s1));
        return BGK_OK;
    };
    C *lin = lamA, *lout = lamB;
    const int nsegs = (int)segs.size();
    for (int k = 0; k < nsegs; k++) {
        if (k >= 2) BGK_HIP(hipStreamWaitEvent(c->s0, ev_tape.e[k % 2], 0));
        // sweeps 1 .. nsegs-2 run beside a re-tape, the first and the last alone
        const int ncu_k = (k >= 1 && k + 1 < nsegs) ? ncu_shared : L.ncu;
        if ((r = adjoint_sweep<C>(c, lin, lout, tape[k % 2], segs[k].s1 - segs[k].s0, ncu_k)))
            return r;
        BGK_HIP(hipEventRecord(ev_sweep.e[k % 2], c->s0));
        if (k + 2 < nsegs && (r = retape(k + 2))) return r;
    }
    hipLaunchKernelGGL(k_grad_u0<C>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->s0, lin,
                       u0, grad, L.nx, L.ny, L.nz, L.ystride, L.zstride);
    BGK_HIP(hipGetLastError());
    BGK_HIP(hipStreamSynchronize(s1));
    BGK_HIP(hipStreamSynchronize(c->s0));
    const double bwd_ms = ms_since(t0);
    if (stats) {
        stats[0] = (double)a.bytes;
        stats[1] = fwd_ms;
        stats[2] = bwd_ms;
        stats[3] = (double)a.recomputed;
    }
    return BGK_OK;
}

template <typename S, typename C>
int forward_t(bgk_ctx *c, const double *u0, int T, double *J) {
    const Lattice &L = c->lat;
    BGK_HIP(launch_init_equilibrium(L, c->cur, u0, L.nsite, 0, L.nsite, c->s0));
    int r = step_periodic(c, T, false);
    return r ? r : reduce_ke(c, c->cur, J);
}

}  // namespace

static int check_adjoint_ctx(bgk_ctx *c, int nsteps) {
    if (c->lat.layout != BGK_LAYOUT_PERIODIC)
        return fail(BGK_EINVAL, "adjoint: periodic layout only");
    if (!c->lat.shifted) return fail(BGK_EINVAL, "adjoint: shifted storage only");
    if (nsteps < 0) return fail(BGK_EINVAL, "adjoint: nsteps must be >= 0");
    return BGK_OK;
}

int adjoint_ke(bgk_ctx *c, const double *u0, int nsteps, int ckpt_every, int levels, double *J,
               double *grad, double stats[4]) {
    int r = check_adjoint_ctx(c, nsteps);
    if (r) return r;
    if (levels < 0 || levels > 2) return fail(BGK_EINVAL, "adjoint: levels must be 0, 1 or 2");
    switch (c->lat.storage) {
    case BGK_STORE_F64: return adjoint_t<double, double>(c, u0, nsteps, ckpt_every, levels, J, grad, stats);
    case BGK_STORE_F32: return adjoint_t<float, float>(c, u0, nsteps, ckpt_every, levels, J, grad, stats);
    case BGK_STORE_F32_MIXED:
        return adjoint_t<float, double>(c, u0, nsteps, ckpt_every, levels, J, grad, stats);
    default: return fail(BGK_EINVAL, "adjoint: storage must be f64, f32 or f32-mixed");
    }
}

int forward_ke(bgk_ctx *c, const double *u0, int nsteps, double *J) {
    int r = check_adjoint_ctx(c, nsteps);
    if (r) return r;
    switch (c->lat.storage) {
    case BGK_STORE_F64: return forward_t<double, double>(c, u0, nsteps, J);
    case BGK_STORE_F32: return forward_t<float, float>(c, u0, nsteps, J);
    case BGK_STORE_F32_MIXED: return forward_t<float, double>(c, u0, nsteps, J);
    case BGK_STORE_F16: return forward_t<half_t, float>(c, u0, nsteps, J);
    default: return forward_t<bf16_t, float>(c, u0, nsteps, J);
    }
}

}  // namespace bgk
