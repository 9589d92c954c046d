// bgk_hit.hip -- seeded homogeneous isotropic turbulence initial velocity.
//
// The reference initialises a Taylor-Green vortex (SRC/init.F90:83-92); the
// HIT field is this library's own definition:
//   * modes: integer wave vectors k with scaled wavenumber
//     kappa = (kx, ky*lx/ly, kz*lx/lz), 0 < |kappa| <= 4 k0, |k_a| < l_a/2
//     (Nyquist planes excluded); energy beyond 4 k0 is < 3e-11 of the total
//   * per mode a counter-based random draw (splitmix64 of seed and k), so the
//     field does not depend on the traversal order or on the lattice size
//   * Rogallo's construction u^(k) = A(k) (alpha e1 + beta e2) with e1, e2
//     orthonormal and perpendicular to kappa (solenoidal),
//     alpha = exp(i th1) cos(phi), beta = exp(i th2) sin(phi),
//     A(k) = |kappa| exp(-(|kappa|/k0)^2) so that the shell energy follows
//     E(k) ~ k^4 exp(-2 (k/k0)^2)
//   * Hermitian symmetry u^(-k) = conj(u^(k)) (real field), no k = 0 mode
//     (zero mean)
//   * amplitudes rescaled through Parseval so that sqrt(<|u|^2>/3) = urms
//   * synthesis u(x) = sum_k u^(k) exp(2 pi i (kx x/lx + ky y/ly + kz z/lz)),
//     site x = 0..lx-1, as three separable passes with tabulated twiddles.
// The host and device syntheses evaluate the same sums in the same order
// with the same twiddle table, so they agree bit for bit.
#include <math.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "bgk_internal.h"

namespace bgk {

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static inline double uniform(uint64_t key, int j) {
    return (double)(splitmix64(key + (uint64_t)j * 0x632BE59BD9B4E019ull) >> 11) *
           (1.0 / 9007199254740992.0);
}

static bool canonical(int kx, int ky, int kz) {
    return kz > 0 || (kz == 0 && (ky > 0 || (ky == 0 && kx > 0)));
}

// coefficient of a canonical mode (3 complex values)
static void canonical_mode(uint64_t seed, int kx, int ky, int kz, double kax, double kay,
                           double kaz, double k0, double out[6]) {
    const uint64_t pack = ((uint64_t)(uint16_t)kx << 32) | ((uint64_t)(uint16_t)ky << 16) |
                          (uint64_t)(uint16_t)kz;
    const uint64_t key = splitmix64(splitmix64(seed) ^ pack);
    const double twopi = 6.283185307179586476925286766559;
    const double th1 = twopi * uniform(key, 0), th2 = twopi * uniform(key, 1);
    const double phi = twopi * uniform(key, 2);
    const double kk = sqrt(kax * kax + kay * kay + kaz * kaz);
    const double A = kk * exp(-(kk / k0) * (kk / k0));
    double e1[3], e2[3];
    const double kp = sqrt(kax * kax + kay * kay);
    if (kp > 0) {
        e1[0] = kay / kp;
        e1[1] = -kax / kp;
        e1[2] = 0;
    } else {
        e1[0] = 1;
        e1[1] = 0;
        e1[2] = 0;
    }
    // e2 = kappa x e1 / |kappa|
    e2[0] = (kay * e1[2] - kaz * e1[1]) / kk;
    e2[1] = (kaz * e1[0] - kax * e1[2]) / kk;
    e2[2] = (kax * e1[1] - kay * e1[0]) / kk;
    const double ar = cos(th1) * cos(phi), ai = sin(th1) * cos(phi);
    const double br = cos(th2) * sin(phi), bi = sin(th2) * sin(phi);
    for (int d = 0; d < 3; d++) {
        out[2 * d] = A * (ar * e1[d] + br * e2[d]);
        out[2 * d + 1] = A * (ai * e1[d] + bi * e2[d]);
    }
}

int hit_spectrum(int lx, int ly, int lz, double k0, double urms, uint64_t seed,
                 HitSpectrum &sp) {
    if (lx < 1 || ly < 1 || lz < 1) return fail(BGK_EINVAL, "lattice sizes must be positive");
    if (!(k0 > 0) || !(urms >= 0)) return fail(BGK_EINVAL, "need k0 > 0 and urms >= 0");
    const double kcut = 4.0 * k0;
    const double sy = (double)lx / ly, sz = (double)lx / lz;
    sp.Kx = std::min((int)floor(kcut), (lx - 1) / 2);
    sp.Ky = std::min((int)floor(kcut / sy), (ly - 1) / 2);
    sp.Kz = std::min((int)floor(kcut / sz), (lz - 1) / 2);
    const int Kx = sp.Kx, Ky = sp.Ky, Kz = sp.Kz;
    const int NX = Kx + 1, NY = 2 * Ky + 1, NZ = 2 * Kz + 1;
    sp.coef.assign((size_t)3 * NZ * NY * NX * 2, 0.0);
    double energy = 0.0;
    for (int kz = -Kz; kz <= Kz; kz++)
        for (int ky = -Ky; ky <= Ky; ky++)
            for (int kx = 0; kx <= Kx; kx++) {
                const double kax = kx, kay = ky * sy, kaz = kz * sz;
                const double kk = sqrt(kax * kax + kay * kay + kaz * kaz);
                if (kk == 0.0 || kk > kcut) continue;
                double m[6];
                if (canonical(kx, ky, kz)) {
                    canonical_mode(seed, kx, ky, kz, kax, kay, kaz, k0, m);
                } else {
                    canonical_mode(seed, -kx, -ky, -kz, -kax, -kay, -kaz, k0, m);
                    for (int d = 0; d < 3; d++) m[2 * d + 1] = -m[2 * d + 1];
                }
                for (int d = 0; d < 3; d++) {
                    const size_t o = ((((size_t)d * NZ + (kz + Kz)) * NY + (ky + Ky)) * NX + kx) * 2;
                    sp.coef[o] = m[2 * d];
                    sp.coef[o + 1] = m[2 * d + 1];
                    // kx > 0 modes stand for their -k partner too
                    const double w = kx > 0 ? 2.0 : 1.0;
                    energy += w * (m[2 * d] * m[2 * d] + m[2 * d + 1] * m[2 * d + 1]);
                }
            }
    // Parseval: <|u|^2> = sum over all modes of |u^|^2
    const double scale = energy > 0 ? urms * sqrt(3.0 / energy) : 0.0;
    for (double &v : sp.coef) v *= scale;
    auto table = [](int n, std::vector<double> &t) {
        t.resize((size_t)2 * n);
        const double twopi = 6.283185307179586476925286766559;
        for (int i = 0; i < n; i++) {
            t[2 * i] = cos(twopi * i / n);
            t[2 * i + 1] = sin(twopi * i / n);
        }
    };
    table(lx, sp.twx);
    table(ly, sp.twy);
    table(lz, sp.twz);
    return BGK_OK;
}

// ----------------------------------------------- shared per-element bodies
// pass 1: P1[d][kx][ky][z] = sum_kz c[d][kz][ky][kx] E_z(kz z)
__host__ __device__ inline void hit_p1(const double *coef, const double *twz, int Kx, int Ky,
                                       int Kz, int lz, int d, int kx, int iky, int z,
                                       double *out) {
    const int NX = Kx + 1, NY = 2 * Ky + 1, NZ = 2 * Kz + 1;
    double re = 0.0, im = 0.0;
    for (int ikz = 0; ikz < NZ; ikz++) {
        const int kz = ikz - Kz;
        const size_t o = ((((size_t)d * NZ + ikz) * NY + iky) * NX + kx) * 2;
        const double cr = coef[o], ci = coef[o + 1];
        int t = (int)(((long long)kz * z) % lz);
        if (t < 0) t += lz;
        const double er = twz[2 * t], ei = twz[2 * t + 1];
        re = re + (cr * er - ci * ei);
        im = im + (cr * ei + ci * er);
    }
    out[0] = re;
    out[1] = im;
}

// pass 2: P2[d][kx][y][z] = sum_ky P1[d][kx][ky][z] E_y(ky y)
__host__ __device__ inline void hit_p2(const double *P1, const double *twy, int Kx, int Ky,
                                       int ly, int lz, int d, int kx, int y, int z, double *out) {
    const int NX = Kx + 1, NY = 2 * Ky + 1;
    double re = 0.0, im = 0.0;
    for (int iky = 0; iky < NY; iky++) {
        const int ky = iky - Ky;
        const size_t o = ((((size_t)d * NX + kx) * NY + iky) * lz + z) * 2;
        const double cr = P1[o], ci = P1[o + 1];
        int t = (int)(((long long)ky * y) % ly);
        if (t < 0) t += ly;
        const double er = twy[2 * t], ei = twy[2 * t + 1];
        re = re + (cr * er - ci * ei);
        im = im + (cr * ei + ci * er);
    }
    out[0] = re;
    out[1] = im;
}

// pass 3: u[d](x,y,z) = Re P2[d][0] + 2 sum_{kx>0} Re(P2[d][kx] E_x(kx x))
__host__ __device__ inline double hit_u(const double *P2, const double *twx, int Kx, int lx,
                                        int ly, int lz, int d, int x, int y, int z) {
    const int NX = Kx + 1;
    const size_t row = (size_t)y + (size_t)ly * z;
    const size_t pl = (size_t)ly * lz;
    double acc = P2[(((size_t)d * NX + 0) * pl + row) * 2];
    for (int kx = 1; kx < NX; kx++) {
        const size_t o = (((size_t)d * NX + kx) * pl + row) * 2;
        const int t = (int)(((long long)kx * x) % lx);
        const double term = P2[o] * twx[2 * t] - P2[o + 1] * twx[2 * t + 1];
        acc = acc + 2.0 * term;
    }
    return acc;
}

template <typename F>
static void parallel_for(long long n, F f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([=]() {
            for (long long i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

void hit_synthesize_host(int lx, int ly, int lz, const HitSpectrum &sp, double *u) {
    const int Kx = sp.Kx, Ky = sp.Ky, Kz = sp.Kz;
    const int NX = Kx + 1, NY = 2 * Ky + 1;
    std::vector<double> P1((size_t)3 * NX * NY * lz * 2), P2((size_t)3 * NX * ly * lz * 2);
    parallel_for((long long)3 * NX * NY, [&](long long i) {
        const int iky = (int)(i % NY), kx = (int)((i / NY) % NX), d = (int)(i / NY / NX);
        for (int z = 0; z < lz; z++)
            hit_p1(sp.coef.data(), sp.twz.data(), Kx, Ky, Kz, lz, d, kx, iky, z,
                   &P1[((((size_t)d * NX + kx) * NY + iky) * lz + z) * 2]);
    });
    parallel_for((long long)3 * NX * lz, [&](long long i) {
        const int z = (int)(i % lz), kx = (int)((i / lz) % NX), d = (int)(i / lz / NX);
        for (int y = 0; y < ly; y++)
            hit_p2(P1.data(), sp.twy.data(), Kx, Ky, ly, lz, d, kx, y, z,
                   &P2[((((size_t)d * NX + kx) * ((size_t)ly * lz)) + y + (size_t)ly * z) * 2]);
    });
    const size_t N = (size_t)lx * ly * lz;
    parallel_for((long long)3 * ly * lz, [&](long long i) {
        const int y = (int)(i % ly), z = (int)((i / ly) % lz), d = (int)(i / ly / lz);
        for (int x = 0; x < lx; x++)
            u[(size_t)d * N + x + (size_t)lx * (y + (size_t)ly * z)] =
                hit_u(P2.data(), sp.twx.data(), Kx, lx, ly, lz, d, x, y, z);
    });
}

// ----------------------------------------------------------- device passes
__global__ void k_hit_p1(const double *coef, const double *twz, int Kx, int Ky, int Kz, int lz,
                         double *P1) {
    const int NX = Kx + 1, NY = 2 * Ky + 1;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)3 * NX * NY * lz;
    if (i >= n) return;
    const int z = (int)(i % lz);
    const long long r = i / lz;
    const int iky = (int)(r % NY), kx = (int)((r / NY) % NX), d = (int)(r / NY / NX);
    hit_p1(coef, twz, Kx, Ky, Kz, lz, d, kx, iky, z,
           &P1[((((size_t)d * NX + kx) * NY + iky) * lz + z) * 2]);
}

__global__ void k_hit_p2(const double *P1, const double *twy, int Kx, int Ky, int ly, int lz,
                         double *P2) {
    const int NX = Kx + 1;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = (long long)3 * NX * ly * lz;
    if (i >= n) return;
    const int y = (int)(i % ly);
    const long long r = i / ly;
    const int z = (int)(r % lz), kx = (int)((r / lz) % NX), d = (int)(r / lz / NX);
    hit_p2(P1, twy, Kx, Ky, ly, lz, d, kx, y, z,
           &P2[((((size_t)d * NX + kx) * ((size_t)ly * lz)) + y + (size_t)ly * z) * 2]);
}

__global__ void k_hit_u(const double *P2, const double *twx, int Kx, int lx, int ly, int lz,
                        double *u) {
    const long long N = (long long)lx * ly * lz;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= 3 * N) return;
    const int d = (int)(i / N);
    const long long s = i % N;
    const int x = (int)(s % lx), y = (int)((s / lx) % ly), z = (int)(s / ((long long)lx * ly));
    u[i] = hit_u(P2, twx, Kx, lx, ly, lz, d, x, y, z);
}

int hit_synthesize_device(bgk_ctx *c, const HitSpectrum &sp, double *u_dev) {
    const Lattice &L = c->lat;
    const int Kx = sp.Kx, Ky = sp.Ky, NX = Kx + 1, NY = 2 * Ky + 1;
    const size_t n1 = (size_t)3 * NX * NY * L.nz * 2, n2 = (size_t)3 * NX * L.ny * L.nz * 2;
    double *coef = nullptr, *tw = nullptr, *P1 = nullptr, *P2 = nullptr;
    const size_t ntw = sp.twx.size() + sp.twy.size() + sp.twz.size();
    int rc = BGK_OK;
    hipError_t e;
    if ((e = hipMalloc(&coef, sizeof(double) * sp.coef.size())) != hipSuccess ||
        (e = hipMalloc(&tw, sizeof(double) * ntw)) != hipSuccess ||
        (e = hipMalloc(&P1, sizeof(double) * n1)) != hipSuccess ||
        (e = hipMalloc(&P2, sizeof(double) * n2)) != hipSuccess) {
        rc = hip_fail(e, "hipMalloc(HIT synthesis)");
    }
    double *twx = tw, *twy = tw + sp.twx.size(), *twz = twy + sp.twy.size();
    if (!rc) {
        e = hipMemcpy(coef, sp.coef.data(), sizeof(double) * sp.coef.size(),
                      hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(twx, sp.twx.data(), sizeof(double) * sp.twx.size(),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(twy, sp.twy.data(), sizeof(double) * sp.twy.size(),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(twz, sp.twz.data(), sizeof(double) * sp.twz.size(),
                          hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy(HIT tables)");
    }
    if (!rc) {
        long long n = (long long)3 * NX * NY * L.nz;
        hipLaunchKernelGGL(k_hit_p1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s0, coef,
                           twz, Kx, Ky, sp.Kz, L.nz, P1);
        n = (long long)3 * NX * L.ny * L.nz;
        hipLaunchKernelGGL(k_hit_p2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s0, P1,
                           twy, Kx, Ky, L.ny, L.nz, P2);
        n = 3 * L.nsite;
        hipLaunchKernelGGL(k_hit_u, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s0, P2,
                           twx, Kx, L.nx, L.ny, L.nz, u_dev);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->s0);
        if (e != hipSuccess) rc = hip_fail(e, "HIT synthesis kernels");
    }
    (void)hipFree(coef);
    (void)hipFree(tw);
    (void)hipFree(P1);
    (void)hipFree(P2);
    return rc;
}

}  // namespace bgk
