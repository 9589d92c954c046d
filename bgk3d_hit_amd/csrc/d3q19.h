// d3q19.h -- D3Q19 lattice constants, storage traits and the per-site BGK
// collision shared by every kernel of the library.
//
// Velocity set and weights: reference SRC/storage.F90:98-133 (p0 = 1/3,
// p1 = 1/18, p2 = 1/36) and SRC/alloca.F90:57-175; population q (1..19) is
// stored at index q-1.  Pull streaming: population q at site x comes from
// site x - c_q (SRC/col_MC.F90:113-131).
//
// The arithmetic of collide_site follows SRC/col_MC.F90:133-219,333-351
// term by term in textual left-to-right order; the library is compiled with
// -ffp-contract=off so that FP64/FP32 results are reproducible bit for bit
// on any IEEE machine that evaluates the same expressions.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bgk {

// c_q for q = 1..19 at index q-1
__host__ __device__ constexpr int cx(int i) {
    constexpr int v[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
    return v[i];
}
__host__ __device__ constexpr int cy(int i) {
    constexpr int v[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
    return v[i];
}
__host__ __device__ constexpr int cz(int i) {
    constexpr int v[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
    return v[i];
}
// weight of population i (0-based): 1/3 rest, 1/18 axis, 1/36 diagonal
__host__ __device__ constexpr double wq(int i) {
    return (cx(i) * cx(i) + cy(i) * cy(i) + cz(i) * cz(i)) == 0
               ? 1.0 / 3.0
               : ((cx(i) * cx(i) + cy(i) * cy(i) + cz(i) * cz(i)) == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
}

// ---------------------------------------------------------------- storage
// Storage element types.  Half and bfloat16 are kept as raw 16-bit words
// in memory and converted with round-to-nearest-even.
struct half_t {
    uint16_t bits;
};
struct bf16_t {
    uint16_t bits;
};

template <typename S>
struct Storage;

template <>
struct Storage<double> {
    using raw = double;
    template <typename C>
    __device__ __forceinline__ static C load(double v) { return (C)v; }
    template <typename C>
    __device__ __forceinline__ static double store(C v) { return (double)v; }
};
template <>
struct Storage<float> {
    using raw = float;
    template <typename C>
    __device__ __forceinline__ static C load(float v) { return (C)v; }
    template <typename C>
    __device__ __forceinline__ static float store(C v) { return (float)v; }
};
template <>
struct Storage<half_t> {
    using raw = half_t;
    template <typename C>
    __device__ __forceinline__ static C load(half_t v) {
        return (C)(float)__builtin_bit_cast(_Float16, v.bits);
    }
    template <typename C>
    __device__ __forceinline__ static half_t store(C v) {
        // Round to f32 first, then to f16 (v_cvt_f16_f32, nearest even).  The
        // empty asm keeps the f32 value opaque: without it the compiler fuses
        // (half)(a*b) into v_fma_mixlo_f16, which rounds a*b once, directly
        // to f16, and differs from rounding the f32 product.
        float f = (float)v;
        asm volatile("" : "+v"(f));
        return half_t{__builtin_bit_cast(uint16_t, (_Float16)f)};
    }
};
template <>
struct Storage<bf16_t> {
    using raw = bf16_t;
    template <typename C>
    __device__ __forceinline__ static C load(bf16_t v) {
        return (C)__builtin_bit_cast(float, (uint32_t)v.bits << 16);
    }
    // Round to f32 first, then to bf16 with one v_cvt_pk_bf16_f32 (nearest
    // even; NaNs keep their high payload bits and become quiet).  This equals
    // the integer form x + 0x7fff + ((x >> 16) & 1) >> 16 (NaN: x >> 16 |
    // 0x40) for every f32 bit pattern, checked exhaustively on gfx950.
    template <typename C>
    __device__ __forceinline__ static bf16_t store(C v) {
        float f = (float)v;
        asm volatile("" : "+v"(f));
        return bf16_t{__builtin_bit_cast(uint16_t, (__bf16)f)};
    }
};

// The storage values of b[0..18] into out[]: for the 16-bit kinds two
// values per conversion instruction (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32:
// each f32 value rounded to nearest even exactly as by the one-value form,
// the pair packed into one register, whose halves the 16-bit stores and the
// widening conversions address directly); element by element otherwise.
// The conversions are written as asm: the compiler lowered the same pair
// conversion written as a two-element vector conversion to one converted
// element stored twice, and the asm also keeps the f32 inputs opaque (see
// Storage<half_t>::store).  store_sel<M> converts only the populations in
// the bit mask M (a pair if either of its two is in M) and leaves the other
// elements of out[] unset; the values it does convert are those of
// store_all.
template <uint32_t M, typename S, typename C>
__device__ __forceinline__ void store_sel(const C (&b)[19], S (&out)[19]) {
    if constexpr (std::is_same<S, half_t>::value || std::is_same<S, bf16_t>::value) {
#pragma unroll
        for (int q = 0; q + 1 < 19; q += 2) {
            if (!((M >> q) & 3u)) continue;
            const float lo = (float)b[q], hi = (float)b[q + 1];
            uint32_t p;
            if constexpr (std::is_same<S, half_t>::value)
                asm("v_cvt_pk_f16_f32 %0, %1, %2" : "=v"(p) : "v"(lo), "v"(hi));
            else
                asm("v_cvt_pk_bf16_f32 %0, %1, %2" : "=v"(p) : "v"(lo), "v"(hi));
            out[q].bits = (uint16_t)p;
            out[q + 1].bits = (uint16_t)(p >> 16);
        }
        if ((M >> 18) & 1u) out[18] = Storage<S>::template store<C>(b[18]);
    } else {
#pragma unroll
        for (int q = 0; q < 19; q++)
            if ((M >> q) & 1u) out[q] = Storage<S>::template store<C>(b[q]);
    }
}
constexpr uint32_t ALL_POPS = (1u << 19) - 1;
template <typename S, typename C>
__device__ __forceinline__ void store_all(const C (&b)[19], S (&out)[19]) {
    store_sel<ALL_POPS, S, C>(b, out);
}
// the populations whose c_x (c_y) is s, as a bit mask
__host__ __device__ constexpr uint32_t pops_cx(int s) {
    uint32_t m = 0;
    for (int q = 0; q < 19; q++)
        if (cx(q) == s) m |= 1u << q;
    return m;
}
__host__ __device__ constexpr uint32_t pops_cy(int s) {
    uint32_t m = 0;
    for (int q = 0; q < 19; q++)
        if (cy(q) == s) m |= 1u << q;
    return m;
}

// ---------------------------------------------------------------- collide
// Lattice constants in the compute kind: the reference casts double
// parameters to mykind (SRC/storage.F90:123-133).
template <typename C>
struct K {
    static constexpr C uno = (C)1.0, tre = (C)3.0, rf = (C)3.0, qf = (C)1.5;
    static constexpr C p0 = (C)(1.0 / 3.0);
    static constexpr C p1 = (C)(1.0 / 18.0);
    static constexpr C p2 = (C)(1.0 / (3.0 * 3.0 * (1.0 + 3.0)));
};

// uno / x, the reference's rhoinv (SRC/col_MC.F90:135), rounded to nearest
// even as the IEEE quotient.  FP32: the hardware reciprocal estimate
// (v_rcp_f32, within one ulp) and one Newton-Raphson correction with fused
// multiply-adds, e = 1 - x y, y + e y (Markstein's correction); it equals
// the IEEE quotient 1.0f / x for every x in [2^-125, 2^126), all
// 2 105 540 608 of them compared on gfx950 (bin/recip_check, run by the GPU
// tests), in 3 instructions where the quotient takes 10 (two scale steps,
// the estimate, five fused steps, the fix-ups).  Any other x (zero,
// denormal, negative, huge, Inf, NaN: a lattice whose density left that
// range has diverged) takes the quotient itself.  FP64: the quotient.
template <typename C>
__device__ __forceinline__ C recip_rn(C x) {
    return K<C>::uno / x;
}
template <>
__device__ __forceinline__ float recip_rn<float>(float x) {
    if (x >= 0x1p-125f && x < 0x1p126f) {
        const float y = __builtin_amdgcn_rcpf(x);
        const float e = __builtin_fmaf(-x, y, 1.0f);
        return __builtin_fmaf(e, y, y);
    }
    return 1.0f / x;
}

// cte0 + y for the quadratic terms of collide_site.  SH (shifted storage,
// cte1 = 1, cte0 = +0): +0 + y equals y bit for bit unless y is -0, and no y
// passed here is -0: y = qf * (a - vsq) with a = (tre * v) * v or tre * v*v,
// a product of two factors of equal sign (>= +0, never -0), vsq a sum of
// squares (>= +0), so a - vsq is never -0 (round to nearest gives -0 only
// for -0 - +0), and qf = 1.5 keeps the sign of a nonzero value without
// underflowing to zero.  The add is then dropped (9 per site).
template <bool SH, typename C>
__device__ __forceinline__ C add_cte0(C cte0, C y) {
    if constexpr (SH)
        return y;
    else
        return cte0 + y;
}

// What the equilibria of a site need once its moments are known: the
// linear terms rf * (velocity combinations), the quadratic terms and the
// density products (SRC/col_MC.F90:133-193).
template <typename C>
struct Moments {
    C vx, vy, vz, vxpy, vxmy, vxpz, vxmz, vypz, vymz;  // times rf
    C qxpy, qxmy, qxpz, qxmz, qypz, qymz, qx, qy, qz, q0;
    C rp0, rp1, rp2, s0, s1, s2;
};

// One site of SRC/col_MC.F90:133-219,333-351, in two parts: collide_moments
// (the density, the velocity and the terms above) and collide_relax<M>
// (the equilibria e_q and the relaxed b_q = x_q - omega (x_q - e_q) of the
// populations in the bit mask M, the others of b[] left unset).
// collide_site is both for all populations.  x[] are the pulled
// populations, b[] receives the post-collision populations (b[18] is the
// rest population the reference updates in place).  cte1 = 1 for shifted
// storage (SRC/col_MC.F90:86-91).
// ux, uy, uz receive the pulled-moment velocity j/rho (before the
// reference scales it by rf for the linear terms).
// SH: compiled for shifted storage only (cte1 = 1 whatever is passed): the
// products cte1 * (...) fold to their second factor and add_cte0 drops the
// additions of cte0 = +0; results are bit-identical to SH = false with
// cte1 = 1.
template <typename C, bool SH = false>
__device__ __forceinline__ Moments<C> collide_moments(const C (&x)[19], C cte1_arg, C &ux, C &uy,
                                                      C &uz) {
    using k = K<C>;
    const C cte1 = SH ? k::uno : cte1_arg;
    const C cte0 = k::uno - cte1;
    const C x01 = x[0], x02 = x[1], x03 = x[2], x04 = x[3], x05 = x[4], x06 = x[5], x07 = x[6];
    const C x08 = x[7], x09 = x[8], x10 = x[9], x11 = x[10], x12 = x[11], x13 = x[12];
    const C x14 = x[13], x15 = x[14], x16 = x[15], x17 = x[16], x18 = x[17], x19 = x[18];

    const C rho = x01 + x02 + x03 + x04 + x05 + x06 + x07 + x08 + x09 + x10 + x11 + x12 + x13 +
                  x14 + x15 + x16 + x17 + x18 + x19 + cte1;
    const C rhoinv = recip_rn<C>(rho);
    C vx = (x01 + x02 + x03 + x04 + x05 - x10 - x11 - x12 - x13 - x14) * rhoinv;
    C vy = (x03 + x07 + x08 + x09 + x12 - x01 - x10 - x16 - x17 - x18) * rhoinv;
    C vz = (x04 + x06 + x07 + x13 + x18 - x02 - x09 - x11 - x15 - x16) * rhoinv;

    ux = vx;
    uy = vy;
    uz = vz;
    Moments<C> m;
    const C vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    const C vsq = vx2 + vy2 + vz2;
    C vxpy = vx + vy;
    m.qxpy = add_cte0<SH>(cte0, k::qf * (k::tre * vxpy * vxpy - vsq));
    C vxmy = vx - vy;
    m.qxmy = add_cte0<SH>(cte0, k::qf * (k::tre * vxmy * vxmy - vsq));
    C vxpz = vx + vz;
    m.qxpz = add_cte0<SH>(cte0, k::qf * (k::tre * vxpz * vxpz - vsq));
    C vxmz = vx - vz;
    m.qxmz = add_cte0<SH>(cte0, k::qf * (k::tre * vxmz * vxmz - vsq));
    C vypz = vy + vz;
    m.qypz = add_cte0<SH>(cte0, k::qf * (k::tre * vypz * vypz - vsq));
    C vymz = vy - vz;
    m.qymz = add_cte0<SH>(cte0, k::qf * (k::tre * vymz * vymz - vsq));
    m.qx = add_cte0<SH>(cte0, k::qf * (k::tre * vx2 - vsq));
    m.qy = add_cte0<SH>(cte0, k::qf * (k::tre * vy2 - vsq));
    m.qz = add_cte0<SH>(cte0, k::qf * (k::tre * vz2 - vsq));
    m.q0 = cte0 + k::qf * (-vsq);  // qf * (-vsq) is -0 when vsq = 0: kept

    m.vx = k::rf * vx;
    m.vy = k::rf * vy;
    m.vz = k::rf * vz;
    m.vxpy = k::rf * vxpy;
    m.vxmy = k::rf * vxmy;
    m.vxpz = k::rf * vxpz;
    m.vxmz = k::rf * vxmz;
    m.vypz = k::rf * vypz;
    m.vymz = k::rf * vymz;

    m.rp0 = rho * k::p0;
    m.rp1 = rho * k::p1;
    m.rp2 = rho * k::p2;
    m.s2 = cte1 * (m.rp2 - k::p2);
    m.s1 = cte1 * (m.rp1 - k::p1);
    m.s0 = cte1 * (m.rp0 - k::p0);
    return m;
}

template <uint32_t M, typename C>
__device__ __forceinline__ void collide_relax(const C (&x)[19], const Moments<C> &m, C omega,
                                              C (&b)[19]) {
    C e[19];
    if (M & (1u << 0)) e[0] = m.rp2 * (+m.vxmy + m.qxmy) + m.s2;
    if (M & (1u << 1)) e[1] = m.rp2 * (+m.vxmz + m.qxmz) + m.s2;
    if (M & (1u << 2)) e[2] = m.rp2 * (+m.vxpy + m.qxpy) + m.s2;
    if (M & (1u << 3)) e[3] = m.rp2 * (+m.vxpz + m.qxpz) + m.s2;
    if (M & (1u << 4)) e[4] = m.rp1 * (+m.vx + m.qx) + m.s1;
    if (M & (1u << 5)) e[5] = m.rp1 * (+m.vz + m.qz) + m.s1;
    if (M & (1u << 6)) e[6] = m.rp2 * (+m.vypz + m.qypz) + m.s2;
    if (M & (1u << 7)) e[7] = m.rp1 * (+m.vy + m.qy) + m.s1;
    if (M & (1u << 8)) e[8] = m.rp2 * (+m.vymz + m.qymz) + m.s2;
    if (M & (1u << 9)) e[9] = m.rp2 * (-m.vxpy + m.qxpy) + m.s2;
    if (M & (1u << 10)) e[10] = m.rp2 * (-m.vxpz + m.qxpz) + m.s2;
    if (M & (1u << 11)) e[11] = m.rp2 * (-m.vxmy + m.qxmy) + m.s2;
    if (M & (1u << 12)) e[12] = m.rp2 * (-m.vxmz + m.qxmz) + m.s2;
    if (M & (1u << 13)) e[13] = m.rp1 * (-m.vx + m.qx) + m.s1;
    if (M & (1u << 14)) e[14] = m.rp1 * (-m.vz + m.qz) + m.s1;
    if (M & (1u << 15)) e[15] = m.rp2 * (-m.vypz + m.qypz) + m.s2;
    if (M & (1u << 16)) e[16] = m.rp1 * (-m.vy + m.qy) + m.s1;
    if (M & (1u << 17)) e[17] = m.rp2 * (-m.vymz + m.qymz) + m.s2;
    if (M & (1u << 18)) e[18] = m.rp0 * (+m.q0) + m.s0;
#pragma unroll
    for (int q = 0; q < 19; q++)
        if ((M >> q) & 1u) b[q] = x[q] - omega * (x[q] - e[q]);
}

template <typename C, bool SH = false>
__device__ __forceinline__ void collide_site(const C (&x)[19], C (&b)[19], C omega, C cte1_arg,
                                             C &ux, C &uy, C &uz) {
    const Moments<C> m = collide_moments<C, SH>(x, cte1_arg, ux, uy, uz);
    collide_relax<(1u << 19) - 1>(x, m, omega, b);
}

// At-site moments of SRC/diagno.F90:96-111 (also the first lines of
// SRC/dissipation.F90:97-114): returns rho, u and |u|^2/2 in the compute kind.
template <typename C>
__device__ __forceinline__ void moments_site(const C (&x)[19], C cte1, C &rho, C &ux, C &uy, C &uz) {
    const C x01 = x[0], x02 = x[1], x03 = x[2], x04 = x[3], x05 = x[4], x06 = x[5], x07 = x[6];
    const C x08 = x[7], x09 = x[8], x10 = x[9], x11 = x[10], x12 = x[11], x13 = x[12];
    const C x14 = x[13], x15 = x[14], x16 = x[15], x17 = x[16], x18 = x[17], x19 = x[18];
    rho = x01 + x02 + x03 + x04 + x05 + x06 + x07 + x08 + x09 + x10 + x11 + x12 + x13 + x14 +
          x15 + x16 + x17 + x18 + x19 + cte1;
    const C rhoinv = recip_rn<C>(rho);
    ux = (x01 + x02 + x03 + x04 + x05 - x10 - x11 - x12 - x13 - x14) * rhoinv;
    uy = (x03 + x07 + x08 + x09 + x12 - x01 - x10 - x16 - x17 - x18) * rhoinv;
    uz = (x04 + x06 + x07 + x13 + x18 - x02 - x09 - x11 - x15 - x16) * rhoinv;
}

// One site of SRC/dissipation.F90:97-249: returns 2 svisc P (P the squared
// Frobenius norm of Pi^neq) and |u|^2/2, both in the compute kind.
template <typename C>
__device__ __forceinline__ void dissipation_site(const C (&x)[19], C cte1, C svisc, C &dsite,
                                                 C &tsite) {
    using k = K<C>;
    const C cte0 = k::uno - cte1;
    const C x01 = x[0], x02 = x[1], x03 = x[2], x04 = x[3], x05 = x[4], x06 = x[5], x07 = x[6];
    const C x08 = x[7], x09 = x[8], x10 = x[9], x11 = x[10], x12 = x[11], x13 = x[12];
    const C x14 = x[13], x15 = x[14], x16 = x[15], x17 = x[16], x18 = x[17], x19 = x[18];
    const C rho = x01 + x02 + x03 + x04 + x05 + x06 + x07 + x08 + x09 + x10 + x11 + x12 + x13 +
                  x14 + x15 + x16 + x17 + x18 + x19 + cte1;
    const C rhoinv = recip_rn<C>(rho);
    C vx = (x01 + x02 + x03 + x04 + x05 - x10 - x11 - x12 - x13 - x14) * rhoinv;
    C vy = (x03 + x07 + x08 + x09 + x12 - x01 - x10 - x16 - x17 - x18) * rhoinv;
    C vz = (x04 + x06 + x07 + x13 + x18 - x02 - x09 - x11 - x15 - x16) * rhoinv;
    const C vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    const C vsq = vx2 + vy2 + vz2;
    tsite = (C)0.5 * vsq;
    C vxpy = vx + vy;
    const C qxpy = cte0 + k::qf * (k::tre * vxpy * vxpy - vsq);
    C vxmy = vx - vy;
    const C qxmy = cte0 + k::qf * (k::tre * vxmy * vxmy - vsq);
    C vxpz = vx + vz;
    const C qxpz = cte0 + k::qf * (k::tre * vxpz * vxpz - vsq);
    C vxmz = vx - vz;
    const C qxmz = cte0 + k::qf * (k::tre * vxmz * vxmz - vsq);
    C vypz = vy + vz;
    const C qypz = cte0 + k::qf * (k::tre * vypz * vypz - vsq);
    C vymz = vy - vz;
    const C qymz = cte0 + k::qf * (k::tre * vymz * vymz - vsq);
    const C qx = cte0 + k::qf * (k::tre * vx2 - vsq);
    const C qy = cte0 + k::qf * (k::tre * vy2 - vsq);
    const C qz = cte0 + k::qf * (k::tre * vz2 - vsq);
    vx = k::rf * vx;
    vy = k::rf * vy;
    vz = k::rf * vz;
    vxpy = k::rf * vxpy;
    vxmy = k::rf * vxmy;
    vxpz = k::rf * vxpz;
    vxmz = k::rf * vxmz;
    vypz = k::rf * vypz;
    vymz = k::rf * vymz;
    const C rp1 = rho * k::p1, rp2 = rho * k::p2;
    const C s2 = cte1 * (rp2 - k::p2), s1 = cte1 * (rp1 - k::p1);
    const C n01 = x01 - (rp2 * (+vxmy + qxmy) + s2);
    const C n02 = x02 - (rp2 * (+vxmz + qxmz) + s2);
    const C n03 = x03 - (rp2 * (+vxpy + qxpy) + s2);
    const C n04 = x04 - (rp2 * (+vxpz + qxpz) + s2);
    const C n05 = x05 - (rp1 * (+vx + qx) + s1);
    const C n06 = x06 - (rp1 * (+vz + qz) + s1);
    const C n07 = x07 - (rp2 * (+vypz + qypz) + s2);
    const C n08 = x08 - (rp1 * (+vy + qy) + s1);
    const C n09 = x09 - (rp2 * (+vymz + qymz) + s2);
    const C n10 = x10 - (rp2 * (-vxpy + qxpy) + s2);
    const C n11 = x11 - (rp2 * (-vxpz + qxpz) + s2);
    const C n12 = x12 - (rp2 * (-vxmy + qxmy) + s2);
    const C n13 = x13 - (rp2 * (-vxmz + qxmz) + s2);
    const C n14 = x14 - (rp1 * (-vx + qx) + s1);
    const C n15 = x15 - (rp1 * (-vz + qz) + s1);
    const C n16 = x16 - (rp2 * (-vypz + qypz) + s2);
    const C n17 = x17 - (rp1 * (-vy + qy) + s1);
    const C n18 = x18 - (rp2 * (-vymz + qymz) + s2);
    const C Pxx = n01 + n02 + n03 + n04 + n05 + n10 + n11 + n12 + n13 + n14;
    const C Pyy = n01 + n03 + n07 + n08 + n09 + n10 + n12 + n16 + n17 + n18;
    const C Pzz = n02 + n04 + n06 + n07 + n09 + n11 + n13 + n15 + n16 + n18;
    const C Pxz = -n02 + n04 + n11 - n13;
    const C Pxy = -n01 + n03 + n10 - n12;
    const C Pyz = +n07 - n09 + n16 - n18;
    const C two = (C)2.0;
    const C Ptotal = Pxx * Pxx + Pyy * Pyy + Pzz * Pzz + (two * Pxy * Pxy) + (two * Pxz * Pxz) +
                     (two * Pyz * Pyz);
    dsite = two * svisc * Ptotal;
}

}  // namespace bgk
