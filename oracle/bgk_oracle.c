/*
 * bgk_oracle.c -- CPU oracle for the BGK3D_HIT time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and timed CPU baseline); see
 * bgk_oracle.h for what each function restates and how it is pinned.
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "bgk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_lattice {
    int l, m, n, kind, shifted;
    double omega;
    size_t ntot;  /* (l+2)(m+2)(n+2) */
    size_t esize; /* bytes per stored value */
    void *a[19];  /* current populations (reference a01..a19)             */
    void *b[19];  /* next populations (reference b01..b18; b19 unused)   */
    void *buf;    /* the step3 message buffer (5 planes)                  */
};

/* ---- storage conversions (round to nearest even) --------------------- */
uint16_t orc_f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x47800000u) return (uint16_t)(sign | 0x7c00u); /* >= 2^16 */
    if (ax < 0x38800000u) {                                   /* < 2^-14: subnormal */
        uint32_t e = ax >> 23;
        int sh = 126 - (int)e;
        if (sh > 24) return (uint16_t)sign;
        uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
        uint32_t r = mant >> sh;
        uint32_t rem = mant & ((1u << sh) - 1u);
        uint32_t halfway = 1u << (sh - 1);
        if (rem > halfway || (rem == halfway && (r & 1u))) r++;
        return (uint16_t)(sign | r);
    }
    uint32_t h = ((ax >> 23) - 112u) << 10 | ((ax & 0x7fffffu) >> 13);
    uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t mt = (uint32_t)h & 0x3ffu;
    uint32_t bits;
    if (e == 0) {
        float v = (float)mt * 5.9604644775390625e-08f; /* 2^-24, exact */
        memcpy(&bits, &v, 4);
        bits |= sign;
    } else if (e == 31) {
        bits = sign | 0x7f800000u | (mt << 13);
    } else {
        bits = sign | ((e + 112u) << 23) | (mt << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

uint16_t orc_f32_to_bf16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}

float orc_bf16_to_f32(uint16_t h)
{
    uint32_t x = (uint32_t)h << 16;
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* ---- kind instantiations --------------------------------------------- */
#define STOR double
#define COMP double
#define LD(v) ((COMP)(v))
#define ST(v) ((STOR)(v))
#define SFX f64
#include "bgk_oracle_body.inc"
#undef STOR
#undef COMP
#undef LD
#undef ST
#undef SFX

#define STOR float
#define COMP float
#define LD(v) ((COMP)(v))
#define ST(v) ((STOR)(v))
#define SFX f32
#include "bgk_oracle_body.inc"
#undef STOR
#undef COMP
#undef LD
#undef ST
#undef SFX

#define STOR float
#define COMP double
#define LD(v) ((COMP)(v))
#define ST(v) ((STOR)(v))
#define SFX f32mix
#include "bgk_oracle_body.inc"
#undef STOR
#undef COMP
#undef LD
#undef ST
#undef SFX

#define STOR uint16_t
#define COMP float
#define LD(v) (orc_f16_to_f32(v))
#define ST(v) (orc_f32_to_f16(v))
#define SFX f16
#include "bgk_oracle_body.inc"
#undef STOR
#undef COMP
#undef LD
#undef ST
#undef SFX

#define STOR uint16_t
#define COMP float
#define LD(v) (orc_bf16_to_f32(v))
#define ST(v) (orc_f32_to_bf16(v))
#define SFX bf16
#include "bgk_oracle_body.inc"
#undef STOR
#undef COMP
#undef LD
#undef ST
#undef SFX

#define DISPATCH(fn, ...)                                                                    \
    switch (s->kind) {                                                                       \
    case ORC_F64: fn##_f64(__VA_ARGS__); break;                                              \
    case ORC_F32: fn##_f32(__VA_ARGS__); break;                                              \
    case ORC_F32MIX: fn##_f32mix(__VA_ARGS__); break;                                        \
    case ORC_F16: fn##_f16(__VA_ARGS__); break;                                              \
    default: fn##_bf16(__VA_ARGS__); break;                                                  \
    }

/* ---- lifecycle (SRC/alloca.F90:178-216) ------------------------------ */
orc_lattice *orc_create(int l, int m, int n, int kind, double omega, int shifted)
{
    if (l < 1 || m < 1 || n < 1 || kind < 0 || kind > 4) return NULL;
    orc_lattice *s = (orc_lattice *)calloc(1, sizeof(*s));
    if (!s) return NULL;
    s->l = l;
    s->m = m;
    s->n = n;
    s->kind = kind;
    s->omega = omega;
    s->shifted = shifted;
    s->ntot = (size_t)(l + 2) * (m + 2) * (n + 2);
    s->esize = kind == ORC_F64 ? 8 : (kind <= ORC_F32MIX ? 4 : 2);
    /* zero-filled: the reference leaves b's halos uninitialised until the
     * first exchange on that array; zero is this oracle's definition */
    for (int q = 0; q < 19; q++) {
        s->a[q] = calloc(s->ntot, s->esize);
        s->b[q] = (q < 18) ? calloc(s->ntot, s->esize) : NULL;
        if (!s->a[q] || (q < 18 && !s->b[q])) {
            orc_destroy(s);
            return NULL;
        }
    }
    size_t maxp = (size_t)(l + 2) * (m + 2);
    if ((size_t)(m + 2) * (n + 2) > maxp) maxp = (size_t)(m + 2) * (n + 2);
    if ((size_t)(l + 2) * (n + 2) > maxp) maxp = (size_t)(l + 2) * (n + 2);
    s->buf = calloc(5 * maxp, s->esize);
    if (!s->buf) {
        orc_destroy(s);
        return NULL;
    }
    return s;
}

void orc_destroy(orc_lattice *s)
{
    if (!s) return;
    for (int q = 0; q < 19; q++) {
        free(s->a[q]);
        free(s->b[q]);
    }
    free(s->buf);
    free(s);
}

void orc_set_omega(orc_lattice *s, double omega) { s->omega = omega; }

void orc_set_populations(orc_lattice *s, const double *f) { DISPATCH(set_pops, s, f); }
void orc_get_populations(const orc_lattice *s, double *f) { DISPATCH(get_pops, s, f); }
void orc_get_halo_array(const orc_lattice *s, int q, double *out) { DISPATCH(get_halo, s, q, out); }
void orc_init_equilibrium(orc_lattice *s, const double *u) { DISPATCH(init_equilibrium, s, u); }
void orc_bcond_step3(orc_lattice *s) { DISPATCH(bcond_step3, s); }

static void swap_ab(orc_lattice *s)
{
    /* SRC/col_MC.F90:365-421: a01..a18 <-> b01..b18, a19 stays */
    for (int q = 0; q < 18; q++) {
        void *t = s->a[q];
        s->a[q] = s->b[q];
        s->b[q] = t;
    }
}

void orc_col_MC(orc_lattice *s)
{
    DISPATCH(col_region, s, 0, 0, 0);
    swap_ab(s);
}

void orc_col_MC_masked(orc_lattice *s, int opt, int border)
{
    DISPATCH(col_region, s, 1, opt, border);
    if (opt == 1) swap_ab(s); /* SRC/col_MC_masked.F90:369 */
}

void orc_step(orc_lattice *s, int nsteps)
{
    for (int t = 0; t < nsteps; t++) {
        orc_bcond_step3(s); /* SRC/bgk3d_HIT.F90:131 */
        orc_col_MC(s);      /* SRC/bgk3d_HIT.F90:140 */
    }
}

void orc_diagno(const orc_lattice *s, int ztot_bug, int acc_mykind, double out[5])
{
    DISPATCH(diagno, s, ztot_bug, acc_mykind, out);
}

void orc_dissipation(const orc_lattice *s, double svisc, double out[2])
{
    DISPATCH(dissipation, s, svisc, out);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads for the following calls (timing the CPU baseline on 1 thread) */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ---- discrete adjoint (not in the reference; FP64) -------------------
 * Forward map per step: g_k(x) = f_k(x - c_k) (pull), f'_i(x) = C_i(g(x))
 * with C_i = g_i - omega (g_i - e_i(g)).  This restatement applies J^T with
 * the generic table form of de_i/dg_k (independent of the GPU's moment
 * form) and stores every state.                                          */
static const int CXg[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
static const int CYg[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
static const int CZg[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
static double Wg(int q)
{
    int s2 = CXg[q] * CXg[q] + CYg[q] * CYg[q] + CZg[q] * CZg[q];
    return s2 == 0 ? 1.0 / 3.0 : (s2 == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
}

static double ke_of_state(const double *f, long N)
{
    double J = 0;
    for (long x = 0; x < N; x++) {
        double rho = 1, jx = 0, jy = 0, jz = 0;
        for (int q = 0; q < 19; q++) {
            double v = f[q * N + x];
            rho += v;
            jx += CXg[q] * v;
            jy += CYg[q] * v;
            jz += CZg[q] * v;
        }
        double ux = jx / rho, uy = jy / rho, uz = jz / rho;
        J += 0.5 * (ux * ux + uy * uy + uz * uz);
    }
    return J;
}

double orc_forward_ke(int l, int m, int n, double omega, int nsteps, const double *u0)
{
    long N = (long)l * m * n;
    orc_lattice *s = orc_create(l, m, n, ORC_F64, omega, 1);
    double *f = (double *)malloc(sizeof(double) * 19 * N);
    orc_init_equilibrium(s, u0);
    orc_step(s, nsteps);
    orc_get_populations(s, f);
    double J = ke_of_state(f, N);
    free(f);
    orc_destroy(s);
    return J;
}

double orc_adjoint_ke(int l, int m, int n, double omega, int nsteps, const double *u0,
                      double *grad)
{
    long N = (long)l * m * n;
    orc_lattice *s = orc_create(l, m, n, ORC_F64, omega, 1);
    double *states = (double *)malloc(sizeof(double) * 19 * N * (size_t)(nsteps + 1));
    double *lam = (double *)malloc(sizeof(double) * 19 * N);
    double *lnew = (double *)malloc(sizeof(double) * 19 * N);
    orc_init_equilibrium(s, u0);
    orc_get_populations(s, states);
    for (int t = 0; t < nsteps; t++) {
        orc_step(s, 1);
        orc_get_populations(s, states + (size_t)(t + 1) * 19 * N);
    }
    const double *fT = states + (size_t)nsteps * 19 * N;
    double J = ke_of_state(fT, N);
    /* lambda^T_i = dJ/df_i = (u.c_i - |u|^2) / rho */
    for (long x = 0; x < N; x++) {
        double rho = 1, jx = 0, jy = 0, jz = 0;
        for (int q = 0; q < 19; q++) {
            double v = fT[q * N + x];
            rho += v;
            jx += CXg[q] * v;
            jy += CYg[q] * v;
            jz += CZg[q] * v;
        }
        double ux = jx / rho, uy = jy / rho, uz = jz / rho;
        double usq = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < 19; q++)
            lam[q * N + x] = (ux * CXg[q] + uy * CYg[q] + uz * CZg[q] - usq) / rho;
    }
    for (int t = nsteps - 1; t >= 0; t--) {
        const double *f = states + (size_t)t * 19 * N;
        for (int k = 0; k < n; k++)
            for (int j = 0; j < m; j++)
                for (int i = 0; i < l; i++) {
                    long x = i + (long)l * (j + (long)m * k);
                    double g[19];
                    long src[19];
                    for (int q = 0; q < 19; q++) {
                        int si = (i - CXg[q] + l) % l, sj = (j - CYg[q] + m) % m,
                            sk = (k - CZg[q] + n) % n;
                        src[q] = si + (long)l * (sj + (long)m * sk);
                        g[q] = f[q * N + src[q]];
                    }
                    double rho = 1, jx = 0, jy = 0, jz = 0;
                    for (int q = 0; q < 19; q++) {
                        rho += g[q];
                        jx += CXg[q] * g[q];
                        jy += CYg[q] * g[q];
                        jz += CZg[q] * g[q];
                    }
                    double ux = jx / rho, uy = jy / rho, uz = jz / rho;
                    double usq = ux * ux + uy * uy + uz * uz;
                    for (int kk = 0; kk < 19; kk++) {
                        double mu = (1.0 - omega) * lam[kk * N + x];
                        double ukc = ux * CXg[kk] + uy * CYg[kk] + uz * CZg[kk];
                        for (int ii = 0; ii < 19; ii++) {
                            double cic = CXg[ii] * CXg[kk] + CYg[ii] * CYg[kk] + CZg[ii] * CZg[kk];
                            double ciu = ux * CXg[ii] + uy * CYg[ii] + uz * CZg[ii];
                            double de = Wg(ii) * (1.0 + 3.0 * cic + 9.0 * ciu * cic -
                                                  4.5 * ciu * ciu - 3.0 * ukc + 1.5 * usq);
                            mu += omega * lam[ii * N + x] * de;
                        }
                        /* push: g_k(x) came from f_k(x - c_k) */
                        lnew[kk * N + src[kk]] = mu;
                    }
                }
        double *tmp = lam;
        lam = lnew;
        lnew = tmp;
    }
    /* df^0_i/du0 = w_i (3 c_i + 9 (c_i.u0) c_i - 3 u0) */
    for (long x = 0; x < N; x++) {
        double ux = u0[x], uy = u0[N + x], uz = u0[2 * N + x];
        double gx = 0, gy = 0, gz = 0;
        for (int q = 0; q < 19; q++) {
            double cu = CXg[q] * ux + CYg[q] * uy + CZg[q] * uz;
            double L = lam[q * N + x] * Wg(q);
            gx += L * (3.0 * CXg[q] + 9.0 * cu * CXg[q] - 3.0 * ux);
            gy += L * (3.0 * CYg[q] + 9.0 * cu * CYg[q] - 3.0 * uy);
            gz += L * (3.0 * CZg[q] + 9.0 * cu * CZg[q] - 3.0 * uz);
        }
        grad[x] = gx;
        grad[N + x] = gy;
        grad[2 * N + x] = gz;
    }
    free(states);
    free(lam);
    free(lnew);
    orc_destroy(s);
    return J;
}
