/*
 * bgk_oracle.h -- CPU oracle for the BGK3D_HIT time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference Fortran (gamati01/BGK3D_HIT, SRC/) used as the parity checker
 * and as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never does.
 *
 * The oracle deliberately mirrors the reference's data model rather than the
 * GPU design: 19 separate arrays per population copy with 1-cell halos
 * (SRC/alloca.F90:178-216, column-major, x fastest), a pointer swap after
 * each collide (SRC/col_MC.F90:365-421), the in-place rest population
 * (SRC/col_MC.F90:351) and halo filling by the six pack / self-exchange /
 * unpack sweeps of SRC/bcond.comm.step3.F90:79-297.  The GPU library instead
 * stores halo-free SoA planes and wraps periodic indices in the kernel, so
 * agreement between the two is a real check of the streaming indices.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   orc_col_MC         pinned: omega=0 exact periodic shift, omega=1 gives the
 *                      generic-table equilibrium of pulled moments, mass and
 *                      momentum conservation, uniform-equilibrium fixed
 *                      point, shear-wave decay rate nu=(1/omega-1/2)/3, and an
 *                      independent dense numpy restatement on 4^3..8^3.
 *   orc_bcond_step3    pinned: ghosts equal periodic images (numpy np.roll).
 *   orc_diagno         pinned: uniform state closed form, numpy restatement.
 *   orc_dissipation    pinned: Pi^neq=0 on uniform equilibrium, numpy
 *                      generic-table restatement of Pi^neq = sum c c (f-feq).
 *   orc_init_equilibrium pinned: moments of the result (rho=1, j=u to O(u^3)).
 *   orc_adjoint_ke     pinned: central finite differences of orc_* forward.
 * The reference's OpenACC reduction order and MPI_Reduce order are not
 * reproducible by any CPU code: parity unpinned at those two boundaries.
 */
#ifndef BGK_ORACLE_H
#define BGK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* storage/compute kinds (reference SRC/precision.F90:13-32 for 0..2) */
enum {
    ORC_F64 = 0,    /* DOUBLE_P: store dp, compute dp                 */
    ORC_F32 = 1,    /* default : store sp, compute sp                 */
    ORC_F32MIX = 2, /* MIXEDPRECISION: store sp, compute dp           */
    ORC_F16 = 3,    /* store IEEE half (shifted), compute sp          */
    ORC_BF16 = 4    /* store bfloat16 (shifted), compute sp           */
};

typedef struct orc_lattice orc_lattice;

orc_lattice *orc_create(int l, int m, int n, int kind, double omega, int shifted);
void orc_destroy(orc_lattice *s);
void orc_set_omega(orc_lattice *s, double omega);

/* f: 19 planes of l*m*n doubles, x fastest, population 1..19 in order.
 * Values are what the reference stores (f_i - w_i when shifted).        */
void orc_set_populations(orc_lattice *s, const double *f);
void orc_get_populations(const orc_lattice *s, double *f);
/* full halo'd current array of one population (0-based q), (l+2)(m+2)(n+2) */
void orc_get_halo_array(const orc_lattice *s, int q, double *out);

/* SRC/init.F90:83-141 equilibrium formula with crho=1 from a velocity
 * field u (3 planes of l*m*n doubles: ux, uy, uz); halos by periodic image */
void orc_init_equilibrium(orc_lattice *s, const double *u);

void orc_bcond_step3(orc_lattice *s);                 /* SRC/bcond.comm.step3.F90:29-316 */
void orc_col_MC(orc_lattice *s);                      /* SRC/col_MC.F90:34-444           */
void orc_col_MC_masked(orc_lattice *s, int opt, int border); /* SRC/col_MC_masked.F90    */
void orc_step(orc_lattice *s, int nsteps);            /* SRC/bgk3d_HIT.F90:127-146 order */

/* SRC/diagno.F90:73-121: out = mean rho, mean ux, mean uy, mean uz, mean |u|^2/2.
 * ztot_bug=1 reproduces SRC/diagno.F90:110 (ztot = ytot + zj).
 * acc_mykind=1 accumulates in the compute kind as the reference's serial loop
 * does; 0 accumulates in double (what the GPU library does).            */
void orc_diagno(const orc_lattice *s, int ztot_bug, int acc_mykind, double out[5]);
/* SRC/dissipation.F90:70-257: out[0] = sum 2 nu P, out[1] = sum |u|^2/2 over
 * pulled moments, reading the current halos as they stand (the reference's
 * stale-halo behaviour).  Call orc_bcond_step3 first for fresh halos.      */
void orc_dissipation(const orc_lattice *s, double svisc, double out[2]);

/* Discrete adjoint (FP64 kinds only): gradient of J = sum_x |u(x)|^2/2 of the
 * state after nsteps steps w.r.t. the initial velocity u0 used by
 * orc_init_equilibrium.  Stores every state (small lattices only).        */
double orc_adjoint_ke(int l, int m, int n, double omega, int nsteps,
                      const double *u0, double *grad);
/* J alone (same definition), FP64 forward from u0 */
double orc_forward_ke(int l, int m, int n, double omega, int nsteps, const double *u0);

/* storage conversions used by the F16/BF16 kinds (round to nearest even) */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float f);
float orc_bf16_to_f32(uint16_t h);

int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
