/*
 * bgk3d_hit.h -- C-ABI of the MI355X (gfx950) BGK3D_HIT time-step hot path.
 *
 * The reference (gamati01/BGK3D_HIT, Fortran + OpenACC) has no FFI: its hot
 * path is a set of external subroutines that share module globals
 * (SRC/storage.F90).  Each entry point below replaces one of them; the
 * Fortran binding is bgk3d_hit_amd/fortran/bgk_capi.F90 and the Python
 * binding is bgk3d_hit_amd/_lib.py (ctypes).
 *
 * Conventions
 *   - every function returns an int status (BGK_OK == 0); the message of the
 *     last failure on the calling thread is read with bgk_last_error.
 *   - host arrays are plain C arrays of doubles: populations are 19 planes of
 *     lx*ly*lz values (x fastest, then y, then z; reference population order
 *     01..19), velocity fields are 3 planes (ux, uy, uz) of lx*ly*lz values.
 *     Population values are what the reference stores: f_i - w_i unless the
 *     context was created with BGK_FLAG_NOSHIFT.
 *   - device state stays resident; the reference's a/b pointer swap
 *     (SRC/col_MC.F90:365-421) becomes a ping-pong index inside the context.
 *   - a context is used by one host thread at a time.
 */
#ifndef BGK3D_HIT_H
#define BGK3D_HIT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bgk_ctx bgk_ctx;

/* status codes */
#define BGK_OK 0
#define BGK_EINVAL 1     /* invalid argument (size, kind, omega, ...)          */
#define BGK_EHIP 2       /* HIP runtime error                                  */
#define BGK_ENOMEM 3     /* device or host allocation failed                   */
#define BGK_ESTATE 4     /* call not valid in the context's current state      */
#define BGK_ENONFINITE 5 /* a reduction found a NaN/Inf (diverged populations) */
#define BGK_ENODEV 6     /* no usable GPU                                      */

/* storage kinds (reference SRC/precision.F90:13-32 for the first three) */
#define BGK_STORE_F64 0       /* DOUBLE_P: store f64, compute f64                 */
#define BGK_STORE_F32 1       /* default : store f32, compute f32                 */
#define BGK_STORE_F32_MIXED 2 /* MIXEDPRECISION: store f32, compute f64           */
#define BGK_STORE_F16 3       /* store shifted f_i - w_i as IEEE half, compute f32 */
#define BGK_STORE_BF16 4      /* store shifted f_i - w_i as bfloat16, compute f32  */

/* layouts */
#define BGK_LAYOUT_PERIODIC 0 /* triply periodic, wrap in the index math          */
#define BGK_LAYOUT_SLAB_Z 1   /* one rank's slab: x,y periodic, 1-site ghost planes
                                 on -z/+z filled by pack / exchange / unpack       */

/* creation flags */
#define BGK_FLAG_NOSHIFT 1         /* reference -DNOSHIFT: store f_i, not f_i - w_i   */
#define BGK_FLAG_REFCOMPAT_HALOS 2 /* keep the stale-halo images the reference's
                                      dissipation reads (periodic layout)          */
#define BGK_FLAG_TIMING 4          /* record HIP events around every kernel launch  */
#define BGK_FLAG_SINGLE_STEP 8     /* bgk_step: one pass over the populations per
                                      time step (no two-step fusion)               */
#define BGK_FLAG_TWO_STEP_ANY_SIZE 16 /* bgk_step: fuse two steps per pass also on
                                      lattices too small for it to pay (tests)     */
#define BGK_FLAG_WIDE_OFFSETS 32   /* 64-bit in-plane offsets in every kernel, the
                                      path of x-y planes above 4 GiB (tests)       */
#define BGK_FLAG_NO_THREE_STEP 64  /* bgk_step: at most two time steps per pass
                                      (no three-step fusion)                       */

/* diagnostic modes */
#define BGK_DIAG_CONSISTENT 0 /* gathers read the current state (fresh periodic images) */
#define BGK_DIAG_REFCOMPAT 1  /* reproduce the reference: dissipation reads the halo
                                 images left by the exchange two steps earlier
                                 (SRC/dissipation.F90:76-94), diagno's 4th column is
                                 ytot + zj(l,m,n) (SRC/diagno.F90:110)           */

/* library version, e.g. 100 for 0.1.0 */
int bgk_version(void);

/* copy the calling thread's last error message into buf (NUL-terminated) */
int bgk_last_error(char *buf, int len);

/* number of visible GPUs (0 when none); does not create a context */
int bgk_device_count(int *count);

/* Allocate the device lattice: replaces alloca (SRC/alloca.F90:178-216) and
 * the relaxation setup of hencol (SRC/hencol.F90:31, omega passed directly).
 * Populations start at zero (the rest state when shifted).                 */
int bgk_create(int lx, int ly, int lz, int storage, int layout, int flags, double omega,
               bgk_ctx **out);
int bgk_destroy(bgk_ctx *ctx);

/* change the relaxation frequency (SRC/hencol.F90:31) */
int bgk_set_omega(bgk_ctx *ctx, double omega);

/* width of the boundary region used by bgk_col_MC_masked in the periodic
 * layout (reference namelist key `border`, SRC/input.F90:53, mask built in
 * SRC/init.F90:44-62); default 10                                          */
int bgk_set_border(bgk_ctx *ctx, int border);

/* info[0..7] = lx, ly, lz, storage, layout, flags, device bytes, steps done */
int bgk_info(bgk_ctx *ctx, int64_t info[8]);

/* upload / download the current populations (19 planes, see conventions);
 * replaces the host-array view the reference's init and I/O use
 * (SRC/init.F90:119-137, SRC/bgk3d_HIT.F90:111-115 data region)          */
int bgk_set_populations(bgk_ctx *ctx, const double *f);
int bgk_get_populations(bgk_ctx *ctx, double *f);

/* population initialisation from a velocity field with rho = 1: the
 * equilibrium formula of init (SRC/init.F90:97-137, crho = 1)             */
int bgk_init_equilibrium(bgk_ctx *ctx, const double *u);

/* Seeded homogeneous-isotropic-turbulence velocity field (not in the
 * reference, whose init is a Taylor-Green vortex, SRC/init.F90:83-92):
 * solenoidal random-phase Fourier field, E(k) ~ k^4 exp(-2 (k/k0)^2),
 * zero mean, rescaled so that sqrt(<|u|^2>/3) = urms.  Host computation into
 * u (3 planes).  bgk_init_hit synthesises the same field on the GPU and
 * applies bgk_init_equilibrium.                                            */
int bgk_hit_velocity(int lx, int ly, int lz, double k0, double urms, uint64_t seed, double *u);
int bgk_init_hit(bgk_ctx *ctx, double k0, double urms, uint64_t seed);

/* Halo path: replaces boundaries (SRC/boundaries.F90:23-65) and its default
 * bcond_comm_step3 (SRC/bcond.comm.step3.F90:29-316).  Slab layout: pack the
 * 5 crossing populations of each z face, exchange (periodic self-exchange
 * of the packed messages), unpack into the ghost planes.  Periodic layout:
 * nothing to exchange (wrap is in the index math); with
 * BGK_FLAG_REFCOMPAT_HALOS the face images are recorded as the reference's
 * ghosts would hold them.                                                   */
int bgk_boundaries(bgk_ctx *ctx);

/* fused pull stream + BGK collide over all sites, then swap: replaces col_MC
 * (SRC/col_MC.F90:34-444)                                                   */
int bgk_col_MC(bgk_ctx *ctx);

/* collide restricted to the bulk (opt = 0) or the boundary region (opt = 1),
 * swap only after opt = 1: replaces col_MC_masked (SRC/col_MC_masked.F90:34-451)
 * without a mask array.  The regions are the reference mask's
 * (SRC/init.F90:44-62) on both layouts: border = sites within `border`
 * (bgk_set_border) of a face of the box, bulk = the rest (none when
 * 2*border >= a side).  On the slab, call bgk_boundaries before opt = 1.  */
int bgk_col_MC_masked(bgk_ctx *ctx, int opt);

/* nsteps of the time loop body boundaries + col_MC (SRC/bgk3d_HIT.F90:127-146
 * without diagnostic).  Where the lattice is large enough (unless
 * BGK_FLAG_SINGLE_STEP): three steps per pass over the populations (not with
 * BGK_FLAG_NO_THREE_STEP; slabs of 12 planes and more), two steps per pass for a remainder of two (two such passes for a
 * remainder of one), single steps otherwise; bit-identical to nsteps calls
 * of bgk_col_MC, stale halo images and slab ghost planes included.  Slab
 * layout: the interior collide overlaps the halo pack/exchange/unpack and
 * boundary collide on a second HIP stream (the structure of
 * SRC/bcond.comm.step9.F90:349 and SRC/do_somethingGPU_overlap.F90:48-58);
 * with n steps per pass the boundary planes take the n steps through plane
 * buffers on that stream, with one exchange per step as before.            */
int bgk_step(bgk_ctx *ctx, int nsteps);

/* at-site moments of the current state: replaces diagno (SRC/diagno.F90:28-154).
 * out = mean rho, mean ux, mean uy, mean uz, mean |u|^2/2 (the reference's
 * loctot(1..5)), FP64 accumulation.  Returns BGK_ENONFINITE on NaN/Inf.     */
int bgk_diagno(bgk_ctx *ctx, int mode, double out[5]);

/* pulled-moment dissipation and kinetic energy of the current state:
 * replaces dissipation (SRC/dissipation.F90:22-280).  out[0] = sum over
 * sites of 2 svisc Pi^neq:Pi^neq, out[1] = sum over sites of |u|^2/2 (the
 * reference's loctot(1..2), before its output normalisation).              */
int bgk_dissipation(bgk_ctx *ctx, double svisc, int mode, double out[2]);

/* diagno and dissipation of the current state in one pass over the
 * populations: out[0..4] = bgk_diagno's out, out[5..6] = bgk_dissipation's,
 * bit-identical to the two calls.  Replaces the pair the reference's
 * diagnostic makes back to back (SRC/diagnostic.F90:47-49); no single
 * reference routine does both.                                              */
int bgk_diagnostics(bgk_ctx *ctx, double svisc, int mode, double out[7]);

/* nsteps time steps (as bgk_step), then bgk_diagnostics of the state they
 * leave in consistent mode (out as bgk_diagnostics').  Replaces the
 * reference's loop body at a diagnostic step (SRC/bgk3d_HIT.F90:127-146:
 * col_MC and boundaries, then diagnostic -> diagno + dissipation,
 * SRC/diagnostic.F90:47-49).  On the periodic lattice without halo images
 * (BGK_FLAG_REFCOMPAT_HALOS clear), where three-step passes run, and for
 * nsteps = 3k + 2 (k >= 0) or 3k + 1 (k >= 1), the last two steps and the
 * reductions run as one pass over the populations; otherwise it is
 * bgk_step + bgk_diagnostics.  The populations are bit-identical to
 * bgk_step's; the sums are the same per-site values summed in another
 * order (equal to bgk_diagnostics' to rounding of the FP64 sums).           */
int bgk_step_diagnostics(bgk_ctx *ctx, int nsteps, double svisc, double out[7]);

/* Allocation-free variant of bgk_diagnostics.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing, synchronizes nothing and needs no
 * context: the caller passes one copy of populations in device memory and a
 * device array out[7], and the pass is queued on `stream` (a hipStream_t,
 * NULL for the null stream).
 *   pop     : one population copy of the periodic layout, storage kind
 *             `storage`, as a context holds it: population q of site
 *             (x, y, z) is element x + 19*lx*y + 19*lx*ly*z + q*lx
 *             (rows of the 19 populations of each y line side by side)
 *   flags   : 0 (shifted storage f_i - w_i) or BGK_FLAG_NOSHIFT
 *   out     : device doubles, out[0..4] as bgk_diagno's (consistent mode),
 *             out[5..6] as bgk_dissipation's; bit-identical to
 *             bgk_diagnostics(ctx, svisc, BGK_DIAG_CONSISTENT, out) on the
 *             same state.  No NaN/Inf check (read out and test it).
 * The reduction partials live in a fixed device array of
 * 2*8*BGK_NOALLOC_MAX_PARTS + 16 doubles (4 MiB) per GPU, kept until the
 * process ends; a lattice needing more slots than BGK_NOALLOC_MAX_PARTS is
 * refused with BGK_EINVAL (one slot per reduction tile of an x-y plane and
 * z chunk: x-y planes of more than ~2^24 sites in FP32 storage, ~2^23 in
 * the other kinds, e.g. 4096 x 4096 FP32 fits at any lz).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (two streams, two host threads) race.                */
#define BGK_NOALLOC_MAX_PARTS 32768
int bgk_diagnostics_noalloc(const void *pop, int lx, int ly, int lz, int storage, int flags,
                            double svisc, double *out, void *stream);

/* Populations of the box [x0,x1) x [y0,y1) x [z0,z1) (0-based, interior
 * sites) into out: 19 planes of (x1-x0)(y1-y0)(z1-z0) doubles, x fastest.
 * Device-side gather replacing the host-array reads of probe_global
 * (SRC/probe_global.F90:55-95, one site) and of the VTK slice writers
 * (SRC/vtk_xy.binary.F90:57-99 and the xz/yz/3d variants, one plane).       */
int bgk_get_box(bgk_ctx *ctx, int x0, int x1, int y0, int y1, int z0, int z1, double *out);

/* Order-independent 64-bit hash of the raw stored bits of the current
 * interior populations (equal hashes <=> bit-identical states, up to hash
 * collisions).  No reference counterpart: the reference compares states
 * through its probe dumps (SRC/probe_global.F90:55-95); this lets a caller
 * check bit identity of two lattices (slab vs periodic, restart) without
 * downloading them.                                                         */
int bgk_checksum(bgk_ctx *ctx, uint64_t *out);

/* wait for all device work of the context */
int bgk_synchronize(bgk_ctx *ctx);

/* With BGK_FLAG_TIMING: t[0] = total ms in collide kernels, t[1] = collide
 * launches, t[2] = ms in reduction kernels, t[3] = reduction launches,
 * t[4] = ms in halo kernels/copies, t[5] = halo launches, t[6] = ms of
 * bgk_step wall time on the device (first to last event).  reset != 0
 * clears the counters after reading.  Replaces the reference's section
 * timers (SRC/mod_tools.F90:11-54, SRC/finalize.F90:44-97).                 */
int bgk_timing(bgk_ctx *ctx, double t[8], int reset);

/* With BGK_FLAG_TIMING: the collide launches of bgk_timing's t[0..1] split by
 * the time steps one launch advances: t[0], t[1] = ms and launches of one
 * step per launch (k_collide, also the slab's boundary planes), t[2], t[3]
 * of two-step passes (k_collide2), t[4], t[5] of three-step passes
 * (k_collide3).  reset != 0 clears every counter (as bgk_timing's does).
 * No counterpart in the reference (its timers are per section).             */
int bgk_timing_passes(bgk_ctx *ctx, double t[6], int reset);

/* With BGK_FLAG_TIMING: t[0], t[1] = ms and launches of the passes of two
 * time steps and the diagnostics of their result (bgk_step_diagnostics),
 * also counted in bgk_timing's collide launches t[0..1].  reset != 0 clears
 * every counter.  No counterpart in the reference.                          */
int bgk_timing_fused(bgk_ctx *ctx, double t[2], int reset);

/* With BGK_FLAG_TIMING: the kernel time of the discrete adjoint's launches
 * (bgk_adjoint_ke): t[0], t[1] = ms and launches of its taped forward passes
 * (the last segment in the forward run and every re-taped segment), t[2],
 * t[3] of its adjoint-step passes.  The re-taping runs on a second stream
 * beside the adjoint steps, so the two add up to more than the backward's
 * wall time.  reset != 0 clears every counter.  No counterpart in the
 * reference.                                                                */
int bgk_timing_adjoint(bgk_ctx *ctx, double t[4], int reset);

/* Discrete adjoint (not in the reference).  J = sum_x |u(x)|^2 / 2 of the
 * state reached after nsteps steps from the equilibrium of u0 (rho = 1);
 * grad = dJ/du0 (3 planes).  Checkpoints every ckpt_every steps with
 * recompute; levels = 1 or 2 (two-level checkpointing), 0 = tape every step.
 * The backward sweeps the segments on the context's main stream while the
 * next segment is re-taped from its checkpoint on the second stream.
 * stats[0] = peak device bytes of the adjoint's buffers, stats[1] = forward
 * ms, stats[2] = backward ms (wall), stats[3] = number of recomputed steps.
 * The context's populations are used as work space (their contents after
 * the call are unspecified).  u0 and grad are host arrays.                  */
int bgk_adjoint_ke(bgk_ctx *ctx, const double *u0, int nsteps, int ckpt_every, int levels,
                   double *J, double *grad, double stats[4]);
/* Same with u0 and grad in device memory (for framework bindings). */
int bgk_adjoint_ke_device(bgk_ctx *ctx, const double *u0_dev, int nsteps, int ckpt_every,
                          int levels, double *J, double *grad_dev, double stats[4]);
/* Forward functional only: J after nsteps from the equilibrium of u0 (host). */
int bgk_forward_ke(bgk_ctx *ctx, const double *u0, int nsteps, double *J);

#ifdef __cplusplus
}
#endif
#endif
