! bgk_capi.F90 -- iso_c_binding interface to libbgk3d_hit.so
! (declarations in include/bgk3d_hit.h).
!
! The reference has no FFI: its hot path is external subroutines sharing
! module globals (SRC/storage.F90).  This module is the binding a Fortran
! host uses instead; bgk3d_HIT.F90 wraps it in subroutines named like the
! reference's (boundaries, col_MC, diagno, dissipation).
module bgk_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: BGK_OK = 0, BGK_EINVAL = 1, BGK_EHIP = 2, &
       BGK_ENOMEM = 3, BGK_ESTATE = 4, BGK_ENONFINITE = 5, BGK_ENODEV = 6
  integer(c_int), parameter :: BGK_STORE_F64 = 0, BGK_STORE_F32 = 1, &
       BGK_STORE_F32_MIXED = 2, BGK_STORE_F16 = 3, BGK_STORE_BF16 = 4
  integer(c_int), parameter :: BGK_LAYOUT_PERIODIC = 0, BGK_LAYOUT_SLAB_Z = 1
  integer(c_int), parameter :: BGK_FLAG_NOSHIFT = 1, BGK_FLAG_REFCOMPAT_HALOS = 2, &
       BGK_FLAG_TIMING = 4, BGK_FLAG_SINGLE_STEP = 8, BGK_FLAG_TWO_STEP_ANY_SIZE = 16, &
       BGK_FLAG_WIDE_OFFSETS = 32, BGK_FLAG_NO_THREE_STEP = 64
  integer(c_int), parameter :: BGK_DIAG_CONSISTENT = 0, BGK_DIAG_REFCOMPAT = 1

  interface
     integer(c_int) function bgk_version() bind(c, name='bgk_version')
       import :: c_int
     end function bgk_version

     integer(c_int) function bgk_last_error(buf, len) bind(c, name='bgk_last_error')
       import :: c_int, c_char
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: len
     end function bgk_last_error

     integer(c_int) function bgk_device_count(count) bind(c, name='bgk_device_count')
       import :: c_int
       integer(c_int), intent(out) :: count
     end function bgk_device_count

     integer(c_int) function bgk_create(lx, ly, lz, storage, layout, flags, omega, ctx) &
          bind(c, name='bgk_create')
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: lx, ly, lz, storage, layout, flags
       real(c_double), value :: omega
       type(c_ptr), intent(out) :: ctx
     end function bgk_create

     integer(c_int) function bgk_destroy(ctx) bind(c, name='bgk_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function bgk_destroy

     integer(c_int) function bgk_set_omega(ctx, omega) bind(c, name='bgk_set_omega')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: omega
     end function bgk_set_omega

     integer(c_int) function bgk_set_border(ctx, border) bind(c, name='bgk_set_border')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: border
     end function bgk_set_border

     integer(c_int) function bgk_info(ctx, info) bind(c, name='bgk_info')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: info(8)
     end function bgk_info

     integer(c_int) function bgk_set_populations(ctx, f) bind(c, name='bgk_set_populations')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: f(*)
     end function bgk_set_populations

     integer(c_int) function bgk_get_populations(ctx, f) bind(c, name='bgk_get_populations')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: f(*)
     end function bgk_get_populations

     integer(c_int) function bgk_init_equilibrium(ctx, u) bind(c, name='bgk_init_equilibrium')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: u(*)
     end function bgk_init_equilibrium

     integer(c_int) function bgk_hit_velocity(lx, ly, lz, k0, urms, seed, u) &
          bind(c, name='bgk_hit_velocity')
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: lx, ly, lz
       real(c_double), value :: k0, urms
       integer(c_int64_t), value :: seed
       real(c_double), intent(out) :: u(*)
     end function bgk_hit_velocity

     integer(c_int) function bgk_init_hit(ctx, k0, urms, seed) bind(c, name='bgk_init_hit')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double), value :: k0, urms
       integer(c_int64_t), value :: seed
     end function bgk_init_hit

     integer(c_int) function bgk_boundaries(ctx) bind(c, name='bgk_boundaries')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function bgk_boundaries

     integer(c_int) function bgk_col_MC(ctx) bind(c, name='bgk_col_MC')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function bgk_col_MC

     integer(c_int) function bgk_col_MC_masked(ctx, opt) bind(c, name='bgk_col_MC_masked')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: opt
     end function bgk_col_MC_masked

     integer(c_int) function bgk_step(ctx, nsteps) bind(c, name='bgk_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function bgk_step

     integer(c_int) function bgk_diagno(ctx, mode, out) bind(c, name='bgk_diagno')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: mode
       real(c_double), intent(out) :: out(5)
     end function bgk_diagno

     integer(c_int) function bgk_dissipation(ctx, svisc, mode, out) &
          bind(c, name='bgk_dissipation')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: svisc
       integer(c_int), value :: mode
       real(c_double), intent(out) :: out(2)
     end function bgk_dissipation

     ! diagno and dissipation in one pass: out(1:5) as bgk_diagno, out(6:7)
     ! as bgk_dissipation (bit-identical to the two calls)
     integer(c_int) function bgk_diagnostics(ctx, svisc, mode, out) &
          bind(c, name='bgk_diagnostics')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: svisc
       integer(c_int), value :: mode
       real(c_double), intent(out) :: out(7)
     end function bgk_diagnostics

     ! nsteps time steps, then bgk_diagnostics (consistent mode) of the state
     ! reached; the last two steps and the sums share one pass where the
     ! lattice allows it
     integer(c_int) function bgk_step_diagnostics(ctx, nsteps, svisc, out) &
          bind(c, name='bgk_step_diagnostics')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       real(c_double), value :: svisc
       real(c_double), intent(out) :: out(7)
     end function bgk_step_diagnostics

     integer(c_int) function bgk_get_box(ctx, x0, x1, y0, y1, z0, z1, out) &
          bind(c, name='bgk_get_box')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: x0, x1, y0, y1, z0, z1
       real(c_double), intent(out) :: out(*)
     end function bgk_get_box

     integer(c_int) function bgk_checksum(ctx, out) bind(c, name='bgk_checksum')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: out
     end function bgk_checksum

     integer(c_int) function bgk_synchronize(ctx) bind(c, name='bgk_synchronize')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function bgk_synchronize

     integer(c_int) function bgk_timing(ctx, t, reset) bind(c, name='bgk_timing')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: t(8)
       integer(c_int), value :: reset
     end function bgk_timing

     integer(c_int) function bgk_timing_passes(ctx, t, reset) bind(c, name='bgk_timing_passes')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: t(6)
       integer(c_int), value :: reset
     end function bgk_timing_passes

     integer(c_int) function bgk_timing_fused(ctx, t, reset) bind(c, name='bgk_timing_fused')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: t(2)
       integer(c_int), value :: reset
     end function bgk_timing_fused

     integer(c_int) function bgk_adjoint_ke(ctx, u0, nsteps, ckpt_every, levels, J, grad, &
          stats) bind(c, name='bgk_adjoint_ke')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: u0(*)
       integer(c_int), value :: nsteps, ckpt_every, levels
       real(c_double), intent(out) :: J
       real(c_double), intent(out) :: grad(*)
       real(c_double), intent(out) :: stats(4)
     end function bgk_adjoint_ke

     integer(c_int) function bgk_forward_ke(ctx, u0, nsteps, J) bind(c, name='bgk_forward_ke')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: u0(*)
       integer(c_int), value :: nsteps
       real(c_double), intent(out) :: J
     end function bgk_forward_ke
  end interface

contains

  ! stop with the library's message when a call fails
  subroutine bgk_check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(1024)
    character(len=1024) :: msg
    integer :: i, rc
    if (status == BGK_OK) return
    rc = bgk_last_error(buf, 1024_c_int)
    msg = ' '
    do i = 1, 1024
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    end do
    write(6, '(a,a,a,i0,a,a)') 'ERROR: ', what, ' failed with status ', status, ': ', trim(msg)
    error stop 1
  end subroutine bgk_check

end module bgk_capi
