! bgk3d_HIT.F90 -- Fortran host driver of the MI355X BGK3D_HIT hot path.
!
! Keeps the reference's program structure and bgk.input semantics
! (SRC/bgk3d_HIT.F90:46-167, SRC/input.F90:41-82) and calls the HIP library
! through bgk_capi.  The hot-path subroutines keep the reference's names
! and call order: boundaries (SRC/boundaries.F90:23), col_MC
! (SRC/col_MC.F90:34), diagnostic -> diagno / dissipation
! (SRC/diagnostic.F90:37-51); output files keep the reference's formats
! (SRC/diagno.F90:132-152, SRC/dissipation.F90:265-278,
! SRC/finalize.F90:75-97) and add full-precision side files.
!
! Namelist &parameters: the reference keys svisc, u0, itfin, ivtim, isignal,
! itsave, icheck, irestart, lx, ly, lz, border, ipad, jpad, proc_x, proc_y,
! proc_z, cteS, plus
!   init    = 'tgv' (reference Taylor-Green vortex, SRC/init.F90:83-141) or
!             'hit' (seeded isotropic turbulence: k0, urms, seed)
!   storage = 'f32' (reference default build), 'f64' (-DDOUBLE_P),
!             'f32mix' (-DMIXEDPRECISION), 'f16', 'bf16'
!   diag    = 'refcompat' (reproduce the reference's diagnostics, stale
!             halos and the ztot column) or 'consistent'
!   omega   = relaxation frequency; when > 0 it replaces 2/(6 svisc + 1)
module bgk_host
  use iso_c_binding
  use bgk_capi
  implicit none
  integer, parameter :: sp = kind(1.0), dp = kind(1.0d0)

  ! reference namelist keys (SRC/input.F90:41-47)
  real(dp) :: svisc = 0.0_dp, u0 = 0.0_dp, cteS = 0.0_dp
  integer :: itfin = 100000, ivtim = 1000000000, isignal = 1000000000, itsave = 0
  integer :: icheck = 100, irestart = 0
  integer :: lx = 0, ly = 0, lz = 0, border = 10, ipad = 0, jpad = 0
  integer :: proc_x = 1, proc_y = 1, proc_z = 1
  ! extensions
  character(len=16) :: init = 'tgv', storage = 'f32', diag = 'refcompat'
  real(dp) :: k0 = 4.0_dp, urms = 0.0_dp, omega = 0.0_dp
  integer(c_int64_t) :: seed = 1234_c_int64_t

  integer :: l, m, n, nprocs = 1
  integer(c_int) :: kind_c, diag_c
  type(c_ptr) :: ctx = c_null_ptr
  ! section timers (SRC/mod_tools.F90:11-54)
  real(dp) :: time_init = 0, time_loop = 0, time_coll = 0, time_dg = 0, time_mp = 0
  integer(8) :: count_rate
end module bgk_host

program bgk_HIT
  use bgk_host
  implicit none
  integer :: itime, itstart
  integer(8) :: c0, c1, e0, e1
  character(len=256) :: deck

  itstart = 1   ! SRC/bgk3d_HIT.F90:63
  deck = 'bgk.input'
  if (command_argument_count() >= 1) call get_command_argument(1, deck)
  call system_clock(c0, count_rate)
  call input(trim(deck))
  call hencol
  call outdat
  call alloca
  call init_fields
  call system_clock(c1)
  time_init = real(c1 - c0, dp) / real(count_rate, dp)

  call system_clock(e0)
  do itime = itstart, itfin                      ! SRC/bgk3d_HIT.F90:127-146
     call boundaries(itime)
     call col_MC(itime)
     call diagnostic(itime)
  end do
  call bgk_check(bgk_synchronize(ctx), 'bgk_synchronize')
  call system_clock(e1)
  time_loop = real(e1 - e0, dp) / real(count_rate, dp)
  call finalize(itstart)
end program bgk_HIT

! SRC/input.F90:41-82
subroutine input(deck)
  use bgk_host
  implicit none
  character(len=*), intent(in) :: deck
  integer :: ios
  namelist /parameters/ svisc, u0, itfin, ivtim, isignal, itsave, icheck, irestart, &
       lx, ly, lz, border, ipad, jpad, proc_x, proc_y, proc_z, cteS, &
       init, storage, diag, k0, urms, seed, omega
  open(15, file=deck, status='old', iostat=ios)
  if (ios /= 0) then
     write(6, *) 'ERROR: cannot open ', deck
     error stop 1
  end if
  read(15, parameters)
  close(15)
  if (proc_x * proc_y * proc_z /= 1) then
     write(6, *) 'ERROR: one GPU per run: proc_x*proc_y*proc_z must be 1', proc_x, proc_y, proc_z
     error stop 1
  end if
  l = lx / proc_x
  m = ly / proc_y
  n = lz / proc_z
  if (l * proc_x /= lx .or. m * proc_y /= ly .or. n * proc_z /= lz .or. l < 1) then
     write(6, *) 'ERROR: global and local size not valid!!', lx, ly, lz
     error stop 1
  end if
  select case (trim(storage))
  case ('f64'); kind_c = BGK_STORE_F64
  case ('f32'); kind_c = BGK_STORE_F32
  case ('f32mix'); kind_c = BGK_STORE_F32_MIXED
  case ('f16'); kind_c = BGK_STORE_F16
  case ('bf16'); kind_c = BGK_STORE_BF16
  case default
     write(6, *) 'ERROR: unknown storage ', trim(storage)
     error stop 1
  end select
  select case (trim(diag))
  case ('refcompat'); diag_c = BGK_DIAG_REFCOMPAT
  case ('consistent'); diag_c = BGK_DIAG_CONSISTENT
  case default
     write(6, *) 'ERROR: unknown diag mode ', trim(diag)
     error stop 1
  end select
  if (trim(init) /= 'tgv' .and. trim(init) /= 'hit') then
     write(6, *) 'ERROR: unknown init ', trim(init)
     error stop 1
  end if
  if (urms <= 0.0_dp) urms = u0
  if (u0 <= 0.0_dp) u0 = urms
end subroutine input

! SRC/hencol.F90:31 in the compute kind of the chosen build
subroutine hencol
  use bgk_host
  implicit none
  real(sp) :: svisc_s
  if (omega > 0.0_dp) then
     svisc = (2.0_dp / omega - 1.0_dp) / 6.0_dp
     return
  end if
  if (kind_c == BGK_STORE_F64 .or. kind_c == BGK_STORE_F32_MIXED) then
     omega = (1.0_dp + 1.0_dp) / ((6.0_dp * svisc) + 1.0_dp)
  else
     svisc_s = real(svisc, sp)
     omega = real((1.0_sp + 1.0_sp) / ((6.0_sp * svisc_s) + 1.0_sp), dp)
  end if
end subroutine hencol

subroutine outdat
  use bgk_host
  implicit none
  open(16, file='bgk.log', status='unknown')
  write(16, *) 'INFO: BGK3D_HIT hot path on MI355X (HIP library version ', bgk_version(), ')'
  write(16, *) 'INFO: lattice ', lx, ly, lz, ' storage ', trim(storage), ' init ', trim(init)
  write(16, *) 'INFO: svisc, omega, u0 ', svisc, omega, u0
  write(16, *) 'INFO: itfin, icheck, isignal ', itfin, icheck, isignal
  write(6, *) 'INFO: lattice ', lx, ly, lz, ' storage ', trim(storage), ' omega ', omega
end subroutine outdat

! SRC/alloca.F90:178-216: the device lattice
subroutine alloca
  use bgk_host
  implicit none
  integer(c_int) :: flags
  flags = 0
  if (diag_c == BGK_DIAG_REFCOMPAT) flags = ior(flags, BGK_FLAG_REFCOMPAT_HALOS)
  call bgk_check(bgk_create(int(l, c_int), int(m, c_int), int(n, c_int), kind_c, &
       BGK_LAYOUT_PERIODIC, flags, real(omega, c_double), ctx), 'bgk_create')
  call bgk_check(bgk_set_border(ctx, int(border, c_int)), 'bgk_set_border')
end subroutine alloca

! SRC/init.F90:83-141 (tgv) or the seeded HIT field, rho = 1
subroutine init_fields
  use bgk_host
  implicit none
  real(dp), allocatable :: f(:, :, :, :)
  ! SRC/init.F90:42: the reference's pi is a default-real literal
  real(dp), parameter :: pi = real(3.141592653589793238462643383279_sp, dp)
  real(dp), parameter :: p0 = 1.0_dp / 3.0_dp, p1 = 1.0_dp / 18.0_dp, p2 = 1.0_dp / 36.0_dp
  real(dp), parameter :: rf = 3.0_dp, qf = 1.5_dp
  integer, parameter :: cxv(19) = (/ 1, 1, 1, 1, 1, 0, 0, 0, 0,-1,-1,-1,-1,-1, 0, 0, 0, 0, 0/)
  integer, parameter :: cyv(19) = (/-1, 0, 1, 0, 0, 0, 1, 1, 1,-1, 0, 1, 0, 0, 0,-1,-1,-1, 0/)
  integer, parameter :: czv(19) = (/ 0,-1, 0, 1, 0, 1, 1, 0,-1, 0,-1, 0, 1, 0,-1,-1, 0, 1, 0/)
  real(dp) :: x, y, z, xj, yj, zj, crho, cvsq, cu, w
  integer :: i, j, k, q

  if (trim(init) == 'hit') then
     call bgk_check(bgk_init_hit(ctx, real(k0, c_double), real(urms, c_double), seed), &
          'bgk_init_hit')
     return
  end if
  allocate(f(l, m, n, 19))
  do k = 1, n
     z = real((real(k, sp) - 0.5_sp) / real(lz, sp), dp)
     do j = 1, m
        y = real((real(j, sp) - 0.5_sp) / real(ly, sp), dp)
        do i = 1, l
           x = real((real(i, sp) - 0.5_sp) / real(lx, sp), dp)
           xj = u0 * sin(2 * pi * x) * cos(2 * pi * y) * cos(2 * pi * z)
           yj = -u0 * cos(2 * pi * x) * sin(2 * pi * y) * cos(2 * pi * z)
           zj = 0.0_dp
           crho = 1.0_dp + ((u0 * u0) / (16.0_dp * 3.0_dp)) * &
                (cos(4 * pi * x) + cos(4 * pi * y)) * (cos(4 * pi * z) + 2)
           cvsq = xj * xj + yj * yj + zj * zj
           do q = 1, 19
              cu = cxv(q) * xj + cyv(q) * yj + czv(q) * zj
              if (q == 19) then
                 w = p0
              else if (abs(cxv(q)) + abs(cyv(q)) + abs(czv(q)) == 1) then
                 w = p1
              else
                 w = p2
              end if
              ! shifted storage: crho * w * (0 + c.u terms)
              f(i, j, k, q) = crho * w * (rf * cu + qf * (3.0_dp * cu * cu - cvsq))
           end do
        end do
     end do
  end do
  call bgk_check(bgk_set_populations(ctx, f), 'bgk_set_populations')
  deallocate(f)
end subroutine init_fields

! SRC/boundaries.F90:23 -> SRC/bcond.comm.step3.F90
subroutine boundaries(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  integer(8) :: c0, c1
  call system_clock(c0)
  call bgk_check(bgk_boundaries(ctx), 'bgk_boundaries')
  call system_clock(c1)
  time_mp = time_mp + real(c1 - c0, dp) / real(count_rate, dp)
  if (itime < 0) write(6, *) itime
end subroutine boundaries

! SRC/col_MC.F90:34-444 (the swap is inside the library)
subroutine col_MC(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  integer(8) :: c0, c1
  call system_clock(c0)
  call bgk_check(bgk_col_MC(ctx), 'bgk_col_MC')
  call bgk_check(bgk_synchronize(ctx), 'bgk_synchronize')
  call system_clock(c1)
  time_coll = time_coll + real(c1 - c0, dp) / real(count_rate, dp)
  if (itime < 0) write(6, *) itime
end subroutine col_MC

! SRC/diagnostic.F90:37-73
subroutine diagnostic(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  integer(8) :: c0, c1
  if (mod(itime, icheck) == 0) then
     call system_clock(c0)
     call diagno(itime)
     call dissipation(itime)
     call system_clock(c1)
     time_dg = time_dg + real(c1 - c0, dp) / real(count_rate, dp)
  end if
  if (mod(itime, isignal) == 0) then
     write(6, '(a,i8,a,es12.4,a)') ' Iteration ', itime, '  mean step time ', &
          time_coll / real(itime, dp), ' s'
  end if
end subroutine diagnostic

! SRC/diagno.F90:28-154
subroutine diagno(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  real(c_double) :: d(5)
  logical, save :: first = .true.
  call bgk_check(bgk_diagno(ctx, diag_c, d), 'bgk_diagno')
  if (first) then
     open(70, file='diagno.dat', status='unknown')
     open(72, file='diagno.full.dat', status='unknown')
     first = .false.
  end if
  ! d = mean rho, mean ux, mean uy, mean uz, mean |u|^2/2 (loctot(1..5))
  write(16, 1001) itime
  write(16, 1002) d(1)
  write(16, 1003) d(2), d(3), d(4), d(5)
  flush(16)
  write(70, 1004) itime, d(2), d(3), d(4), d(1), d(5)
  flush(70)
  write(72, '(i8,5(1x,es24.16))') itime, d(2), d(3), d(4), d(1), d(5)
  flush(72)
1001 format(" Timestep ", i8)
1002 format("       mean rho ", 1(e14.6, 1x))
1003 format("       mean vel ", 4(e14.6, 1x))
1004 format(i8, 5(e14.6, 1x))
end subroutine diagno

! SRC/dissipation.F90:22-280
subroutine dissipation(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  real(c_double) :: s(2)
  real(sp) :: dx_s
  real(dp) :: dt, tke
  logical, save :: first = .true.
  call bgk_check(bgk_dissipation(ctx, real(svisc, c_double), diag_c, s), 'bgk_dissipation')
  if (first) then
     open(71, file='dissipation.dat', status='unknown')
     open(73, file='dissipation.full.dat', status='unknown')
     first = .false.
  end if
  ! SRC/dissipation.F90:65-66 evaluates dx in default real
  dx_s = (2.0_sp * 3.141592_sp) / real(lx, sp)
  dt = real(dx_s, dp) * u0
  tke = s(2) / real(nprocs, dp)
  write(71, 1004) itime, real(itime, dp) * dt, tke / real(l, dp) / real(m, dp) / real(n, dp) / &
       (u0 * u0), s(1)
  flush(71)
  write(73, '(i8,4(1x,es24.16))') itime, real(itime, dp) * dt, s(2), s(1)
  flush(73)
1004 format(i8, 4(e14.6, 1x))
end subroutine dissipation

! SRC/finalize.F90:44-103
subroutine finalize(itstart)
  use bgk_host
  implicit none
  integer, intent(in) :: itstart
  real(dp) :: glups
  real(c_double) :: t(8)
  ! SRC/finalize.F90:83-85 (itfin - itstart steps in the numerator)
  glups = real(lx, dp) * real(ly, dp) * real(lz, dp) * real(itfin - itstart, dp) / &
       (time_loop * 1000 * 1000 * 1000)
  open(69, file='bgk.perf', status='unknown')
  write(69, 9999)
  write(69, 1100) time_init
  write(69, 1101) time_loop
  write(69, 1102) time_coll
  write(69, 1103) time_dg
  write(69, 1115) time_mp
  write(69, 9999)
  write(69, 3000) glups
  write(69, 9999)
  close(69)
  write(6, 9999)
  write(6, 2000) time_loop
  write(6, 3000) glups
  write(6, 9999)
  call bgk_check(bgk_timing(ctx, t, 1_c_int), 'bgk_timing')
  call bgk_check(bgk_destroy(ctx), 'bgk_destroy')
  close(16)
9999 format(" #--------------------------------")
1100 format(" # init   time", 1(e14.6, 1x))
1101 format(" # loop   time", 1(e14.6, 1x))
1102 format(" # comp.  time", 1(e14.6, 1x))
1103 format(" # diag.  time", 1(e14.6, 1x))
1115 format(" # MPI    time", 1(e14.6, 1x))
2000 format(" # loop time (s) ", e14.6)
3000 format(" # GLUPs ", e14.6)
end subroutine finalize
