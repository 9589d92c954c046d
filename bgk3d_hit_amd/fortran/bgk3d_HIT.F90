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
! and gives meaning to the reference's unused restart keys: itsave > 0 writes
! restart.bgk every itsave steps; irestart = 1 starts from restart.bgk.
! Observability as the reference's diagnostic: probe_global (fort.88, the
! populations of one site, SRC/probe_global.F90:55-95) every icheck steps
! and the VTK velocity slices tec_{xy,xz,yz}.RRRRR.TTTTTTTT.vtk
! (SRC/vtk_xy.binary.F90) every ivtim steps and at step 0.
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
  ! section timers (SRC/mod_tools.F90:11-54): time_* from system_clock,
  ! time_*1 from the clock of the day (the reference's two columns)
  real(dp) :: time_init = 0, time_loop = 0, time_coll = 0, time_dg = 0, time_mp = 0
  real(dp) :: time_init1 = 0, time_loop1 = 0, time_coll1 = 0, time_dg1 = 0, time_mp1 = 0
  integer(8) :: count_rate
contains
  ! seconds of the day from date_and_time: the reference's second clock
  ! (subroutine time, SRC/mod_tools.F90:46-54)
  function tod() result(t)
    real(dp) :: t
    integer :: v(8)
    call date_and_time(values=v)
    t = 3600.0_dp * v(5) + 60.0_dp * v(6) + v(7) + v(8) / 1000.0_dp
  end function tod

  ! one line in the form of the date command to time.log (the reference
  ! runs system("date > time.log") at start, SRC/input.F90:39, and
  ! system("date >> time.log") at the end, SRC/finalize.F90:41-43)
  subroutine date_line(append)
    logical, intent(in) :: append
    character(len=3), parameter :: wd(0:6) = ['Sun', 'Mon', 'Tue', 'Wed', 'Thu', 'Fri', 'Sat']
    character(len=3), parameter :: mo(12) = ['Jan', 'Feb', 'Mar', 'Apr', 'May', 'Jun', &
         'Jul', 'Aug', 'Sep', 'Oct', 'Nov', 'Dec']
    integer, parameter :: off(12) = [0, 3, 2, 5, 0, 3, 5, 1, 4, 6, 2, 4]
    integer :: v(8), y, dow
    character(len=5) :: zone
    call date_and_time(zone=zone, values=v)
    y = v(1)
    if (v(2) < 3) y = y - 1
    dow = mod(y + y / 4 - y / 100 + y / 400 + off(v(2)) + v(3), 7)  ! Sakamoto
    if (append) then
       open(97, file='time.log', status='unknown', position='append')
    else
       open(97, file='time.log', status='replace')
    end if
    write(97, '(a3,1x,a3,1x,i2,1x,i2.2,":",i2.2,":",i2.2,1x,a5,1x,i4)') wd(dow), mo(v(2)), &
         v(3), v(5), v(6), v(7), zone, v(1)
    close(97)
  end subroutine date_line
end module bgk_host

program bgk_HIT
  use bgk_host
  implicit none
  integer :: itime, itstart, nrun
  integer(8) :: c0, c1, e0, e1
  real(dp) :: t0, t1
  character(len=256) :: deck
  logical, external :: step_event

  itstart = 1   ! SRC/bgk3d_HIT.F90:63
  deck = 'bgk.input'
  if (command_argument_count() >= 1) call get_command_argument(1, deck)
  call system_clock(c0, count_rate)
  t0 = tod()
  call input(trim(deck))
  call hencol
  call outdat
  call alloca
  call init_fields
  if (irestart == 1) call read_restart(itstart)
  call system_clock(c1)
  time_init = real(c1 - c0, dp) / real(count_rate, dp)
  time_init1 = tod() - t0

  call vtk_slices(itstart - 1)                    ! SRC/bgk3d_HIT.F90:118-120
  call system_clock(e0)
  t0 = tod()
  itime = itstart
  do while (itime <= itfin)                      ! SRC/bgk3d_HIT.F90:127-146
     call boundaries(itime)
     ! the steps up to the next one followed by output (or itfin) in one
     ! call, which the library runs as three- and two-step passes where it can
     nrun = 1
     do while (itime + nrun - 1 < itfin .and. .not. step_event(itime + nrun - 1))
        nrun = nrun + 1
     end do
     if (nrun > 1) then
        call col_MCn(itime, nrun)
        itime = itime + nrun - 1
     else
        call col_MC(itime)
     end if
     call diagnostic(itime)
     if (itsave > 0) then
        if (mod(itime, itsave) == 0) call write_restart(itime)
     end if
     itime = itime + 1
  end do
  call bgk_check(bgk_synchronize(ctx), 'bgk_synchronize')
  call system_clock(e1)
  time_loop = real(e1 - e0, dp) / real(count_rate, dp)
  time_loop1 = tod() - t0
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
  call date_line(.false.)  ! SRC/input.F90:39
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
  ! the rank's task log (SRC/setup_MPI.F90:260-323): the corners of its
  ! block and its neighbours, all rank 0 here (one rank, periodic)
  open(38, file='task.000000.log', status='unknown')
  write(38, *) 0, 0, 0
  write(38, *) 0, 0, n
  write(38, *) " "
  write(38, *) l, 0, 0
  write(38, *) l, 0, n
  write(38, *) " "
  write(38, *) l, m, 0
  write(38, *) l, m, n
  write(38, *) " "
  write(38, *) 0, m, 0
  write(38, *) 0, m, n
  write(38, *) " "
  write(38, *) 0, 0, 0
  write(38, *) 0, 0, n
  write(38, *) "# my front task is ", 0
  write(38, *) "# my rear task is  ", 0
  write(38, *) "# my left task is  ", 0
  write(38, *) "# my right task is ", 0
  write(38, *) "# my up task is    ", 0
  write(38, *) "# my down task is  ", 0
end subroutine outdat

! SRC/alloca.F90:178-216: the device lattice
subroutine alloca
  use bgk_host
  implicit none
  integer(c_int) :: flags
  ! kernel events per launch: the task log reports the collide launches by
  ! the time steps each advanced
  flags = BGK_FLAG_TIMING
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
  real(dp) :: t0
  call system_clock(c0)
  t0 = tod()
  call bgk_check(bgk_boundaries(ctx), 'bgk_boundaries')
  call system_clock(c1)
  time_mp = time_mp + real(c1 - c0, dp) / real(count_rate, dp)
  time_mp1 = time_mp1 + (tod() - t0)
  if (itime < 0) write(6, *) itime
end subroutine boundaries

! SRC/col_MC.F90:34-444 (the swap is inside the library)
subroutine col_MC(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  integer(8) :: c0, c1
  real(dp) :: t0
  call system_clock(c0)
  t0 = tod()
  call bgk_check(bgk_col_MC(ctx), 'bgk_col_MC')
  call bgk_check(bgk_synchronize(ctx), 'bgk_synchronize')
  call system_clock(c1)
  time_coll = time_coll + real(c1 - c0, dp) / real(count_rate, dp)
  time_coll1 = time_coll1 + (tod() - t0)
  if (itime < 0) write(6, *) itime
end subroutine col_MC

! Steps itime .. itime+nrun-1 (their boundaries included: nothing to
! exchange on the periodic lattice; with refcompat halos the library keeps
! the halo images of every step itself)
subroutine col_MCn(itime, nrun)
  use bgk_host
  implicit none
  integer, intent(in) :: itime, nrun
  integer(8) :: c0, c1
  real(dp) :: t0
  call system_clock(c0)
  t0 = tod()
  call bgk_check(bgk_step(ctx, int(nrun, c_int)), 'bgk_step')
  call bgk_check(bgk_synchronize(ctx), 'bgk_synchronize')
  call system_clock(c1)
  time_coll = time_coll + real(c1 - c0, dp) / real(count_rate, dp)
  time_coll1 = time_coll1 + (tod() - t0)
  if (itime < 0) write(6, *) itime
end subroutine col_MCn

! whether step itime is followed by output (diagnostics, signal, restart)
logical function step_event(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  step_event = mod(itime, icheck) == 0 .or. mod(itime, isignal) == 0
  if (itsave > 0) step_event = step_event .or. mod(itime, itsave) == 0
end function step_event

! SRC/diagnostic.F90:37-73
subroutine diagnostic(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  integer(8) :: c0, c1
  real(dp) :: t0
  real(c_double) :: ds(7)
  if (mod(itime, icheck) == 0) then
     call system_clock(c0)
     t0 = tod()
     ! diagno and dissipation (SRC/diagnostic.F90:47-49) from one pass over
     ! the populations, bit-identical to the two calls
     call bgk_check(bgk_diagnostics(ctx, real(svisc, c_double), diag_c, ds), &
          'bgk_diagnostics')
     call write_diagno(itime, ds(1:5))
     call write_dissipation(itime, ds(6:7))
     call probe_global(itime, lz / 2, ly / 2, lz / 2, 88)  ! SRC/diagnostic.F90:50
     if (mod(itime, ivtim) == 0) call vtk_slices(itime)
     call system_clock(c1)
     time_dg = time_dg + real(c1 - c0, dp) / real(count_rate, dp)
     time_dg1 = time_dg1 + (tod() - t0)
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
  call bgk_check(bgk_diagno(ctx, diag_c, d), 'bgk_diagno')
  call write_diagno(itime, d)
end subroutine diagno

! diagno's outputs (SRC/diagno.F90:132-152)
subroutine write_diagno(itime, d)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  real(c_double), intent(in) :: d(5)
  logical, save :: first = .true.
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
end subroutine write_diagno

! SRC/dissipation.F90:22-280
subroutine dissipation(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  real(c_double) :: s(2)
  call bgk_check(bgk_dissipation(ctx, real(svisc, c_double), diag_c, s), 'bgk_dissipation')
  call write_dissipation(itime, s)
end subroutine dissipation

! dissipation's outputs (SRC/dissipation.F90:262-278)
subroutine write_dissipation(itime, s)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  real(c_double), intent(in) :: s(2)
  real(sp) :: dx_s
  real(dp) :: dt, tke
  logical, save :: first = .true.
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
end subroutine write_dissipation

! SRC/probe_global.F90:55-95: the 19 populations of global site (i0,j0,k0)
subroutine probe_global(itime, ip, jp, kp, unit)
  use bgk_host
  implicit none
  integer, intent(in) :: itime, ip, jp, kp, unit
  real(c_double) :: f(19)
  if (ip < 1 .or. ip > l .or. jp < 1 .or. jp > m .or. kp < 1 .or. kp > n) return
  call bgk_check(bgk_get_box(ctx, int(ip - 1, c_int), int(ip, c_int), int(jp - 1, c_int), &
       int(jp, c_int), int(kp - 1, c_int), int(kp, c_int), f), 'bgk_get_box')
  write(unit, *) itime
  if (kind_c == BGK_STORE_F64 .or. kind_c == BGK_STORE_F32_MIXED) then
     write(unit, *) f(1), f(2), f(3), f(4)
     write(unit, *) f(5), f(6), f(7), f(8)
     write(unit, *) f(9), f(10), f(11), f(12)
     write(unit, *) f(13), f(14), f(15), f(16)
     write(unit, *) f(17), f(18), f(19)
  else
     write(unit, *) real(f(1:4), sp)
     write(unit, *) real(f(5:8), sp)
     write(unit, *) real(f(9:12), sp)
     write(unit, *) real(f(13:16), sp)
     write(unit, *) real(f(17:19), sp)
  end if
  write(unit, *) " "
  flush(unit)
end subroutine probe_global

! SRC/vtk_xy.binary.F90, vtk_xz, vtk_yz at the reference's planes n/2, m/2, l/2
subroutine vtk_slices(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  if (ivtim <= 0 .or. ivtim >= 1000000000) return
  call vtk_plane(itime, 'xy', 1, l, 1, m, n / 2, n / 2)
  call vtk_plane(itime, 'xz', 1, l, m / 2, m / 2, 1, n)
  call vtk_plane(itime, 'yz', l / 2, l / 2, 1, m, 1, n)
end subroutine vtk_slices

! one legacy-VTK STRUCTURED_POINTS slice: ASCII header, then big-endian
! float32 velocity triplets u/den, w/den, v/den with x fastest
subroutine vtk_plane(itime, tag, xa, xb, ya, yb, za, zb)
  use bgk_host
  implicit none
  integer, intent(in) :: itime, xa, xb, ya, yb, za, zb
  character(len=2), intent(in) :: tag
  character(len=25) :: file_name
  real(c_double), allocatable :: f(:, :)
  real(sp) :: u, w, v, den
  real(dp) :: ud, wd, vd, dend
  integer :: ni, nj, nk, np, p
  if (xa < 1 .or. ya < 1 .or. za < 1) return
  ni = xb - xa + 1
  nj = yb - ya + 1
  nk = zb - za + 1
  np = ni * nj * nk
  allocate(f(np, 19))
  call bgk_check(bgk_get_box(ctx, int(xa - 1, c_int), int(xb, c_int), int(ya - 1, c_int), &
       int(yb, c_int), int(za - 1, c_int), int(zb, c_int), f), 'bgk_get_box')
  file_name = 'tec_' // tag // '.yyyyy.xxxxxxxx.vtk'
  write(file_name(8:12), '(i5.5)') 0
  write(file_name(14:21), '(i8.8)') itime
  open(52, file=file_name, status='unknown')
  write(52, '(A26)') '# vtk DataFile Version 2.0'
  write(52, '(A5)') 'Campo'
  write(52, '(A6)') 'BINARY'
  write(52, '(A25)') 'DATASET STRUCTURED_POINTS'
  write(52, '(A11,I10,A1,I10,A1,I10)') 'DIMENSIONS ', ni, ' ', nj, ' ', nk
  write(52, '(A7,I10,A1,I10,A1,I10)') 'ORIGIN ', xa, ' ', ya, ' ', za
  write(52, '(A8,I10,A1,I10,A1,I10)') 'SPACING ', 1, ' ', 1, ' ', 1
  write(52, '(A10,I10)') 'POINT_DATA ', np
  write(52, '(A23)') 'VECTORS velocity float'
  close(52)
  open(52, file=file_name, status='old', position='append', form='unformatted', &
       access='stream', convert='big_endian')
  do p = 1, np
     if (kind_c == BGK_STORE_F64) then
        ! DOUBLE_P build: sums in double, rounded to the float32 output
        ud = f(p,1) + f(p,2) + f(p,3) + f(p,4) + f(p,5) - f(p,10) - f(p,11) - f(p,12) - &
             f(p,13) - f(p,14)
        wd = f(p,3) + f(p,7) + f(p,8) + f(p,9) + f(p,12) - f(p,1) - f(p,10) - f(p,16) - &
             f(p,17) - f(p,18)
        vd = f(p,4) + f(p,6) + f(p,7) + f(p,13) + f(p,18) - f(p,2) - f(p,9) - f(p,11) - &
             f(p,15) - f(p,16)
        dend = sum(f(p, :)) + 1.0_dp
        u = real(ud, sp); w = real(wd, sp); v = real(vd, sp); den = real(dend, sp)
     else
        ! single-precision storage: sums of the stored values in float32
        u = real(f(p,1), sp) + real(f(p,2), sp) + real(f(p,3), sp) + real(f(p,4), sp) + &
            real(f(p,5), sp) - real(f(p,10), sp) - real(f(p,11), sp) - real(f(p,12), sp) - &
            real(f(p,13), sp) - real(f(p,14), sp)
        w = real(f(p,3), sp) + real(f(p,7), sp) + real(f(p,8), sp) + real(f(p,9), sp) + &
            real(f(p,12), sp) - real(f(p,1), sp) - real(f(p,10), sp) - real(f(p,16), sp) - &
            real(f(p,17), sp) - real(f(p,18), sp)
        v = real(f(p,4), sp) + real(f(p,6), sp) + real(f(p,7), sp) + real(f(p,13), sp) + &
            real(f(p,18), sp) - real(f(p,2), sp) - real(f(p,9), sp) - real(f(p,11), sp) - &
            real(f(p,15), sp) - real(f(p,16), sp)
        den = sum(real(f(p, :), sp)) + 1.0_sp
     end if
     write(52) u / den, w / den, v / den
  end do
  close(52)
  deallocate(f)
end subroutine vtk_plane

! restart file: 'BGK3DRST', lx, ly, lz, itime, storage kind, then the 19
! population planes as doubles (exact for every storage kind)
subroutine write_restart(itime)
  use bgk_host
  implicit none
  integer, intent(in) :: itime
  real(c_double), allocatable :: f(:)
  allocate(f(19_8 * int(l, 8) * int(m, 8) * int(n, 8)))
  call bgk_check(bgk_get_populations(ctx, f), 'bgk_get_populations')
  open(60, file='restart.bgk', status='replace', form='unformatted', access='stream')
  write(60) 'BGK3DRST', int(lx, 4), int(ly, 4), int(lz, 4), int(itime, 4), int(kind_c, 4)
  write(60) f
  close(60)
  deallocate(f)
end subroutine write_restart

subroutine read_restart(itstart)
  use bgk_host
  implicit none
  integer, intent(out) :: itstart
  real(c_double), allocatable :: f(:)
  character(len=8) :: magic
  integer(4) :: rx, ry, rz, rit, rkind
  integer :: ios
  open(60, file='restart.bgk', status='old', form='unformatted', access='stream', iostat=ios)
  if (ios /= 0) then
     write(6, *) 'ERROR: irestart=1 but restart.bgk cannot be opened'
     error stop 1
  end if
  read(60) magic, rx, ry, rz, rit, rkind
  if (magic /= 'BGK3DRST' .or. rx /= lx .or. ry /= ly .or. rz /= lz) then
     write(6, *) 'ERROR: restart.bgk does not match the lattice', rx, ry, rz
     error stop 1
  end if
  allocate(f(19_8 * int(l, 8) * int(m, 8) * int(n, 8)))
  read(60) f
  close(60)
  call bgk_check(bgk_set_populations(ctx, f), 'bgk_set_populations')
  deallocate(f)
  itstart = rit + 1
  write(6, *) 'INFO: restarted from step ', rit
end subroutine read_restart

! SRC/finalize.F90:23-118
subroutine finalize(itstart)
  use bgk_host
  implicit none
  integer, intent(in) :: itstart
  real(dp) :: glups, glups1, knorm, r(2)
  real(c_double) :: t(8), tp(6), tf(2)
  call date_line(.true.)
  ! per-rank section timers and ratios (SRC/finalize.F90:45-63); one rank
  ! here, so task.000000.log
  write(38, *) "# Time for section "
  write(38, 1100) time_init, time_init1
  write(38, 1101) time_loop, time_loop1
  write(38, 1102) time_coll, time_coll1
  write(38, 1103) time_dg, time_dg1
  write(38, 1115) time_mp, time_mp1
  write(38, 9999)
  write(38, *) "# Ratio "
  write(38, 1202) ratio(time_coll, time_loop), ratio(time_coll1, time_loop1)
  write(38, 1203) ratio(time_dg, time_loop), ratio(time_dg1, time_loop1)
  write(38, 1215) ratio(time_mp, time_loop), ratio(time_mp1, time_loop1)
  r(1) = ratio(time_loop - time_coll - time_dg - time_mp, time_loop)
  r(2) = ratio(time_loop1 - time_coll1 - time_dg1 - time_mp1, time_loop1)
  write(38, 1216) r
  write(38, 9999)
  ! exchange time and bandwidth per axis (SRC/finalize.F90:58-63): the one
  ! GPU exchanges nothing across x and y (periodic wrap in the index math),
  ! and its z exchange (slab ghosts, reference halo images) is the
  ! boundaries time
  knorm = real(storage_bytes(), dp) / (1024.0_dp * 1024.0_dp)
  write(38, *) "# Z-MPI time, BW (MB/s) -->", time_mp, &
       ratio(real(m + 2, dp) * real(l + 2, dp) * real(itfin, dp) * knorm, time_mp)
  write(38, *) "# Y-MPI time, BW (MB/s) -->", 0.0_dp, 0.0_dp
  write(38, *) "# X-MPI time, BW (MB/s) -->", 0.0_dp, 0.0_dp
  ! collide launches (and their kernel ms) by the steps one launch advances:
  ! 1, 2, 3, and 2 with the diagnostics of the result fused in
  call bgk_check(bgk_timing_passes(ctx, tp, 0_c_int), 'bgk_timing_passes')
  call bgk_check(bgk_timing_fused(ctx, tf, 0_c_int), 'bgk_timing_fused')
  write(38, 1300) nint(tp(2)), nint(tp(4)), nint(tp(6)), nint(tf(2))
  write(38, 1301) tp(1), tp(3), tp(5), tf(1)
  close(38)
  ! SRC/finalize.F90:75-90 (GLUPs of bgk.perf on the second clock, of
  ! stdout on the first, itfin - itstart steps in the numerator)
  glups = real(lx, dp) * real(ly, dp) * real(lz, dp) * real(itfin - itstart, dp) / &
       (time_loop * 1000 * 1000 * 1000)
  glups1 = real(lx, dp) * real(ly, dp) * real(lz, dp) * real(itfin - itstart, dp) / &
       (max(time_loop1, 1e-9_dp) * 1000 * 1000 * 1000)
  open(69, file='bgk.perf', status='unknown')
  write(69, 9999)
  write(69, 1100) time_init, time_init1
  write(69, 1101) time_loop, time_loop1
  write(69, 1102) time_coll, time_coll1
  write(69, 1103) time_dg, time_dg1
  write(69, 1115) time_mp, time_mp1
  write(69, 9999)
  write(69, 3000) glups1
  write(69, 1216) r
  write(69, 9999)
  close(69)
  write(6, 9999)
  write(6, *) "INFO: That's all Folks!!!!!! "
  write(6, 2000) time_loop, time_loop1
  write(6, 3000) glups
  write(6, *) "INFO: That's all Folks!!!!!! "
  write(6, 9999)
  call bgk_check(bgk_timing(ctx, t, 1_c_int), 'bgk_timing')
  call bgk_check(bgk_destroy(ctx), 'bgk_destroy')
  close(16)
9999 format(" #--------------------------------")
1100 format(" # init   time", 2(e14.6, 1x))
1101 format(" # loop   time", 2(e14.6, 1x))
1102 format(" # comp.  time", 2(e14.6, 1x))
1103 format(" # diag.  time", 2(e14.6, 1x))
1115 format(" # MPI    time", 2(e14.6, 1x))
1202 format(" # Ratio Coll ", 2(f7.3, 1x))
1203 format(" # Ratio Dg.  ", 2(f7.3, 1x))
1215 format(" # Ratio MPI  ", 2(f7.3, 1x))
1216 format(" # Check      ", 2(f7.3, 1x))
2000 format(" # loop   time", 2(e14.6, 1x))
3000 format(" # GLUPs      ", 1(e14.6))
1300 format(" # Passes 1/2/3/2+diag steps", 4(i10, 1x))
1301 format(" # Pass ms 1/2/3/2+diag steps", 4(e14.6, 1x))
contains
  real(dp) function ratio(a, b)
    real(dp), intent(in) :: a, b
    ratio = 0.0_dp
    if (b > 0.0_dp) ratio = a / b
  end function ratio
  integer function storage_bytes()
    select case (trim(storage))
    case ('f64'); storage_bytes = 8
    case ('f16', 'bf16'); storage_bytes = 2
    case default; storage_bytes = 4
    end select
  end function storage_bytes
end subroutine finalize
