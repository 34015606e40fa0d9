! ttgpu_mod.f90 -- ISO_C_BINDING layer over libttgpu.so (include/ttgpu.h).
!
! Module ttgpu_c mirrors the C structs and entry points.  Module ttgpu_host
! provides parameterless subroutines with the reference's names and calling
! style (they read DELT and III from the host module, as ttcode's routines
! read fluid_var_init): ADVECT3D (ppadv.f:525), FORGAS3D (ppfor.f:142),
! COUPLED_ENERGY_EQS (ppenergy.f:2), BOUNDS (ppbou.f:1), PRESSG (ppphy.f:1),
! SCHRITT (ppsub.f:1), SOLVE_WITH_SOR (ppsor.f:2).  A non-zero positive
! return from the library stops the program, as the reference's CLEAN_STOP
! would; negative returns (iteration caps) are reported and ignored, as the
! reference does.
module ttgpu_c
  use iso_c_binding
  implicit none

  type, bind(C) :: tt_params
     integer(c_int32_t) :: nx, ny, nz, bcxmin, ibdrad, iflimtyp, f_incident
     integer(c_int32_t) :: ikaptyp, isigtyp, isoth, itmax, igrav
     real(c_double) :: xmin, xmax, ymin, ymax, zmin, zmax, gamma, mu_gas
     real(c_double) :: zkapmult, kap0, dkap0, kap_rho0, kap_t0, kap_a, kap_b
     real(c_double) :: tirr, fbottom, d_t, omrot, sorparam, epsmat, fdelt, dtmax
     real(c_double) :: gamma_opc, mstar2
  end type tt_params

  type, bind(C) :: tt_energy_stats
     integer(c_int32_t) :: energy_iter, sor_iter_total, sor_iter_last, sor_iter_first
     integer(c_int32_t) :: fp32_sweeps, fallbacks, learned_used, learned_skipped
     real(c_double) :: max_dq, max_de, guess_scale_first, guess_r0_first
     integer(c_int32_t) :: first_start, first_choice
     real(c_double) :: ms_build, ms_sor, ms_update, ms_sor_first
  end type tt_energy_stats

  type, bind(C) :: tt_viscosity
     integer(c_int32_t) :: ivis, ivistyp, ieta, idiss
     real(c_double) :: xnue
     integer(c_int32_t) :: iartvis, pad0
  end type tt_viscosity

  type, bind(C) :: tt_sor_stats
     integer(c_int32_t) :: sor_iter, outer_iter, fp32_sweeps, guess_used
     real(c_double) :: rhsnorm, rnorm, r0_guess, r0_cold, guess_scale
  end type tt_sor_stats

  ! field ids (include/ttgpu.h)
  integer(c_int), parameter :: TT_T = 1, TT_RH = 2, TT_V = 3, TT_G = 4, TT_H = 5
  integer(c_int), parameter :: TT_ER = 6, TT_DELTA_ER = 8, TT_GRAVX = 32, TT_MU_STAR = 81

  interface
     integer(c_int) function tt_init(p) bind(C, name='tt_init')
       import :: c_int, tt_params
       type(tt_params), intent(in) :: p
     end function tt_init
     subroutine tt_finalize() bind(C, name='tt_finalize')
     end subroutine tt_finalize
     type(c_ptr) function tt_last_error() bind(C, name='tt_last_error')
       import :: c_ptr
     end function tt_last_error
     integer(c_int) function tt_upload(field, host) bind(C, name='tt_upload')
       import :: c_int, c_ptr
       integer(c_int), value :: field
       type(c_ptr), value :: host
     end function tt_upload
     integer(c_int) function tt_download(field, host) bind(C, name='tt_download')
       import :: c_int, c_ptr
       integer(c_int), value :: field
       type(c_ptr), value :: host
     end function tt_download
     integer(c_int) function tt_bounds() bind(C, name='tt_bounds')
       import :: c_int
     end function tt_bounds
     integer(c_int) function tt_pressg() bind(C, name='tt_pressg')
       import :: c_int
     end function tt_pressg
     ! the reference's FreedmanXKAP_plnk/_ross/_abs(42,18) after
     ! READFREEDMANOPCTABLE pass straight through (same memory order)
     integer(c_int) function tt_set_opacity_table(kp, kr, ka) bind(C, name='tt_set_opacity_table')
       import :: c_int, c_double
       real(c_double), intent(in) :: kp(42, 18), kr(42, 18), ka(42, 18)
     end function tt_set_opacity_table
     ! ColumnXKAP_abs(42,18,8) after READCOLUMNDEPTHOPCTABLE
     integer(c_int) function tt_set_column_table(ka) bind(C, name='tt_set_column_table')
       import :: c_int, c_double
       real(c_double), intent(in) :: ka(42, 18, 8)
     end function tt_set_column_table
     ! BurrowsXKAP_plnk/_ross/_abs(50,50) after READBURROWSOPCTABLE
     integer(c_int) function tt_set_burrows_table(kp, kr, ka) bind(C, name='tt_set_burrows_table')
       import :: c_int, c_double
       real(c_double), intent(in) :: kp(50, 50), kr(50, 50), ka(50, 50)
     end function tt_set_burrows_table
     integer(c_int) function tt_kappa() bind(C, name='tt_kappa')
       import :: c_int
     end function tt_kappa
     ! IVIS, IVISTYP, IETA, IDISS and XNUE of input_init (ppmod.f:42-71)
     integer(c_int) function tt_set_viscosity(v) bind(C, name='tt_set_viscosity')
       import :: c_int, tt_viscosity
       type(tt_viscosity), intent(in) :: v
     end function tt_set_viscosity
     integer(c_int) function tt_get_dt_artvis(dt) bind(C, name='tt_get_dt_artvis')
       import :: c_int, c_double
       real(c_double), intent(out) :: dt
     end function tt_get_dt_artvis
     integer(c_int) function tt_advect3d(delt) bind(C, name='tt_advect3d')
       import :: c_int, c_double
       real(c_double), value :: delt
     end function tt_advect3d
     integer(c_int) function tt_forgas3d(delt) bind(C, name='tt_forgas3d')
       import :: c_int, c_double
       real(c_double), value :: delt
     end function tt_forgas3d
     integer(c_int) function tt_set_step(iii) bind(C, name='tt_set_step')
       import :: c_int
       integer(c_int), value :: iii
     end function tt_set_step
     integer(c_int) function tt_coupled_energy(delt, iii, st) bind(C, name='tt_coupled_energy')
       import :: c_int, c_double, tt_energy_stats
       real(c_double), value :: delt
       integer(c_int), value :: iii
       type(tt_energy_stats), intent(out) :: st
     end function tt_coupled_energy
     integer(c_int) function tt_solve_sor(fill, st) bind(C, name='tt_solve_sor')
       import :: c_int, tt_sor_stats
       integer(c_int), value :: fill
       type(tt_sor_stats), intent(out) :: st
     end function tt_solve_sor
     integer(c_int) function tt_schritt(delt) bind(C, name='tt_schritt')
       import :: c_int, c_double
       real(c_double), intent(out) :: delt
     end function tt_schritt
  end interface
end module ttgpu_c

module ttgpu_host
  use iso_c_binding
  use ttgpu_c
  implicit none
  ! the host copy of the step scalars the reference keeps in fluid_var_init
  double precision :: DELT = 0.d0
  integer :: III = 0
  integer :: energy_iter = 0, sor_iter = 0
contains
  subroutine tt_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(tt_last_error(), msg, [256])
    n = 0
    do while (n < 256)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    write(*, '(A,A,I4,A,256A1)') what, ': rc=', rc, ' ', msg(1:n)
    if (rc > 0) stop 1
  end subroutine tt_check

  subroutine BOUNDS
    call tt_check(tt_set_step(int(III, c_int)), 'BOUNDS')
    call tt_check(tt_bounds(), 'BOUNDS')
  end subroutine BOUNDS

  subroutine PRESSG
    call tt_check(tt_pressg(), 'PRESSG')
  end subroutine PRESSG

  subroutine ADVECT3D
    call tt_check(tt_advect3d(DELT), 'ADVECT3D')
  end subroutine ADVECT3D

  subroutine FORGAS3D
    call tt_check(tt_forgas3d(DELT), 'FORGAS3D')
  end subroutine FORGAS3D

  subroutine COUPLED_ENERGY_EQS
    type(tt_energy_stats) :: st
    call tt_check(tt_coupled_energy(DELT, int(III, c_int), st), 'COUPLED_ENERGY_EQS')
    energy_iter = st%energy_iter
    sor_iter = st%sor_iter_last
  end subroutine COUPLED_ENERGY_EQS

  subroutine SOLVE_WITH_SOR(print_max_values)
    logical, intent(in) :: print_max_values
    type(tt_sor_stats) :: st
    integer(c_int) :: rc
    rc = tt_solve_sor(1_c_int, st)
    if (rc /= 0 .and. print_max_values) call tt_check(rc, 'SOLVE_WITH_SOR')
    sor_iter = st%sor_iter
  end subroutine SOLVE_WITH_SOR

  subroutine SCHRITT
    real(c_double) :: d
    call tt_check(tt_schritt(d), 'SCHRITT')
    DELT = d
  end subroutine SCHRITT

  ! upload / download one module array in its declared bounds
  subroutine tt_put3(field, a)
    integer(c_int), intent(in) :: field
    double precision, intent(in), target, contiguous :: a(:,:,:)
    call tt_check(tt_upload(field, c_loc(a)), 'tt_upload')
  end subroutine tt_put3

  subroutine tt_get3(field, a)
    integer(c_int), intent(in) :: field
    double precision, intent(inout), target, contiguous :: a(:,:,:)
    call tt_check(tt_download(field, c_loc(a)), 'tt_download')
  end subroutine tt_get3

  subroutine tt_put2(field, a)
    integer(c_int), intent(in) :: field
    double precision, intent(in), target, contiguous :: a(:,:)
    call tt_check(tt_upload(field, c_loc(a)), 'tt_upload')
  end subroutine tt_put2
end module ttgpu_host
