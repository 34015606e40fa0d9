! tt_driver.f90 -- a Fortran host that runs ttcode's step loop (pp.f:483-821
! for IADV=IFOR=1, IRAD=3: ADVECT3D, FORGAS3D, COUPLED_ENERGY_EQS, BOUNDS,
! BOUNDS, PRESSG) on the GPU through the ISO_C_BINDING layer, with module
! arrays in the reference's declared bounds.
!   tt_driver <dir>   reads <dir>/params.nml and <dir>/in_<name>.bin,
!                     writes <dir>/out_<name>.bin and <dir>/out_scalars.txt
program tt_driver
  use iso_c_binding
  use ttgpu_c
  use ttgpu_host
  implicit none
  type(tt_params), target :: p
  integer :: h_nx, h_ny, h_nz, h_itmax, h_iii, h_nsteps
  double precision :: h_epsmat, h_delt, h_tirr, h_kap0, h_dkap0, h_rho0, h_t0, h_a, h_b
  namelist /ttdrv/ h_nx, h_ny, h_nz, h_itmax, h_iii, h_nsteps, h_epsmat, h_delt, h_tirr, &
       h_kap0, h_dkap0, h_rho0, h_t0, h_a, h_b
  double precision, allocatable, target :: T(:,:,:), RH(:,:,:), V(:,:,:), G(:,:,:), H(:,:,:)
  double precision, allocatable, target :: ER(:,:,:), gravx(:,:,:)
  character(len=512) :: dir
  integer :: u, s, nx, ny, nz

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/params.nml', status='old')
  read(u, nml=ttdrv)
  close(u)
  nx = h_nx; ny = h_ny; nz = h_nz
  p%nx = nx; p%ny = ny; p%nz = nz
  p%bcxmin = 12; p%ibdrad = 7; p%iflimtyp = 2; p%f_incident = 2; p%ikaptyp = 26
  p%isigtyp = 0; p%isoth = 0; p%itmax = h_itmax; p%igrav = 1; p%mstar2 = 0.d0
  p%xmin = 7149199872.d0; p%xmax = 1.1d0 * 7149199872.d0
  p%ymin = 0.d0; p%ymax = 1.d0; p%zmin = -0.5d0; p%zmax = 0.5d0
  p%gamma = 1.4d0; p%mu_gas = 2.3d0; p%gamma_opc = 0.1d0
  p%zkapmult = h_kap0; p%kap0 = h_kap0; p%dkap0 = h_dkap0; p%kap_rho0 = h_rho0; p%kap_t0 = h_t0
  p%kap_a = h_a; p%kap_b = h_b; p%tirr = h_tirr; p%fbottom = 0.d0; p%d_t = 0.d0; p%omrot = 0.d0
  p%sorparam = 1.7d0; p%epsmat = h_epsmat; p%fdelt = 0.4d0; p%dtmax = 1.d12

  ! module arrays with the reference's bounds (ppmod.f:320-328)
  allocate(T(-1:nx+2,-1:ny+2,-1:nz+2), RH(-1:nx+2,-1:ny+2,-1:nz+2), ER(-1:nx+2,-1:ny+2,-1:nz+2))
  allocate(V(0:nx+2,-1:ny+2,-1:nz+2), G(-1:nx+2,0:ny+2,-1:nz+2), H(-1:nx+2,-1:ny+2,0:nz+2))
  allocate(gravx(-1:nx+2,-1:ny+2,-1:nz+2))
  call load(T, 'T'); call load(RH, 'RH'); call load(V, 'V'); call load(G, 'G')
  call load(H, 'H'); call load(ER, 'ER'); call load(gravx, 'gravx')

  call tt_check(tt_init(p), 'tt_init')
  call tt_put3(TT_T, T); call tt_put3(TT_RH, RH); call tt_put3(TT_V, V)
  call tt_put3(TT_G, G); call tt_put3(TT_H, H); call tt_put3(TT_ER, ER)
  call tt_put3(TT_GRAVX, gravx)

  DELT = h_delt
  do s = 1, h_nsteps
     III = h_iii + s - 1
     call ADVECT3D
     call FORGAS3D
     call COUPLED_ENERGY_EQS
     call BOUNDS
     call BOUNDS
     call PRESSG
  end do

  call tt_get3(TT_T, T); call tt_get3(TT_RH, RH); call tt_get3(TT_V, V)
  call tt_get3(TT_G, G); call tt_get3(TT_H, H); call tt_get3(TT_ER, ER)
  call dump(T, 'T'); call dump(RH, 'RH'); call dump(V, 'V'); call dump(G, 'G')
  call dump(H, 'H'); call dump(ER, 'ER')
  open(newunit=u, file=trim(dir)//'/out_scalars.txt', status='replace')
  write(u, '(A,1X,I0)') 'energy_iter', energy_iter
  write(u, '(A,1X,I0)') 'SOR_ITER', sor_iter
  close(u)
  call tt_finalize()

contains
  subroutine load(a, name)
    double precision, intent(inout) :: a(:,:,:)
    character(*), intent(in) :: name
    integer :: lu
    open(newunit=lu, file=trim(dir)//'/in_'//name//'.bin', access='stream', form='unformatted', status='old')
    read(lu) a
    close(lu)
  end subroutine load
  subroutine dump(a, name)
    double precision, intent(in) :: a(:,:,:)
    character(*), intent(in) :: name
    integer :: lu
    open(newunit=lu, file=trim(dir)//'/out_'//name//'.bin', access='stream', form='unformatted', status='replace')
    write(lu) a
    close(lu)
  end subroutine dump
end program tt_driver
