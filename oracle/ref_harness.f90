! ref_harness.f90 -- drives the reference Fortran routines of ttcode's hot path
! on one rank to produce golden fixtures.  TEST INFRASTRUCTURE ONLY: built by
! oracle/Makefile.ref into oracle/_ref/ from the reference sources where they
! lie; never shipped, never on the product path.
!
! It mirrors the set-up of pp.f:124-200 (MPI, allocate*var, MPI_CART_CREATE,
! SETUP3DMPIGRID, GITTER) but takes the module switches from a namelist
! instead of ipp, loads the state from raw Fortran-order binaries
! (in_<name>.bin) and runs one command:
!   gitter | bounds | kappa | advect | forgas | energy | sor | adjust | step |
!   diffusion | printout | restart
! For IKAPTYP 4 and 5 the reference's own READFREEDMANOPCTABLE reads its table
! file (Freedman6117.dat or Freedman6117_noTIVO.dat) from the working directory,
! for IKAPTYP 9-12 and 14 READBURROWSOPCTABLE its (BURROWS6117_wTiVO.dat or
! BURROWS6117_woutTiVO.dat, planet_num 1).
! then writes every array of interest to out_<name>.bin and the scalars to
! out_scalars.txt.  CALCMUSTAR lives in ppsub.f, which needs the absent hypre
! include; mu_star is therefore an input (in_mu_star.bin).
program tt_ref_harness
  use input_init
  use mpi_var_init
  use mpi_var_alloc
  use mpi_grid_init
  use fluid_var_init
  use fluid_var_alloc
  use force_var_init
  use force_var_alloc
  use rad_var_init
  use rad_var_alloc
  use grid_var_init
  use grid_var_alloc
  use deltaER_var_init
  use deltaER_var_alloc
  use sor_var_init
  use sor_var_alloc
  use scalar_var_init
  use scalar_var_alloc
  use dif_var_init
  use dif_var_alloc
  use global_var_init
  use global_var_alloc
  use artvis_var_init
  use artvis_var_alloc
  use global_constants, only: FD_c => FD, ARAD_c => ARAD, GRAV, MSOL, DAY, PI
  implicit none
  include "mpif.h"

  integer :: h_nx, h_ny, h_nz, h_bcxmin, h_ibdrad, h_iflimtyp, h_f_incident
  integer :: h_ikaptyp, h_isigtyp, h_itmax, h_iii, h_nsteps
  double precision :: h_xmin, h_xmax, h_ymin, h_ymax, h_zmin, h_zmax
  double precision :: h_gamma, h_mu_gas, h_zkapmult, h_tirr, h_fbottom, h_d_t, h_gamma_opc
  double precision :: h_omrot, h_sorparam, h_epsmat, h_fdelt, h_dtmax, h_delt
  double precision :: h_xnue, h_pr_num
  integer :: h_prefill_kappa, h_num_iter, h_ivis, h_ivistyp, h_ieta, h_idiss, h_iartvis
  double precision :: h_zeit, h_tinitbottom, h_tinittop, h_rhinittop, h_rhinitbottom
  double precision :: h_hydrotime, h_maccreated, h_mstar2
  integer :: h_igrav
  logical :: have_sin
  character(len=32) :: cmd
  character(len=512) :: dir
  namelist /ttin/ h_nx, h_ny, h_nz, h_bcxmin, h_ibdrad, h_iflimtyp, &
       h_f_incident, h_ikaptyp, h_isigtyp, h_itmax, h_iii, h_nsteps, &
       h_xmin, h_xmax, h_ymin, h_ymax, h_zmin, h_zmax, h_gamma, h_mu_gas, &
       h_zkapmult, h_tirr, h_fbottom, h_d_t, h_omrot, h_sorparam, h_epsmat, h_gamma_opc, &
       h_fdelt, h_dtmax, h_delt, h_xnue, h_pr_num, h_prefill_kappa, cmd, &
       h_zeit, h_num_iter, h_tinitbottom, h_tinittop, h_rhinittop, h_rhinitbottom, &
       h_hydrotime, h_maccreated, h_ivis, h_ivistyp, h_ieta, h_idiss, h_iartvis, &
       h_igrav, h_mstar2
  integer :: u, n, s
  integer :: adj_iter(50)
  integer(8) :: clk0, clk1, clk_rate
  double precision :: step_ms(1000)

  call get_command_argument(1, dir)
  h_nsteps = 1
  h_zeit = 0.d0; h_num_iter = 0; h_tinitbottom = 0.d0; h_tinittop = 0.d0
  h_rhinittop = 0.d0; h_rhinitbottom = 0.d0; h_hydrotime = 0.d0; h_maccreated = 0.d0
  h_ivis = 0; h_ivistyp = 1; h_ieta = 4; h_idiss = 0; h_iartvis = 0; h_gamma_opc = 0.1d0
  h_prefill_kappa = 0
  h_igrav = 1; h_mstar2 = 0.d0
  h_xnue = 1.d9
  h_pr_num = 1.d0
  open(newunit=u, file=trim(dir)//'/params.nml', status='old')
  read(u, nml=ttin)
  close(u)

  call MPI_INIT(ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, myid, ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, numprocs, ierr)

  ! input_init, as readinitfile (ppmod.f:42-71) would set it for these runs
  ISTEPMAX = 1; IREA = 10000; FDELT = h_fdelt; DTMAX = h_dtmax
  iprstep = 0; FILESTART = 0
  ietot = 0; isoth = 0; isoden = 0
  IADV = 1; IFOR = 1; IGRAV = h_igrav; IVIS = h_ivis; IRAD = 3; INUC = 0; ISCAL = 1; IDIF = 2
  IRHO = 0; RHMIN = 1.d-11
  ITEMP = 0; TEMPMAX = 1.d4; TEMPMIN = 1.d1
  IVEL = 0; VELMAX = 6.d5; VELDAMP = 0.d0; DAMP_STEPS = 100000
  IVISTYP = h_ivistyp; XNUE = h_xnue; IDISS = h_idiss; ALPHA = 1.d-4; IETA = h_ieta; IARTVIS = h_iartvis; PR_NUM = h_pr_num
  IGASTYP = 1; GAMMA = h_gamma; mu_gas = h_mu_gas; ICSTYP = 0; aspect = 0.04d0
  IKAPTYP = h_ikaptyp; ISIGTYP = h_isigtyp; ZKAPMULT = h_zkapmult; gamma_OPC = h_gamma_opc
  IROTATE = 0; PROT = 3.52d0; ramp_start = 0.d0; ramp_full = 0.d0; windAMP = 1.d0
  IBDRAD = h_ibdrad; IFLIMTYP = h_iflimtyp; F_INCIDENT = h_f_incident
  BCXMIN = h_bcxmin; Fbot = h_fbottom
  SORPARAM = h_sorparam; ITMAX = h_itmax; EPSMAT = h_epsmat
  iperturb = 0; perturb_typ = 0; perturb_iter = 100
  Porb = 3.52d0; ecc = 0.d0; ecc_start = 0.d0
  iFbotramp = 0; Framp_steps = 3000; Fbot_final = 2000.d0
  NCOSYS = 2; inc_poles = 0; POLES = .false.
  MODTYP = 2; MODVER = 3; PLANET_NUM = 1
  STARTYP = 0.d0; RSTAR = 1.155d0; TSTAR = 6070.d0; MSTAR1 = 1.119d0; MSTAR2 = h_mstar2
  TIRR = h_tirr; RH0 = 0.d0
  NUMDIM = 3
  global_NX = h_nx; XMIN = h_xmin; XMAX = h_xmax; DXMIN = 0.d0
  global_NY = h_ny; YMIN = h_ymin; YMAX = h_ymax; DYMIN = 0.d0
  global_NZ = h_nz; ZMIN = h_zmin; ZMAX = h_zmax; DZMIN = 0.d0

  call allocatempivar(MPI_STATUS_SIZE)
  call allocatefluidvar
  call allocateforcevar
  call allocateradvar
  call allocategridvar
  call allocatedeltaERvar
  call allocatesorvar
  call allocatescalarvar
  call allocatedifvar
  if (IARTVIS > 0) then
     ! pp.f:145 allocates the fluxes undefined; zero them so that the entries
     ! NR_ARTIFICIAL_VISC reads without writing (upperbnd varying between
     ! columns) are defined; dt_artvis = DTMAX as pp.f:439 sets it
     call allocateartvisvar
     QviscX = 0.d0; QviscY = 0.d0; QviscZ = 0.d0
     dt_artvis = DTMAX
  end if
  ISCAL = 0
  call zero_modules()

  call MPI_CART_CREATE(MPI_COMM_WORLD, proc_ndims, proc_dims, periods, &
       reorder, COMM_CART, ierr)
  call MPI_CART_COORDS(COMM_CART, myid, proc_ndims, proc_coords, ierr)
  call SETUP3DMPIGRID

  III = h_iii
  num_iter = 0
  DELT = h_delt
  OMROT = h_omrot
  D_T = h_d_t
  D_RH = 0.d0
  FBottom = h_fbottom
  hydrotime = h_hydrotime
  ! the semi-major axis as pp.f:204-206 computes it (the wavelength-binned
  ! STELLARINPUT_HEATING divides by it)
  semi = (GRAV*MSOL*(mstar1+mstar2)*((Porb*DAY)**2.d0)/(4.0*PI*PI))**(1.d0/3.d0)
  call GITTER
  cv = 0.d0
  call GASINIT

  if (trim(cmd) /= 'gitter') then
     call load3(T, 'T')
     call load3(RH, 'RH')
     call load3(V, 'V')
     call load3(G, 'G')
     call load3(H, 'H')
     call load3(ER, 'ER')
     if (IGRAV == 4) then
        ! the point-mass gravity FGRAV computes (pp.f:461)
        call FGRAV
     else
        call load3(gravx, 'gravx')
     end if
     call load2(mu_star, 'mu_star')
     if (h_prefill_kappa /= 0) then
        call load3(xkapR, 'xkapR')
        call load3(xkapP, 'xkapP')
        call load3(xkapA, 'xkapA')
        call load3(dxkapP_dT, 'dxkapP_dT')
        call load3(dxkapA_dT, 'dxkapA_dT')
     end if
     ! BOUNDXMAX(2) version 9 at III > 1 reads STELLARINPUT from the energy
     ! solve of the step before: take it as an input when one is given
     inquire(file=trim(dir)//'/in_STELLARINPUT.bin', exist=have_sin)
     if (have_sin) call load3(STELLARINPUT, 'STELLARINPUT')
  end if

  if (IKAPTYP >= 4 .and. IKAPTYP <= 7) call READFREEDMANOPCTABLE
  if (IKAPTYP == 6 .or. IKAPTYP == 7) call READCOLUMNDEPTHOPCTABLE
  if ((IKAPTYP >= 9 .and. IKAPTYP <= 12) .or. IKAPTYP == 14 .or. &
      (IKAPTYP >= 20 .and. IKAPTYP <= 25)) call READBURROWSOPCTABLE

  select case (trim(cmd))
  case ('gitter')
  case ('bounds')
     call BOUNDS
  case ('kappa')
     ! KAPPA reads upperbnd, which BOUNDS sets (PASSUPPERBND)
     call BOUNDS
     call KAPPA
     call SUBSIG
  case ('advect')
     call ADVECT3D
  case ('forgas')
     ! FORGAS3D reads the face densities ADVECT3D leaves behind (it calls
     ! QDENS itself only when IADV=0, ppfor.f:167)
     call BOUNDS
     call QDENS
     call FORGAS3D
  case ('energy')
     call COUPLED_ENERGY_EQS
  case ('sor')
     call load3(EVAL1, 'EVAL1'); call load3(EVAL2, 'EVAL2')
     call load3(EVAL3, 'EVAL3'); call load3(EVAL4, 'EVAL4')
     call load3(EVAL5, 'EVAL5'); call load3(EVAL6, 'EVAL6')
     call load3(EVAL7, 'EVAL7'); call load3(EVAL8, 'EVAL8')
     call load3(EVAL_1_6, 'EVAL_1_6'); call load3(RHS_ER, 'RHS_ER')
     call load2i(upperbnd, 'upperbnd')
     call SOLVE_WITH_SOR(.false.)
  case ('adjust')
     call load3(EVAL1, 'EVAL1'); call load3(EVAL2, 'EVAL2')
     call load3(EVAL3, 'EVAL3'); call load3(EVAL4, 'EVAL4')
     call load3(EVAL5, 'EVAL5'); call load3(EVAL6, 'EVAL6')
     call load3(EVAL7, 'EVAL7'); call load3(EVAL8, 'EVAL8')
     call load3(EVAL_1_6, 'EVAL_1_6'); call load3(RHS_ER, 'RHS_ER')
     call load2i(upperbnd, 'upperbnd')
     ! SOR_ITER at each omega ADJUST_SOR_PARAM tries (its own counts are
     ! local to it), then the routine itself picks SORPARAM
     SORPARAM = 1.01
     do s = 1, 50
        call SOLVE_WITH_SOR(.false.)
        adj_iter(s) = SOR_ITER
        SORPARAM = SORPARAM + 0.02
     end do
     SORPARAM = h_sorparam
     call ADJUST_SOR_PARAM
  case ('step')
     ! each step's wall time (system_clock) goes to out_scalars.txt
     call system_clock(count_rate=clk_rate)
     do s = 1, h_nsteps
        III = h_iii + s - 1
        call system_clock(clk0)
        call ADVECT3D
        call FORGAS3D
        call COUPLED_ENERGY_EQS
        call BOUNDS
        call BOUNDS
        call PRESSG
        call system_clock(clk1)
        if (s <= 1000) step_ms(s) = 1.d3 * dble(clk1 - clk0) / dble(clk_rate)
     end do
  case ('diffusion')
     call BOUNDS
     pass_sc(:,:,:) = T(:,:,:)
     call DIFFUSION(1)
     T(:,:,:) = pass_sc(:,:,:)
  case ('printout')
     ! PRINTOUT(1) (ppprt.f:13-124) after one energy solve, with the
     ! namelist's time, counters and initial-profile scalars.  The arrays it
     ! writes that this path does not compute (viscous dissipation, the
     ! convective flux diagnostic) are zero, as pp.f:454 starts DISFN.
     DISFN = 0.d0
     conv_flux = 0.d0
     call COUPLED_ENERGY_EQS
     call PRESSG
     call SOUND
     ZEIT = h_zeit; NUM_ITER = h_num_iter
     TinitBottom = h_tinitbottom; Tinittop = h_tinittop
     rhinittop = h_rhinittop; rhinitbottom = h_rhinitbottom
     hydrotime = h_hydrotime; Maccreated = h_maccreated
     open(12, file=trim(dir)//'/SNAP_000.0001', form='unformatted', status='replace')
     call PRINTOUT(1)
     close(12)
  case ('stepprint')
     ! one step in the pp.f order, then PRINTOUT(1) of the state it leaves:
     ! conv_flux holds that step's ADVECT3D fluxes (ppadv.f:857-865) and
     ! DISFN its FORGAS3D dissipation (ppvisc.f, with IVIS and IDISS); CS
     ! as SCHRITT's SOUND leaves it before the next step prints (pp.f:565-598)
     III = h_iii
     call ADVECT3D
     call FORGAS3D
     call COUPLED_ENERGY_EQS
     call BOUNDS
     call BOUNDS
     call PRESSG
     call SOUND
     ZEIT = h_zeit; NUM_ITER = h_num_iter
     TinitBottom = h_tinitbottom; Tinittop = h_tinittop
     rhinittop = h_rhinittop; rhinitbottom = h_rhinitbottom
     hydrotime = h_hydrotime; Maccreated = h_maccreated
     open(12, file=trim(dir)//'/SNAP_000.0001', form='unformatted', status='replace')
     call PRINTOUT(1)
     close(12)
  case ('restart')
     ! READRESTART (ppinit.f:1543-1600) of the file AKTTD in the working
     ! directory; the global arrays it fills are written out as read
     call allocateglobalvar
     call READRESTART
  case default
     print *, 'unknown cmd ', trim(cmd)
     stop 2
  end select

  ! ---- outputs
  call dump1(xxa, 'xxa'); call dump1(xxb, 'xxb'); call dump1(dxa, 'dxa'); call dump1(dxb, 'dxb')
  call dump1(xya, 'xya'); call dump1(xyb, 'xyb'); call dump1(dya, 'dya'); call dump1(dyb, 'dyb')
  call dump1(xza, 'xza'); call dump1(xzb, 'xzb'); call dump1(dza, 'dza'); call dump1(dzb, 'dzb')
  call dump1(ddx0kep, 'ddx0kep'); call dump1(ddx1kep, 'ddx1kep')
  call dump1(surxa, 'surxa'); call dump1(surxb, 'surxb'); call dump1(volxa, 'volxa'); call dump1(volxb, 'volxb')
  call dump1(surza, 'surza'); call dump1(surzb, 'surzb'); call dump1(volza, 'volza'); call dump1(volzb, 'volzb')
  call dump1(ccza, 'ccza'); call dump1(geozg, 'geozg'); call dump1(geozga, 'geozga')
  call dump1(geoxg, 'geoxg'); call dump1(geoxga, 'geoxga')
  if (trim(cmd) /= 'gitter') then
     call dump3(T, 'T'); call dump3(RH, 'RH'); call dump3(V, 'V'); call dump3(G, 'G')
     call dump3(H, 'H'); call dump3(ER, 'ER'); call dump3(ER_OLD, 'ER_OLD')
     call dump3(delta_ER, 'delta_ER'); call dump3(GD, 'GD'); call dump3(PG, 'PG')
     call dump3(RHQX, 'RHQX'); call dump3(RHQY, 'RHQY'); call dump3(RHQZ, 'RHQZ')
     call dump3(xkapR, 'xkapR'); call dump3(dxkapP_dT, 'dxkapP_dT')
     call dump3(xkapP, 'xkapP'); call dump3(xkapA, 'xkapA'); call dump3(dxkapA_dT, 'dxkapA_dT')
     call dump3(sig, 'sig')
     call dump3(eta, 'eta'); call dump3(divtx, 'divtx'); call dump3(divty, 'divty')
     call dump3(divtz, 'divtz'); call dump3(DISFN, 'DISFN')
     call dump3(STELLARINPUT, 'STELLARINPUT'); call dump3(STELLARTAU, 'STELLARTAU')
     call dump3(DTAUS, 'DTAUS')
     call dump3(FLIMX, 'FLIMX'); call dump3(FLIMY, 'FLIMY'); call dump3(FLIMZ, 'FLIMZ')
     call dump3(difrx, 'difrx'); call dump3(difry, 'difry'); call dump3(difrz, 'difrz')
     call dump3(flux_x, 'flux_x'); call dump3(flux_y, 'flux_y'); call dump3(flux_z, 'flux_z')
     call dump3(EVAL1, 'EVAL1'); call dump3(EVAL7, 'EVAL7'); call dump3(EVAL8, 'EVAL8')
     call dump3(EVAL_1_6, 'EVAL_1_6'); call dump3(Jacob_eta, 'Jacob_eta')
     call dump3(FQ, 'FQ'); call dump3(FR, 'FR'); call dump3(RHS_ER, 'RHS_ER')
     call dump3(SOR_A, 'SOR_A'); call dump3(SOR_B, 'SOR_B'); call dump3(SOR_C, 'SOR_C')
     call dump3(SOR_D, 'SOR_D'); call dump3(SOR_E, 'SOR_E'); call dump3(SOR_F, 'SOR_F')
     call dump3(SOR_G, 'SOR_G')
     call dump3(centcorx, 'centcorx'); call dump3(centcorz, 'centcorz')
     call dump3(conv_flux(1,1:locNX,1:locNY,1:locNZ), 'conv_flux1')
     call dump3(conv_flux(2,1:locNX,1:locNY,1:locNZ), 'conv_flux2')
     call dump3(conv_flux(3,1:locNX,1:locNY,1:locNZ), 'conv_flux3')
     call dump3(CS, 'CS')
     call dump3(gravx, 'gravx')
     call dump2i(upperbnd, 'upperbnd')
  end if
  open(newunit=u, file=trim(dir)//'/out_scalars.txt', status='replace')
  write(u, '(A,1X,I0)') 'SOR_ITER', SOR_ITER
  write(u, '(A,1X,I0)') 'energy_iter', energy_iter
  write(u, '(A,1X,ES25.17)') 'global_RNORM', global_RNORM
  write(u, '(A,1X,ES25.17)') 'global_RHSNORM', global_RHSNORM
  write(u, '(A,1X,ES25.17)') 'FD', FD_c
  write(u, '(A,1X,ES25.17)') 'ARAD', ARAD_c
  write(u, '(A,1X,ES25.17)') 'SORPARAM', SORPARAM
  if (trim(cmd) == 'step') then
     do s = 1, min(h_nsteps, 1000)
        write(u, '(A,I0,1X,ES25.17)') 'step_ms_', s, step_ms(s)
     end do
  end if
  if (IARTVIS > 0) write(u, '(A,1X,ES25.17)') 'dt_artvis', dt_artvis
  if (trim(cmd) == 'restart') then
     write(u, '(A,1X,ES25.17)') 'ZEIT', ZEIT
     write(u, '(A,1X,ES25.17)') 'DELT', DELT
     write(u, '(A,1X,I0)') 'NUM_ITER', NUM_ITER
     write(u, '(A,1X,ES25.17)') 'TinitBottom', TinitBottom
     write(u, '(A,1X,ES25.17)') 'Tinittop', Tinittop
     write(u, '(A,1X,ES25.17)') 'rhinittop', rhinittop
     write(u, '(A,1X,ES25.17)') 'rhinitbottom', rhinitbottom
     write(u, '(A,1X,ES25.17)') 'hydrotime', hydrotime
     write(u, '(A,1X,ES25.17)') 'Maccreated', Maccreated
     call dump3(global_T, 'rs_T'); call dump3(global_ER, 'rs_ER'); call dump3(global_RH, 'rs_RH')
     call dump3(global_V, 'rs_V'); call dump3(global_G, 'rs_G'); call dump3(global_H, 'rs_H')
     call dump2i(global_upperbnd, 'rs_upperbnd')
  end if
  if (trim(cmd) == 'adjust') then
     do s = 1, 50
        write(u, '(A,I2.2,1X,I0)') 'ADJ_ITER_', s, adj_iter(s)
     end do
  end if
  close(u)
  call MPI_FINALIZE(ierr)

contains

  subroutine zero_modules()
    T = 0.d0; ER = 0.d0; ER_OLD = 0.d0; RH = 0.d0; V = 0.d0; G = 0.d0; H = 0.d0
    GD = 0.d0; PG = 0.d0; CS = 0.d0; CV = 0.d0; XMUE = 0.d0
    RHQX = 0.d0; RHQY = 0.d0; RHQZ = 0.d0; STELLARINPUT = 0.d0; upperbnd = 0
    centcorx = 0.d0; centcory = 0.d0; centcorz = 0.d0
    gravx = 0.d0; gravy = 0.d0; gravz = 0.d0; divtx = 0.d0; divty = 0.d0; divtz = 0.d0
    eta = 0.d0; DISFN = 0.d0; conv_flux = 0.d0
    FLIMX = 0.d0; FLIMY = 0.d0; FLIMZ = 0.d0; difrx = 0.d0; difry = 0.d0; difrz = 0.d0
    flux_x = 0.d0; flux_y = 0.d0; flux_z = 0.d0
    xkapR = 0.d0; xkapP = 0.d0; xkapA = 0.d0; gam_kap = 0.d0
    dxkapP_dT = 0.d0; dxkapA_dT = 0.d0; sig = 0.d0; delta_ER = 0.d0
    STELLARTAU = 0.d0; DTAUS = 0.d0; mu_star = 0.d0; saveRRX = 0.d0
    RHS_ER = 0.d0; EVAL1 = 0.d0; EVAL2 = 0.d0; EVAL3 = 0.d0; EVAL4 = 0.d0
    EVAL5 = 0.d0; EVAL6 = 0.d0; EVAL7 = 0.d0; EVAL8 = 0.d0; EVAL_1_6 = 0.d0
    Jacob_eta = 0.d0; FQ = 0.d0; FR = 0.d0
    SOR_A = 0.d0; SOR_B = 0.d0; SOR_C = 0.d0; SOR_D = 0.d0; SOR_E = 0.d0
    SOR_F = 0.d0; SOR_G = 0.d0
    RHS_DIF = 0.d0; DVAL1 = 0.d0; DVAL2 = 0.d0; DVAL3 = 0.d0; DVAL4 = 0.d0
    DVAL5 = 0.d0; DVAL6 = 0.d0; DVAL_1_6 = 0.d0
    difrx_DIF = 0.d0; difry_DIF = 0.d0; difrz_DIF = 0.d0; pass_sc = 0.d0
  end subroutine zero_modules

  subroutine load3(a, name)
    double precision, intent(inout) :: a(:,:,:)
    character(*), intent(in) :: name
    integer :: lu
    open(newunit=lu, file=trim(dir)//'/in_'//name//'.bin', access='stream', &
         form='unformatted', status='old')
    read(lu) a
    close(lu)
  end subroutine load3

  subroutine load2(a, name)
    double precision, intent(inout) :: a(:,:)
    character(*), intent(in) :: name
    integer :: lu
    open(newunit=lu, file=trim(dir)//'/in_'//name//'.bin', access='stream', &
         form='unformatted', status='old')
    read(lu) a
    close(lu)
  end subroutine load2

  subroutine load2i(a, name)
    integer, intent(inout) :: a(:,:)
    character(*), intent(in) :: name
    integer :: lu
    open(newunit=lu, file=trim(dir)//'/in_'//name//'.bin', access='stream', &
         form='unformatted', status='old')
    read(lu) a
    close(lu)
  end subroutine load2i

  subroutine dump3(a, name)
    double precision, intent(in) :: a(:,:,:)
    character(*), intent(in) :: name
    integer :: lu
    open(newunit=lu, file=trim(dir)//'/out_'//name//'.bin', access='stream', &
         form='unformatted', status='replace')
    write(lu) a
    close(lu)
  end subroutine dump3

  subroutine dump2i(a, name)
    integer, intent(in) :: a(:,:)
    character(*), intent(in) :: name
    integer :: lu
    open(newunit=lu, file=trim(dir)//'/out_'//name//'.bin', access='stream', &
         form='unformatted', status='replace')
    write(lu) a
    close(lu)
  end subroutine dump2i

  subroutine dump1(a, name)
    double precision, intent(in) :: a(:)
    character(*), intent(in) :: name
    integer :: lu
    open(newunit=lu, file=trim(dir)//'/out_'//name//'.bin', access='stream', &
         form='unformatted', status='replace')
    write(lu) a
    close(lu)
  end subroutine dump1

end program tt_ref_harness
