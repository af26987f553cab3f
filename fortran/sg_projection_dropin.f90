! Drop-in bodies for the reference's projection calls.
!
!   sg_projection_init        once per grid, after Grid_Preprocess
!                             (Cutcell_Mod.f90:18-32) has built the cut cells
!   sg_poisson_equation_solver  replaces the body of PoissonEquationSolver
!                             (ProjectionP_Mod.f90:26-99): same inputs
!                             (TPred%u/v, PU%Dp, PV%Dp, PCell%vof/phi, vb),
!                             same output (Proj%Pp with its ghost cells)
!   sg_update_puv_correction  replaces the correction block of UpdatePUV
!                             (ComputePUV_Mod.f90:59-217) and
!                             VariablesInternalCellCondition (:227-245)
!
!   sg_transport_init         once per grid after sg_projection_init: P-cell
!                             centres and the solid line (PCell%nxS, phiS)
!   sg_coupled_ls_vof         replaces the body of Coupled_LS_VOF
!                             (Clsvof_Mod.f90:730-959) for solids at rest
!   sg_initial_clsvof_uv      replaces Initial_ClsVofUV (Clsvof_Mod.f90:402-710)
!   sg_predictor_geometry     U / V cell geometry of NewCellFace
!                             (Cutcell_Mod.f90:399-725), once per grid
!   sg_predictor_uv           replaces the body of Predictor_UV
!                             (PredictorUV_mod.f90:28-551)
!
! The reference's `pause` on a failed check becomes `error stop` with the
! library's message.
module sg_projection_dropin
  use iso_c_binding
  use simcogpu
  implicit none
  private
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: ni_ctx = 0, nj_ctx = 0
  public :: sg_projection_init, sg_poisson_equation_solver, sg_update_puv_correction, sg_projection_finalize
  public :: sg_transport_init, sg_coupled_ls_vof, sg_initial_clsvof_uv, sg_predictor_geometry, sg_predictor_uv

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= SG_OK) then
      write(*, '(a,a,a,i0,a,a)') 'simcogpu: ', what, ' failed (', rc, '): ', sg_error_message()
      error stop 1
    end if
  end subroutine check

  subroutine sg_projection_init(Isize, Jsize, dx, dy, udx, vdy, params, vofS, nyS, WlLh, aE, aW, aN, aS, &
                                uflag, vflag, vofSU, vofSV)
    integer, intent(in) :: Isize, Jsize
    ! PGrid%dx(:,1), PGrid%dy(1,:), UGrid%dx(:,1), VGrid%dy(1,:)
    real(c_double), intent(in), target :: dx(Isize), dy(Jsize), udx(Isize), vdy(Jsize)
    type(sg_params), intent(in) :: params
    real(c_double), intent(in), target :: vofS(Isize, Jsize), nyS(Isize, Jsize), WlLh(Isize, Jsize)
    real(c_double), intent(in), target :: aE(0:Isize+1, 0:Jsize+1), aW(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in), target :: aN(0:Isize+1, 0:Jsize+1), aS(0:Isize+1, 0:Jsize+1)
    ! 1 where UCell%Posnu/=-1 .and. UCell%MoExCell/=1 (resp. V), else 0
    integer(c_int), intent(in), target :: uflag(Isize, Jsize), vflag(Isize, Jsize)
    real(c_double), intent(in), target :: vofSU(Isize, Jsize), vofSV(Isize, Jsize)
    if (c_associated(ctx)) call sg_projection_finalize()
    call check(sg_ctx_create(int(Isize, c_int), int(Jsize, c_int), c_loc(dx), c_loc(dy), c_loc(udx), &
                             c_loc(vdy), params, ctx), 'sg_ctx_create')
    ni_ctx = Isize
    nj_ctx = Jsize
    call check(sg_set_geometry(ctx, c_loc(vofS), c_loc(nyS), c_loc(WlLh), c_loc(aE), c_loc(aW), c_loc(aN), &
                               c_loc(aS), c_loc(uflag), c_loc(vflag), c_loc(vofSU), c_loc(vofSV)), 'sg_set_geometry')
  end subroutine sg_projection_init

  subroutine sg_poisson_equation_solver(Isize, Jsize, u, v, DpU, DpV, vof, phi, vb, Pp, info)
    integer, intent(in) :: Isize, Jsize
    real(c_double), intent(in), target :: u(0:Isize+1, 0:Jsize+1), v(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in), target :: DpU(0:Isize+1, 0:Jsize+1), DpV(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in), target :: vof(Isize, Jsize), phi(Isize, Jsize)
    real(c_double), intent(in) :: vb
    real(c_double), intent(inout), target :: Pp(0:Isize+1, 0:Jsize+1)
    type(sg_solve_info), intent(out) :: info
    if (Isize /= ni_ctx .or. Jsize /= nj_ctx) then
      write(*, *) 'simcogpu: grid size differs from sg_projection_init'
      error stop 1
    end if
    call check(sg_poisson_solve(ctx, c_loc(u), c_loc(v), c_loc(DpU), c_loc(DpV), c_loc(vof), c_loc(phi), &
                                real(vb, c_double), c_loc(Pp), info), 'PoissonEquationSolver')
  end subroutine sg_poisson_equation_solver

  ! TVar%u/v/p in/out; Pred%u/v, Proj%Pp, PU%Dp, PV%Dp, PCell%vof/phi in;
  ! mres = TVar%mres out.  Uses the geometry given to sg_projection_init.
  subroutine sg_update_puv_correction(Isize, Jsize, u, v, p, upred, vpred, Pp, DpU, DpV, vof, phi, vb, mres)
    integer, intent(in) :: Isize, Jsize
    real(c_double), intent(inout), target :: u(0:Isize+1, 0:Jsize+1), v(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(inout), target :: p(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in), target :: upred(0:Isize+1, 0:Jsize+1), vpred(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in), target :: Pp(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in), target :: DpU(0:Isize+1, 0:Jsize+1), DpV(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in), target :: vof(Isize, Jsize), phi(Isize, Jsize)
    real(c_double), intent(in) :: vb
    real(c_double), intent(out), target :: mres(Isize, Jsize)
    call check(sg_upload(ctx, SG_F_U, c_loc(u)), 'upload u')
    call check(sg_upload(ctx, SG_F_V, c_loc(v)), 'upload v')
    call check(sg_upload(ctx, SG_F_P, c_loc(p)), 'upload p')
    call check(sg_upload(ctx, SG_F_UPRED, c_loc(upred)), 'upload upred')
    call check(sg_upload(ctx, SG_F_VPRED, c_loc(vpred)), 'upload vpred')
    call check(sg_upload(ctx, SG_F_PP, c_loc(Pp)), 'upload Pp')
    call check(sg_upload(ctx, SG_F_DPU, c_loc(DpU)), 'upload DpU')
    call check(sg_upload(ctx, SG_F_DPV, c_loc(DpV)), 'upload DpV')
    call check(sg_upload(ctx, SG_F_VOF, c_loc(vof)), 'upload vof')
    call check(sg_upload(ctx, SG_F_PHI, c_loc(phi)), 'upload phi')
    call check(sg_velocity_correct_resident(ctx, real(vb, c_double)), 'UpdatePUV correction')
    call check(sg_download(ctx, SG_F_U, c_loc(u)), 'download u')
    call check(sg_download(ctx, SG_F_V, c_loc(v)), 'download v')
    call check(sg_download(ctx, SG_F_P, c_loc(p)), 'download p')
    call check(sg_download(ctx, SG_F_MRES, c_loc(mres)), 'download mres')
  end subroutine sg_update_puv_correction

  subroutine sg_transport_init(Isize, Jsize, xc, yc, nxS, phiS)
    integer, intent(in) :: Isize, Jsize
    ! PGrid%x(:,1), PGrid%y(1,:); PCell%nxS, PCell%phiS
    real(c_double), intent(in), target :: xc(Isize), yc(Jsize), nxS(Isize, Jsize), phiS(Isize, Jsize)
    call check_size(Isize, Jsize)
    call check(sg_set_centres(ctx, c_loc(xc), c_loc(yc)), 'sg_set_centres')
    call check(sg_upload(ctx, SG_F_NXS, c_loc(nxS)), 'upload nxS')
    call check(sg_upload(ctx, SG_F_PHIS, c_loc(phiS)), 'upload phiS')
  end subroutine sg_transport_init

  ! PCell%vof, phi in/out, PCell%nx, ny out; TVar%u, v in; NondiT = time
  subroutine sg_coupled_ls_vof(Isize, Jsize, vof, phi, nx, ny, u, v, dt, time, params)
    integer, intent(in) :: Isize, Jsize
    real(c_double), intent(inout), target :: vof(Isize, Jsize), phi(Isize, Jsize)
    real(c_double), intent(out), target :: nx(Isize, Jsize), ny(Isize, Jsize)
    real(c_double), intent(in), target :: u(0:Isize+1, 0:Jsize+1), v(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in) :: dt, time
    type(sg_clsvof_params), intent(in) :: params
    call check_size(Isize, Jsize)
    call check(sg_clsvof_advect(ctx, c_loc(vof), c_loc(phi), c_loc(nx), c_loc(ny), c_loc(u), c_loc(v), &
                                real(dt, c_double), real(time, c_double), params), 'Coupled_LS_VOF')
  end subroutine sg_coupled_ls_vof

  ! PCell%vof, phi, nx, ny in (nxS/phiS from sg_transport_init, the U/V
  ! solid fraction and line from sg_predictor_geometry); U/VCell%vof, phi,
  ! nx, ny out
  subroutine sg_initial_clsvof_uv(Isize, Jsize, vof, phi, nx, ny, uvof, uphi, unx, uny, vvof, vphi, vnx, vny)
    integer, intent(in) :: Isize, Jsize
    real(c_double), intent(in), target :: vof(Isize, Jsize), phi(Isize, Jsize), nx(Isize, Jsize), ny(Isize, Jsize)
    real(c_double), intent(out), target :: uvof(Isize, Jsize), uphi(Isize, Jsize), unx(Isize, Jsize), &
                                           uny(Isize, Jsize)
    real(c_double), intent(out), target :: vvof(Isize, Jsize), vphi(Isize, Jsize), vnx(Isize, Jsize), &
                                           vny(Isize, Jsize)
    call check_size(Isize, Jsize)
    call check(sg_upload(ctx, SG_F_VOF, c_loc(vof)), 'upload vof')
    call check(sg_upload(ctx, SG_F_PHI, c_loc(phi)), 'upload phi')
    call check(sg_upload(ctx, SG_F_NX, c_loc(nx)), 'upload nx')
    call check(sg_upload(ctx, SG_F_NY, c_loc(ny)), 'upload ny')
    call check(sg_interface_to_uv_resident(ctx), 'Initial_ClsVofUV')
    call check(sg_download(ctx, SG_F_UVOF, c_loc(uvof)), 'download uvof')
    call check(sg_download(ctx, SG_F_UPHI, c_loc(uphi)), 'download uphi')
    call check(sg_download(ctx, SG_F_UNX, c_loc(unx)), 'download unx')
    call check(sg_download(ctx, SG_F_UNY, c_loc(uny)), 'download uny')
    call check(sg_download(ctx, SG_F_VVOF, c_loc(vvof)), 'download vvof')
    call check(sg_download(ctx, SG_F_VPHI, c_loc(vphi)), 'download vphi')
    call check(sg_download(ctx, SG_F_VNX, c_loc(vnx)), 'download vnx')
    call check(sg_download(ctx, SG_F_VNY, c_loc(vny)), 'download vny')
  end subroutine sg_initial_clsvof_uv

  ! which = 0: UCell/UGrid, 1: VCell/VGrid.  Cell arrays: vofS, nxS, nyS,
  ! phiS, WlLh, delh, Cell_Cent(:,:,1:2) as ccx/ccy, MoExCell, active
  ! (Posnu /= -1); ghosted: the edge apertures, EtaE/N, AlE/N, DAlE/N, SxE,
  ! SyN and FCE/FCN(:,:,1:2)
  subroutine sg_predictor_geometry(which, Isize, Jsize, vofS, nxS, nyS, phiS, WlLh, delh, ccx, ccy, moex, active, &
                                   aE, aW, aN, aS, EtaE, EtaN, AlE, AlN, DAlE, DAlN, SxE, SyN, fcex, fcey, fcnx, fcny)
    integer, intent(in) :: which, Isize, Jsize
    real(c_double), intent(in), target, dimension(Isize, Jsize) :: vofS, nxS, nyS, phiS, WlLh, delh, ccx, ccy
    integer(c_int), intent(in), target, dimension(Isize, Jsize) :: moex, active
    real(c_double), intent(in), target, dimension(0:Isize+1, 0:Jsize+1) :: aE, aW, aN, aS, EtaE, EtaN, AlE, AlN, &
         DAlE, DAlN, SxE, SyN, fcex, fcey, fcnx, fcny
    type(sg_stag_geometry) :: g
    call check_size(Isize, Jsize)
    g = sg_stag_geometry(c_loc(vofS), c_loc(nxS), c_loc(nyS), c_loc(phiS), c_loc(WlLh), c_loc(delh), c_loc(ccx), &
                         c_loc(ccy), c_loc(moex), c_loc(active), c_loc(aE), c_loc(aW), c_loc(aN), c_loc(aS), &
                         c_loc(EtaE), c_loc(EtaN), c_loc(AlE), c_loc(AlN), c_loc(DAlE), c_loc(DAlN), c_loc(SxE), &
                         c_loc(SyN), c_loc(fcex), c_loc(fcey), c_loc(fcnx), c_loc(fcny))
    call check(sg_set_staggered_geometry(ctx, int(which, c_int), g), 'sg_set_staggered_geometry')
  end subroutine sg_predictor_geometry

  ! TVar%u, v and PCellO%vof, phi, nx, ny (time n) in; UCellO/VCellO%vof
  ! (time n) and UCell/VCell%vof (after the transport) in; Flux_n1 (both
  ! planes), PU%Dp, PV%Dp in/out; Pred%u, v out.  itt as in the reference
  ! (1 = first step: first-order fluxes, no Adams-Bashforth).
  subroutine sg_predictor_uv(Isize, Jsize, u, v, vofo, phio, nxo, nyo, uvofo, vvofo, uvof, vvof, fluxn1, DpU, DpV, &
                             upred, vpred, dt, vb, itt, info)
    integer, intent(in) :: Isize, Jsize
    real(c_double), intent(in), target :: u(0:Isize+1, 0:Jsize+1), v(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in), target, dimension(Isize, Jsize) :: vofo, phio, nxo, nyo, uvofo, vvofo, uvof, vvof
    real(c_double), intent(inout), target :: fluxn1(0:Isize+1, 0:Jsize+1, 2)
    real(c_double), intent(inout), target :: DpU(0:Isize+1, 0:Jsize+1), DpV(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(inout), target :: upred(0:Isize+1, 0:Jsize+1), vpred(0:Isize+1, 0:Jsize+1)
    real(c_double), intent(in) :: dt, vb
    integer(c_int64_t), intent(in) :: itt
    type(sg_predictor_info), intent(out) :: info
    integer(c_int) :: first
    call check_size(Isize, Jsize)
    call check(sg_upload(ctx, SG_F_U, c_loc(u)), 'upload u')
    call check(sg_upload(ctx, SG_F_V, c_loc(v)), 'upload v')
    call check(sg_upload(ctx, SG_F_VOFO, c_loc(vofo)), 'upload vofo')
    call check(sg_upload(ctx, SG_F_PHIO, c_loc(phio)), 'upload phio')
    call check(sg_upload(ctx, SG_F_NXO, c_loc(nxo)), 'upload nxo')
    call check(sg_upload(ctx, SG_F_NYO, c_loc(nyo)), 'upload nyo')
    call check(sg_upload(ctx, SG_F_UVOFO, c_loc(uvofo)), 'upload uvofo')
    call check(sg_upload(ctx, SG_F_VVOFO, c_loc(vvofo)), 'upload vvofo')
    call check(sg_upload(ctx, SG_F_UVOF, c_loc(uvof)), 'upload uvof')
    call check(sg_upload(ctx, SG_F_VVOF, c_loc(vvof)), 'upload vvof')
    call check(sg_upload(ctx, SG_F_FLUXN1U, c_loc(fluxn1(0, 0, 1))), 'upload Flux_n1(:,:,1)')
    call check(sg_upload(ctx, SG_F_FLUXN1V, c_loc(fluxn1(0, 0, 2))), 'upload Flux_n1(:,:,2)')
    call check(sg_upload(ctx, SG_F_DPU, c_loc(DpU)), 'upload DpU')
    call check(sg_upload(ctx, SG_F_DPV, c_loc(DpV)), 'upload DpV')
    call check(sg_upload(ctx, SG_F_UPRED, c_loc(upred)), 'upload upred')
    call check(sg_upload(ctx, SG_F_VPRED, c_loc(vpred)), 'upload vpred')
    first = 0
    if (itt == 1) first = 1
    call check(sg_predictor_resident(ctx, real(dt, c_double), real(vb, c_double), first, info), 'Predictor_UV')
    call check(sg_download(ctx, SG_F_FLUXN1U, c_loc(fluxn1(0, 0, 1))), 'download Flux_n1(:,:,1)')
    call check(sg_download(ctx, SG_F_FLUXN1V, c_loc(fluxn1(0, 0, 2))), 'download Flux_n1(:,:,2)')
    call check(sg_download(ctx, SG_F_DPU, c_loc(DpU)), 'download DpU')
    call check(sg_download(ctx, SG_F_DPV, c_loc(DpV)), 'download DpV')
    call check(sg_download(ctx, SG_F_UPRED, c_loc(upred)), 'download upred')
    call check(sg_download(ctx, SG_F_VPRED, c_loc(vpred)), 'download vpred')
  end subroutine sg_predictor_uv

  subroutine check_size(Isize, Jsize)
    integer, intent(in) :: Isize, Jsize
    if (.not. c_associated(ctx) .or. Isize /= ni_ctx .or. Jsize /= nj_ctx) then
      write(*, *) 'simcogpu: call sg_projection_init for this grid first'
      error stop 1
    end if
  end subroutine check_size

  subroutine sg_projection_finalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = sg_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine sg_projection_finalize

end module sg_projection_dropin
