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

  subroutine sg_projection_finalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = sg_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine sg_projection_finalize

end module sg_projection_dropin
