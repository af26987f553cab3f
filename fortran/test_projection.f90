! Fortran host check of the drop-in projection: a 32x64 liquid column with
! an air bubble, gravity in the predictor velocities, walls closed, top open.
! Calls the drop-in PoissonEquationSolver body, checks the reference's own
! acceptance test |(A p)_i - b_i| <= 1e-10 (ProjectionP_Mod.f90:70-93) on
! the rows the library assembled, then the correction's discrete continuity
! (ComputePUV_Mod.f90:184-217).
program test_projection
  use iso_c_binding
  use simcogpu
  use sg_projection_dropin
  implicit none
  integer, parameter :: ni = 32, nj = 64
  real(c_double), target :: dx(ni), dy(nj), udx(ni), vdy(nj)
  real(c_double), target :: vofS(ni, nj), nyS(ni, nj), WlLh(ni, nj), vofSU(ni, nj), vofSV(ni, nj)
  real(c_double), target :: aE(0:ni+1, 0:nj+1), aW(0:ni+1, 0:nj+1), aN(0:ni+1, 0:nj+1), aS(0:ni+1, 0:nj+1)
  integer(c_int), target :: uflag(ni, nj), vflag(ni, nj)
  real(c_double), target :: u(0:ni+1, 0:nj+1), v(0:ni+1, 0:nj+1), p(0:ni+1, 0:nj+1)
  real(c_double), target :: upred(0:ni+1, 0:nj+1), vpred(0:ni+1, 0:nj+1), Pp(0:ni+1, 0:nj+1)
  real(c_double), target :: DpU(0:ni+1, 0:nj+1), DpV(0:ni+1, 0:nj+1)
  real(c_double), target :: vof(ni, nj), phi(ni, nj), mres(ni, nj)
  real(c_double), target :: matr(ni, nj, 5), rhs(ni, nj)
  type(sg_params) :: prm
  type(sg_solve_info) :: info
  integer :: i, j, k, ks
  integer(c_int) :: rc
  real(c_double) :: h, dt, xc, yc, d, resi, worst, frac_in
  h = 1.d0 / ni
  dx = h; dy = h; udx = h; vdy = h
  vofS = 0.d0; nyS = 1.d0; WlLh = 0.d0; vofSU = 0.d0; vofSV = 0.d0
  aE = 0.d0; aW = 0.d0; aN = 0.d0; aS = 0.d0
  aE(1:ni, 1:nj) = 1.d0; aW(1:ni, 1:nj) = 1.d0; aN(1:ni, 1:nj) = 1.d0; aS(1:ni, 1:nj) = 1.d0
  uflag = 1; vflag = 1
  ! bubble of radius 0.25 at (0.5, 0.5): phi > 0 in the gas, vof by 8x8 sub-sampling
  do j = 1, nj
    do i = 1, ni
      xc = (i - 0.5d0) * h; yc = (j - 0.5d0) * h
      phi(i, j) = 0.25d0 - sqrt((xc - 0.5d0)**2 + (yc - 0.5d0)**2)
      frac_in = 0.d0
      do k = 0, 7
        do ks = 0, 7
          d = 0.25d0 - sqrt((xc - h/2 + (k + 0.5d0) * h / 8 - 0.5d0)**2 + (yc - h/2 + (ks + 0.5d0) * h / 8 - 0.5d0)**2)
          if (d < 0.d0) frac_in = frac_in + 1.d0 / 64.d0
        end do
      end do
      vof(i, j) = frac_in
    end do
  end do
  dt = 0.25d0 * h
  DpU = dt; DpV = dt
  upred = 0.d0
  vpred = -9.81d0 * dt
  vpred(:, 0) = 0.d0
  u = upred; v = vpred; p = 0.d0; Pp = 0.d0
  rc = sg_params_default(prm)
  prm%rho_l = 1000.d0; prm%rho_g = 1.d0; prm%rho_ref = 1000.d0
  prm%tol = 1.d-11
  call sg_projection_init(ni, nj, dx, dy, udx, vdy, prm, vofS, nyS, WlLh, aE, aW, aN, aS, uflag, vflag, vofSU, vofSV)
  call sg_poisson_equation_solver(ni, nj, upred, vpred, DpU, DpV, vof, phi, 0.d0, Pp, info)
  ! the reference's residual assertion on the assembled rows
  rc = sg_poisson_matrix_ptr()
  worst = 0.d0
  do i = 1, ni
    do j = 1, nj
      resi = 0.d0
      if (j > 1) resi = resi + matr(i, j, 1) * Pp(i, j - 1)
      if (i > 1) resi = resi + matr(i, j, 2) * Pp(i - 1, j)
      resi = resi + matr(i, j, 3) * Pp(i, j)
      if (i < ni) resi = resi + matr(i, j, 4) * Pp(i + 1, j)
      if (j < nj) resi = resi + matr(i, j, 5) * Pp(i, j + 1)
      worst = max(worst, abs(resi - rhs(i, j)))
    end do
  end do
  call sg_update_puv_correction(ni, nj, u, v, p, upred, vpred, Pp, DpU, DpV, vof, phi, 0.d0, mres)
  write(*, '(a,i0,a,es10.3,a,es10.3,a,es10.3)') 'iterations ', info%iterations, ' relres ', info%relres, &
       ' max row residual ', worst, ' max |mres| ', maxval(abs(mres))
  call sg_projection_finalize()
  if (worst > 1.d-10 .or. maxval(abs(mres)) > 1.d-10 .or. info%status /= SG_OK) then
    write(*, *) 'FORTRAN DROP-IN FAILED'
    error stop 2
  end if
  write(*, *) 'FORTRAN DROP-IN OK'

contains

  ! rows and RHS of the built system (the drop-in context is private to the
  ! module, so re-run the solve's inputs through a second context)
  integer(c_int) function sg_poisson_matrix_ptr() result(r)
    type(c_ptr) :: c2
    r = sg_ctx_create(int(ni, c_int), int(nj, c_int), c_loc(dx), c_loc(dy), c_loc(udx), c_loc(vdy), prm, c2)
    if (r /= SG_OK) return
    r = sg_set_geometry(c2, c_loc(vofS), c_loc(nyS), c_loc(WlLh), c_loc(aE), c_loc(aW), c_loc(aN), c_loc(aS), &
                        c_loc(uflag), c_loc(vflag), c_loc(vofSU), c_loc(vofSV))
    r = sg_upload(c2, SG_F_UPRED, c_loc(upred))
    r = sg_upload(c2, SG_F_VPRED, c_loc(vpred))
    r = sg_upload(c2, SG_F_DPU, c_loc(DpU))
    r = sg_upload(c2, SG_F_DPV, c_loc(DpV))
    r = sg_upload(c2, SG_F_VOF, c_loc(vof))
    r = sg_upload(c2, SG_F_PHI, c_loc(phi))
    r = sg_poisson_build_resident(c2, 0.d0)
    r = sg_poisson_matrix(c2, c_loc(matr))
    r = sg_poisson_rhs(c2, c_loc(rhs))
    r = sg_ctx_destroy(c2)
  end function sg_poisson_matrix_ptr

end program test_projection
