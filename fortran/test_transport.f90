! Fortran host check of the drop-in transport and predictor bodies on a
! 32x32 closed box (h = 1/32, no solid):
!  1. Coupled_LS_VOF: an air bubble in liquid turned by a discretely
!     divergence-free vortex (stream function on the cell corners): the total
!     liquid volume is kept to 1e-6 relative and vof stays in [0, 1] to
!     round-off (the explicit sweep is not clipped, Clsvof_Mod.f90:844-850);
!  2. Initial_ClsVofUV: a horizontal surface at y = 0.4: a U cell spans the
!     rows of its P cells, so it takes their fraction; a V cell spans
!     yc_j..yc_j+h, so it takes clamp((0.4 - yc_j)/h, 0, 1);
!  3. Predictor_UV for liquid at rest under gravity: the V rows away from the
!     walls gain exactly -g dt, U stays 0, and Dp = dt.
program test_transport
  use iso_c_binding
  use simcogpu
  use sg_projection_dropin
  implicit none
  integer, parameter :: ni = 32, nj = 32
  real(c_double), parameter :: pi = 3.14159265358979323846d0, g = 9.81d0
  real(c_double), target :: dx(ni), dy(nj), udx(ni), vdy(nj), xc(ni), yc(nj)
  real(c_double), target, dimension(ni, nj) :: vofS, nyS, nxS, phiS, WlLh, vofSU, vofSV, vof, phi, nx, ny
  real(c_double), target, dimension(ni, nj) :: uvof, uphi, unx, uny, vvof, vphi, vnx, vny, vofo
  real(c_double), target, dimension(ni, nj) :: sz, sone, sfar, sdelh
  integer(c_int), target :: uflag(ni, nj), vflag(ni, nj), moex(ni, nj), active(ni, nj)
  real(c_double), target, dimension(0:ni+1, 0:nj+1) :: aE, aW, aN, aS, u, v, DpU, DpV, upred, vpred, psi
  real(c_double), target, dimension(0:ni+1, 0:nj+1) :: eta, one, zero, sx, half_h
  real(c_double), target :: fluxn1(0:ni+1, 0:nj+1, 2)
  type(sg_params) :: prm
  type(sg_clsvof_params) :: cp
  type(sg_predictor_info) :: pinfo
  integer :: i, j, k, ks, w
  integer(c_int) :: rc
  real(c_double) :: h, dt, d, vol0, vol1, err_uv, err_g, err_dp, err_u
  logical :: ok
  h = 1.d0 / ni
  dx = h; dy = h; udx = h; vdy = h
  do i = 1, ni
    xc(i) = (i - 0.5d0) * h
  end do
  do j = 1, nj
    yc(j) = (j - 0.5d0) * h
  end do
  vofS = 0.d0; nyS = 1.d0; nxS = 0.d0; phiS = 1.d3; WlLh = 0.d0; vofSU = 0.d0; vofSV = 0.d0
  aE = 0.d0; aW = 0.d0; aN = 0.d0; aS = 0.d0
  aE(1:ni, 1:nj) = 1.d0; aW(1:ni, 1:nj) = 1.d0; aN(1:ni, 1:nj) = 1.d0; aS(1:ni, 1:nj) = 1.d0
  uflag = 1; vflag = 1
  rc = sg_params_default(prm)
  prm%rho_l = 1000.d0; prm%rho_g = 1.d0; prm%rho_ref = 1000.d0
  prm%nu_l = 1.d-6; prm%nu_g = 1.d-5; prm%nu_ref = 1.d-6; prm%Re = 1.d6; prm%Fr = 1.d0 / sqrt(g)
  call sg_projection_init(ni, nj, dx, dy, udx, vdy, prm, vofS, nyS, WlLh, aE, aW, aN, aS, uflag, vflag, vofSU, vofSV)
  call sg_transport_init(ni, nj, xc, yc, nxS, phiS)
  ok = .true.

  ! ---- 1. transport ----------------------------------------------------------
  do j = 1, nj
    do i = 1, ni
      phi(i, j) = 0.2d0 - sqrt((xc(i) - 0.5d0)**2 + (yc(j) - 0.45d0)**2)
      vof(i, j) = 0.d0
      do k = 0, 7
        do ks = 0, 7
          d = 0.2d0 - sqrt((xc(i) - h/2 + (k + 0.5d0) * h / 8 - 0.5d0)**2 + &
                           (yc(j) - h/2 + (ks + 0.5d0) * h / 8 - 0.45d0)**2)
          if (d < 0.d0) vof(i, j) = vof(i, j) + 1.d0 / 64.d0
        end do
      end do
    end do
  end do
  ! psi(i,j) at the corner (i*h, j*h), zero on the walls
  do j = 0, nj
    do i = 0, ni
      psi(i, j) = 0.5d0 / pi * sin(pi * i * h)**2 * sin(pi * j * h)**2
    end do
  end do
  u = 0.d0; v = 0.d0
  do j = 1, nj
    do i = 0, ni
      u(i, j) = (psi(i, j) - psi(i, j - 1)) / h
    end do
  end do
  do j = 0, nj
    do i = 1, ni
      v(i, j) = -(psi(i, j) - psi(i - 1, j)) / h
    end do
  end do
  vol0 = sum(vof)
  dt = 0.25d0 * h
  rc = sg_clsvof_params_default(cp)
  do k = 1, 3
    call sg_coupled_ls_vof(ni, nj, vof, phi, nx, ny, u, v, dt, (k - 1) * dt, cp)
  end do
  vol1 = sum(vof)
  write(*, '(a,es10.3,a,2es11.3)') 'transport: rel volume change ', abs(vol1 - vol0) / vol0, &
       '  vof range ', minval(vof), maxval(vof)
  if (abs(vol1 - vol0) / vol0 > 1.d-6 .or. minval(vof) < -1.d-12 .or. maxval(vof) > 1.d0 + 1.d-12) ok = .false.

  ! ---- staggered geometry (no solid) ----------------------------------------
  sz = 0.d0; sone = 1.d0; sfar = 1.d3; sdelh = 1.d3
  moex = 0; active = 1
  eta = 0.5d0; one = 1.d0; zero = 0.d0; sx = h; half_h = 0.5d0 * h
  do w = 0, 1
    call sg_predictor_geometry(w, ni, nj, sz, sz, sone, sfar, sz, sdelh, sz, sz, moex, active, &
                               aE, aW, aN, aS, eta, eta, one, one, zero, zero, sx, sx, half_h, zero, zero, half_h)
  end do

  ! ---- 2. interface on the U / V cells ----------------------------------------
  do j = 1, nj
    do i = 1, ni
      vof(i, j) = max(0.d0, min(1.d0, (0.4d0 - (j - 1) * h) / h))
      phi(i, j) = yc(j) - 0.4d0
    end do
  end do
  nx = 0.d0; ny = 1.d0
  call sg_initial_clsvof_uv(ni, nj, vof, phi, nx, ny, uvof, uphi, unx, uny, vvof, vphi, vnx, vny)
  err_uv = 0.d0
  do j = 1, nj
    do i = 1, ni
      err_uv = max(err_uv, abs(uvof(i, j) - vof(i, j)))
      if (j < nj) err_uv = max(err_uv, abs(vvof(i, j) - max(0.d0, min(1.d0, (0.4d0 - yc(j)) / h))))
      err_uv = max(err_uv, abs(uphi(i, j) - phi(i, j)), abs(uny(i, j) - 1.d0))
    end do
  end do
  write(*, '(a,es10.3)') 'interface to U/V: max error ', err_uv
  if (err_uv > 1.d-12) ok = .false.

  ! ---- 3. predictor: liquid at rest under gravity -----------------------------
  vof = 1.d0; phi = -1.d0; nx = 0.d0; ny = 1.d0; vofo = 1.d0
  uvof = 1.d0; vvof = 1.d0
  u = 0.d0; v = 0.d0; fluxn1 = 0.d0; DpU = 0.d0; DpV = 0.d0; upred = 0.d0; vpred = 0.d0
  dt = 0.004d0
  call sg_predictor_uv(ni, nj, u, v, vofo, phi, nx, ny, uvof, vvof, uvof, vvof, fluxn1, DpU, DpV, upred, vpred, &
                       dt, 0.d0, 1_c_int64_t, pinfo)
  err_g = maxval(abs(vpred(4:ni-3, 4:nj-4) / (-g * dt) - 1.d0))
  err_u = maxval(abs(upred(1:ni-1, 1:nj)))
  err_dp = max(maxval(abs(DpU(1:ni-1, 1:nj) - dt)), maxval(abs(DpV(1:ni, 1:nj-1) - dt)))
  write(*, '(a,es10.3,a,es10.3,a,es10.3,a,2i4)') 'predictor: gravity error ', err_g, '  |u| ', err_u, &
       '  Dp error ', err_dp, '  iterations ', pinfo%iterations_u, pinfo%iterations_v
  if (err_g > 1.d-12 .or. err_u > 1.d-15 .or. err_dp > 1.d-15) ok = .false.

  call sg_projection_finalize()
  if (.not. ok) then
    write(*, *) 'FORTRAN TRANSPORT DROP-IN FAILED'
    error stop 2
  end if
  write(*, *) 'FORTRAN TRANSPORT DROP-IN OK'
end program test_transport
