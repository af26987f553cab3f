! iso_c_binding interface to libsimcogpu (include/simcogpu.h).
!
! A Fortran host (the reference's Solver_Mod / ProjectionP_Mod callers) uses
! this module to call the MI355X projection path.  Arrays are passed as
! c_loc of their first element; the layouts are the reference's own:
! cell arrays (1:Isize,1:Jsize), ghosted arrays (0:Isize+1,0:Jsize+1).
module simcogpu
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: SG_OK = 0, SG_NOT_CONVERGED = 1, SG_NAN = 2, SG_GUARD = 3
  integer(c_int), parameter, public :: SG_PREC_FP64 = 0, SG_PREC_MIXED = 1
  ! enum sg_field
  integer(c_int), parameter, public :: SG_F_VOF = 0, SG_F_PHI = 1, SG_F_VOFS = 2, SG_F_NYS = 3, &
       SG_F_WLLH = 4, SG_F_NXS = 5, SG_F_PHIS = 6, SG_F_NX = 7, SG_F_NY = 8, SG_F_MRES = 9, &
       SG_F_VOFSU = 10, SG_F_VOFSV = 11, SG_F_U = 12, SG_F_V = 13, SG_F_P = 14, SG_F_UPRED = 15, &
       SG_F_VPRED = 16, SG_F_PP = 17, SG_F_DPU = 18, SG_F_DPV = 19, SG_F_AE = 20, SG_F_AW = 21, &
       SG_F_AN = 22, SG_F_AS = 23, SG_F_UFLAG = 24, SG_F_VFLAG = 25, SG_F_VOFO = 26, SG_F_PHIO = 27, &
       SG_F_NXO = 28, SG_F_NYO = 29, SG_F_UVOF = 30, SG_F_VVOF = 31, SG_F_UVOFO = 32, SG_F_VVOFO = 33, &
       SG_F_FLUXN1U = 34, SG_F_FLUXN1V = 35, SG_F_UPHI = 36, SG_F_UNX = 37, SG_F_UNY = 38, &
       SG_F_VPHI = 39, SG_F_VNX = 40, SG_F_VNY = 41

  type, bind(C), public :: sg_params
    real(c_double) :: rho_l, rho_g, rho_ref
    real(c_double) :: nu_l, nu_g, nu_ref, Re, Fr
    integer(c_int) :: bc_w, bc_e, bc_s, bc_n
    integer(c_int) :: precision
    integer(c_int) :: max_iter
    real(c_double) :: tol
    integer(c_int) :: mg_pre, mg_post
    integer(c_int) :: warm_start
    integer(c_int) :: check_every
    integer(c_int) :: mg_coarse_krylov
    integer(c_int) :: momentum_method
  end type sg_params

  type, bind(C), public :: sg_solve_info
    integer(c_int) :: iterations
    integer(c_int) :: status
    real(c_double) :: relres
    real(c_double) :: max_row_residual
    real(c_double) :: ms
    real(c_double) :: setup_ms
    real(c_double) :: warm_ms
    real(c_double) :: warm_relres
    real(c_double) :: warm_energy
  end type sg_solve_info

  ! U (which = 0) or V (which = 1) cell geometry: c_loc of each array, cell
  ! arrays (1:Isize,1:Jsize) then ghosted arrays (0:Isize+1,0:Jsize+1)
  type, bind(C), public :: sg_stag_geometry
    type(c_ptr) :: vofS, nxS, nyS, phiS, WlLh, delh, ccx, ccy
    type(c_ptr) :: moex, active
    type(c_ptr) :: aE, aW, aN, aS, EtaE, EtaN, AlE, AlN, DAlE, DAlN, SxE, SyN
    type(c_ptr) :: fcex, fcey, fcnx, fcny
  end type sg_stag_geometry

  type, bind(C), public :: sg_predictor_info
    integer(c_int) :: iterations_u, iterations_v
    real(c_double) :: relres_u, relres_v
    real(c_double) :: ms
  end type sg_predictor_info

  type, bind(C), public :: sg_clsvof_params
    integer(c_int) :: nsub
    real(c_double) :: depth, amp, kw, omega
  end type sg_clsvof_params

  ! a solid shape for sg_geometry_rebuild (kind: 0 none, 1 box, 2 circle, 3 ellipse, 4 boom)
  type, bind(C), public :: sg_solid_shape
    integer(c_int) :: kind
    real(c_double) :: cx, cy, a, b, c
  end type sg_solid_shape

  ! TrackingParticles (Particles.f90:71-205) constants; see sg_particle_params in simcogpu.h
  type, bind(C), public :: sg_particle_params
    real(c_double) :: rho_l, rho_g, nu_l, nu_g, rho_p, gx, gy, lref, uref, us, vs
  end type sg_particle_params

  public :: sg_set_staggered_geometry, sg_predictor_resident, sg_predictor_rows, sg_interface_to_uv_resident
  public :: sg_clsvof_params_default, sg_set_centres, sg_clsvof_advect_resident, sg_clsvof_advect
  public :: sg_last_error, sg_params_default, sg_ctx_create, sg_ctx_destroy, sg_set_params
  public :: sg_upload, sg_download, sg_synchronize, sg_set_geometry
  public :: sg_poisson_build_resident, sg_poisson_solve_resident, sg_poisson_solve
  public :: sg_poisson_matrix, sg_poisson_rhs, sg_poisson_apply, sg_velocity_correct_resident
  public :: sg_particles_set, sg_particles_get, sg_particles_track
  public :: sg_state_bytes, sg_state_save, sg_state_load
  public :: sg_diagnostics_record, sg_diagnostics_read
  public :: sg_geometry_rebuild, sg_object_force, sg_geometry_snapshot, sg_inter_new_var
  public :: sg_set_moving_solid, sg_get_moving_solid
  public :: sg_error_message

  interface
    function sg_last_error() bind(C, name="sg_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function sg_params_default(p) bind(C, name="sg_params_default") result(rc)
      import :: c_int, sg_params
      type(sg_params), intent(out) :: p
      integer(c_int) :: rc
    end function
    function sg_ctx_create(ni, nj, dx, dy, udx, vdy, p, ctx) bind(C, name="sg_ctx_create") result(rc)
      import :: c_int, c_ptr, sg_params
      integer(c_int), value :: ni, nj
      type(c_ptr), value :: dx, dy, udx, vdy
      type(sg_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function
    function sg_ctx_destroy(ctx) bind(C, name="sg_ctx_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function sg_set_params(ctx, p) bind(C, name="sg_set_params") result(rc)
      import :: c_int, c_ptr, sg_params
      type(c_ptr), value :: ctx
      type(sg_params), intent(in) :: p
      integer(c_int) :: rc
    end function
    function sg_upload(ctx, field, host) bind(C, name="sg_upload") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: field
      integer(c_int) :: rc
    end function
    function sg_download(ctx, field, host) bind(C, name="sg_download") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: field
      integer(c_int) :: rc
    end function
    function sg_synchronize(ctx) bind(C, name="sg_synchronize") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function sg_set_geometry(ctx, vofS, nyS, WlLh, aE, aW, aN, aS, uflag, vflag, vofSU, vofSV) &
        bind(C, name="sg_set_geometry") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, vofS, nyS, WlLh, aE, aW, aN, aS, uflag, vflag, vofSU, vofSV
      integer(c_int) :: rc
    end function
    function sg_poisson_build_resident(ctx, vb) bind(C, name="sg_poisson_build_resident") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: vb
      integer(c_int) :: rc
    end function
    function sg_poisson_solve_resident(ctx, info) bind(C, name="sg_poisson_solve_resident") result(rc)
      import :: c_int, c_ptr, sg_solve_info
      type(c_ptr), value :: ctx
      type(sg_solve_info), intent(out) :: info
      integer(c_int) :: rc
    end function
    function sg_poisson_solve(ctx, u_star, v_star, DpU, DpV, vof, phi, vb, Pp, info) &
        bind(C, name="sg_poisson_solve") result(rc)
      import :: c_int, c_ptr, c_double, sg_solve_info
      type(c_ptr), value :: ctx, u_star, v_star, DpU, DpV, vof, phi, Pp
      real(c_double), value :: vb
      type(sg_solve_info), intent(out) :: info
      integer(c_int) :: rc
    end function
    function sg_poisson_matrix(ctx, matr) bind(C, name="sg_poisson_matrix") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, matr
      integer(c_int) :: rc
    end function
    function sg_poisson_rhs(ctx, rhs) bind(C, name="sg_poisson_rhs") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rhs
      integer(c_int) :: rc
    end function
    function sg_poisson_apply(ctx, x, y) bind(C, name="sg_poisson_apply") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, x, y
      integer(c_int) :: rc
    end function
    function sg_velocity_correct_resident(ctx, vb) bind(C, name="sg_velocity_correct_resident") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: vb
      integer(c_int) :: rc
    end function
    function sg_set_staggered_geometry(ctx, which, g) bind(C, name="sg_set_staggered_geometry") result(rc)
      import :: c_int, c_ptr, sg_stag_geometry
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      type(sg_stag_geometry), intent(in) :: g
      integer(c_int) :: rc
    end function
    function sg_predictor_resident(ctx, dt, vb, first_step, info) bind(C, name="sg_predictor_resident") result(rc)
      import :: c_int, c_ptr, c_double, sg_predictor_info
      type(c_ptr), value :: ctx
      real(c_double), value :: dt, vb
      integer(c_int), value :: first_step
      type(sg_predictor_info), intent(out) :: info
      integer(c_int) :: rc
    end function
    function sg_predictor_rows(ctx, matr_u, rhs_u, matr_v, rhs_v) bind(C, name="sg_predictor_rows") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, matr_u, rhs_u, matr_v, rhs_v
      integer(c_int) :: rc
    end function
    function sg_interface_to_uv_resident(ctx) bind(C, name="sg_interface_to_uv_resident") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function sg_clsvof_params_default(p) bind(C, name="sg_clsvof_params_default") result(rc)
      import :: c_int, sg_clsvof_params
      type(sg_clsvof_params), intent(out) :: p
      integer(c_int) :: rc
    end function
    function sg_set_centres(ctx, xc, yc) bind(C, name="sg_set_centres") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, xc, yc
      integer(c_int) :: rc
    end function
    function sg_clsvof_advect_resident(ctx, dt, time, p) bind(C, name="sg_clsvof_advect_resident") result(rc)
      import :: c_int, c_ptr, c_double, sg_clsvof_params
      type(c_ptr), value :: ctx
      real(c_double), value :: dt, time
      type(sg_clsvof_params), intent(in) :: p
      integer(c_int) :: rc
    end function
    function sg_clsvof_advect(ctx, vof, phi, nx, ny, u, v, dt, time, p) bind(C, name="sg_clsvof_advect") result(rc)
      import :: c_int, c_ptr, c_double, sg_clsvof_params
      type(c_ptr), value :: ctx, vof, phi, nx, ny, u, v
      real(c_double), value :: dt, time
      type(sg_clsvof_params), intent(in) :: p
      integer(c_int) :: rc
    end function
    function sg_particles_set(ctx, n, x, y, u, v, d) bind(C, name="sg_particles_set") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, x, y, u, v, d
      integer(c_int), value :: n
      integer(c_int) :: rc
    end function
    function sg_particles_get(ctx, n, x, y, u, v) bind(C, name="sg_particles_get") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, x, y, u, v
      integer(c_int), value :: n
      integer(c_int) :: rc
    end function
    function sg_geometry_rebuild(ctx, shape) bind(C, name="sg_geometry_rebuild") result(rc)
      import :: c_int, c_ptr, sg_solid_shape
      type(c_ptr), value :: ctx
      type(sg_solid_shape), intent(in) :: shape
      integer(c_int) :: rc
    end function
    function sg_object_force(ctx, vs, force, clp, clf) bind(C, name="sg_object_force") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: vs
      real(c_double), intent(out) :: force, clp, clf
      integer(c_int) :: rc
    end function
    function sg_set_moving_solid(ctx, shape, us, vs, asx, asy) bind(C, name="sg_set_moving_solid") result(rc)
      import :: c_int, c_ptr, c_double, sg_solid_shape
      type(c_ptr), value :: ctx
      type(sg_solid_shape), intent(in) :: shape
      real(c_double), value :: us, vs, asx, asy
      integer(c_int) :: rc
    end function
    function sg_get_moving_solid(ctx, shape, us, vs) bind(C, name="sg_get_moving_solid") result(rc)
      import :: c_int, c_ptr, c_double, sg_solid_shape
      type(c_ptr), value :: ctx
      type(sg_solid_shape), intent(out) :: shape
      real(c_double), intent(out) :: us, vs
      integer(c_int) :: rc
    end function
    function sg_geometry_snapshot(ctx) bind(C, name="sg_geometry_snapshot") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function sg_inter_new_var(ctx, vb) bind(C, name="sg_inter_new_var") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: vb
      integer(c_int) :: rc
    end function
    function sg_diagnostics_record(ctx) bind(C, name="sg_diagnostics_record") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function sg_diagnostics_read(ctx, max_rows, rows, n_rows, n_lost) bind(C, name="sg_diagnostics_read") &
        result(rc)
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx, rows
      integer(c_int), value :: max_rows
      integer(c_int), intent(out) :: n_rows
      integer(c_long_long), intent(out) :: n_lost
      integer(c_int) :: rc
    end function
    function sg_state_bytes(ctx, bytes) bind(C, name="sg_state_bytes") result(rc)
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), intent(out) :: bytes
      integer(c_int) :: rc
    end function
    function sg_state_save(ctx, buf, bytes) bind(C, name="sg_state_save") result(rc)
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, buf
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function
    function sg_state_load(ctx, buf, bytes) bind(C, name="sg_state_load") result(rc)
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, buf
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function
    function sg_particles_track(ctx, dt, p, flagged) bind(C, name="sg_particles_track") result(rc)
      import :: c_int, c_ptr, c_double, sg_particle_params
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      type(sg_particle_params), intent(in) :: p
      integer(c_int), intent(out) :: flagged
      integer(c_int) :: rc
    end function
  end interface

contains

  ! the library's last error message as a Fortran string
  function sg_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = sg_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (s(n + 1) /= c_null_char .and. n < 4096)
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function sg_error_message

end module simcogpu
