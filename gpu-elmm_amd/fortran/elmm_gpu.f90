! elmm_gpu.f90 - Fortran binding of the MI355X LES step library
! (include/elmm_gpu.h) through iso_c_binding.
!
! A Fortran host keeps its own allocatables with the reference's bounds
! (U, V, W, Scalar: -2:n+3; Pr, Viscosity: -1:n+2, src/main.f90:229-285) and
! moves them to and from the device with elmm_put / elmm_get, which pass the
! array's lower bound and extents.  elmm_step replaces
!   call TMarchRK3(U, V, W, Pr, Temperature, Moisture, Scalar, delta)
! of the time loop (src/main.f90:112-114).
module elmm_gpu
  use iso_c_binding
  implicit none

  private
  public :: elmm_config, elmm_cuboid, elmm_point_source, elmm_diag
  public :: elmm_case_build, elmm_case_build_device, elmm_case_free, elmm_create_case, elmm_destroy, elmm_init, elmm_step
  public :: elmm_upload, elmm_download, elmm_divergence
  public :: elmm_put, elmm_get, elmm_check, elmm_error_message
  public :: ELMM_WE, ELMM_EA, ELMM_SO, ELMM_NO, ELMM_BO, ELMM_TO
  public :: ELMM_BC_NOSLIP, ELMM_BC_FREESLIP, ELMM_BC_PERIODIC, ELMM_BC_DIRICHLET, ELMM_BC_NEUMANN, ELMM_BC_CONSTFLUX
  public :: ELMM_FP64, ELMM_FP32_REFINED, ELMM_POISSON_FFT
  public :: ELMM_F_U, ELMM_F_V, ELMM_F_W, ELMM_F_PR, ELMM_F_VISC, ELMM_F_TDIFF, ELMM_F_SCALAR

  ! sides (1-based index into the 6-element arrays is side+1)
  integer(c_int32_t), parameter :: ELMM_WE = 0, ELMM_EA = 1, ELMM_SO = 2, ELMM_NO = 3, ELMM_BO = 4, ELMM_TO = 5
  ! boundary codes, the reference's BC_* numbers (src/parameters.f90:249)
  integer(c_int32_t), parameter :: ELMM_BC_NOSLIP = 1, ELMM_BC_FREESLIP = 2, ELMM_BC_PERIODIC = 3, &
                                   ELMM_BC_DIRICHLET = 4, ELMM_BC_NEUMANN = 5, &
                                   ELMM_BC_CONSTFLUX = 6
  integer(c_int32_t), parameter :: ELMM_FP64 = 0, ELMM_FP32_REFINED = 1
  integer(c_int32_t), parameter :: ELMM_POISSON_FFT = 2
  integer(c_int32_t), parameter :: ELMM_F_U = 0, ELMM_F_V = 1, ELMM_F_W = 2, ELMM_F_PR = 3, ELMM_F_VISC = 4, &
                                   ELMM_F_TDIFF = 5, ELMM_F_SCALAR = 16

  type, bind(C) :: elmm_config
    integer(c_int32_t) :: nx, ny, nz
    real(c_double) :: dx, dy, dz
    real(c_double) :: x0, y0, z0
    integer(c_int32_t) :: stretched_z
    type(c_ptr) :: zw = c_null_ptr
    integer(c_int32_t) :: btype(6)
    real(c_double) :: z0_wall(6)
    integer(c_int32_t) :: scal_btype(6)
    real(c_double) :: side_scal(6)
    real(c_double) :: nu, nu_scalar, c_smag, filter_ratio, dt, pr_gradient_x, pr_gradient_y
    integer(c_int32_t) :: advection_method, sgs_model, wall_model, nscalars, poisson_solver
    real(c_double) :: poisson_tol
    integer(c_int32_t) :: cfl_check
    real(c_double) :: cfl_max
    integer(c_int32_t) :: precision, ib_mass_sources, refine_iters
    integer(c_int32_t) :: variable_dt, delta_diag
    real(c_double) :: cfl_target, dt_max, dt_min, start_time, end_time
    integer(c_int32_t) :: autotune
    integer(c_int32_t) :: mg_cycles_max
    integer(c_int32_t) :: enable_buoyancy
    real(c_double) :: grav_acc, temperature_ref
    integer(c_int32_t) :: temp_btype(6)
    real(c_double) :: side_temp(6)
    integer(c_int32_t) :: enable_moisture
    integer(c_int32_t) :: moist_btype(6)
    integer(c_int32_t) :: moist_pad
    real(c_double) :: moisture_ref
    real(c_double) :: side_moist(6)
    integer(c_int32_t) :: reserved(2)
  end type

  type, bind(C) :: elmm_cuboid
    real(c_double) :: xmin, xmax, ymin, ymax, zmin, zmax, z0
    real(c_double) :: temperature_flux = 0, moisture_flux = 0
  end type

  type, bind(C) :: elmm_point_source
    real(c_double) :: x, y, z, flux
    integer(c_int32_t) :: scalar, pad
  end type

  type, bind(C) :: elmm_diag
    real(c_double) :: delta, cfl, max_div, poisson_res
    integer(c_int32_t) :: refinements, steps_done
    real(c_double) :: stored_div, dt, time
    integer(c_int32_t) :: reserved(4)
  end type

  interface
    integer(c_int) function elmm_case_build(cfg, n_cuboids, cuboids, n_sources, sources, case) &
        bind(C, name="elmm_case_build")
      import
      type(elmm_config), intent(in) :: cfg
      integer(c_int32_t), value :: n_cuboids
      type(elmm_cuboid), intent(in) :: cuboids(*)
      integer(c_int32_t), value :: n_sources
      type(elmm_point_source), intent(in) :: sources(*)
      type(c_ptr), intent(out) :: case
    end function

    ! the same case, its solid-body masks and lists built on the GPU
    integer(c_int) function elmm_case_build_device(cfg, n_cuboids, cuboids, n_sources, sources, case) &
        bind(C, name="elmm_case_build_device")
      import
      type(elmm_config), intent(in) :: cfg
      integer(c_int32_t), value :: n_cuboids
      type(elmm_cuboid), intent(in) :: cuboids(*)
      integer(c_int32_t), value :: n_sources
      type(elmm_point_source), intent(in) :: sources(*)
      type(c_ptr), intent(out) :: case
    end function

    subroutine elmm_case_free(case) bind(C, name="elmm_case_free")
      import
      type(c_ptr), value :: case
    end subroutine

    integer(c_int) function elmm_create_case(cfg, case, h) bind(C, name="elmm_create_case")
      import
      type(elmm_config), intent(in) :: cfg
      type(c_ptr), value :: case
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function elmm_destroy(h) bind(C, name="elmm_destroy")
      import
      type(c_ptr), value :: h
    end function

    integer(c_int) function elmm_init(h) bind(C, name="elmm_init")
      import
      type(c_ptr), value :: h
    end function

    integer(c_int) function elmm_step(h, nsteps, diag) bind(C, name="elmm_step")
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
      type(elmm_diag), intent(out) :: diag
    end function

    integer(c_int) function elmm_upload(h, field, lb, ext, src) bind(C, name="elmm_upload")
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: field, lb
      integer(c_int32_t), intent(in) :: ext(3)
      real(c_double), intent(in) :: src(*)
    end function

    integer(c_int) function elmm_download(h, field, lb, ext, dst) bind(C, name="elmm_download")
      import
      type(c_ptr), value :: h
      integer(c_int32_t), value :: field, lb
      integer(c_int32_t), intent(in) :: ext(3)
      real(c_double), intent(inout) :: dst(*)
    end function

    integer(c_int) function elmm_divergence(h, max_div) bind(C, name="elmm_divergence")
      import
      type(c_ptr), value :: h
      real(c_double), intent(out) :: max_div
    end function

    type(c_ptr) function elmm_last_error() bind(C, name="elmm_last_error")
      import
    end function
  end interface

contains

  ! Upload a host array whose lower bound is lb in every direction.
  subroutine elmm_put(h, field, a, lb)
    type(c_ptr), intent(in) :: h
    integer(c_int32_t), intent(in) :: field, lb
    real(c_double), intent(in), contiguous :: a(:,:,:)
    integer(c_int32_t) :: ext(3)
    ext = int(shape(a), c_int32_t)
    call elmm_check(elmm_upload(h, field, lb, ext, a), "elmm_upload")
  end subroutine

  subroutine elmm_get(h, field, a, lb)
    type(c_ptr), intent(in) :: h
    integer(c_int32_t), intent(in) :: field, lb
    real(c_double), intent(inout), contiguous :: a(:,:,:)
    integer(c_int32_t) :: ext(3)
    ext = int(shape(a), c_int32_t)
    call elmm_check(elmm_download(h, field, lb, ext, a), "elmm_download")
  end subroutine

  function elmm_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    type(c_ptr) :: p
    p = elmm_last_error()
    msg = ""
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

  ! Stops the program with the library's message on a nonzero status (the
  ! reference's error_stop behaviour, src/stop.f90:25-79).
  subroutine elmm_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc /= 0) then
      write(*, '(a,a,i0,a,a)') what, " failed (", rc, "): ", elmm_error_message()
      error stop 1
    end if
  end subroutine

end module elmm_gpu
