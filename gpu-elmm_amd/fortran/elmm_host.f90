! elmm_host.f90 - namelist-driven Fortran host of the MI355X LES step.
!
! Mirrors the reference's main program (src/main.f90): set-up from the
! configuration, allocation of the state with the reference's bounds
! (src/main.f90:229-285), the initial-condition tail (elmm_init,
! src/initial.f90:4284-4342) and the time loop with one call per step in
! place of TMarchRK3 (src/main.f90:112-128).  Usage:
!   elmm_host case.nml
! Namelist groups (a subset of the reference's grid.conf, main.conf,
! boundconds.conf, les.conf and time_stepping.conf inputs):
!   &grid        nx, ny, nz, dx, dy, dz, x0, y0, z0, zgrid_file
!   &physics     nu, nu_scalar, c_smag, z0_bottom, z0_top, top_bc, nscalars, precision,
!                advection_method, pr_gradient_x
! zgrid_file (optional): the z levels zW(0..nz) one per line, as the
! reference's zgrid.txt (src/initial.f90:254-279); it sets nz and a stretched
! z grid (which needs advection_method = 2).
!   &time_stepping dt, nsteps, report_every
!   &obstacles   n_cuboids, cuboids(7, n) = xmin xmax ymin ymax zmin zmax z0,
!                cuboid_fluxes(2, n) = surface heat and moisture flux (default 0),
!                device_lists (masks and lists built on the GPU; default .true.)
!   &point_sources n_sources, sources(5, n) = x y z flux scalar(1-based)
!   &io          init_file, out_file, u0
! init_file (optional): stream of float64 U, V, W (lb -2), Pr (lb -1),
! Scalar(:,:,:,s) (lb -2); out_file receives U, V, W, Pr, Scalar the same way.
program elmm_host
  use iso_c_binding
  use elmm_gpu
  implicit none

  integer :: nx = 32, ny = 32, nz = 16, nscalars = 1, precision = 0, top_bc = 2
  real(c_double) :: dx = 1, dy = 1, dz = 1, x0 = 0, y0 = 0, z0 = 0
  ! molecular_viscosity and molecular_diffusivity defaults (src/parameters.f90:128-129: 1/70000 and
  ! 1/70000/0.7 with 0.7 a default-real literal)
  real(c_double) :: nu = 1.0_c_double / 70000, nu_scalar = (1.0_c_double / 70000) / real(0.7, c_double)
  real(c_double) :: c_smag = 0.122_c_double
  real(c_double) :: z0_bottom = 0.01_c_double, dt = 0.05_c_double, u0 = 1
  integer :: nsteps = 10, report_every = 5, n_cuboids = 0, n_sources = 0
  logical :: device_lists = .true.
  integer, parameter :: maxobj = 256
  real(c_double) :: cuboids(7, maxobj) = 0, sources(5, maxobj) = 0, cuboid_fluxes(2, maxobj) = 0
  character(len=512) :: init_file = "", out_file = "", zgrid_file = ""
  integer :: advection_method = 4
  real(c_double) :: z0_top = 0, pr_gradient_x = 0, rtmp
  real(c_double), allocatable, target :: zwl(:)
  namelist /grid/ nx, ny, nz, dx, dy, dz, x0, y0, z0, zgrid_file
  namelist /physics/ nu, nu_scalar, c_smag, z0_bottom, z0_top, top_bc, nscalars, precision, advection_method, &
                     pr_gradient_x
  namelist /time_stepping/ dt, nsteps, report_every
  namelist /obstacles/ n_cuboids, cuboids, cuboid_fluxes, device_lists
  namelist /point_sources/ n_sources, sources
  namelist /io/ init_file, out_file, u0

  type(elmm_config) :: cfg
  type(elmm_cuboid), allocatable :: cub(:)
  type(elmm_point_source), allocatable :: src(:)
  type(elmm_diag) :: diag
  type(c_ptr) :: case, h
  real(c_double), allocatable :: U(:,:,:), V(:,:,:), W(:,:,:), Pr(:,:,:), Scalar(:,:,:,:)
  character(len=512) :: path
  integer :: unit, ios, i, s, done, chunk

  call get_command_argument(1, path)
  if (len_trim(path) == 0) then
    write(*, *) "usage: elmm_host case.nml | --sizes"
    error stop 2
  end if
  if (trim(path) == "--sizes") then
    ! interoperable sizes, checked against the C header by the test suite
    write(*, '(a,4(1x,i0))') "sizes", c_sizeof(cfg), c_sizeof(elmm_cuboid(0, 0, 0, 0, 0, 0, 0, 0, 0)), &
        c_sizeof(elmm_point_source(0, 0, 0, 0, 0, 0)), c_sizeof(diag)
    stop
  end if
  open(newunit=unit, file=trim(path), status="old", action="read")
  read(unit, nml=grid, iostat=ios); rewind(unit)
  read(unit, nml=physics, iostat=ios); rewind(unit)
  read(unit, nml=time_stepping, iostat=ios); rewind(unit)
  read(unit, nml=obstacles, iostat=ios); rewind(unit)
  read(unit, nml=point_sources, iostat=ios); rewind(unit)
  read(unit, nml=io, iostat=ios)
  close(unit)

  ! z levels from a file (zgridfromfile): nz = number of values - 1
  cfg%stretched_z = 0
  if (len_trim(zgrid_file) > 0) then
    open(newunit=unit, file=trim(zgrid_file), status="old", action="read")
    i = -1
    do
      read(unit, *, iostat=ios) rtmp
      if (ios /= 0) exit
      i = i + 1
    end do
    rewind(unit)
    nz = i
    allocate(zwl(0:nz))
    do i = 0, nz
      read(unit, *) zwl(i)
    end do
    close(unit)
    z0 = zwl(0)
    cfg%stretched_z = 1
    cfg%zw = c_loc(zwl)
  end if

  ! periodic x/y, log-law no-slip bottom, free-slip (or no-slip) top
  cfg%nx = nx; cfg%ny = ny; cfg%nz = nz
  cfg%dx = dx; cfg%dy = dy; cfg%dz = dz
  cfg%x0 = x0; cfg%y0 = y0; cfg%z0 = z0
  cfg%btype = [ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_NOSLIP, &
               int(top_bc, c_int32_t)]
  cfg%z0_wall = [0.0_c_double, 0.0_c_double, 0.0_c_double, 0.0_c_double, z0_bottom, z0_top]
  cfg%scal_btype = [ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_NEUMANN, &
                    ELMM_BC_NEUMANN]
  cfg%side_scal = 0
  cfg%nu = nu; cfg%nu_scalar = nu_scalar; cfg%c_smag = c_smag; cfg%filter_ratio = 1
  cfg%dt = dt; cfg%pr_gradient_x = pr_gradient_x; cfg%pr_gradient_y = 0
  cfg%advection_method = int(advection_method, c_int32_t); cfg%sgs_model = 1; cfg%wall_model = 1
  cfg%nscalars = nscalars; cfg%poisson_solver = ELMM_POISSON_FFT; cfg%poisson_tol = 1e-10_c_double
  cfg%cfl_check = 1; cfg%cfl_max = 0
  cfg%precision = precision; cfg%ib_mass_sources = 1; cfg%refine_iters = 0
  cfg%variable_dt = 0; cfg%delta_diag = 0
  cfg%cfl_target = 0.3_c_double; cfg%dt_max = dt; cfg%dt_min = 0; cfg%start_time = 0; cfg%end_time = 0
  cfg%autotune = 1
  cfg%mg_cycles_max = 0
  cfg%enable_buoyancy = 0; cfg%grav_acc = 9.81_c_double; cfg%temperature_ref = 295._c_double
  cfg%temp_btype = [ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_NEUMANN, &
                    ELMM_BC_NEUMANN]
  cfg%side_temp = 0
  cfg%enable_moisture = 0; cfg%moist_pad = 0; cfg%moisture_ref = 0.01_c_double
  cfg%moist_btype = [ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_PERIODIC, ELMM_BC_NEUMANN, &
                     ELMM_BC_NEUMANN]
  cfg%side_moist = 0
  cfg%reserved = 0

  allocate(cub(max(1, n_cuboids)), src(max(1, n_sources)))
  do i = 1, n_cuboids
    cub(i) = elmm_cuboid(cuboids(1, i), cuboids(2, i), cuboids(3, i), cuboids(4, i), cuboids(5, i), &
                         cuboids(6, i), cuboids(7, i), cuboid_fluxes(1, i), cuboid_fluxes(2, i))
  end do
  do i = 1, n_sources
    src(i) = elmm_point_source(sources(1, i), sources(2, i), sources(3, i), sources(4, i), &
                               int(sources(5, i)) - 1, 0)
  end do
  if (device_lists) then
    call elmm_check(elmm_case_build_device(cfg, int(n_cuboids, c_int32_t), cub, int(n_sources, c_int32_t), &
                                           src, case), "elmm_case_build_device")
  else
    call elmm_check(elmm_case_build(cfg, int(n_cuboids, c_int32_t), cub, int(n_sources, c_int32_t), src, case), &
                    "elmm_case_build")
  end if
  call elmm_check(elmm_create_case(cfg, case, h), "elmm_create")

  ! the reference's bounds: velocities and scalars -2:n+3, pressure -1:n+2
  allocate(U(-2:nx+3, -2:ny+3, -2:nz+3), V(-2:nx+3, -2:ny+3, -2:nz+3), W(-2:nx+3, -2:ny+3, -2:nz+3))
  allocate(Pr(-1:nx+2, -1:ny+2, -1:nz+2), Scalar(-2:nx+3, -2:ny+3, -2:nz+3, max(1, nscalars)))
  if (len_trim(init_file) > 0) then
    open(newunit=unit, file=trim(init_file), access="stream", form="unformatted", status="old", action="read")
    read(unit) U, V, W, Pr
    if (nscalars > 0) read(unit) Scalar(:,:,:,1:nscalars)
    close(unit)
  else
    U = 0; V = 0; W = 0; Pr = 0; Scalar = 0
    U(1:nx, 1:ny, 1:nz) = u0
  end if
  call elmm_put(h, ELMM_F_U, U, -2)
  call elmm_put(h, ELMM_F_V, V, -2)
  call elmm_put(h, ELMM_F_W, W, -2)
  call elmm_put(h, ELMM_F_PR, Pr, -1)
  do s = 1, nscalars
    call elmm_put(h, ELMM_F_SCALAR + s - 1, Scalar(:,:,:,s), -2)
  end do
  call elmm_check(elmm_init(h), "elmm_init")

  done = 0
  do while (done < nsteps)
    chunk = min(max(report_every, 1), nsteps - done)
    call elmm_check(elmm_step(h, int(chunk, c_int32_t), diag), "elmm_step")
    done = done + chunk
    write(*, '(a,i0,a,es12.4,a,es12.4)') "step ", done, "  CFL ", diag%cfl, "  max|div| ", diag%max_div
  end do

  call elmm_get(h, ELMM_F_U, U, -2)
  call elmm_get(h, ELMM_F_V, V, -2)
  call elmm_get(h, ELMM_F_W, W, -2)
  call elmm_get(h, ELMM_F_PR, Pr, -1)
  do s = 1, nscalars
    call elmm_get(h, ELMM_F_SCALAR + s - 1, Scalar(:,:,:,s), -2)
  end do
  if (len_trim(out_file) > 0) then
    open(newunit=unit, file=trim(out_file), access="stream", form="unformatted", status="replace", action="write")
    write(unit) U, V, W, Pr
    if (nscalars > 0) write(unit) Scalar(:,:,:,1:nscalars)
    close(unit)
  end if
  call elmm_check(elmm_destroy(h), "elmm_destroy")
  call elmm_case_free(case)
end program elmm_host
