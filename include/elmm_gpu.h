/*
 * elmm_gpu.h - C ABI of the MI355X-native ELMM LES time step.
 *
 * The library advances ELMM's staggered-grid LES state on one GPU.  The
 * entry points replace the reference's CPU dynamics loop:
 *
 *   elmm_step      <-> TMarchRK3 (src/tsteps.f90:24-243) called from the
 *                      time loop in src/main.f90:112-128
 *   elmm_init      <-> the initial-condition tail of InitialConditions
 *                      (src/initial.f90:4284-4342)
 *   elmm_case_*    <-> the list/mask set-up of InitBoundaryConditions
 *                      (src/initial.f90:4507-4681, src/wallmodels.f90:449-1113,
 *                       src/immersedboundary.f90:1475-1768)
 *   elmm_upload /  <-> the Fortran allocatables U,V,W,Pr,Scalar,Viscosity,TDiff
 *   elmm_download      owned by the main program (src/main.f90:229-285)
 *
 * All functions return 0 on success and a negative ELMM_E* code on error;
 * elmm_last_error() returns the message of the last failure on the calling
 * thread.  Arrays cross the boundary in Fortran order (x fastest), with the
 * lower bound and extents stated by the caller, so a Fortran host passes its
 * own allocatables through iso_c_binding without copies of its own.
 */
#ifndef ELMM_GPU_H
#define ELMM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sides, 0-based (src/parameters.f90:233: We=1, Ea=2, So=3, No=4, Bo=5, To=6). */
enum { ELMM_WE = 0, ELMM_EA = 1, ELMM_SO = 2, ELMM_NO = 3, ELMM_BO = 4, ELMM_TO = 5 };

/* Boundary codes: identical numbers to src/parameters.f90:260-261 BC_*. */
enum {
  ELMM_BC_NOSLIP = 1,
  ELMM_BC_FREESLIP = 2,
  ELMM_BC_PERIODIC = 3,
  ELMM_BC_DIRICHLET = 4,
  ELMM_BC_NEUMANN = 5,
  ELMM_BC_CONSTFLUX = 6   /* temperature and moisture: a surface flux on a wall
                             with wall-model points (Monin-Obukhov law)     */
};

/* Thermal class of a wall-model point (WMPoint%prescribed_temperature,
 * %temperature_flux; src/wallmodels.f90:592-602, 821-831): */
enum {
  ELMM_WM_NEUTRAL = 0,    /* no surface heat flux: the neutral wall laws    */
  ELMM_WM_SURFACE_TEMP = 1, /* prescribed surface temperature (tval, K)       */
  ELMM_WM_SURFACE_FLUX = 2  /* prescribed surface heat flux (tval, K m/s)     */
};

/* Precision forms of the step. */
enum {
  ELMM_FP64 = 0,          /* every field and every operation in FP64          */
  ELMM_FP32_REFINED = 1   /* FP32 fields and stencils; FP64 divergence/RHS,
                             Phi, residual, corrections and global sums; the
                             pressure solve either one FP64 spectral solve
                             (refine_iters = 0) or FP32 spectral solves inside
                             FP64 iterative refinement (refine_iters > 0)     */
};

/* Pressure solvers (src/pressure.f90:17-20 numbering). */
enum { ELMM_POISSON_FFT = 2, ELMM_POISSON_MG = 3 };

/* Field ids for upload/download. */
enum {
  ELMM_F_U = 0, ELMM_F_V = 1, ELMM_F_W = 2, ELMM_F_PR = 3,
  ELMM_F_VISC = 4, ELMM_F_TDIFF = 5, ELMM_F_PHI = 6, ELMM_F_Q = 7,
  ELMM_F_MOISTURE = 14,    /* with enable_moisture */
  ELMM_F_TEMPERATURE = 15, /* with enable_buoyancy */
  ELMM_F_SCALAR = 16      /* + scalar index (0-based) */
};

/* Error codes. */
enum {
  ELMM_OK = 0,
  ELMM_E_ARG = -1,        /* invalid argument or unsupported configuration   */
  ELMM_E_HIP = -2,        /* HIP runtime failure                             */
  ELMM_E_NOMEM = -3,      /* allocation failed                               */
  ELMM_E_DIVERGED = -4,   /* NaN detected or CFL above the limit (main.f90:154-169) */
  ELMM_E_STATE = -5       /* call out of order (e.g. step before init)       */
};

/*
 * Case configuration: the subset of ELMM's namelists/.conf inputs that the
 * time step consumes (src/parameters.f90:34-287, src/initial.f90:56-1193).
 */
typedef struct elmm_config {
  int32_t nx, ny, nz;        /* Prnx, Prny, Prnz                              */
  double  dx, dy, dz;        /* dxmin, dymin, dzmin (dz unused if stretched)  */
  double  x0, y0, z0;        /* domain origin (im_xmin, im_ymin, im_zmin)     */
  int32_t stretched_z;       /* 0 GRID_UNIFORM, 1 GRID_VARIABLE_Z             */
  const double *zw;          /* zW(0:nz) when stretched_z (nz+1 values)       */
  int32_t btype[6];          /* Btype: velocity boundary per side             */
  double  z0_wall[6];        /* z0W..z0T roughness lengths                    */
  int32_t scal_btype[6];     /* ScalBtype                                     */
  double  side_scal[6];      /* sideScal (flux for Neumann, value Dirichlet)  */
  double  nu;                /* molecular_viscosity                           */
  double  nu_scalar;         /* molecular_diffusivity                         */
  double  c_smag;            /* C_Smagorinsky                                 */
  double  filter_ratio;      /* filter_ratios(filtertype)                     */
  double  dt;                /* time_stepping%dt_constant                     */
  double  pr_gradient_x;     /* uniform mean pressure gradient, 0 = off       */
  double  pr_gradient_y;
  int32_t advection_method;  /* 4: CD4div (Morinishi), 2: CD2 variable z      */
  int32_t sgs_model;         /* 0 none, 1 Smagorinsky                         */
  int32_t wall_model;        /* wallmodeltype: 0 off, 1 log law               */
  int32_t nscalars;          /* num_of_scalars (passive)                      */
  int32_t poisson_solver;    /* ELMM_POISSON_FFT / ELMM_POISSON_MG            */
  double  poisson_tol;       /* residual target of iterative solves           */
  int32_t cfl_check;         /* time_stepping%enable_CFL_check                */
  double  cfl_max;           /* CFL_max, 0 = no check                         */
  int32_t precision;         /* ELMM_FP64 / ELMM_FP32_REFINED                 */
  int32_t ib_mass_sources;   /* enable_ibm_mass_sources                       */
  int32_t refine_iters;      /* ELMM_FP32_REFINED: FP32 solves of the FP64
                                residual after a first FP32 solve; 0 = one
                                FP64 spectral solve instead (default)        */
  int32_t variable_dt;       /* time_stepping%variable_time_steps: dt =
                                min(cfl_target / max, dt_max) every step
                                (src/dynamics.f90:519-530)                    */
  int32_t delta_diag;        /* debuglevel > 0 .or. steady == 1: compute the
                                step's delta (src/tsteps.f90:188-207)         */
  double  cfl_target;        /* time_stepping%CFL (variable dt)               */
  double  dt_max;            /* time_stepping%dt_max (variable dt)            */
  double  dt_min;            /* time_stepping%dt_min: a variable dt below it
                                from step 3 on is divergence (main.f90:153-159) */
  double  start_time;        /* time_stepping%start_time                      */
  double  end_time;          /* time_stepping%end_time: a variable dt is cut to
                                reach it (src/dynamics.f90:548-549); 0 = none */
  int32_t autotune;          /* 1: time the candidate tile shapes and z chunks
                                of the z-marching kernels on the first step
                                and keep the fastest (cached per grid shape);
                                replaces InitTiles (src/tiling.f90:16-111)    */
  int32_t mg_cycles_max;     /* ELMM_POISSON_MG: V-cycles at most per solve
                                (maxPoissoniter); 0 = 50                      */
  int32_t enable_buoyancy;   /* 1: temperature as an active scalar: its RK3
                                stage (src/scalars.f90:124-211), the buoyancy
                                force on W (BuoyancyForce, 2nd order,
                                src/dynamics.f90:608-700) and TDiff recomputed
                                every stage (ComputeTDiff,
                                src/scalar_diffusion.f90:983-1009); uniform
                                grids, adiabatic wall model                   */
  double  grav_acc;          /* grav_acc                                      */
  double  temperature_ref;   /* temperature_ref                               */
  int32_t temp_btype[6];     /* TempBtype (periodic x/y; Dirichlet or Neumann
                                in z)                                        */
  double  side_temp[6];      /* sideTemp (value or gradient)                  */
  int32_t enable_moisture;   /* 1 (with enable_buoyancy): moisture as an
                                active scalar: its RK3 stage (src/scalars.f90:
                                202-208), Moisture = moisture_ref in solid
                                cells after it, and the virtual temperature
                                T (1 + 0.61 q) in the buoyancy force
                                (src/dynamics.f90:620-641, 695-710, 766-770);
                                no liquid water                               */
  int32_t moist_btype[6];    /* MoistBtype (periodic x/y; Dirichlet or Neumann
                                in z)                                        */
  int32_t moist_pad;         /* (alignment; 0)                                */
  double  moisture_ref;      /* moisture_ref                                  */
  double  side_moist[6];     /* sideMoist (value or gradient)                 */
  int32_t reserved[2];
} elmm_config;

/*
 * Masks and integer point lists consumed by the step (read-only view).
 * 3D masks are full ghosted arrays (lower bound -2, extents n+6 in each
 * direction, Fortran order).  Points are 1-based (i,j,k) triples as in the
 * reference.  Momentum wall-model lists are indexed [component*6 + dir] with
 * component 0=U,1=V,2=W and dir 0..5 = xm,xp,ym,yp,zm,zp
 * (src/wallmodels.f90:146-149).
 */
typedef struct elmm_lists {
  const int32_t *prtype, *utype, *vtype, *wtype;  /* Prtype/Utype/Vtype/Wtype */
  const uint8_t *scfl_mask[3];                    /* Scfl{x,y,z}_mask, 1=flux */

  int32_t        wm_n[18];
  const int32_t *wm_ijk[18];     /* 3 per point                               */
  const double  *wm_dist[18];    /* distx,disty,distz per point               */
  const double  *wm_z0[18];

  int32_t        wmp_n;          /* WMPoints (scalar points)                  */
  const int32_t *wmp_ijk;
  const double  *wmp_dist;
  const double  *wmp_z0;

  int32_t        scfl_n[3];      /* Scfl{x,y,z}_points                        */
  const int32_t *scfl_ijk[3];

  int32_t        ib_n[4];        /* U,V,W,ScalFl IB points (CSR)              */
  const int32_t *ib_ijk[4];
  const int32_t *ib_ptr[4];      /* n+1 offsets into src/coef                 */
  const int32_t *ib_src_ijk[4];
  const double  *ib_coef[4];

  int32_t        null_n[3];      /* Unull, Vnull, Wnull                       */
  const int32_t *null_ijk[3];

  int32_t        src_n;          /* scalar volume (point) sources             */
  const int32_t *src_ijk;
  const int32_t *src_scalar;     /* 0-based scalar index                      */
  const double  *src_flux;       /* flux / cell volume                        */

  /* thermal class (ELMM_WM_*) and value of each wall-model point; null:
     every point neutral */
  const int8_t  *wm_tclass[18];
  const double  *wm_tval[18];
  const int8_t  *wmp_tclass;
  const double  *wmp_tval;
  /* surface moisture (WMPoint%moisture) and initial moisture flux
     (%moisture_flux) of each wall-model point; null: zero */
  const double  *wm_qval[18];
  const double  *wm_qflux[18];
  const double  *wmp_qval;
  const double  *wmp_qflux;
} elmm_lists;

/* Per-call diagnostics of elmm_step. */
typedef struct elmm_diag {
  double  delta;          /* sum of mean |U-U2| over components / dt (tsteps.f90:188-209) */
  double  cfl;            /* CFL of the last step when cfl_check           */
  double  max_div;        /* max |div U - Q| after the last projection,
                             evaluated in FP64 on the corrected velocity
                             before it is stored in the field type         */
  double  poisson_res;    /* max residual of the last pressure solve       */
  int32_t refinements;    /* refinement sweeps of the last solve           */
  int32_t steps_done;
  double  stored_div;     /* max |div U - Q| of the stored field (equals
                             max_div in the FP64 form)                      */
  double  dt;             /* time step of the last step                    */
  double  time;           /* simulated time after the call (main.f90:116-127) */
  int32_t reserved[4];
} elmm_diag;

typedef struct elmm_case elmm_case;
typedef struct elmm_handle elmm_handle;

/* Axis-aligned solid cuboid, metres: xmin,xmax,ymin,ymax,zmin,zmax, the
 * roughness length z0 of its walls (0 = smooth wall law), and the surface
 * heat flux (K m/s) and moisture flux of its walls (rough walls only; with
 * buoyancy and moisture).  Equivalent to a POLYHEDRON of six planes in the
 * reference's obstacle file (src/geometricshapes.f90:2655-2760) wrapped as
 * SolidBody(z0, temperature_flux, moisture_flux) (src/solidbodies.f90:
 * 439-454; the fluxes reach its wall-model points, src/immersedboundary.f90:
 * 111-125, 260-285). */
typedef struct elmm_cuboid {
  double xmin, xmax, ymin, ymax, zmin, zmax, z0;
  double temperature_flux, moisture_flux;
} elmm_cuboid;

/* Scalar point source: position (m), flux (units/s), scalar index. */
typedef struct elmm_point_source { double x, y, z, flux; int32_t scalar; int32_t pad; } elmm_point_source;

/* ---------------------------------------------------------------------------
 * Case set-up (host only, no GPU needed).
 * Builds Prtype/Utype/Vtype/Wtype, the wall-model lists, the scalar
 * wall-face lists, the IB interpolation lists, the nullified points and the
 * point sources for the given configuration and solids.
 *   replaces: InitBoundaryConditions obstacle/list section
 *             (src/initial.f90:4646-4705) with GetSolidBodiesBC
 *             (src/immersedboundary.f90:1750-1768), GetOutsideBoundariesWM
 *             (src/wallmodels.f90:449-904), InitWMMasks (911-1113),
 *             SetNullifiedPoints (src/initial.f90:4861-4962),
 *             InitPointSources (src/pointsources.f90:44-66)
 * ------------------------------------------------------------------------- */
int  elmm_case_build(const elmm_config *cfg,
                     int32_t n_cuboids, const elmm_cuboid *cuboids,
                     int32_t n_sources, const elmm_point_source *sources,
                     elmm_case **out);
/* elmm_case_build with the solid-body masks and lists (FindInsideCells,
 * FindNeighbouringCells, InitImBoundaries, GetSolidBodiesWM(_UVW),
 * SetNullifiedPoints, the scalar part of InitWMMasks) built on the calling
 * thread's current HIP device; the result is the same as elmm_case_build's,
 * byte for byte and in the same order, and lives in host memory like it.
 *   replaces: as elmm_case_build; FindInsideCells (src/solidbodies.f90:
 *             75-210), FindNeighbouringCells (src/immersedboundary.f90:
 *             1475-1506), InitImBoundaries / TVelIBPoint_Create
 *             (1650-1747, 690-1020), GetSolidBodiesWM(_UVW) (14-308),
 *             InitWMMasks (src/wallmodels.f90:911-1113) and
 *             SetNullifiedPoints (src/initial.f90:4861-4962) */
int  elmm_case_build_device(const elmm_config *cfg,
                            int32_t n_cuboids, const elmm_cuboid *cuboids,
                            int32_t n_sources, const elmm_point_source *sources,
                            elmm_case **out);
int  elmm_case_view(const elmm_case *c, elmm_lists *view);
void elmm_case_free(elmm_case *c);

/* ---------------------------------------------------------------------------
 * Device-resident solver.
 * ------------------------------------------------------------------------- */

/* Allocate the device state of one case and copy masks/lists to HBM.
 * replaces: AllocateGlobals (src/main.f90:229-285) and the first-call
 * allocations of TMarchRK3 (src/tsteps.f90:54-77). */
int elmm_create(const elmm_config *cfg, const elmm_lists *lists, elmm_handle **out);

/* elmm_create with the lists of a case from elmm_case_build (the case may be
 * freed afterwards); the form a Fortran host binds, which then never needs
 * the layout of elmm_lists.  replaces: as elmm_create. */
int elmm_create_case(const elmm_config *cfg, const elmm_case *c, elmm_handle **out);

/* Copy one field between a host array and HBM.  The host array has lower
 * bound lb in every direction and extents ext[3]; the overlap with the
 * field's ghosted device extent (-2..n+3) is copied. */
int elmm_upload(elmm_handle *h, int32_t field, int32_t lb, const int32_t ext[3], const double *src);
int elmm_download(elmm_handle *h, int32_t field, int32_t lb, const int32_t ext[3], double *dst);

/* Initial-condition tail: ghost fills, initial projection with coef=1,
 * SGS viscosity, frozen TDiff, wall-model seed.
 * replaces: src/initial.f90:4284-4342 */
int elmm_init(elmm_handle *h);

/* Advance nsteps RK3 steps with constant dt; synchronous at return.
 * replaces: TMarchRK3 (src/tsteps.f90:24-243) + main loop time update
 * (src/main.f90:112-128). diag may be NULL. */
int elmm_step(elmm_handle *h, int32_t nsteps, elmm_diag *diag);

/* Enqueue nsteps without waiting (device stream order); elmm_wait joins. */
int elmm_step_async(elmm_handle *h, int32_t nsteps);
int elmm_wait(elmm_handle *h, elmm_diag *diag);

/* Maximum |div U| of the current state (PostPoisson check_divergence,
 * src/pressure.f90:884-989), computed in FP64 on the device. */
int elmm_divergence(elmm_handle *h, double *max_div);

/* Friction velocity u* of every scalar wall-model point (WMPoints%ustar,
 * the state ComputeViscsWM keeps, src/wallmodels.f90:2050-2152), in the
 * order of elmm_lists.wmp_ijk: *n receives the count; with ustar not NULL,
 * cap >= *n values are copied (synchronises the handle's stream). */
int elmm_wm_ustar(elmm_handle *h, int32_t cap, double *ustar, int32_t *n);

/* elmm_divergence without device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops).  Its block maxima live in a
 * file-scope device array of 2 MiB per device (held until the process
 * ends), so grids of at most 67,108,864 cells (larger: ELMM_E_ARG).  Calls
 * that overlap on one device race on that array. */
int elmm_divergence_noalloc(elmm_handle *h, double *max_div);

int elmm_destroy(elmm_handle *h);

/* 1 in *idle when every step enqueued on the handle has completed, else 0
 * (non-blocking).  No counterpart in the reference (synchronous loop). */
int elmm_query(elmm_handle *h, int32_t *idle);

/* Device bytes elmm_create allocates for a configuration: fields, masks,
 * spectral work array, reduction partials, the set-up's transient mark array
 * and the allocation granularity of its buffers (the surface lists, which
 * depend on the case, are excluded; the server adds an estimate of them).
 * Used to admit cases against an HBM budget. */
int elmm_footprint(const elmm_config *cfg, int64_t *bytes);
/* elmm_footprint plus an estimate of the case's surface lists on the device:
 * the bytes the serving form admits a request by.  No counterpart in the
 * reference. */
int elmm_case_footprint(const elmm_config *cfg, const elmm_case *c, int64_t *bytes);

/* ---------------------------------------------------------------------------
 * Field files (host only): the reference's legacy VTK files, RECTILINEAR_GRID,
 * BINARY, big-endian float32.  Arrays are ghosted host arrays (lower bound
 * -2, extents n+6, Fortran order), as elmm_download returns them.
 *   elmm_vtk_write: out.vtk (p, visc when given, cell-centred VECTORS u),
 *     scalars.vtk (scalar01..), U.vtk, V.vtk, W.vtk in directory dir.
 *     replaces: OutputOut, OutputScalars, OutputUVW
 *               (src/outputs.f90:1560-1854, 2260-2490)
 *   elmm_vtk_read: the restart reader; p (0 when absent), the scalars, and
 *     U, V, W from U/V/W.vtk, else from out.vtk's vectors.
 *     replaces: ReadInitialConditions (src/initial.f90:3575-3811)
 * ------------------------------------------------------------------------- */
int elmm_vtk_write(const char *dir, const elmm_config *cfg, const double *U, const double *V, const double *W,
                   const double *Pr, const double *visc, int32_t nscal, const double *const *scalars);
int elmm_vtk_read(const char *dir, const elmm_config *cfg, double *U, double *V, double *W, double *Pr,
                  int32_t nscal, double *const *scalars);

/* Asynchronous field frame: copies the current U, V, W, Pr and scalars to
 * pinned host memory in stream order (after every step enqueued before the
 * call) and writes the VTK files of elmm_vtk_write into dir from a writer
 * thread, while later steps run.  elmm_frame_wait joins the writer and
 * returns its status.  replaces: the frame writers (src/frames.f90:369-420,
 * pthreads.c). */
int elmm_frame_async(elmm_handle *h, const char *dir);
int elmm_frame_wait(elmm_handle *h);

/* ---------------------------------------------------------------------------
 * Serving form: continuous batching of independent cases on one GPU.  No
 * counterpart in the reference (one case per process, src/main.f90).
 * Requests are admitted in arrival order while their footprint fits the HBM
 * budget and a slot is free; each active case keeps its fields resident on
 * its own stream and advances chunk_steps per elmm_server_progress call when
 * its previous chunk has finished, so the cases' kernels overlap on the
 * device.  Finished cases copy their fields to the caller's output buffers
 * (full ghosted arrays, lower bound -2, extents n+6) and free their memory.
 * The case (c) and the output buffers must stay valid until the request has
 * finished; the initial fields are copied at submit.
 * ------------------------------------------------------------------------- */
typedef struct elmm_server elmm_server;
typedef struct elmm_request {
  const elmm_config *cfg;
  const elmm_case   *c;
  int32_t nsteps;
  int32_t pad;
  const double *U, *V, *W, *Pr;        /* initial state, lb -2, extents n+6     */
  const double *const *scalars;        /* nscalars arrays or NULL (zeros)       */
  double *out_U, *out_V, *out_W, *out_Pr;
  double *const *out_scalars;          /* NULL: not returned                    */
} elmm_request;
typedef struct elmm_request_status {
  int32_t state;          /* 0 queued, 1 running, 2 finished, -1 failed       */
  int32_t status;         /* ELMM_E_* of a failed request                      */
  int32_t steps_done;
  int32_t pad;
  double  latency_s;      /* completion - submit, seconds                      */
  double  max_div;        /* max |div U - Q| of the final state                */
} elmm_request_status;
int elmm_server_create(int64_t hbm_budget, int32_t max_active, int32_t chunk_steps, elmm_server **out);
int elmm_server_submit(elmm_server *s, const elmm_request *r, int64_t *id);
int elmm_server_progress(elmm_server *s, int32_t *n_active, int32_t *n_queued);
int elmm_server_status(elmm_server *s, int64_t id, elmm_request_status *st);
int elmm_server_stats(elmm_server *s, int64_t *cell_steps, int64_t *hbm_in_use);
int elmm_server_destroy(elmm_server *s);

/* ---------------------------------------------------------------------------
 * Kernel timers: HIP events around each kernel group on the handle's
 * stream, accumulated per group while enabled.
 *   replaces: the stage timers printed with debugparam>1
 *             (src/tsteps.f90:234-238, src/pressure.f90:252-264) and
 *             poisson_solver_time (src/poisson.f90:125-138)
 * ------------------------------------------------------------------------- */
enum {
  ELMM_T_SGS = 0, ELMM_T_WALL = 1, ELMM_T_MOMENTUM = 2, ELMM_T_MOMENTUM_CORR = 3,
  ELMM_T_SCALAR = 4, ELMM_T_SCALAR_CORR = 5, ELMM_T_GHOST = 6, ELMM_T_IB = 7,
  ELMM_T_POISSON_X_FWD = 8, ELMM_T_POISSON_Y = 9, ELMM_T_POISSON_Z = 10, ELMM_T_POISSON_X_INV = 11,
  ELMM_T_PROJECTION = 12, ELMM_T_REDUCE = 13, ELMM_T_OTHER = 14, ELMM_T_REFINE = 15, ELMM_T_NGROUPS = 16
};
typedef struct elmm_timers {
  double  ms[ELMM_T_NGROUPS];        /* summed event time per group          */
  int64_t launches[ELMM_T_NGROUPS];  /* kernel launches per group            */
} elmm_timers;
int elmm_timers_enable(elmm_handle *h, int32_t enable);

/* Tile shape (x by y columns per block) and z-chunk lengths of the
 * z-marching kernels a handle uses; tuned = 1 once the autotuner has chosen
 * them.  replaces: the tile sizes of InitTiles (src/tiling.f90:16-111). */
typedef struct elmm_tiles {
  int32_t mom_tx, mom_ty, mom_zchunk;   /* fused momentum kernel (0: per-cell kernels) */
  int32_t sgs_zchunk, scalar_zchunk;    /* SGS and scalar z-marches (0: per-cell)      */
  int32_t tuned;
  int32_t reserved[2];
} elmm_tiles;
int elmm_get_tiles(elmm_handle *h, elmm_tiles *out);
int elmm_timers_read(elmm_handle *h, elmm_timers *out, int32_t reset);
const char *elmm_timer_name(int32_t group);

/* Plain device copy of nbytes, reps times (plus one untimed warm-up), on the
 * default stream; *gbps = bytes read + written per second.  The reference
 * point of the HBM traffic counters and of the achievable copy bandwidth.
 * No counterpart in the reference. */
int elmm_copy_probe(int64_t nbytes, int32_t reps, double *gbps);

/* Version and last error of the calling thread. */
const char *elmm_version(void);
const char *elmm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ELMM_GPU_H */
