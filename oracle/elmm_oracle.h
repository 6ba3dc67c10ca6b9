/*
 * elmm_oracle.h - CPU FP64 restatement of ELMM's LES time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness oracle for the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or the timed CPU baseline.  The
 * product path never links or calls it.
 *
 * The functions restate the reference's Fortran procedures in their
 * operation order (each function cites the file:line it follows), compiled
 * with -ffp-contract=off so that no multiply-add is fused.  Scatter loops that
 * the reference runs under OpenMP with atomics or races are executed in the
 * single-thread order (OMP_NUM_THREADS=1 semantics).
 *
 * Parity pins (tests/test_oracle_pins.py): Poisson solve (FD2 residual,
 * eigenfunction, dense solve), Smagorinsky (linear shear, constant field),
 * CD4 advection (Morinishi's textbook form on periodic grids, constant
 * field), diffusion (Fourier mode), rough- and smooth-wall laws (closed form,
 * fixed point), kappa limiter (closed form, constant field), ghost fills
 * (mirror rules, idempotence), projection (divergence), pressure reference
 * (top-layer mean), scalar budget (conservation, point source), RK3
 * (coefficient identities, uniform forcing).
 *
 * Parity unpinned (no independent check; agreement with the GPU library is
 * the only evidence): TimeStepLength's CFL scan, the immersed-boundary
 * runtime (IBMomentum gathers, IBMassSources + Bound_Q fold, NullInterior,
 * the ScalFl viscosity gather), DivergenceWM / ScalarExtra, and the
 * even/odd ordering of the z pass of KappaScalar.
 */
#ifndef ELMM_ORACLE_H
#define ELMM_ORACLE_H

#include "../include/elmm_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

oracle *oracle_create(const elmm_config *cfg, const elmm_lists *lists);
void    oracle_destroy(oracle *o);
int     oracle_upload(oracle *o, int32_t field, int32_t lb, const int32_t ext[3], const double *src);
int     oracle_download(oracle *o, int32_t field, int32_t lb, const int32_t ext[3], double *dst);
int     oracle_init(oracle *o);
int     oracle_step(oracle *o, int32_t nsteps, elmm_diag *diag);
double  oracle_divergence(oracle *o);
const char *oracle_last_error(void);

/* Oracle-only field ids (intermediate arrays). */
enum { ORACLE_F_USTAR = 8, ORACLE_F_VSTAR = 9, ORACLE_F_WSTAR = 10, ORACLE_F_RHS = 11, ORACLE_F_SADV = 32 };

/* Single operators on the current state, for operator-level tests. */
void oracle_op_bound_uvw(oracle *o);           /* BoundUVW(U,V,W)              */
void oracle_op_advection(oracle *o);           /* CD4div* (CD?_variable_z) -> Ustar.. */
void oracle_op_diffusion(oracle *o);           /* diffusion incl. WM -> Ustar.. */
void oracle_op_kappa(oracle *o, int32_t sc);   /* KappaScalar -> Sadv(sc)       */
void oracle_op_sgs(oracle *o);                 /* SubgridStresses              */
void oracle_op_poisson(oracle *o, const double *rhs_interior, double *phi_interior);
double oracle_op_residual(oracle *o, const double *phi_ghosted, const double *rhs_ghosted);
/* the logarithm both the oracle and the GPU library use in the smooth-wall
   law (positive normal x) */
double oracle_det_log(double x);
/* exp, atan and x**y (x >= 0) with fixed operations, evaluated identically
   by the GPU library (Cody-Waite exp, reduced Taylor atan) */
double oracle_det_exp(double x);
double oracle_det_atan(double x);
double oracle_det_pow(double x, double y);
/* Monin-Obukhov pieces (src/wallmodels.f90:1546-1716): PsiM_MO, the
   modified PsiM/PsiH with k_U = 0.4, WM_MO_Flux_ustar, and
   WM_MO_Dirichlet_ustar_tfl with z0H = z0 (ustar, tflux updated in place) */
double oracle_psim_mo(double zeta);
double oracle_psim_mo_mod(double zeta);
double oracle_psih_mo_mod(double zeta);
double oracle_mo_flux_ustar(double vel, double dist, double ustar, double z0, double tflux, double tref, double g,
                            double nu);
void oracle_mo_dirichlet(double *ustar, double *tflux, double dist, double z0, double vel, double tempdif, double nu,
                         double kappa, double g, double tref);
/* WM_MO_Dirichlet_ustar_tfl_mfl with z0H = z0 (ustar, tflux, qflux updated) */
void oracle_mo_dirichlet_moist(double *ustar, double *tflux, double *qflux, double dist, double z0, double vel,
                               double temp, double surf_temp, double moist, double surf_moist, double nu,
                               double kappa, double g, double tref);
void oracle_op_wm_fluxes(oracle *o);            /* ComputeUVWFluxesWM           */
void oracle_op_ib_momentum(oracle *o);          /* IBMomentum(U,V,W)            */
void oracle_op_mass_sources(oracle *o);         /* IBMassSources + Bound_Q -> Q */
void oracle_op_null(oracle *o);                 /* NullInterior(U,V,W)          */
void oracle_op_ib_visc(oracle *o);              /* serial ScalFl gather on Visc */
int32_t oracle_op_wm_div(oracle *o, double *out); /* DivergenceWM -> out (n)    */
void oracle_op_wm_visc(oracle *o);              /* ComputeViscsWM (u*, Visc)    */
void oracle_op_buoyancy(oracle *o);             /* BuoyancyForce on Wstar       */
void oracle_op_temp_fix(oracle *o);             /* temperature in solid cells   */
void oracle_op_bound_temperature(oracle *o);    /* BoundTemperature             */
void oracle_op_moist_fill(oracle *o);           /* moisture in solid cells      */
void oracle_op_bound_moisture(oracle *o);       /* BoundMoisture                */
int32_t oracle_op_poisson_mg(oracle *o, const double *rhs_interior, double *phi_interior, double *res);
void oracle_mg_tables(int nz, double ax, double ay, const double *az, const double *cz, double *den, double *e);
int32_t oracle_wm_ustar(oracle *o, double *out); /* WMPoints%ustar -> out (n)   */
void oracle_op_scalar_extra(oracle *o, int32_t sc); /* ScalarExtra on Sadv(sc)  */
void oracle_op_bound_pr(oracle *o);             /* Bound_Pr(Pr)                 */
void oracle_op_bound_scalar(oracle *o, int32_t sc); /* BoundScalar(S(sc))       */
void oracle_op_bound_visc(oracle *o);           /* BoundViscosity(Visc)         */
/* fluxp/fluxm of momentum wall-model list `list` (comp*6+dir); returns the
   number of points (buffers may be NULL to query it) */
int32_t oracle_wm_list(oracle *o, int32_t list, double *fluxp, double *fluxm);

#ifdef __cplusplus
}
#endif
#endif
