// kernels.hpp - host-side launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "elmm_gpu.h"

namespace elmm {

// ghost fills (k_ghost.hip)
template <typename T> void launch_bound_uvw(const Geo &g, T *U, T *V, T *W, hipStream_t st);
// Ghost cells of component c's BoundUVW shell (count), and the ghost cells
// whose source is one of the n interior cells ijk (device, 3 ints each):
// written to list as (ghost offset, source offset | neg << 62) pairs, at most
// cap; mark is a zeroed byte per cell (left zeroed).  Returns the number of
// entries (synchronises st), -1 on a runtime error.
long ghost_shell_cells(const Geo &g, int c);
int ghost_delta_build(const Geo &g, int c, const int32_t *ijk, int n, uint8_t *mark, long long *list, int cap,
                      int *count, hipStream_t st);
// Re-applies the ghost rules of the concatenated lists (n[0] for U, n[1] for
// V, n[2] for W).
template <typename T>
void launch_ghost_delta(T *U, T *V, T *W, const long long *list, const int n[3], hipStream_t st);
template <typename T> void launch_bound_visc(const Geo &g, T *F, hipStream_t st);
template <typename T> void launch_bound_scalar(const Geo &g, T *F, hipStream_t st);
template <typename T> void launch_bound_pr(const Geo &g, T *P, hipStream_t st);

// Thermal state of wall-model points for the Monin-Obukhov laws
// (src/wallmodels.f90:1546-1966): with buoyancy, a point of class
// ELMM_WM_SURFACE_TEMP (surface temperature tval) takes the Dirichlet law,
// one whose persistent heat flux tflux is positive the flux law.
struct MOParams {
  int buoy;                // 0: neutral laws only
  const int8_t *tclass;    // ELMM_WM_* per point
  const double *tval;      // surface temperature per point
  double *tflux;           // WMPoint%temperature_flux, in/out
  double g, tref, kappa;   // grav_acc, temperature_ref, molecular diffusivity
  int moist;               // with moisture: the moist Dirichlet law
  const double *qval;      // surface moisture per point
  double *qflux;           // WMPoint%moisture_flux, in/out
};

// Wall-model points of the 18 momentum lists, concatenated (SoA).
struct WMPoints {
  int n;
  const int32_t *ijk;      // 3 per point
  const int32_t *list;     // list id 0..17 = comp*6+dir
  const double *dist;      // 3 per point
  const double *z0;
  double *ustar;           // in/out state
  double *fluxp;           // out: flux scaled to the cell beyond the face
  double *fluxm;           // out: to the cell before it (= fluxp on uniform z)
  const double *logdz0;    // log(|dist|/z0) per point (static, from the host libm)
  MOParams mo;             // Monin-Obukhov laws (with buoyancy)
};

// Scalar wall-model points (WMPoints of the reference).
struct WMScalarPoints {
  int n;
  const int32_t *ijk;
  double *div;
};

// WMPoints with their persistent friction velocity (ComputeViscsWM).
struct WMUstarPoints {
  int n;
  const int32_t *ijk;     // 3 per point
  const double *dist;     // 3 per point
  const double *z0;
  const double *logdz0;   // log(|dist| / z0) where z0 > 0 (host libm)
  double *ustar;
  MOParams mo;
  const double *af;       // WMPoint%area_factor
  double *heat;           // out: tflux * af (the temperature's wall source)
  double *mheat;          // out: qflux * af (the moisture's wall source)
};

// IB list on the device (CSR).
struct IBDev {
  int n;
  const int32_t *tgt;      // 3 per point
  const int32_t *ptr;      // n+1
  const int32_t *src;      // 3 per tap
  const double *coef;
};

// Smagorinsky filter width filter_ratio (dxPr dyPr dzPr)^(1/3) tabulated on
// the host: tab[idx[(i-1) + nx (j-1)] * ld + k] (idx null: one column class;
// tab null: evaluated on the device).
// rcp: reciprocals for the FP64 strain quotients, all satisfying div_rn's
// condition: rdx = RN(1/dxmin), rdy = RN(1/dymin), rdz[k] = RN(1/dzPr(k)),
// rz2[k] = RN(1/(2 (zPr(k+1) - zPr(k-1)))) (null when some does not qualify).
struct SgsWidth {
  const double *tab;
  const uint8_t *idx;
  int ld;
  double rdx = 0, rdy = 0;
  const double *rdz = nullptr, *rz2 = nullptr;
  // FP32 level reciprocals RN(1 / dzPr(k)), RN(1 / (2 (zPr(k+1) - zPr(k-1))))
  // in single precision (the precision form's z-march)
  const float *rdzf = nullptr, *rz2f = nullptr;
  bool rcp() const { return rdz != nullptr; }
};

// momentum (k_momentum.hip); T = field type (double or float)
// Number of CFL partial maxima launch_sgs writes (one per block).
int sgs_partials(const Geo &g, const SgsWidth &w);
template <typename T>
void launch_sgs(const Geo &g, const T *U, const T *V, const T *W, T *Visc, double nu, double c_smag,
                double filter_ratio, SgsWidth width, int sgs_model, double *cfl_partial, int cfl_blocks,
                hipStream_t st);
template <typename T>
void launch_wm_fluxes(const Geo &g, const T *U, const T *V, const T *W, const T *Temp, const T *Moist, WMPoints wm,
                      double nu, int variable_z, hipStream_t st);
// Multigrid level (k_multigrid.hip): dense nx x ny x nz arrays, x fastest.
struct MgLevel {
  int nx, ny, nz;
  double ax, ay;  // 1/dx^2, 1/dy^2 of the level
  double *phi, *rhs;
};
// Work arrays and host tables of the exact coarsest-level solve.
struct MgCoarse {
  double2 *Y, *X;      // nx ny nz complex each
  double *cp;          // nx ny nz
  const double *cx, *sx, *cy, *sy;  // cos, -sin of 2 pi q / N
  const double *lam;   // -4 sin^2(pi m/nx) ax - 4 sin^2(pi n/ny) ay per mode
};
template <typename T>
void launch_mg_rhs(const Geo &g, const T *U, const T *V, const T *W, const double *Q, double dt2_rec, double *rhs0,
                   double *RHS, hipStream_t st);
void launch_mg_to_phi(const Geo &g, const double *phi0, double *Phi, hipStream_t st);
void launch_mg_zero(double *a, long n, hipStream_t st);
void launch_mg_sweep(const MgLevel &L, const double *az, const double *den, const double *e, int color,
                     hipStream_t st);
int mg_resmax_blocks(const MgLevel &L);
void launch_mg_resmax(const MgLevel &L, const double *az, const double *cz, double *partial, hipStream_t st);
void launch_mg_restrict(const MgLevel &F, const MgLevel &Cs, const double *az, const double *cz, hipStream_t st);
void launch_mg_prolong(const MgLevel &Cs, const MgLevel &F, hipStream_t st);
void launch_mg_coarse(const MgLevel &L, const MgCoarse &C, const double *az, const double *cz, const double *dzPr,
                      const double *dzW, hipStream_t st);

template <typename R>
void launch_temp_solid_fix(const Geo &g, const int32_t *ijk, int n, const int32_t *prtype, R *T, hipStream_t st);
// Moisture = ref in the solid cells (scalar code CODE_SOLID) of -2..n+3
template <typename R>
void launch_moist_fill(const Geo &g, const uint8_t *code, R *q, double ref, hipStream_t st);
template <typename T>
void launch_wm_ustar(const Geo &g, const T *U, const T *V, const T *W, const T *Temp, const T *Moist,
                     WMUstarPoints wp, double nu, T *Visc, hipStream_t st);
template <typename T>
void launch_momentum(const Geo &g, int comp, const T *U, const T *V, const T *W, const T *Visc, const T *Pr, T *Ustar,
                     T *U2, const uint8_t *code, StageConst sc, hipStream_t st);
// fused three-component momentum kernel, z-marching (k_momentum_zm.hip)
bool momentum_zm_supported(const Geo &g);
size_t momentum_zm_lds_bytes(int elem);
enum { kMomTile32x8 = 0, kMomTile64x4 = 1, kMomTile64x8 = 3 };
template <typename T>
void launch_momentum_zm(const Geo &g, const T *U, const T *V, const T *W, const T *Visc, const T *Pr, const T *Temp,
                        const T *Moist, T *const Us[3], T *const U2[3], const uint8_t *const code[3], StageConst sc,
                        int kchunk, int tile, hipStream_t st);
template <typename T>
void launch_momentum_corr(const Geo &g, int comp, const T *U, const T *V, const T *W, const T *Visc, const T *Pr,
                          T *Ustar, T *U2, const uint8_t *code, StageConst sc, TermList tl, const double *fluxp,
                          const double *fluxm, hipStream_t st);
// The three components' corrections in one launch (same arithmetic as
// launch_momentum_corr per component).
template <typename T>
void launch_momentum_corr3(const Geo &g, const T *U, const T *V, const T *W, const T *Visc, const T *Pr,
                           T *const Ustar[3], T *const U2[3], StageConst sc, const TermList tl[3],
                           const double *fluxp, const double *fluxm, hipStream_t st);
// IBMomentum of the three components: interpolation of every point (tmp
// holds ib[0].n + ib[1].n + ib[2].n values), then the scatter.
template <typename T>
void launch_ib_momentum(const Geo &g, const IBDev ib[3], T *U, T *V, T *W, double *tmp, hipStream_t st);
template <typename T>
void launch_ib_visc_serial(const Geo &g, IBDev ib, T *Visc, const int32_t *level_ptr, int nlevels, hipStream_t st);
template <typename T>
void launch_null(const Geo &g, const int n[3], const int32_t *const ijk[3], T *U, T *V, T *W, hipStream_t st);
template <typename T>
void launch_mass_sources(const Geo &g, TermList tl, const T *U, const T *V, const T *W, double *Q, hipStream_t st);

// scalar (k_scalar.hip); R = field type (double or float)
template <typename R>
void launch_wm_div(const Geo &g, WMScalarPoints p, const int32_t *prtype, const R *U, const R *V, const R *W,
                   hipStream_t st);
template <typename R>
void launch_scalar(const Geo &g, const R *U, const R *V, const R *W, const R *TDiff, const R *S, R *Snew, R *Sadv,
                   const uint8_t *code, const uint8_t *mx, const uint8_t *my, const uint8_t *mz, StageConst sc,
                   hipStream_t st);
// z-marching form for uniform grids (same per-cell arithmetic as launch_scalar)
template <typename R>
void launch_scalar_zm(const Geo &g, const R *U, const R *V, const R *W, const R *TDiff, const R *S, R *Snew, R *Sadv,
                      const uint8_t *code, const uint8_t *mx, const uint8_t *my, const uint8_t *mz, StageConst sc,
                      int kchunk, hipStream_t st);
template <typename R>
void launch_scalar_corr(const Geo &g, const R *U, const R *V, const R *W, const R *TDiff, const R *S, R *Snew,
                        R *Sadv, const uint8_t *code, const uint8_t *mx, const uint8_t *my, const uint8_t *mz,
                        StageConst sc, TermList tl, const double *src_flux, const double *wm_div, hipStream_t st);
template <typename R>
void launch_tdiff_init(const Geo &g, const R *Visc, R *TDiff, double nu, double nu_scalar, hipStream_t st);

// pressure (k_pressure.hip).  R = spectral type of the solve (double: FP64
// form; float: inner solves of the precision form), T = field type.
template <typename R> struct cplx_of;
template <> struct cplx_of<double> { using type = double2; };
template <> struct cplx_of<float> { using type = float2; };
template <typename R> using C2 = typename cplx_of<R>::type;

template <typename R>
struct PoissonPlan {
  int nx, ny, nz, M;            // M = nx/2+1 x-modes per row
  int Mp;                       // row pitch: M rounded up to 128 bytes
  C2<R> *X;                     // spectral work array Mp*ny*nz (m fastest)
  const C2<R> *twx, *twy;       // exp(-2 pi i t/N), t < N/2
  const double *lamx, *lamy;    // eigenvalues of the periodic directions (FP64 in both forms)
  const double *az, *cz;        // z coupling coefficients per k (0-based)
  const double *dzPr, *dzW;     // grid spacings indexed like the reference (k = 1..nz)
  int zpipe = -1;               // z solve load pipelining: -1 by grid size, 0 off, 1 on
};
template <typename T, typename R>
void launch_fft_x_fwd(const Geo &g, const PoissonPlan<R> &pp, const T *U, const T *V, const T *W, const double *Q,
                      const uint32_t *qmask, const double *rhs, const double *phi, double dt2_rec, double *rhs_out,
                      hipStream_t st);
template <typename R> void launch_fft_y(const PoissonPlan<R> &pp, int inverse, hipStream_t st);
// Mixed-radix (2^a 3^b) forms of the transform passes (k_fft_mr.hip); the
// launchers above hand over to them when nx or ny is not a power of two.
template <typename T, typename R>
void launch_fft_x_fwd_mr(const Geo &g, const PoissonPlan<R> &pp, const T *U, const T *V, const T *W, const double *Q,
                         const uint32_t *qmask, const double *rhs, const double *phi, double dt2_rec,
                         double *rhs_out, hipStream_t st);
template <typename R> void launch_fft_y_mr(const PoissonPlan<R> &pp, int inverse, hipStream_t st);
template <typename R>
void launch_fft_x_inv_mr(const Geo &g, const PoissonPlan<R> &pp, double *Phi, int accumulate, hipStream_t st);
// n = 2^a 3^b; a power of two
inline bool fft_len_ok(int n) {
  if (n < 1) return false;
  while (n % 3 == 0) n /= 3;
  return (n & (n - 1)) == 0;
}
inline bool fft_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
template <typename R> void launch_poisson_solve_z(const Geo &g, const PoissonPlan<R> &pp, hipStream_t st);
template <typename R>
void launch_fft_x_inv(const Geo &g, const PoissonPlan<R> &pp, double *Phi, int accumulate, hipStream_t st);
template <typename T>
void launch_top_sum(const Geo &g, const T *Pr, const double *Phi, double *partial, int nblocks, hipStream_t st);
template <typename T>
void launch_projection(const Geo &g, T *U, T *V, T *W, T *Pr, const double *Phi, const double *partial, int nblocks,
                       double dt2, hipStream_t st);
template <typename T>
void launch_max_div(const Geo &g, const T *U, const T *V, const T *W, const double *Q, double *partial, int nblocks,
                    hipStream_t st);
void launch_max_final(const double *partial, int n, double *out, hipStream_t st);
// Step diagnostics kept in device memory (k_monitor.hip).
struct StepMon {
  double cfl_last;        // CFL number of the last step (max * dt)
  double umax_last;       // TimeStepLength's maximum of the last step
  double delta;           // delta of the step in progress
  double delta_last;      // delta of the last completed step
  long long step;         // steps counted by the CFL monitor
  long long first_bad;    // first step (1-based) with CFL > CFL_max, -1 none
};
void launch_cfl_monitor(const double *partial, int nblocks, double dt, double cfl_max, StepMon *mon,
                        hipStream_t st);
int delta_blocks(const Geo &g);
template <typename T>
void launch_delta(const Geo &g, const T *U, const T *U2, const T *V, const T *V2, const T *W, const T *W2,
                  double *partial, int nblocks, int stage, double dt, StepMon *mon, hipStream_t st);

void launch_residual(const Geo &g, const double *Phi, const double *RHS, double *partial, int nblocks,
                     hipStream_t st);
void launch_refine_residual(const Geo &g, const double *Phi, const double *RHS, const double *az, const double *cz,
                            double *r_out, double *partial, int nblocks, hipStream_t st);

}  // namespace elmm
