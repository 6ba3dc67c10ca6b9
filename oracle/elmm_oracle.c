/*
 * elmm_oracle.c - CPU FP64 restatement of ELMM's TMarchRK3 step.
 *
 * TEST INFRASTRUCTURE ONLY (see elmm_oracle.h).  Compile with
 *   gcc -O2 -fopenmp -ffp-contract=off -fno-fast-math
 * so the arithmetic is the reference's, operation by operation.
 *
 * Storage: every 3D field uses the index space -2..n+3 in each direction
 * (n = Prnx/Prny/Prnz), Fortran order.  This contains each reference array's
 * own bounds (U: -2:Unx+3, Pr/Phi/Viscosity: -1:Prn+2, RHS/Q: 0:Prn+1).
 */
#include "elmm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];
static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
const char *oracle_last_error(void) { return g_err; }

/* RK3 coefficients, src/parameters.f90:312-322. */
static const double RK_alpha[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0};
static const double RK_beta[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double RK_rho[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};

enum { We = 0, Ea = 1, So = 2, No = 3, Bo = 4, To = 5 };
enum { NOSLIP = 1, FREESLIP = 2, PERIODIC = 3, DIRICHLET = 4, NEUMANN = 5, CONSTFLUX = 6 };

typedef struct {
  int n;
  int32_t *ijk;
  double *dist, *z0, *ustar, *fluxp, *fluxm;
  int8_t *tclass;  /* ELMM_WM_* (prescribed_temperature, temperature_flux) */
  double *tval;    /* surface temperature or heat flux */
  double *tflux;   /* WMPoint%temperature_flux, persistent */
  double *qval;    /* WMPoint%moisture (surface moisture) */
  double *qflux;   /* WMPoint%moisture_flux, persistent */
} wm_list;

struct oracle {
  elmm_config c;
  int nx, ny, nz;
  int Un[3], Vn[3], Wn[3];
  long sy, sz, ntot;
  int variable_z;
  double dxmin, dymin, dzmin;
  /* 1-D coordinate arrays, stored with offset 3 (index -3 at position 0) */
  double *xU, *yV, *zW, *xPr, *yPr, *zPr, *dxPr, *dyPr, *dzPr, *dxU, *dyV, *dzW;
  double *U, *V, *W, *U2, *V2, *W2, *Us, *Vs, *Ws;
  double *Pr, *Visc, *TDiff, *Q, *Phi, *RHS;
  double **S, **S2, **Sadv;
  int32_t *Prtype, *Utype, *Vtype, *Wtype;
  uint8_t *scflmask[3];
  wm_list wm[18];
  wm_list wmp;
  double *wmp_div;
  double *wmp_af;   /* WMPoint%area_factor */
  int temp_bt[6];   /* TempBtype after GetOutsideBoundariesWM (0 on MO walls) */
  int moist_bt[6];  /* MoistBtype after it (0 on walls with a moisture flux) */
  int scfl_n[3];
  int32_t *scfl_ijk[3];
  int ib_n[4];
  int32_t *ib_ijk[4], *ib_ptr[4], *ib_src[4];
  double *ib_coef[4];
  int null_n[3];
  int32_t *null_ijk[3];
  int src_n;
  int32_t *src_ijk, *src_sc;
  double *src_flux;
  int initialized, first_step_done;
  double last_res, last_div;
  int last_mg_cycles;
  /* time stepping (src/main.f90:81-169) */
  double time, dt_last, delta_last, cfl_last;
  long long steps_total;
};

#define C1D(a, i) ((a)[(i) + 3])
#define IX(o, i, j, k) ((long)((i) + 2) + (long)((j) + 2) * (o)->sy + (long)((k) + 2) * (o)->sz)
#define A3(o, a, i, j, k) ((a)[IX(o, i, j, k)])

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

/* ---------------------------------------------------------------- grid --- */
/* Coordinates as in InitBoundaryConditions, src/initial.f90:4395-4502. */
static void init_grid(oracle *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  o->xU = xcalloc(nx + 8, sizeof(double));
  o->yV = xcalloc(ny + 8, sizeof(double));
  o->zW = xcalloc(nz + 8, sizeof(double));
  o->xPr = xcalloc(nx + 8, sizeof(double));
  o->yPr = xcalloc(ny + 8, sizeof(double));
  o->zPr = xcalloc(nz + 8, sizeof(double));
  o->dxPr = xcalloc(nx + 8, sizeof(double));
  o->dyPr = xcalloc(ny + 8, sizeof(double));
  o->dzPr = xcalloc(nz + 8, sizeof(double));
  o->dxU = xcalloc(nx + 8, sizeof(double));
  o->dyV = xcalloc(ny + 8, sizeof(double));
  o->dzW = xcalloc(nz + 8, sizeof(double));
  o->dxmin = o->c.dx;
  o->dymin = o->c.dy;
  for (int i = -3; i <= nx + 3; i++) C1D(o->xU, i) = i * o->dxmin + o->c.x0;
  for (int j = -3; j <= ny + 3; j++) C1D(o->yV, j) = j * o->dymin + o->c.y0;
  if (o->variable_z) {
    for (int k = 0; k <= nz; k++) C1D(o->zW, k) = o->c.zw[k];
    /* bottom extension (src/initial.f90:4418-4421 for a non-periodic bottom) */
    if (o->c.btype[Bo] == PERIODIC) {
      for (int j = -1; j >= -3; j--)
        C1D(o->zW, j) = C1D(o->zW, 0) - (C1D(o->zW, nz - 1) - C1D(o->zW, nz - 1 + j));
    } else {
      for (int j = -1; j >= -3; j--) C1D(o->zW, j) = C1D(o->zW, 0) - (C1D(o->zW, -j) - C1D(o->zW, 0));
    }
    if (o->c.btype[To] == PERIODIC) {
      for (int j = nz + 1; j <= nz + 3; j++)
        C1D(o->zW, j) = C1D(o->zW, nz - 1) + (C1D(o->zW, j - nz + 1) - C1D(o->zW, 0));
    } else {
      for (int j = nz + 1; j <= nz + 3; j++)
        C1D(o->zW, j) = C1D(o->zW, nz - 1) + (C1D(o->zW, nz - 1) - C1D(o->zW, nz - 1 - (j - nz + 1)));
    }
  } else {
    o->dzmin = o->c.dz;
    for (int k = -3; k <= nz + 3; k++) C1D(o->zW, k) = k * o->dzmin + o->c.z0;
  }
  for (int i = -2; i <= nx + 3; i++) {
    C1D(o->xPr, i) = (C1D(o->xU, i - 1) + C1D(o->xU, i)) / 2.0;
    C1D(o->dxPr, i) = C1D(o->xU, i) - C1D(o->xU, i - 1);
  }
  for (int j = -2; j <= ny + 3; j++) {
    C1D(o->yPr, j) = (C1D(o->yV, j - 1) + C1D(o->yV, j)) / 2.0;
    C1D(o->dyPr, j) = C1D(o->yV, j) - C1D(o->yV, j - 1);
  }
  for (int k = -2; k <= nz + 3; k++) {
    C1D(o->zPr, k) = (C1D(o->zW, k - 1) + C1D(o->zW, k)) / 2.0;
    C1D(o->dzPr, k) = C1D(o->zW, k) - C1D(o->zW, k - 1);
  }
  for (int i = -2; i <= nx + 2; i++) C1D(o->dxU, i) = C1D(o->xPr, i + 1) - C1D(o->xPr, i);
  for (int j = -2; j <= ny + 2; j++) C1D(o->dyV, j) = C1D(o->yPr, j + 1) - C1D(o->yPr, j);
  for (int k = -2; k <= nz + 2; k++) C1D(o->dzW, k) = C1D(o->zPr, k + 1) - C1D(o->zPr, k);
  if (o->variable_z) {
    double m = C1D(o->dzPr, -2);
    for (int k = -2; k <= nz + 3; k++)
      if (C1D(o->dzPr, k) < m) m = C1D(o->dzPr, k);
    o->dzmin = m;
  }
}

/* ------------------------------------------------------------ creation --- */
static double *field(oracle *o) { return xcalloc(o->ntot, sizeof(double)); }

static void copy_wm(wm_list *d, int n, const int32_t *ijk, const double *dist, const double *z0, const int8_t *tc,
                    const double *tv, const double *qv, const double *qf) {
  d->n = n;
  d->qval = xcalloc((size_t)n + 1, sizeof(double));
  d->qflux = xcalloc((size_t)n + 1, sizeof(double));
  for (int p = 0; p < n && qv; p++) {
    d->qval[p] = qv[p];
    d->qflux[p] = qf[p];
  }
  d->tclass = xcalloc((size_t)n + 1, sizeof(int8_t));
  d->tval = xcalloc((size_t)n + 1, sizeof(double));
  d->tflux = xcalloc((size_t)n + 1, sizeof(double));
  for (int p = 0; p < n && tc; p++) {
    d->tclass[p] = tc[p];
    d->tval[p] = tv[p];
    /* temperature_flux = sideTemp for a flux wall, else 0 (wallmodels.f90:592-596) */
    d->tflux[p] = tc[p] == ELMM_WM_SURFACE_FLUX ? tv[p] : 0.0;
  }
  d->ijk = xcalloc(3 * (size_t)n, sizeof(int32_t));
  d->dist = xcalloc(3 * (size_t)n, sizeof(double));
  d->z0 = xcalloc((size_t)n, sizeof(double));
  d->ustar = xcalloc((size_t)n, sizeof(double));
  d->fluxp = xcalloc((size_t)n, sizeof(double));
  d->fluxm = xcalloc((size_t)n, sizeof(double));
  if (n) {
    memcpy(d->ijk, ijk, 3 * (size_t)n * sizeof(int32_t));
    memcpy(d->dist, dist, 3 * (size_t)n * sizeof(double));
    memcpy(d->z0, z0, (size_t)n * sizeof(double));
  }
  for (int p = 0; p < n; p++) d->ustar[p] = 1.0; /* WMpointUVW%ustar = 1 (wallmodels.f90:79) */
}

/* ComputeAreaFactor, src/wallmodels.f90:170-203: the wall area over the
   cell volume from the dominant component of the point's wall normal
   (norm2 as the square root of the sum of squares). */
static double area_factor(const oracle *o, const double *d) {
  if (-d[0] == 0 && -d[1] == 0 && -d[2] == 0) return 0;
  const double V = o->dxmin * o->dymin * o->dzmin;
  const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double n0 = fabs(-d[0] / nrm), n1 = fabs(-d[1] / nrm), n2 = fabs(-d[2] / nrm);
  double on, dd;
  if (n0 > n1 && n0 > n2) {
    on = n0;
    dd = sqrt(o->dymin * o->dzmin);
  } else if (n1 > n0 && n1 > n2) {
    on = n1;
    dd = sqrt(o->dxmin * o->dzmin);
  } else {
    on = n2;
    dd = sqrt(o->dxmin * o->dymin);
  }
  const double l = dd / on;
  const double S = dd * l;
  return S / V;
}

oracle *oracle_create(const elmm_config *cfg, const elmm_lists *L) {
  if (!cfg || !L) {
    set_err("null argument");
    return NULL;
  }
  if (cfg->nx < 4 || cfg->ny < 4 || cfg->nz < 4) {
    set_err("grid too small");
    return NULL;
  }
  for (int s = 0; s < 6; s++) {
    int b = cfg->btype[s];
    if (b != NOSLIP && b != FREESLIP && b != PERIODIC) {
      set_err("velocity boundary type %d unsupported on side %d", b, s);
      return NULL;
    }
  }
  if ((cfg->btype[We] == PERIODIC) != (cfg->btype[Ea] == PERIODIC) ||
      (cfg->btype[So] == PERIODIC) != (cfg->btype[No] == PERIODIC) ||
      (cfg->btype[Bo] == PERIODIC) != (cfg->btype[To] == PERIODIC)) {
    set_err("periodic boundaries must be paired");
    return NULL;
  }
  if (cfg->btype[We] != PERIODIC || cfg->btype[So] != PERIODIC || cfg->btype[Bo] == PERIODIC) {
    set_err("pressure solver supports periodic x/y with walls in z");
    return NULL;
  }
  oracle *o = xcalloc(1, sizeof(oracle));
  o->c = *cfg;
  o->time = cfg->start_time;
  o->nx = cfg->nx;
  o->ny = cfg->ny;
  o->nz = cfg->nz;
  o->variable_z = cfg->stretched_z != 0;
  if (o->variable_z) {
    double *zw = xcalloc(cfg->nz + 1, sizeof(double));
    memcpy(zw, cfg->zw, (cfg->nz + 1) * sizeof(double));
    o->c.zw = zw;
  }
  /* extents, src/initial.f90:871-896 */
  o->Un[0] = cfg->btype[Ea] == PERIODIC ? o->nx : o->nx - 1;
  o->Un[1] = o->ny;
  o->Un[2] = o->nz;
  o->Vn[0] = o->nx;
  o->Vn[1] = cfg->btype[No] == PERIODIC ? o->ny : o->ny - 1;
  o->Vn[2] = o->nz;
  o->Wn[0] = o->nx;
  o->Wn[1] = o->ny;
  o->Wn[2] = cfg->btype[To] == PERIODIC ? o->nz : o->nz - 1;
  o->sy = o->nx + 6;
  o->sz = (long)(o->nx + 6) * (o->ny + 6);
  o->ntot = o->sz * (o->nz + 6);
  init_grid(o);

  o->U = field(o);
  o->V = field(o);
  o->W = field(o);
  o->U2 = field(o);
  o->V2 = field(o);
  o->W2 = field(o);
  o->Us = field(o);
  o->Vs = field(o);
  o->Ws = field(o);
  o->Pr = field(o);
  o->Visc = field(o);
  o->TDiff = field(o);
  o->Q = field(o);
  o->Phi = field(o);
  o->RHS = field(o);
  /* scalar slots: the passive scalars, then the temperature (buoyancy),
     then the moisture */
  int ns = cfg->nscalars + (cfg->enable_buoyancy ? 1 : 0) + (cfg->enable_buoyancy && cfg->enable_moisture ? 1 : 0);
  o->S = xcalloc(ns + 1, sizeof(double *));
  o->S2 = xcalloc(ns + 1, sizeof(double *));
  o->Sadv = xcalloc(ns + 1, sizeof(double *));
  for (int s = 0; s < ns; s++) {
    o->S[s] = field(o);
    o->S2[s] = field(o);
    o->Sadv[s] = field(o);
  }
  o->Prtype = xcalloc(o->ntot, sizeof(int32_t));
  o->Utype = xcalloc(o->ntot, sizeof(int32_t));
  o->Vtype = xcalloc(o->ntot, sizeof(int32_t));
  o->Wtype = xcalloc(o->ntot, sizeof(int32_t));
  memcpy(o->Prtype, L->prtype, o->ntot * sizeof(int32_t));
  memcpy(o->Utype, L->utype, o->ntot * sizeof(int32_t));
  memcpy(o->Vtype, L->vtype, o->ntot * sizeof(int32_t));
  memcpy(o->Wtype, L->wtype, o->ntot * sizeof(int32_t));
  for (int d = 0; d < 3; d++) {
    o->scflmask[d] = xcalloc(o->ntot, 1);
    memcpy(o->scflmask[d], L->scfl_mask[d], o->ntot);
  }
  for (int l = 0; l < 18; l++)
    copy_wm(&o->wm[l], L->wm_n[l], L->wm_ijk[l], L->wm_dist[l], L->wm_z0[l], L->wm_tclass[l], L->wm_tval[l],
            L->wm_qval[l], L->wm_qflux[l]);
  copy_wm(&o->wmp, L->wmp_n, L->wmp_ijk, L->wmp_dist, L->wmp_z0, L->wmp_tclass, L->wmp_tval, L->wmp_qval,
          L->wmp_qflux);
  o->wmp_div = xcalloc(L->wmp_n + 1, sizeof(double));
  o->wmp_af = xcalloc(L->wmp_n + 1, sizeof(double));
  for (int p = 0; p < L->wmp_n; p++) o->wmp_af[p] = area_factor(o, &o->wmp.dist[3 * p]);
  /* GetOutsideBoundariesWM sets TempBtype = 0 (the mirror ghost rule) on a
     no-slip wall with a Dirichlet or flux temperature (wallmodels.f90:618-620,
     678-680) */
  for (int q = 0; q < 6; q++) {
    o->temp_bt[q] = o->c.temp_btype[q];
    if (q >= Bo && o->c.enable_buoyancy && o->c.wall_model && o->c.btype[q] == NOSLIP &&
        (o->c.temp_btype[q] == DIRICHLET || o->c.temp_btype[q] == CONSTFLUX))
      o->temp_bt[q] = 0;
    o->moist_bt[q] = o->c.moist_btype[q];
    if (q >= Bo && o->c.enable_buoyancy && o->c.enable_moisture && o->c.wall_model && o->c.btype[q] == NOSLIP &&
        o->c.moist_btype[q] == CONSTFLUX)
      o->moist_bt[q] = 0;
  }
  for (int d = 0; d < 3; d++) {
    o->scfl_n[d] = L->scfl_n[d];
    o->scfl_ijk[d] = xcalloc(3 * (size_t)L->scfl_n[d] + 1, sizeof(int32_t));
    if (L->scfl_n[d]) memcpy(o->scfl_ijk[d], L->scfl_ijk[d], 3 * (size_t)L->scfl_n[d] * sizeof(int32_t));
  }
  for (int c = 0; c < 4; c++) {
    int n = L->ib_n[c];
    o->ib_n[c] = n;
    o->ib_ijk[c] = xcalloc(3 * (size_t)n + 1, sizeof(int32_t));
    o->ib_ptr[c] = xcalloc((size_t)n + 1, sizeof(int32_t));
    int nt = n ? L->ib_ptr[c][n] : 0;
    o->ib_src[c] = xcalloc(3 * (size_t)nt + 1, sizeof(int32_t));
    o->ib_coef[c] = xcalloc((size_t)nt + 1, sizeof(double));
    if (n) {
      memcpy(o->ib_ijk[c], L->ib_ijk[c], 3 * (size_t)n * sizeof(int32_t));
      memcpy(o->ib_ptr[c], L->ib_ptr[c], ((size_t)n + 1) * sizeof(int32_t));
      memcpy(o->ib_src[c], L->ib_src_ijk[c], 3 * (size_t)nt * sizeof(int32_t));
      memcpy(o->ib_coef[c], L->ib_coef[c], (size_t)nt * sizeof(double));
    }
  }
  for (int c = 0; c < 3; c++) {
    o->null_n[c] = L->null_n[c];
    o->null_ijk[c] = xcalloc(3 * (size_t)L->null_n[c] + 1, sizeof(int32_t));
    if (L->null_n[c]) memcpy(o->null_ijk[c], L->null_ijk[c], 3 * (size_t)L->null_n[c] * sizeof(int32_t));
  }
  o->src_n = L->src_n;
  o->src_ijk = xcalloc(3 * (size_t)L->src_n + 1, sizeof(int32_t));
  o->src_sc = xcalloc((size_t)L->src_n + 1, sizeof(int32_t));
  o->src_flux = xcalloc((size_t)L->src_n + 1, sizeof(double));
  if (L->src_n) {
    memcpy(o->src_ijk, L->src_ijk, 3 * (size_t)L->src_n * sizeof(int32_t));
    memcpy(o->src_sc, L->src_scalar, (size_t)L->src_n * sizeof(int32_t));
    memcpy(o->src_flux, L->src_flux, (size_t)L->src_n * sizeof(double));
  }
  return o;
}

static void free_wm(wm_list *l) {
  free(l->qval);
  free(l->qflux);
  free(l->tclass);
  free(l->tval);
  free(l->tflux);
  free(l->ijk);
  free(l->dist);
  free(l->z0);
  free(l->ustar);
  free(l->fluxp);
  free(l->fluxm);
}

void oracle_destroy(oracle *o) {
  if (!o) return;
  double *arrs[] = {o->xU, o->yV, o->zW, o->xPr, o->yPr, o->zPr, o->dxPr, o->dyPr, o->dzPr, o->dxU, o->dyV,
                    o->dzW, o->U, o->V, o->W, o->U2, o->V2, o->W2, o->Us, o->Vs, o->Ws, o->Pr, o->Visc,
                    o->TDiff, o->Q, o->Phi, o->RHS};
  for (size_t i = 0; i < sizeof arrs / sizeof arrs[0]; i++) free(arrs[i]);
  for (int s = 0; s < o->c.nscalars + (o->c.enable_buoyancy ? 1 : 0) + (o->c.enable_buoyancy && o->c.enable_moisture ? 1 : 0);
       s++) {
    free(o->S[s]);
    free(o->S2[s]);
    free(o->Sadv[s]);
  }
  free(o->S);
  free(o->S2);
  free(o->Sadv);
  free(o->Prtype);
  free(o->Utype);
  free(o->Vtype);
  free(o->Wtype);
  for (int d = 0; d < 3; d++) {
    free(o->scflmask[d]);
    free(o->scfl_ijk[d]);
    free(o->null_ijk[d]);
  }
  for (int l = 0; l < 18; l++) free_wm(&o->wm[l]);
  free_wm(&o->wmp);
  free(o->wmp_div);
  free(o->wmp_af);
  for (int c = 0; c < 4; c++) {
    free(o->ib_ijk[c]);
    free(o->ib_ptr[c]);
    free(o->ib_src[c]);
    free(o->ib_coef[c]);
  }
  free(o->src_ijk);
  free(o->src_sc);
  free(o->src_flux);
  if (o->variable_z) free((void *)o->c.zw);
  free(o);
}

static double *field_ptr(oracle *o, int32_t f) {
  switch (f) {
    case ELMM_F_U: return o->U;
    case ELMM_F_V: return o->V;
    case ELMM_F_W: return o->W;
    case ELMM_F_PR: return o->Pr;
    case ELMM_F_VISC: return o->Visc;
    case ELMM_F_TDIFF: return o->TDiff;
    case ELMM_F_PHI: return o->Phi;
    case ELMM_F_Q: return o->Q;
    case ORACLE_F_USTAR: return o->Us;
    case ORACLE_F_VSTAR: return o->Vs;
    case ORACLE_F_WSTAR: return o->Ws;
    case ORACLE_F_RHS: return o->RHS;
    default:
      if (f >= ELMM_F_SCALAR && f < ELMM_F_SCALAR + o->c.nscalars) return o->S[f - ELMM_F_SCALAR];
      if (f == ELMM_F_TEMPERATURE && o->c.enable_buoyancy) return o->S[o->c.nscalars];
      if (f == ELMM_F_MOISTURE && o->c.enable_buoyancy && o->c.enable_moisture) return o->S[o->c.nscalars + 1];
      if (f >= ORACLE_F_SADV && f < ORACLE_F_SADV + o->c.nscalars) return o->Sadv[f - ORACLE_F_SADV];
  }
  return NULL;
}

static int xfer(oracle *o, int32_t f, int32_t lb, const int32_t ext[3], double *host, int up) {
  double *a = field_ptr(o, f);
  if (!a) {
    set_err("bad field %d", f);
    return ELMM_E_ARG;
  }
  int n[3] = {o->nx, o->ny, o->nz};
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = lb > -2 ? lb : -2;
    int top = lb + ext[d] - 1;
    hi[d] = top < n[d] + 3 ? top : n[d] + 3;
  }
  long hsy = ext[0], hsz = (long)ext[0] * ext[1];
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) {
        long h = (i - lb) + (j - lb) * hsy + (k - lb) * hsz;
        if (up)
          A3(o, a, i, j, k) = host[h];
        else
          host[h] = A3(o, a, i, j, k);
      }
  return 0;
}
int oracle_upload(oracle *o, int32_t f, int32_t lb, const int32_t ext[3], const double *src) {
  return xfer(o, f, lb, ext, (double *)src, 1);
}
int oracle_download(oracle *o, int32_t f, int32_t lb, const int32_t ext[3], double *dst) {
  return xfer(o, f, lb, ext, dst, 0);
}

/* ------------------------------------------------------- ghost fills --- */

/* Fills of one velocity component along one axis; the reference's rules per
 * boundary type (src/boundaries.f90:461-572 west, 783-894 east, and the same
 * pattern for y 1122-1233/1323-1434 and z 1533-1647/1676-1789). */
static void vel_low(double *a, long base, long st, int n, int bt, int normal) {
#define E(m) a[base + (long)(m) * st]
  if (bt == PERIODIC) {
    E(0) = E(n);
    E(-1) = E(n - 1);
    E(-2) = E(n - 2);
  } else if (normal) { /* noslip and freeslip: normal component */
    E(0) = 0;
    E(-1) = -E(1);
    E(-2) = -E(2);
  } else if (bt == NOSLIP) {
    E(0) = -E(1);
    E(-1) = -E(2);
    E(-2) = -E(3);
  } else { /* FREESLIP tangential */
    E(0) = E(1);
    E(-1) = E(1);
    E(-2) = E(1);
  }
}
static void vel_high(double *a, long base, long st, int n, int bt, int normal) {
  if (bt == PERIODIC) {
    E(n + 1) = E(1);
    E(n + 2) = E(2);
    E(n + 3) = E(3);
  } else if (normal) {
    E(n + 1) = 0;
    E(n + 2) = -E(n);
    E(n + 3) = -E(n - 1);
  } else if (bt == NOSLIP) {
    E(n + 1) = -E(n);
    E(n + 2) = -E(n - 1);
    E(n + 3) = -E(n - 2);
  } else {
    E(n + 1) = E(n);
    E(n + 2) = E(n);
    E(n + 3) = E(n);
  }
#undef E
}

/* periodic_corners_edges, src/boundaries.f90:189-394 */
static void periodic_corners_edges(oracle *o, double *U, const int *n) {
  int nx = n[0], ny = n[1], nz = n[2];
  int px = o->c.btype[Ea] == PERIODIC, py = o->c.btype[No] == PERIODIC, pz = o->c.btype[To] == PERIODIC;
  if (px && py && pz) {
    for (int sk = 0; sk < 2; sk++)
      for (int sj = 0; sj < 2; sj++)
        for (int si = 0; si < 2; si++) {
          int k0 = sk ? nz + 1 : -2, j0 = sj ? ny + 1 : -2, i0 = si ? nx + 1 : -2;
          int dk = sk ? -nz : nz, dj = sj ? -ny : ny, di = si ? -nx : nx;
          for (int k = k0; k < k0 + 3; k++)
            for (int j = j0; j < j0 + 3; j++)
              for (int i = i0; i < i0 + 3; i++) A3(o, U, i, j, k) = A3(o, U, i + di, j + dj, k + dk);
        }
  }
  if (px && py) {
    for (int sj = 0; sj < 2; sj++)
      for (int si = 0; si < 2; si++) {
        int j0 = sj ? ny + 1 : -2, i0 = si ? nx + 1 : -2;
        int dj = sj ? -ny : ny, di = si ? -nx : nx;
        for (int k = 1; k <= nz; k++)
          for (int j = j0; j < j0 + 3; j++)
            for (int i = i0; i < i0 + 3; i++) A3(o, U, i, j, k) = A3(o, U, i + di, j + dj, k);
      }
  }
  if (px && pz) {
    for (int sk = 0; sk < 2; sk++)
      for (int si = 0; si < 2; si++) {
        int k0 = sk ? nz + 1 : -2, i0 = si ? nx + 1 : -2;
        int dk = sk ? -nz : nz, di = si ? -nx : nx;
        for (int k = k0; k < k0 + 3; k++)
          for (int j = 1; j <= ny; j++)
            for (int i = i0; i < i0 + 3; i++) A3(o, U, i, j, k) = A3(o, U, i + di, j, k + dk);
      }
  }
  if (py && pz) {
    for (int sk = 0; sk < 2; sk++)
      for (int sj = 0; sj < 2; sj++) {
        int k0 = sk ? nz + 1 : -2, j0 = sj ? ny + 1 : -2;
        int dk = sk ? -nz : nz, dj = sj ? -ny : ny;
        for (int k = k0; k < k0 + 3; k++)
          for (int j = j0; j < j0 + 3; j++)
            for (int i = 1; i <= nx; i++) A3(o, U, i, j, k) = A3(o, U, i, j + dj, k + dk);
      }
  }
}

/* BoundUVW_do_BC, src/boundaries.f90:398-1791 (periodic, noslip, freeslip). */
static void bound_uvw(oracle *o, double *U, double *V, double *W) {
  double *F[3] = {U, V, W};
  const int *N[3] = {o->Un, o->Vn, o->Wn};
  for (int c = 0; c < 3; c++) periodic_corners_edges(o, F[c], N[c]);
  /* x faces: j,k interior */
  for (int c = 0; c < 3; c++) {
    const int *n = N[c];
    for (int k = 1; k <= n[2]; k++)
      for (int j = 1; j <= n[1]; j++) {
        long b = IX(o, 0, j, k);
        vel_low(F[c], b, 1, n[0], o->c.btype[We], c == 0);
        vel_high(F[c], b, 1, n[0], o->c.btype[Ea], c == 0);
      }
  }
  /* y faces: i = -2..n+3, k interior */
  for (int c = 0; c < 3; c++) {
    const int *n = N[c];
    for (int k = 1; k <= n[2]; k++)
      for (int i = -2; i <= n[0] + 3; i++) {
        long b = IX(o, i, 0, k);
        vel_low(F[c], b, o->sy, n[1], o->c.btype[So], c == 1);
        vel_high(F[c], b, o->sy, n[1], o->c.btype[No], c == 1);
      }
  }
  /* z faces: i = -2..n+3, j = -2..n+3 */
  for (int c = 0; c < 3; c++) {
    const int *n = N[c];
    for (int j = -2; j <= n[1] + 3; j++)
      for (int i = -2; i <= n[0] + 3; i++) {
        long b = IX(o, i, j, 0);
        vel_low(F[c], b, o->sz, n[2], o->c.btype[Bo], c == 2);
        vel_high(F[c], b, o->sz, n[2], o->c.btype[To], c == 2);
      }
  }
}

/* CommonBase, src/scalarboundaries.f90:18-279 (Dirichlet, periodic, Neumann). */
static void bound_scalar_bc(oracle *o, double *a, const int *bt, const double *side) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
#define S_(i, j, k) A3(o, a, i, j, k)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      if (bt[We] == DIRICHLET) {
        S_(-2, j, k) = side[We] - (S_(3, j, k) - side[We]);
        S_(-1, j, k) = side[We] - (S_(2, j, k) - side[We]);
        S_(0, j, k) = side[We] - (S_(1, j, k) - side[We]);
      } else if (bt[We] == PERIODIC) {
        S_(-2, j, k) = S_(nx - 2, j, k);
        S_(-1, j, k) = S_(nx - 1, j, k);
        S_(0, j, k) = S_(nx, j, k);
      } else if (bt[We] == NEUMANN) {
        S_(0, j, k) = S_(1, j, k) - side[We] * o->dxmin;
        S_(-1, j, k) = S_(1, j, k) - side[We] * 2 * o->dxmin;
        S_(-2, j, k) = S_(1, j, k) - side[We] * 3 * o->dxmin;
      }
      if (bt[Ea] == DIRICHLET) {
        S_(nx + 1, j, k) = side[Ea] - (S_(nx, j, k) - side[Ea]);
        S_(nx + 2, j, k) = side[Ea] - (S_(nx - 1, j, k) - side[Ea]);
        S_(nx + 3, j, k) = side[Ea] - (S_(nx - 2, j, k) - side[Ea]);
      } else if (bt[Ea] == PERIODIC) {
        S_(nx + 1, j, k) = S_(1, j, k);
        S_(nx + 2, j, k) = S_(2, j, k);
        S_(nx + 3, j, k) = S_(3, j, k);
      } else if (bt[Ea] == NEUMANN) {
        S_(nx + 1, j, k) = S_(nx, j, k) + side[Ea] * o->dxmin;
        S_(nx + 2, j, k) = S_(nx, j, k) + side[Ea] * 2 * o->dxmin;
        S_(nx + 3, j, k) = S_(nx, j, k) + side[Ea] * 3 * o->dxmin;
      }
    }
  for (int k = 1; k <= nz; k++)
    for (int i = -1; i <= nx + 2; i++) {
      if (bt[So] == DIRICHLET) {
        S_(i, -2, k) = side[So] - (S_(i, 3, k) - side[So]);
        S_(i, -1, k) = side[So] - (S_(i, 2, k) - side[So]);
        S_(i, 0, k) = side[So] - (S_(i, 1, k) - side[So]);
      } else if (bt[So] == PERIODIC) {
        S_(i, -2, k) = S_(i, ny - 2, k);
        S_(i, -1, k) = S_(i, ny - 1, k);
        S_(i, 0, k) = S_(i, ny, k);
      } else if (bt[So] == NEUMANN) {
        S_(i, 0, k) = S_(i, 1, k) - side[So] * o->dymin;
        S_(i, -1, k) = S_(i, 1, k) - side[So] * 2 * o->dymin;
        S_(i, -2, k) = S_(i, 1, k) - side[So] * 3 * o->dymin;
      }
      if (bt[No] == DIRICHLET) {
        S_(i, ny + 1, k) = side[No] - (S_(i, ny, k) - side[No]);
        S_(i, ny + 2, k) = side[No] - (S_(i, ny - 1, k) - side[No]);
        S_(i, ny + 3, k) = side[No] - (S_(i, ny - 2, k) - side[No]);
      } else if (bt[No] == PERIODIC) {
        S_(i, ny + 1, k) = S_(i, 1, k);
        S_(i, ny + 2, k) = S_(i, 2, k);
        S_(i, ny + 3, k) = S_(i, 3, k);
      } else if (bt[No] == NEUMANN) {
        S_(i, ny + 1, k) = S_(i, ny, k) + side[No] * o->dymin;
        S_(i, ny + 2, k) = S_(i, ny, k) + side[No] * 2 * o->dymin;
        S_(i, ny + 3, k) = S_(i, ny, k) + side[No] * 3 * o->dymin;
      }
    }
  for (int j = -1; j <= ny + 2; j++)
    for (int i = -1; i <= nx + 2; i++) {
      if (bt[Bo] == DIRICHLET) {
        S_(i, j, -2) = side[Bo] - (S_(i, j, 3) - side[Bo]);
        S_(i, j, -1) = side[Bo] - (S_(i, j, 2) - side[Bo]);
        S_(i, j, 0) = side[Bo] - (S_(i, j, 1) - side[Bo]);
      } else if (bt[Bo] == PERIODIC) { /* reversed order as written (scalarboundaries.f90:213) */
        S_(i, j, -2) = S_(i, j, nz);
        S_(i, j, -1) = S_(i, j, nz - 1);
        S_(i, j, 0) = S_(i, j, nz - 2);
      } else if (bt[Bo] == NEUMANN) {
        S_(i, j, 0) = S_(i, j, 1) - side[Bo] * C1D(o->dzW, 0);
        S_(i, j, -1) = S_(i, j, 1) - side[Bo] * (C1D(o->zPr, 1) - C1D(o->zPr, -1));
        S_(i, j, -2) = S_(i, j, 1) - side[Bo] * (C1D(o->zPr, 1) - C1D(o->zPr, -2));
      } else if (bt[Bo] == 0) { /* any other type: the mirror (scalarboundaries.f90:233-239) */
        S_(i, j, -2) = S_(i, j, 3);
        S_(i, j, -1) = S_(i, j, 2);
        S_(i, j, 0) = S_(i, j, 1);
      }
      if (bt[To] == DIRICHLET) {
        S_(i, j, nz + 1) = side[To] - (S_(i, j, nz) - side[To]);
        S_(i, j, nz + 2) = side[To] - (S_(i, j, nz - 1) - side[To]);
        S_(i, j, nz + 3) = side[To] - (S_(i, j, nz - 2) - side[To]);
      } else if (bt[To] == PERIODIC) {
        /* arr(i,j,nz+1:3) = arr(i,j,1:3) is an empty section for nz > 2
           (scalarboundaries.f90:251) */
      } else if (bt[To] == NEUMANN) {
        S_(i, j, nz + 1) = S_(i, j, nz) + side[To] * C1D(o->dzW, nz + 1);
        S_(i, j, nz + 2) = S_(i, j, nz) + side[To] * (C1D(o->zPr, nz + 2) - C1D(o->zPr, nz));
        S_(i, j, nz + 3) = S_(i, j, nz) + side[To] * (C1D(o->zPr, nz + 3) - C1D(o->zPr, nz));
      } else if (bt[To] == 0) { /* the mirror (scalarboundaries.f90:271-277) */
        S_(i, j, nz + 1) = S_(i, j, nz);
        S_(i, j, nz + 2) = S_(i, j, nz - 1);
        S_(i, j, nz + 3) = S_(i, j, nz - 2);
      }
    }
#undef S_
}

/* BoundScalar (ScalBtype, sideScal) and BoundTemperature (TempBtype,
   sideTemp; src/scalarboundaries.f90:296-308): the same rules. */
static void bound_scalar(oracle *o, double *a) { bound_scalar_bc(o, a, o->c.scal_btype, o->c.side_scal); }
static void bound_temperature(oracle *o, double *a) { bound_scalar_bc(o, a, o->temp_bt, o->c.side_temp); }
/* BoundMoisture, src/scalarboundaries.f90:311-323: the same rules with
   MoistBtype and sideMoist */
static void bound_moisture(oracle *o, double *a) { bound_scalar_bc(o, a, o->moist_bt, o->c.side_moist); }

/* BoundViscosity, src/scalarboundaries.f90:326-431 */
static void bound_visc(oracle *o, double *a) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  int px = o->c.btype[We] == PERIODIC, py = o->c.btype[So] == PERIODIC, pz = o->c.btype[Bo] == PERIODIC;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      A3(o, a, 0, j, k) = px ? A3(o, a, nx, j, k) : A3(o, a, 1, j, k);
      A3(o, a, nx + 1, j, k) = px ? A3(o, a, 1, j, k) : A3(o, a, nx, j, k);
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 0; i <= nx + 1; i++) {
      A3(o, a, i, 0, k) = py ? A3(o, a, i, ny, k) : A3(o, a, i, 1, k);
      A3(o, a, i, ny + 1, k) = py ? A3(o, a, i, 1, k) : A3(o, a, i, ny, k);
    }
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 0; i <= nx + 1; i++) {
      A3(o, a, i, j, 0) = pz ? A3(o, a, i, j, nz) : A3(o, a, i, j, 1);
      A3(o, a, i, j, nz + 1) = pz ? A3(o, a, i, j, 1) : A3(o, a, i, j, nz);
    }
}

/* Bound_Phi, src/boundaries.f90:1817-1948 (PrBtype periodic or Neumann). */
static void bound_phi(oracle *o, double *a) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  int px = o->c.btype[We] == PERIODIC, py = o->c.btype[So] == PERIODIC, pz = o->c.btype[Bo] == PERIODIC;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      if (px) {
        A3(o, a, -1, j, k) = A3(o, a, nx - 1, j, k);
        A3(o, a, 0, j, k) = A3(o, a, nx, j, k);
        A3(o, a, nx + 1, j, k) = A3(o, a, 1, j, k);
        A3(o, a, nx + 2, j, k) = A3(o, a, 2, j, k);
      } else {
        A3(o, a, -1, j, k) = A3(o, a, 2, j, k);
        A3(o, a, 0, j, k) = A3(o, a, 1, j, k);
        A3(o, a, nx + 1, j, k) = A3(o, a, nx, j, k);
        A3(o, a, nx + 2, j, k) = A3(o, a, nx - 1, j, k);
      }
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      if (py) {
        A3(o, a, i, -1, k) = A3(o, a, i, ny - 1, k);
        A3(o, a, i, 0, k) = A3(o, a, i, ny, k);
        A3(o, a, i, ny + 1, k) = A3(o, a, i, 1, k);
        A3(o, a, i, ny + 2, k) = A3(o, a, i, 2, k);
      } else {
        A3(o, a, i, -1, k) = A3(o, a, i, 2, k);
        A3(o, a, i, 0, k) = A3(o, a, i, 1, k);
        A3(o, a, i, ny + 1, k) = A3(o, a, i, ny, k);
        A3(o, a, i, ny + 2, k) = A3(o, a, i, ny - 1, k);
      }
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      if (pz) {
        A3(o, a, i, j, -1) = A3(o, a, i, j, nz - 1);
        A3(o, a, i, j, 0) = A3(o, a, i, j, nz);
        A3(o, a, i, j, nz + 1) = A3(o, a, i, j, 1);
        A3(o, a, i, j, nz + 2) = A3(o, a, i, j, 2);
      } else {
        A3(o, a, i, j, -1) = A3(o, a, i, j, 2);
        A3(o, a, i, j, 0) = A3(o, a, i, j, 1);
        A3(o, a, i, j, nz + 1) = A3(o, a, i, j, nz);
        A3(o, a, i, j, nz + 2) = A3(o, a, i, j, nz - 1);
      }
    }
}

/* Bound_Pr, src/boundaries.f90:1951-2090 (periodic or extrapolation). */
static void bound_pr(oracle *o, double *a) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  int px = o->c.btype[We] == PERIODIC, py = o->c.btype[So] == PERIODIC, pz = o->c.btype[Bo] == PERIODIC;
#define P_(i, j, k) A3(o, a, i, j, k)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      if (px) {
        P_(-1, j, k) = P_(nx - 1, j, k);
        P_(0, j, k) = P_(nx, j, k);
        P_(nx + 1, j, k) = P_(1, j, k);
        P_(nx + 2, j, k) = P_(2, j, k);
      } else {
        P_(0, j, k) = P_(3, j, k) - 3 * (P_(2, j, k) - P_(1, j, k));
        P_(-1, j, k) = P_(2, j, k);
        P_(nx + 1, j, k) = P_(nx - 2, j, k) - 3 * (P_(nx - 1, j, k) - P_(nx, j, k));
        P_(nx + 2, j, k) = P_(nx - 1, j, k);
      }
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      if (py) {
        P_(i, -1, k) = P_(i, ny - 1, k);
        P_(i, 0, k) = P_(i, ny, k);
        P_(i, ny + 1, k) = P_(i, 1, k);
        P_(i, ny + 2, k) = P_(i, 2, k);
      } else {
        P_(i, 0, k) = P_(i, 3, k) - 3 * (P_(i, 2, k) - P_(i, 1, k));
        P_(i, -1, k) = P_(i, 2, k);
        P_(i, ny + 1, k) = P_(i, ny - 2, k) - 3 * (P_(i, ny - 1, k) - P_(i, ny, k));
        P_(i, ny + 2, k) = P_(i, ny - 1, k);
      }
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      if (pz) {
        P_(i, j, -1) = P_(i, j, nz - 1);
        P_(i, j, 0) = P_(i, j, nz);
        P_(i, j, nz + 1) = P_(i, j, 1);
        P_(i, j, nz + 2) = P_(i, j, 2);
      } else {
        P_(i, j, 0) = P_(i, j, 3) - 3 * (P_(i, j, 2) - P_(i, j, 1));
        P_(i, j, -1) = P_(i, j, 2);
        P_(i, j, nz + 1) = P_(i, j, nz - 2) - 3 * (P_(i, j, nz - 1) - P_(i, j, nz));
        P_(i, j, nz + 2) = P_(i, j, nz - 1);
      }
    }
#undef P_
}

/* Bound_Q, src/boundaries.f90:2094-2192 */
static void bound_q(oracle *o, double *q) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  int px = o->c.btype[We] == PERIODIC, py = o->c.btype[So] == PERIODIC, pz = o->c.btype[Bo] == PERIODIC;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      if (px)
        A3(o, q, nx, j, k) = A3(o, q, 0, j, k) + A3(o, q, nx, j, k);
      else
        A3(o, q, 1, j, k) = A3(o, q, 1, j, k) + A3(o, q, 0, j, k);
    }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      if (px)
        A3(o, q, 1, j, k) = A3(o, q, 1, j, k) + A3(o, q, nx + 1, j, k);
      else
        A3(o, q, nx, j, k) = A3(o, q, nx, j, k) + A3(o, q, nx + 1, j, k);
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      if (py)
        A3(o, q, i, ny, k) = A3(o, q, i, ny, k) + A3(o, q, i, 0, k);
      else
        A3(o, q, i, 1, k) = A3(o, q, i, 1, k) + A3(o, q, i, 0, k);
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      if (py)
        A3(o, q, i, 1, k) = A3(o, q, i, 1, k) + A3(o, q, i, ny + 1, k);
      else
        A3(o, q, i, ny, k) = A3(o, q, i, ny, k) + A3(o, q, i, ny + 1, k);
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      if (pz)
        A3(o, q, i, j, nz) = A3(o, q, i, j, nz) + A3(o, q, i, j, 0);
      else
        A3(o, q, i, j, 1) = A3(o, q, i, j, 1) + A3(o, q, i, j, 0);
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      if (pz)
        A3(o, q, i, j, 1) = A3(o, q, i, j, 1) + A3(o, q, i, j, nz + 1);
      else
        A3(o, q, i, j, nz) = A3(o, q, i, j, nz) + A3(o, q, i, j, nz + 1);
    }
}

/* ------------------------------------------------------------- SGS ------ */
/* NuSmag + StrainIJ + Strainu, src/subgrid.f90:55-67, 170-209 */
static double nu_smag(oracle *o, int i, int j, int k, const double *U, const double *V, const double *W) {
  double D[3][3] = {{0}};
  double dx = o->dxmin, dy = o->dymin;
  double zp1 = C1D(o->zPr, k + 1), zm1 = C1D(o->zPr, k - 1);
  D[0][0] = (A3(o, U, i, j, k) - A3(o, U, i - 1, j, k)) / dx;
  D[1][1] = (A3(o, V, i, j, k) - A3(o, V, i, j - 1, k)) / dy;
  D[2][2] = (A3(o, W, i, j, k) - A3(o, W, i, j, k - 1)) / C1D(o->dzPr, k);
  D[0][1] = (A3(o, U, i, j + 1, k) + A3(o, U, i - 1, j + 1, k) - A3(o, U, i, j - 1, k) - A3(o, U, i - 1, j - 1, k)) /
            (4 * dy);
  D[0][2] = (A3(o, U, i, j, k + 1) + A3(o, U, i - 1, j, k + 1) - A3(o, U, i, j, k - 1) - A3(o, U, i - 1, j, k - 1)) /
            (2 * (zp1 - zm1));
  D[1][0] = (A3(o, V, i + 1, j, k) + A3(o, V, i + 1, j - 1, k) - A3(o, V, i - 1, j, k) - A3(o, V, i - 1, j - 1, k)) /
            (4 * dx);
  D[1][2] = (A3(o, V, i, j, k + 1) + A3(o, V, i, j - 1, k + 1) - A3(o, V, i, j, k - 1) - A3(o, V, i, j - 1, k - 1)) /
            (2 * (zp1 - zm1));
  D[2][0] = (A3(o, W, i + 1, j, k) + A3(o, W, i + 1, j, k - 1) - A3(o, W, i - 1, j, k) - A3(o, W, i - 1, j, k - 1)) /
            (4 * dx);
  D[2][1] = (A3(o, W, i, j + 1, k) + A3(o, W, i, j + 1, k - 1) - A3(o, W, i, j - 1, k) - A3(o, W, i, j - 1, k - 1)) /
            (4 * dy);
  /* Strainu: sum over jj outer, ii inner of S(ii,jj)^2 with S = (D + D^T)/2.
     D is indexed here D[row=ii][col=jj]. */
  double s = 0;
  for (int jj = 0; jj < 3; jj++)
    for (int ii = 0; ii < 3; ii++) {
      double Sij = (D[ii][jj] + D[jj][ii]) / 2;
      s = s + Sij * Sij;
    }
  double Sbar = sqrt(2.0 * s);
  double width = o->c.filter_ratio * pow(C1D(o->dxPr, i) * C1D(o->dyPr, j) * C1D(o->dzPr, k), 1.0 / 3.0);
  double wc = width * o->c.c_smag;
  return Sbar * (wc * wc);
}

/* SGS_Smag, src/subgrid.f90:30-50 (incl. add of molecular_viscosity to the
   whole array) */
static void sgs_smag(oracle *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  if (o->c.sgs_model == 1) {
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) A3(o, o->Visc, i, j, k) = nu_smag(o, i, j, k, o->U, o->V, o->W);
    if (o->c.nu > 0)
      for (long p = 0; p < o->ntot; p++) o->Visc[p] = o->Visc[p] + o->c.nu;
  } else {
    for (long p = 0; p < o->ntot; p++) o->Visc[p] = o->c.nu;
  }
}

/* ------------------------------------------------------ wall model ------ */
static double norm3(const double v[3]) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* vel_and_dist, src/wallmodels.f90:1382-1391 */
static void vel_and_dist(double tan[3], double *dist, const double u[3], const double wallu[3], const double d[3]) {
  *dist = norm3(d);
  for (int c = 0; c < 3; c++) tan[c] = u[c] - wallu[c];
  double dot = tan[0] * d[0] + tan[1] * d[1] + tan[2] * d[2];
  double d2 = (*dist) * (*dist);
  for (int c = 0; c < 3; c++) tan[c] = tan[c] - dot * d[c] / d2;
}

/* WMFlatUstar, src/wallmodels.f90:1190-1240 */
/* Natural logarithm of a positive normal x, evaluated identically by the
   oracle and the GPU library (libm's log differs between the two by up to an
   ulp): x = 2^k m with m in [sqrt(2)/2, sqrt(2)), f = m - 1 (exact),
   s = f / (2 + f), log(1 + f) = 2 atanh(s) = f - (hfsq - s (hfsq + R(s^2)))
   with hfsq = f^2 / 2 and R(z) = sum_{n=1..11} 2 z^n / (2n + 1), the Taylor
   series of the atanh expansion (truncation below 1e-19 for |s| <= 0.172);
   log x = k ln2_hi - ((hfsq - (s (hfsq + R) + k ln2_lo)) - f), ln2_hi holding
   31 significant bits so that k ln2_hi is exact.  This is the reduction of
   P. T. P. Tang (ACM TOMS 16, 1990) with a plain Taylor polynomial. */
double oracle_det_log(double x) {
  uint64_t b;
  memcpy(&b, &x, sizeof b);
  int k = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  memcpy(&m, &b, sizeof m);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    k = k + 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 2.0 / 23.0;
  R = 2.0 / 21.0 + z * R;
  R = 2.0 / 19.0 + z * R;
  R = 2.0 / 17.0 + z * R;
  R = 2.0 / 15.0 + z * R;
  R = 2.0 / 13.0 + z * R;
  R = 2.0 / 11.0 + z * R;
  R = 2.0 / 9.0 + z * R;
  R = 2.0 / 7.0 + z * R;
  R = 2.0 / 5.0 + z * R;
  R = 2.0 / 3.0 + z * R;
  R = z * R;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
  return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}

static double wm_flat_ustar(double ustar, double vel, double dist, double nu) {
  const double eps = 1e-4;
  const double zpl_lam = 5.0, zpl_turb = 30.0;
  if (nu <= 0) return 0;
  double ustar_lam = sqrt(nu * vel / dist);
  if ((dist * ustar_lam / nu) < zpl_lam) return ustar_lam;
  int i = 0;
  for (;;) {
    i = i + 1;
    double ustar2 = ustar;
    double z_pl = dist * ustar2 / nu;
    if (z_pl <= zpl_lam)
      ustar = sqrt(nu * vel / dist);
    else if (z_pl < zpl_turb)
      ustar = vel / (5 * oracle_det_log(z_pl) - (double)3.05f);
    else
      ustar = vel / (oracle_det_log(z_pl) / 0.4 + 5.5);
    if (fabs(ustar - ustar2) / fabs(ustar) < eps) break;
    if (i >= 50) {
      ustar = 0;
      break;
    }
  }
  return ustar;
}

/* WMRoughUstar, src/wallmodels.f90:1312-1333 */
static double wm_rough_ustar(double ustar, double vel, double dist, double z0, double nu) {
  const double yplcrit = 11.225;
  if (nu > 0) {
    ustar = sqrt(nu * vel / dist);
    if (dist * ustar / nu < yplcrit * 1.1) return ustar;
  }
  if (dist <= z0 * 2.0) {
    if (nu > 0)
      ustar = wm_flat_ustar(ustar, vel, dist, nu);
    else
      ustar = 0;
  } else {
    ustar = vel * 0.41 / log(dist / z0);
  }
  return ustar;
}

/* exp(x) with fixed operations, evaluated identically by the device (host
   and device libm differ by an ulp): x = k ln2 + r with k = rint(x / ln2)
   and r = (x - k ln2_hi) - k ln2_lo (ln2_hi of 33 significant bits, so
   k ln2_hi is exact), |r| <= ln2/2; exp(r) by its Taylor polynomial to
   r^13/13! (truncation below 4e-18 relative) in Horner form; times 2^k
   (exact).  The reduction of W. J. Cody and W. Waite (1980).  Within a few
   ulps of libm (pinned in tests). */
double oracle_det_exp(double x) {
  const double inv_ln2 = 0x1.71547652b82fep+0;
  const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
  const double k = rint(x * inv_ln2);
  const double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = 1.0 / 6227020800.0;
  p = 1.0 / 479001600.0 + r * p;
  p = 1.0 / 39916800.0 + r * p;
  p = 1.0 / 3628800.0 + r * p;
  p = 1.0 / 362880.0 + r * p;
  p = 1.0 / 40320.0 + r * p;
  p = 1.0 / 5040.0 + r * p;
  p = 1.0 / 720.0 + r * p;
  p = 1.0 / 120.0 + r * p;
  p = 1.0 / 24.0 + r * p;
  p = 1.0 / 6.0 + r * p;
  p = 0.5 + r * p;
  p = 1.0 + r * p;
  p = 1.0 + r * p;
  return ldexp(p, (int)k);
}

/* atan(x) with fixed operations, the same on the device: |x| > 1 by
   atan(a) = pi/2 - atan(1/a); a > tan(pi/12) by atan(a) = pi/6 +
   atan((a sqrt3 - 1) / (a + sqrt3)); then |t| <= tan(pi/12) = 0.268 by the
   Taylor series t sum_{n=0..15} (-1)^n t^2n / (2n+1) (truncation below
   1e-19). */
double oracle_det_atan(double x) {
  const double pi2 = 0x1.921fb54442d18p+0, pi6 = 0x1.0c152382d7365p-1;
  const double sqrt3 = 0x1.bb67ae8584caap+0, tan12 = 0x1.126145e9ecd58p-2;
  double a = fabs(x), off = 0;
  int inv = 0;
  if (a > 1) {
    a = 1 / a;
    inv = 1;
  }
  if (a > tan12) {
    a = (a * sqrt3 - 1) / (a + sqrt3);
    off = pi6;
  }
  const double z = a * a;
  double p = -1.0 / 31.0;
  p = 1.0 / 29.0 + z * p;
  p = -1.0 / 27.0 + z * p;
  p = 1.0 / 25.0 + z * p;
  p = -1.0 / 23.0 + z * p;
  p = 1.0 / 21.0 + z * p;
  p = -1.0 / 19.0 + z * p;
  p = 1.0 / 17.0 + z * p;
  p = -1.0 / 15.0 + z * p;
  p = 1.0 / 13.0 + z * p;
  p = -1.0 / 11.0 + z * p;
  p = 1.0 / 9.0 + z * p;
  p = -1.0 / 7.0 + z * p;
  p = 1.0 / 5.0 + z * p;
  p = -1.0 / 3.0 + z * p;
  p = 1.0 + z * p;
  double r = off + a * p;
  if (inv) r = pi2 - r;
  return x < 0 ? -r : r;
}

/* x**y for x >= 0, y > 0: exp(y log x) by the two functions above (0 for
   x = 0) */
double oracle_det_pow(double x, double y) {
  if (x == 0) return 0;
  return oracle_det_exp(y * oracle_det_log(x));
}

/* ------------------------------------------------ Monin-Obukhov laws --- */
static const double MO_PI = 0x1.921fb54442d18p+1; /* pi = acos(-1.0_knd), parameters.f90:43 */

/* PsiM_MO, src/wallmodels.f90:1546-1556 */
static double psim_mo(double zeta) {
  if (zeta < 0) {
    const double x = oracle_det_pow(1 - 15.0 * zeta, 0.25);
    const double c = (1 + x) / 2.0;
    return (oracle_det_log(((1 + x * x) / 2.0) * (c * c)) - 2.0 * oracle_det_atan(x)) + MO_PI / 2;
  }
  return -(4.8 * zeta);
}

/* PsiM_MO_mod, src/wallmodels.f90:1572-1582 */
static double psim_mo_mod(double zeta, double kU) {
  if (zeta < 0) {
    const double x = oracle_det_pow(1 - (6.0 * zeta) * kU, 0.25);
    const double c = (1 + x) / 2.0;
    return (oracle_det_log(((1 + x * x) / 2.0) * (c * c)) - 2.0 * oracle_det_atan(x)) + MO_PI / 2;
  }
  return -(3 * oracle_det_pow(zeta * kU, 5.0 / 6.0));
}

/* PsiH_MO_mod, src/wallmodels.f90:1585-1595 */
static double psih_mo_mod(double zeta, double kU) {
  if (zeta < 0) {
    const double x = oracle_det_pow(1 - (6.0 * zeta) * kU, 0.25);
    return 2.0 * oracle_det_log((1 + x * x) / 2.0);
  }
  return -(2.5 * oracle_det_pow(zeta * kU, 4.0 / 5.0));
}

/* Obukhov_zL, src/wallmodels.f90:1598-1602 */
static double obukhov_zl(double ustar, double tflux, double tref, double g, double z) {
  return (z * ((0.4 * (g / tref)) * tflux)) / (-((ustar * ustar) * ustar));
}

/* WM_MO_Flux_ustar, src/wallmodels.f90:1606-1660 (its eps = 1e-3, 1E-5 and
   1E-4 are default-real literals) */
static double mo_flux_ustar(double vel, double dist, double ustar, double z0, double tflux, double tref, double g,
                            double nu) {
  const double eps = (double)1e-3f, yplcrit = 11.225;
  if (dist <= z0) {
    if (nu > 0) {
      if ((dist * ustar / nu) < yplcrit)
        ustar = sqrt(nu * vel / dist);
      else
        ustar = vel / (oracle_det_log(fabs(ustar * dist / nu)) / 0.4 + 5.2);
    } else {
      ustar = 0;
    }
    return ustar;
  }
  int i = 0;
  double zL = 0, Psi = 0;
  for (;;) {
    i = i + 1;
    const double ustar2 = ustar;
    double m = dist / z0 - Psi;
    if (!(m > (double)1e-5f)) m = (double)1e-5f;
    double t = vel * 0.4 / oracle_det_log(m);
    if (!(t > 0)) t = 0;
    ustar = ustar + (t - ustar) / 2;
    if (ustar < (double)1e-4f) {
      zL = -10000;
    } else {
      const double zL2 = obukhov_zl(ustar, tflux, tref, g, dist);
      zL = zL + (zL2 - zL) / 2;
    }
    Psi = psim_mo(zL);
    double den = fabs(ustar);
    if (!(den > 1e-3)) den = 1e-3;
    if (fabs(ustar - ustar2) / den < eps) break;
    if (i >= 50) {
      ustar = 0;
      break;
    }
  }
  return ustar;
}

/* WM_MO_Dirichlet_ustar_tfl, src/wallmodels.f90:1664-1716 (k_T = 0.47 of
   Zilitinkevich and Esau; 1.1 a default-real literal).  logz0 = log(dist/z0)
   and logz0h = log(dist/z0H) (libm). */
static void mo_dirichlet_ustar_tfl(double *ustar_io, double *tflux_io, double dist, double z0, double logz0,
                                   double logz0h, double vel, double tempdif, double nu, double kappa, double g,
                                   double tref) {
  const double eps = 1e-3, yplcrit = 11.225, kU = 0.4, kT = 0.47;
  double ustar = wm_rough_ustar(*ustar_io, vel, dist, z0, nu), tflux = *tflux_io;
  double dist_plus;
  if (nu > 0) {
    const double ustar_lam = sqrt(nu * vel / dist);
    dist_plus = dist * ustar_lam / nu;
  } else {
    dist_plus = DBL_MAX;
  }
  if (dist <= z0 || dist_plus < (double)1.1f * yplcrit) {
    tflux = -((kappa * tempdif) / dist);
  } else {
    double zL = 0, zL0 = -10000;
    int i = 0;
    for (;;) {
      i = i + 1;
      zL = -(((dist * g) * tflux) / ((kU * tref) * ((ustar * ustar) * ustar)));
      const double psi_m = psim_mo_mod(zL, kU), psi_h = psih_mo_mod(zL, kU);
      ustar = (vel * kU) / (logz0 - psi_m);
      tflux = -(((tempdif * ustar) * kT) / (logz0h - psi_h));
      double den = fabs(zL);
      if (!(den > 1e-4)) den = 1e-4;
      if (i > 1 && fabs(zL - zL0) / den < eps) break;
      if (i >= 50 || zL > 10000) break;
      zL0 = zL;
    }
    if (i >= 50 || zL > 10000) {
      ustar = sqrt(nu * vel / dist);
      tflux = -((kappa * tempdif) / dist);
    }
  }
  *ustar_io = ustar;
  *tflux_io = tflux;
}

/* WM_MO_Dirichlet_ustar_tfl_mfl, src/wallmodels.f90:1721-1784: the
   Dirichlet law with the virtual heat flux (no saturation, e = 0.61) in the
   Obukhov length and the moisture flux by the same stability function
   (k_Q = k_T); its laminar branch is dist <= z0 only, and its fallback
   resets u* and the heat flux but keeps the moisture flux. */
static void mo_dirichlet_moist(double *ustar_io, double *tflux_io, double *qflux_io, double dist, double z0,
                               double logz0, double logz0h, double vel, double temp, double surf_temp, double moist,
                               double surf_moist, double nu, double kappa, double g, double tref) {
  const double eps = 1e-3, kU = 0.4, kT = 0.47, kQ = 0.47, e_coef = 0.61;
  const double tempdif = temp - surf_temp, moistdif = moist - surf_moist;
  double ustar = wm_rough_ustar(*ustar_io, vel, dist, z0, nu), tflux = *tflux_io, qflux = *qflux_io;
  if (dist <= z0) {
    tflux = -((kappa * tempdif) / dist);
    qflux = -((kappa * moistdif) / dist);
  } else {
    double zL = 0, zL0 = -10000;
    int i = 0;
    for (;;) {
      i = i + 1;
      const double virt = tflux * (1 + e_coef * moist) + (e_coef * temp) * qflux;
      zL = -(((dist * g) * virt) / ((kU * tref) * ((ustar * ustar) * ustar)));
      const double psi_m = psim_mo_mod(zL, kU), psi_h = psih_mo_mod(zL, kU), psi_e = psi_h;
      ustar = (vel * kU) / (logz0 - psi_m);
      tflux = -(((tempdif * ustar) * kT) / (logz0h - psi_h));
      qflux = -(((moistdif * ustar) * kQ) / (logz0h - psi_e));
      double den = fabs(zL);
      if (!(den > 1e-4)) den = 1e-4;
      if (i > 1 && fabs(zL - zL0) / den < eps) break;
      if (i >= 50 || zL > 10000) break;
      zL0 = zL;
    }
    if (i >= 50 || zL > 10000) {
      ustar = sqrt(nu * vel / dist);
      tflux = -((kappa * tempdif) / dist);
    }
  }
  *ustar_io = ustar;
  *tflux_io = tflux;
  *qflux_io = qflux;
}

void oracle_mo_dirichlet_moist(double *ustar, double *tflux, double *qflux, double dist, double z0, double vel,
                               double temp, double surf_temp, double moist, double surf_moist, double nu,
                               double kappa, double g, double tref) {
  const double lz = log(dist / z0);
  mo_dirichlet_moist(ustar, tflux, qflux, dist, z0, lz, lz, vel, temp, surf_temp, moist, surf_moist, nu, kappa, g,
                     tref);
}

/* exported for the pins of the Monin-Obukhov laws */
double oracle_psim_mo(double zeta) { return psim_mo(zeta); }
double oracle_psim_mo_mod(double zeta) { return psim_mo_mod(zeta, 0.4); }
double oracle_psih_mo_mod(double zeta) { return psih_mo_mod(zeta, 0.4); }
double oracle_mo_flux_ustar(double vel, double dist, double ustar, double z0, double tflux, double tref, double g,
                            double nu) {
  return mo_flux_ustar(vel, dist, ustar, z0, tflux, tref, g, nu);
}
void oracle_mo_dirichlet(double *ustar, double *tflux, double dist, double z0, double vel, double tempdif, double nu,
                         double kappa, double g, double tref) {
  const double lz = log(dist / z0);
  mo_dirichlet_ustar_tfl(ustar, tflux, dist, z0, lz, lz, vel, tempdif, nu, kappa, g, tref);
}

/* local_velocity, src/wallmodels.f90:2356-2399 */
static void local_velocity(oracle *o, int comp, int i, int j, int k, double v[3]) {
  const double *U = o->U, *V = o->V, *W = o->W;
  if (comp == 0) {
    v[0] = A3(o, U, i, j, k);
    v[1] = (A3(o, V, i + 1, j, k) + A3(o, V, i, j, k) + A3(o, V, i + 1, j - 1, k) + A3(o, V, i, j - 1, k)) / 4.0;
    v[2] = (A3(o, W, i + 1, j, k) + A3(o, W, i, j, k) + A3(o, W, i + 1, j, k - 1) + A3(o, W, i, j, k - 1)) / 4.0;
  } else if (comp == 1) {
    v[0] = (A3(o, U, i, j + 1, k) + A3(o, U, i, j, k) + A3(o, U, i - 1, j + 1, k) + A3(o, U, i - 1, j, k)) / 4.0;
    v[1] = A3(o, V, i, j, k);
    v[2] = (A3(o, W, i, j + 1, k) + A3(o, W, i, j, k) + A3(o, W, i, j + 1, k - 1) + A3(o, W, i, j, k - 1)) / 4.0;
  } else {
    v[0] = (A3(o, U, i, j, k + 1) + A3(o, U, i, j, k) + A3(o, U, i - 1, j, k + 1) + A3(o, U, i - 1, j, k)) / 4.0;
    v[1] = (A3(o, V, i, j, k + 1) + A3(o, V, i, j, k) + A3(o, V, i, j - 1, k + 1) + A3(o, V, i, j - 1, k)) / 4.0;
    v[2] = A3(o, W, i, j, k);
  }
}

/* ComputeUVWFluxesWM + fluxes, src/wallmodels.f90:2159-2337 (no prescribed
   ustar, wall velocity 0): with buoyancy, a point with a prescribed surface
   temperature takes the Monin-Obukhov Dirichlet law and one with a positive
   surface heat flux the flux law, else the neutral laws. */
static void wm_uvw_fluxes(oracle *o) {
  if (!o->c.wall_model) return;
  const double nu = o->c.nu;
  const int buoy = o->c.enable_buoyancy;
  const double *T = buoy ? o->S[o->c.nscalars] : NULL;
  const double *Mq = buoy && o->c.enable_moisture ? o->S[o->c.nscalars + 1] : NULL;
  const double eps = sqrt(2.220446049250313e-16);
  double drec[6] = {1 / o->dxmin, 1 / o->dxmin, 1 / o->dymin, 1 / o->dymin, 1 / o->dzmin, 1 / o->dzmin};
  for (int l = 0; l < 18; l++) {
    wm_list *L = &o->wm[l];
    int comp = l / 6, dir = l % 6; /* dir 0..5 = MINUSX..PLUSZ - 1 */
    for (int p = 0; p < L->n; p++) {
      int xi = L->ijk[3 * p], yj = L->ijk[3 * p + 1], zk = L->ijk[3 * p + 2];
      double d[3] = {L->dist[3 * p], L->dist[3 * p + 1], L->dist[3 * p + 2]};
      double vel[3], tan[3], dist, wallu[3] = {0, 0, 0};
      local_velocity(o, comp, xi, yj, zk, vel);
      vel_and_dist(tan, &dist, vel, wallu, d);
      double v = norm3(tan);
      double ustar = L->ustar[p];
      if (L->z0[p] > 0) {
        if (buoy && L->tclass[p] == ELMM_WM_SURFACE_TEMP) {
          /* WM_MO_DirichletStress (wallmodels.f90:1922-1938) with the
             temperature at the velocity point (local_value, 2340-2353) */
          const long off = comp == 0 ? 1 : comp == 1 ? o->sy : o->sz;
          const double temp = (T[IX(o, xi, yj, zk) + off] + T[IX(o, xi, yj, zk)]) / 2;
          const double lz = log(dist / L->z0[p]);
          if (Mq) /* WM_MO_DirichletStress_moist (1941-1963) */
            mo_dirichlet_moist(&ustar, &L->tflux[p], &L->qflux[p], dist, L->z0[p], lz, lz, v, temp, L->tval[p],
                               (Mq[IX(o, xi, yj, zk) + off] + Mq[IX(o, xi, yj, zk)]) / 2, L->qval[p], nu,
                               o->c.nu_scalar, o->c.grav_acc, o->c.temperature_ref);
          else
            mo_dirichlet_ustar_tfl(&ustar, &L->tflux[p], dist, L->z0[p], lz, lz, v, temp - L->tval[p], nu,
                                   o->c.nu_scalar, o->c.grav_acc, o->c.temperature_ref);
        } else if (buoy && L->tflux[p] > 0) {
          /* WM_MO_FluxStress (1836-1854) */
          if (v != 0)
            ustar = mo_flux_ustar(v, dist, ustar, L->z0[p], L->tflux[p], o->c.temperature_ref, o->c.grav_acc, nu);
          else
            ustar = 0;
        } else if (v != 0) {
          ustar = wm_rough_ustar(ustar, v, dist, L->z0[p], nu);
        }
      } else {
        if (v != 0) ustar = wm_flat_ustar(ustar, v, dist, nu);
      }
      if (ustar < 0) ustar = 0;
      L->ustar[p] = ustar;
      double mag = v;
      /* fluxp scales the flux to the cell beyond the face, fluxm to the
         cell before it; they differ only on stretched z grids */
      double fluxp, fluxm;
      if (mag > eps) {
        if (o->variable_z && dir == 4) {
          fluxp = fluxm = (tan[comp] / mag) * (ustar * ustar);
          fluxp = comp < 2 ? fluxp / C1D(o->dzPr, zk) : fluxp / C1D(o->dzW, zk);
          fluxm = comp < 2 ? fluxm / C1D(o->dzPr, zk - 1) : fluxm / C1D(o->dzW, zk - 1);
        } else if (o->variable_z && dir == 5) {
          fluxp = fluxm = (tan[comp] / mag) * (ustar * ustar);
          fluxp = comp < 2 ? fluxp / C1D(o->dzPr, zk + 1) : fluxp / C1D(o->dzW, zk + 1);
          fluxm = comp < 2 ? fluxm / C1D(o->dzPr, zk) : fluxm / C1D(o->dzW, zk);
        } else {
          fluxp = fluxm = (tan[comp] / mag) * (ustar * ustar) * drec[dir];
        }
        if ((dir + 1) % 2 == 1) {
          fluxp = -fluxp;
          fluxm = -fluxm;
        }
      } else {
        fluxp = fluxm = 0;
      }
      L->fluxp[p] = fluxp;
      L->fluxm[p] = fluxm;
    }
  }
}

/* DivergenceWM (2nd order), src/wallmodels.f90:1153-1166 */
static void wm_divergence(oracle *o) {
  for (int p = 0; p < o->wmp.n; p++) {
    int xi = o->wmp.ijk[3 * p], yj = o->wmp.ijk[3 * p + 1], zk = o->wmp.ijk[3 * p + 2];
    double div = 0;
    if (A3(o, o->Prtype, xi + 1, yj, zk) <= 0) div = div + A3(o, o->U, xi, yj, zk) / o->dxmin;
    if (A3(o, o->Prtype, xi - 1, yj, zk) <= 0) div = div - A3(o, o->U, xi - 1, yj, zk) / o->dxmin;
    if (A3(o, o->Prtype, xi, yj + 1, zk) <= 0) div = div + A3(o, o->V, xi, yj, zk) / o->dymin;
    if (A3(o, o->Prtype, xi, yj - 1, zk) <= 0) div = div - A3(o, o->V, xi, yj - 1, zk) / o->dymin;
    if (A3(o, o->Prtype, xi, yj, zk + 1) <= 0) div = div + A3(o, o->W, xi, yj, zk) / C1D(o->dzPr, zk);
    if (A3(o, o->Prtype, xi, yj, zk - 1) <= 0) div = div - A3(o, o->W, xi, yj, zk - 1) / C1D(o->dzPr, zk);
    o->wmp_div[p] = div;
  }
}

/* --------------------------------------------------------- IB gathers --- */
/* TIBPoint_Interpolate, src/immersedboundary.f90:637-651 */
static double ib_interp(oracle *o, int c, int p, const double *F) {
  double s = 0;
  for (int t = o->ib_ptr[c][p]; t < o->ib_ptr[c][p + 1]; t++) {
    const int32_t *q = &o->ib_src[c][3 * t];
    s = s + o->ib_coef[c][t] * A3(o, F, q[0], q[1], q[2]);
  }
  return s;
}

/* IBMomentum, src/tsteps.f90:368-418: gather all, then scatter, per component */
static void ib_momentum(oracle *o, double *U, double *V, double *W) {
  double *F[3] = {U, V, W};
  for (int c = 0; c < 3; c++) {
    int n = o->ib_n[c];
    if (!n) continue;
    double *tmp = xcalloc(n, sizeof(double));
    for (int p = 0; p < n; p++) tmp[p] = ib_interp(o, c, p, F[c]);
    for (int p = 0; p < n; p++) {
      const int32_t *q = &o->ib_ijk[c][3 * p];
      A3(o, F[c], q[0], q[1], q[2]) = tmp[p];
    }
    free(tmp);
  }
}

/* IBMassSources (2nd order), src/tsteps.f90:425-538, sequential order */
static void ib_mass_sources(oracle *o, const double *U, const double *V, const double *W) {
  for (long p = 0; p < o->ntot; p++) o->Q[p] = 0;
  const double *F[3] = {U, V, W};
  const double *dl[3] = {o->dxPr, o->dyPr, o->dzPr};
  for (int c = 0; c < 3; c++)
    for (int p = 0; p < o->ib_n[c]; p++) {
      int q[3] = {o->ib_ijk[c][3 * p], o->ib_ijk[c][3 * p + 1], o->ib_ijk[c][3 * p + 2]};
      double val = A3(o, F[c], q[0], q[1], q[2]);
      A3(o, o->Q, q[0], q[1], q[2]) = A3(o, o->Q, q[0], q[1], q[2]) + val / C1D(dl[c], q[c]);
      int r[3] = {q[0], q[1], q[2]};
      r[c] += 1;
      A3(o, o->Q, r[0], r[1], r[2]) = A3(o, o->Q, r[0], r[1], r[2]) - val / C1D(dl[c], q[c] + 1);
    }
  bound_q(o, o->Q);
}

/* NullInterior, src/tsteps.f90:543-567 */
static void null_interior(oracle *o) {
  double *F[3] = {o->U, o->V, o->W};
  for (int c = 0; c < 3; c++)
    for (int p = 0; p < o->null_n[c]; p++) {
      const int32_t *q = &o->null_ijk[c][3 * p];
      A3(o, F[c], q[0], q[1], q[2]) = 0;
    }
}

/* ComputeViscsWM, src/wallmodels.f90:2050-2152 (no prescribed u*, wall
   velocity 0, wallmodeltype 1): per scalar wall-model point the
   cell-centred velocity from two-point averages (2069-2071), then where
   z0 > 0 with buoyancy the Monin-Obukhov Dirichlet law for a prescribed
   surface temperature (which updates the point's temperature_flux) or the
   flux law for a positive surface heat flux, else WMRoughVisc (1362-1379,
   via WMRoughStress 1336-1359); where z0 = 0 WMFlatVisc (1268-1304, via
   WMFlatStress 1243-1265).  WMPoints%ustar is the
   persistent state (the flat law iterates from it); the viscosity written
   at the point is returned in visc when write_visc is set (the caller in
   SubgridStresses lets SubgridModel overwrite it, as the reference does). */
static void wm_visc_points(oracle *o, int write_visc) {
  if (!o->c.wall_model) return;
  const double nu = o->c.nu;
  const int buoy = o->c.enable_buoyancy;
  const double *T = buoy ? o->S[o->c.nscalars] : NULL;
  const double *Mq = buoy && o->c.enable_moisture ? o->S[o->c.nscalars + 1] : NULL;
  wm_list *L = &o->wmp;
  for (int p = 0; p < L->n; p++) {
    const int xi = L->ijk[3 * p], yj = L->ijk[3 * p + 1], zk = L->ijk[3 * p + 2];
    const double d[3] = {L->dist[3 * p], L->dist[3 * p + 1], L->dist[3 * p + 2]};
    double vel[3], tan[3], dist;
    const double wallu[3] = {0, 0, 0};
    vel[0] = (A3(o, o->U, xi, yj, zk) + A3(o, o->U, xi - 1, yj, zk)) / 2.0;
    vel[1] = (A3(o, o->V, xi, yj, zk) + A3(o, o->V, xi, yj - 1, zk)) / 2.0;
    vel[2] = (A3(o, o->W, xi, yj, zk) + A3(o, o->W, xi, yj, zk - 1)) / 2.0;
    double ustar = L->ustar[p], visc;
    if (L->z0[p] > 0 && buoy && L->tclass[p] == ELMM_WM_SURFACE_TEMP) {
      /* WM_MO_Dirichlet (wallmodels.f90:1858-1884) */
      vel_and_dist(tan, &dist, vel, wallu, d);
      const double v = norm3(tan);
      const double lz = log(dist / L->z0[p]);
      if (Mq) /* WM_MO_Dirichlet_moist (1887-1919) */
        mo_dirichlet_moist(&ustar, &L->tflux[p], &L->qflux[p], dist, L->z0[p], lz, lz, v, A3(o, T, xi, yj, zk),
                           L->tval[p], A3(o, Mq, xi, yj, zk), L->qval[p], nu, o->c.nu_scalar, o->c.grav_acc,
                           o->c.temperature_ref);
      else
        mo_dirichlet_ustar_tfl(&ustar, &L->tflux[p], dist, L->z0[p], lz, lz, v, A3(o, T, xi, yj, zk) - L->tval[p],
                               nu, o->c.nu_scalar, o->c.grav_acc, o->c.temperature_ref);
      if (ustar < 0) ustar = 0;
      if (v > 0 || L->tflux[p] != 0) {
        const double a = ustar * ustar * dist / v;
        visc = a > nu ? a : nu;
      } else {
        visc = nu > 0 ? nu : 0;
      }
    } else if (L->z0[p] > 0 && buoy && L->tflux[p] > 0) {
      /* WM_MO_FLUX (wallmodels.f90:1792-1832; gravity replaced by the
         cosine of the wall normal with the vertical, as the reference does) */
      vel_and_dist(tan, &dist, vel, wallu, d);
      const double v = norm3(tan);
      if (v != 0) {
        const double g_cos = fabs(d[2]) / dist;
        ustar = mo_flux_ustar(v, dist, ustar, L->z0[p], L->tflux[p], o->c.temperature_ref, g_cos, nu);
        if (ustar < 0) ustar = 0;
        if (ustar * ustar * dist / v > nu)
          visc = ustar * ustar * dist / v;
        else
          visc = nu > 0 ? nu : 0;
      } else {
        ustar = 0;
        visc = nu > 0 ? nu : 0;
      }
    } else if (L->z0[p] > 0) {
      vel_and_dist(tan, &dist, vel, wallu, d);
      const double v = norm3(tan);
      if (v != 0) ustar = wm_rough_ustar(ustar, v, dist, L->z0[p], nu);
      if (ustar < 0) ustar = 0;
      if (v > 0 && ustar * ustar * dist / v > nu)
        visc = ustar * ustar * dist / v;
      else if (nu > 0)
        visc = nu;
      else
        visc = 0;
    } else if (nu <= 0) {
      ustar = 0;
      visc = 0;
    } else {
      vel_and_dist(tan, &dist, vel, wallu, d);
      const double v = norm3(tan);
      if (v != 0) ustar = wm_flat_ustar(ustar, v, dist, nu);
      if (ustar < 0) ustar = 0;
      if (v > 0)
        visc = dist * ustar / nu > 1 ? ustar * ustar * dist / v : nu;
      else
        visc = nu;
    }
    L->ustar[p] = ustar;
    if (write_visc) A3(o, o->Visc, xi, yj, zk) = visc;
  }
}

/* ComputeTDiff, src/scalar_diffusion.f90:983-1009 (constPr_sgs = 0.6, a
   default-real literal); with buoyancy, every stage. */
static void compute_tdiff(oracle *o) {
  const double prs = (double)0.6f;
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++)
        A3(o, o->TDiff, i, j, k) = o->c.nu > 0 ? (A3(o, o->Visc, i, j, k) - o->c.nu) / prs + o->c.nu_scalar
                                               : A3(o, o->Visc, i, j, k) / prs;
}

/* SubgridStresses, src/dynamics.f90:869-908 (ComputeTDiff with buoyancy). */
static void subgrid_stresses(oracle *o) {
  wm_visc_points(o, 1);
  sgs_smag(o);
  wm_uvw_fluxes(o);
  bound_visc(o, o->Visc);
  /* serial loop over ScalFlIBPoints (dynamics.f90:899-902) */
  for (int p = 0; p < o->ib_n[3]; p++) {
    const int32_t *q = &o->ib_ijk[3][3 * p];
    A3(o, o->Visc, q[0], q[1], q[2]) = ib_interp(o, 3, p, o->Visc);
  }
  if (o->c.enable_buoyancy) compute_tdiff(o);
  if (o->c.nscalars > 0 || o->c.enable_buoyancy) bound_visc(o, o->TDiff);
}

/* ------------------------------------------------- momentum advection --- */
static inline double mask4(double a, double b, double c, double d) {
  /* sum(i_mask4 * X(m-2:m+1)) with i_mask4 = [-1, 9, 9, -1], left to right */
  return ((-a + 9 * b) + 9 * c) + -d;
}

/* CD4divU, src/advection.f90:334-491 */
static double cd4_u(oracle *o, int i, int j, int k) {
  const double *U = o->U, *V = o->V, *W = o->W;
  double acc = 0, dU;
  int t = A3(o, o->Utype, i, j, k);
#define Ux(m) A3(o, U, m, j, k)
  if (t == 0) {
    double uv1a, uv1b, uv3a, uv3b, Ui;
    Ui = mask4(Ux(i - 1), Ux(i), Ux(i + 1), Ux(i + 2)); /* face i+1 */
    uv1b = Ui * (Ux(i) + Ux(i + 1));
    Ui = mask4(Ux(i - 2), Ux(i - 1), Ux(i), Ux(i + 1)); /* face i */
    uv1a = Ui * (Ux(i - 1) + Ux(i));
    Ui = mask4(Ux(i), Ux(i + 1), Ux(i + 2), Ux(i + 3)); /* face i+2 */
    uv3b = Ui * (Ux(i) + Ux(i + 3));
    Ui = mask4(Ux(i - 3), Ux(i - 2), Ux(i - 1), Ux(i)); /* face i-1 */
    uv3a = Ui * (Ux(i - 3) + Ux(i));
    dU = 9 * (uv1b - uv1a);
    dU = dU - (uv3b - uv3a) / 3;
    acc = acc + dU / o->dxmin;
  } else if (t < 0) {
    dU = 64 * ((Ux(i + 1) + Ux(i)) * (Ux(i + 1) + Ux(i)) - (Ux(i) + Ux(i - 1)) * (Ux(i) + Ux(i - 1)));
    acc = acc + dU / o->dxmin;
  }
#undef Ux
  /* y: Vint(m) = mask4(V(i-1:i+2, m)), UV1(m) = Vint(m)*(U(j=m)+U(m+1)) */
#define VI(m) mask4(A3(o, V, i - 1, m, k), A3(o, V, i, m, k), A3(o, V, i + 1, m, k), A3(o, V, i + 2, m, k))
#define Uy(m) A3(o, U, i, m, k)
  if (t == 0) {
    double uv1b = VI(j) * (Uy(j) + Uy(j + 1));
    double uv1a = VI(j - 1) * (Uy(j - 1) + Uy(j));
    double uv3b = VI(j + 1) * (Uy(j) + Uy(j + 3));
    double uv3a = VI(j - 2) * (Uy(j - 3) + Uy(j));
    dU = 9 * (uv1b - uv1a);
    dU = dU - (uv3b - uv3a) / 3;
    acc = acc + dU / o->dymin;
  } else if (t < 0) {
    dU = 64 * ((Uy(j + 1) + Uy(j)) * (A3(o, V, i + 1, j, k) + A3(o, V, i, j, k)) -
               (Uy(j) + Uy(j - 1)) * (A3(o, V, i + 1, j - 1, k) + A3(o, V, i, j - 1, k)));
    acc = acc + dU / o->dymin;
  }
#undef VI
#undef Uy
#define WI(m) mask4(A3(o, W, i - 1, j, m), A3(o, W, i, j, m), A3(o, W, i + 1, j, m), A3(o, W, i + 2, j, m))
#define Uz(m) A3(o, U, i, j, m)
  if (t == 0) {
    double uv1b = WI(k) * (Uz(k) + Uz(k + 1));
    double uv1a = WI(k - 1) * (Uz(k - 1) + Uz(k));
    double uv3b = WI(k + 1) * (Uz(k) + Uz(k + 3));
    double uv3a = WI(k - 2) * (Uz(k - 3) + Uz(k));
    dU = 9 * (uv1b - uv1a);
    dU = dU - (uv3b - uv3a) / 3;
    acc = acc + dU / o->dzmin;
  } else if (t < 0) {
    dU = 64 * ((Uz(k + 1) + Uz(k)) * (A3(o, W, i + 1, j, k) + A3(o, W, i, j, k)) -
               (Uz(k) + Uz(k - 1)) * (A3(o, W, i + 1, j, k - 1) + A3(o, W, i, j, k - 1)));
    acc = acc + dU / o->dzmin;
  }
#undef WI
#undef Uz
  return acc * (-1.0 / 256);
}

/* CD4divV, src/advection.f90:500-658 */
static double cd4_v(oracle *o, int i, int j, int k) {
  const double *U = o->U, *V = o->V, *W = o->W;
  double acc = 0, dV;
  int t = A3(o, o->Vtype, i, j, k);
#define UI(m) mask4(A3(o, U, m, j - 1, k), A3(o, U, m, j, k), A3(o, U, m, j + 1, k), A3(o, U, m, j + 2, k))
#define Vx(m) A3(o, V, m, j, k)
  if (t == 0) {
    double uv1b = UI(i) * (Vx(i) + Vx(i + 1));
    double uv1a = UI(i - 1) * (Vx(i - 1) + Vx(i));
    double uv3b = UI(i + 1) * (Vx(i) + Vx(i + 3));
    double uv3a = UI(i - 2) * (Vx(i - 3) + Vx(i));
    dV = 9 * (uv1b - uv1a);
    dV = dV - (uv3b - uv3a) / 3;
    acc = acc + dV / o->dxmin;
  } else if (t < 0) {
    dV = 64 * ((Vx(i + 1) + Vx(i)) * (A3(o, U, i, j + 1, k) + A3(o, U, i, j, k)) -
               (Vx(i) + Vx(i - 1)) * (A3(o, U, i - 1, j + 1, k) + A3(o, U, i - 1, j, k)));
    acc = acc + dV / o->dxmin;
  }
#undef UI
#undef Vx
#define Vy(m) A3(o, V, i, m, k)
  if (t == 0) {
    double Vi;
    Vi = mask4(Vy(j - 1), Vy(j), Vy(j + 1), Vy(j + 2));
    double uv1b = Vi * (Vy(j) + Vy(j + 1));
    Vi = mask4(Vy(j - 2), Vy(j - 1), Vy(j), Vy(j + 1));
    double uv1a = Vi * (Vy(j - 1) + Vy(j));
    Vi = mask4(Vy(j), Vy(j + 1), Vy(j + 2), Vy(j + 3));
    double uv3b = Vi * (Vy(j) + Vy(j + 3));
    Vi = mask4(Vy(j - 3), Vy(j - 2), Vy(j - 1), Vy(j));
    double uv3a = Vi * (Vy(j - 3) + Vy(j));
    dV = 9 * (uv1b - uv1a);
    dV = dV - (uv3b - uv3a) / 3;
    acc = acc + dV / o->dymin;
  } else if (t < 0) {
    dV = 64 * ((Vy(j + 1) + Vy(j)) * (Vy(j + 1) + Vy(j)) - (Vy(j) + Vy(j - 1)) * (Vy(j) + Vy(j - 1)));
    acc = acc + dV / o->dymin;
  }
#undef Vy
#define WI(m) mask4(A3(o, W, i, j - 1, m), A3(o, W, i, j, m), A3(o, W, i, j + 1, m), A3(o, W, i, j + 2, m))
#define Vz(m) A3(o, V, i, j, m)
  if (t == 0) {
    double uv1b = WI(k) * (Vz(k) + Vz(k + 1));
    double uv1a = WI(k - 1) * (Vz(k - 1) + Vz(k));
    double uv3b = WI(k + 1) * (Vz(k) + Vz(k + 3));
    double uv3a = WI(k - 2) * (Vz(k - 3) + Vz(k));
    dV = 9 * (uv1b - uv1a);
    dV = dV - (uv3b - uv3a) / 3;
    acc = acc + dV / o->dzmin;
  } else if (t < 0) {
    dV = 64 * ((Vz(k + 1) + Vz(k)) * (A3(o, W, i, j + 1, k) + A3(o, W, i, j, k)) -
               (Vz(k) + Vz(k - 1)) * (A3(o, W, i, j + 1, k - 1) + A3(o, W, i, j, k - 1)));
    acc = acc + dV / o->dzmin;
  }
#undef WI
#undef Vz
  return acc * (-1.0 / 256);
}

/* CD4divW, src/advection.f90:675-832 */
static double cd4_w(oracle *o, int i, int j, int k) {
  const double *U = o->U, *V = o->V, *W = o->W;
  double acc = 0, dW;
  int t = A3(o, o->Wtype, i, j, k);
#define UI(m) mask4(A3(o, U, m, j, k - 1), A3(o, U, m, j, k), A3(o, U, m, j, k + 1), A3(o, U, m, j, k + 2))
#define Wx(m) A3(o, W, m, j, k)
  if (t == 0) {
    double uv1b = UI(i) * (Wx(i) + Wx(i + 1));
    double uv1a = UI(i - 1) * (Wx(i - 1) + Wx(i));
    double uv3b = UI(i + 1) * (Wx(i) + Wx(i + 3));
    double uv3a = UI(i - 2) * (Wx(i - 3) + Wx(i));
    dW = 9 * (uv1b - uv1a);
    dW = dW - (uv3b - uv3a) / 3;
    acc = acc + dW / o->dxmin;
  } else if (t < 0) {
    dW = 64 * ((Wx(i + 1) + Wx(i)) * (A3(o, U, i, j, k + 1) + A3(o, U, i, j, k)) -
               (Wx(i) + Wx(i - 1)) * (A3(o, U, i - 1, j, k + 1) + A3(o, U, i - 1, j, k)));
    acc = acc + dW / o->dxmin;
  }
#undef UI
#undef Wx
#define VI(m) mask4(A3(o, V, i, m, k - 1), A3(o, V, i, m, k), A3(o, V, i, m, k + 1), A3(o, V, i, m, k + 2))
#define Wy(m) A3(o, W, i, m, k)
  if (t == 0) {
    double uv1b = VI(j) * (Wy(j) + Wy(j + 1));
    double uv1a = VI(j - 1) * (Wy(j - 1) + Wy(j));
    double uv3b = VI(j + 1) * (Wy(j) + Wy(j + 3));
    double uv3a = VI(j - 2) * (Wy(j - 3) + Wy(j));
    dW = 9 * (uv1b - uv1a);
    dW = dW - (uv3b - uv3a) / 3;
    acc = acc + dW / o->dymin;
  } else if (t < 0) {
    dW = 64 * ((Wy(j + 1) + Wy(j)) * (A3(o, V, i, j, k + 1) + A3(o, V, i, j, k)) -
               (Wy(j) + Wy(j - 1)) * (A3(o, V, i, j - 1, k) + A3(o, V, i, j - 1, k + 1)));
    acc = acc + dW / o->dymin;
  }
#undef VI
#undef Wy
#define Wz(m) A3(o, W, i, j, m)
  if (t == 0) {
    double Wi;
    Wi = mask4(Wz(k - 1), Wz(k), Wz(k + 1), Wz(k + 2));
    double uv1b = Wi * (Wz(k) + Wz(k + 1));
    Wi = mask4(Wz(k - 2), Wz(k - 1), Wz(k), Wz(k + 1));
    double uv1a = Wi * (Wz(k - 1) + Wz(k));
    Wi = mask4(Wz(k), Wz(k + 1), Wz(k + 2), Wz(k + 3));
    double uv3b = Wi * (Wz(k) + Wz(k + 3));
    Wi = mask4(Wz(k - 3), Wz(k - 2), Wz(k - 1), Wz(k));
    double uv3a = Wi * (Wz(k - 3) + Wz(k));
    dW = 9 * (uv1b - uv1a);
    dW = dW - (uv3b - uv3a) / 3;
    acc = acc + dW / o->dzmin;
  } else if (t < 0) {
    dW = 64 * ((Wz(k + 1) + Wz(k)) * (Wz(k + 1) + Wz(k)) - (Wz(k) + Wz(k - 1)) * (Wz(k) + Wz(k - 1)));
    acc = acc + dW / o->dzmin;
  }
#undef Wz
  return acc * (-1.0 / 256);
}

/* ------------------------------------------------ momentum diffusion --- */
/* Face fluxes of MomentumDiffusion_nobranch_2ord (momentum_diffusion.f90:
   169-288); flux F_d(c) sits on the upper face of c in direction d and the
   bulk adds +F_d(c) - F_d(c - e_d), d = x, y, z. */
static double flux_u(oracle *o, int d, int i, int j, int k) {
  const double *nu = o->Visc, *U = o->U;
  double rx = 1.0 / (o->dxmin * o->dxmin), ry = 1.0 / (o->dymin * o->dymin), rz = 1.0 / (o->dzmin * o->dzmin);
  if (d == 0) return A3(o, nu, i + 1, j, k) * (A3(o, U, i + 1, j, k) - A3(o, U, i, j, k)) * rx;
  if (d == 1)
    return 0.25 * (A3(o, nu, i + 1, j + 1, k) + A3(o, nu, i + 1, j, k) + A3(o, nu, i, j + 1, k) + A3(o, nu, i, j, k)) *
           (A3(o, U, i, j + 1, k) - A3(o, U, i, j, k)) * ry;
  return 0.25 * (A3(o, nu, i + 1, j, k + 1) + A3(o, nu, i + 1, j, k) + A3(o, nu, i, j, k + 1) + A3(o, nu, i, j, k)) *
         (A3(o, U, i, j, k + 1) - A3(o, U, i, j, k)) * rz;
}
static double flux_v(oracle *o, int d, int i, int j, int k) {
  const double *nu = o->Visc, *V = o->V;
  double rx = 1.0 / (o->dxmin * o->dxmin), ry = 1.0 / (o->dymin * o->dymin), rz = 1.0 / (o->dzmin * o->dzmin);
  if (d == 0)
    return 0.25 * (A3(o, nu, i + 1, j + 1, k) + A3(o, nu, i + 1, j, k) + A3(o, nu, i, j + 1, k) + A3(o, nu, i, j, k)) *
           (A3(o, V, i + 1, j, k) - A3(o, V, i, j, k)) * rx;
  if (d == 1) return A3(o, nu, i, j + 1, k) * (A3(o, V, i, j + 1, k) - A3(o, V, i, j, k)) * ry;
  return 0.25 * (A3(o, nu, i, j + 1, k + 1) + A3(o, nu, i, j + 1, k) + A3(o, nu, i, j, k + 1) + A3(o, nu, i, j, k)) *
         (A3(o, V, i, j, k + 1) - A3(o, V, i, j, k)) * rz;
}
static double flux_w(oracle *o, int d, int i, int j, int k) {
  const double *nu = o->Visc, *W = o->W;
  double rx = 1.0 / (o->dxmin * o->dxmin), ry = 1.0 / (o->dymin * o->dymin), rz = 1.0 / (o->dzmin * o->dzmin);
  if (d == 0)
    return 0.25 * (A3(o, nu, i + 1, j, k + 1) + A3(o, nu, i + 1, j, k) + A3(o, nu, i, j, k + 1) + A3(o, nu, i, j, k)) *
           (A3(o, W, i + 1, j, k) - A3(o, W, i, j, k)) * rx;
  if (d == 1)
    return 0.25 * (A3(o, nu, i, j + 1, k + 1) + A3(o, nu, i, j, k + 1) + A3(o, nu, i, j + 1, k) + A3(o, nu, i, j, k)) *
           (A3(o, W, i, j + 1, k) - A3(o, W, i, j, k)) * ry;
  return A3(o, nu, i, j, k + 1) * (A3(o, W, i, j, k + 1) - A3(o, W, i, j, k)) * rz;
}
typedef double (*flux_fn)(oracle *, int, int, int, int);
static const flux_fn FLUX[3] = {flux_u, flux_v, flux_w};

static void momentum_diffusion(oracle *o) {
  double *S[3] = {o->Us, o->Vs, o->Ws};
  const int *N[3] = {o->Un, o->Vn, o->Wn};
  for (int c = 0; c < 3; c++) {
    const int *n = N[c];
    flux_fn F = FLUX[c];
    double *s = S[c];
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= n[2]; k++)
      for (int j = 1; j <= n[1]; j++)
        for (int i = 1; i <= n[0]; i++) {
          double a = A3(o, s, i, j, k);
          a = a + F(o, 0, i, j, k);
          a = a - F(o, 0, i - 1, j, k);
          a = a + F(o, 1, i, j, k);
          a = a - F(o, 1, i, j - 1, k);
          a = a + F(o, 2, i, j, k);
          a = a - F(o, 2, i, j, k - 1);
          A3(o, s, i, j, k) = a;
        }
    if (!o->c.wall_model) continue;
    /* wall-model face-flux replacement, wmfluxes-nobranch-inc.f90:49-208:
       per list (xm, xp, ym, yp, zm, zp) two passes over the points */
    for (int dir = 0; dir < 6; dir++) {
      wm_list *L = &o->wm[c * 6 + dir];
      int ax = dir / 2, plus = dir % 2;
      for (int pass = 0; pass < 2; pass++)
        for (int p = 0; p < L->n; p++) {
          int q[3] = {L->ijk[3 * p], L->ijk[3 * p + 1], L->ijk[3 * p + 2]};
          int f[3] = {q[0], q[1], q[2]}; /* lower cell of the face */
          int t[3] = {q[0], q[1], q[2]}; /* written cell */
          if (!plus) f[ax] -= 1;
          double sign;
          if (!plus) {
            if (pass == 1) t[ax] -= 1;
            sign = pass == 0 ? 1.0 : -1.0;
          } else {
            if (pass == 0) t[ax] += 1;
            sign = pass == 0 ? 1.0 : -1.0;
          }
          double fl = F(o, ax, f[0], f[1], f[2]);
          double *w = &A3(o, s, t[0], t[1], t[2]);
          if (sign > 0) {
            *w = *w + fl;
            *w = *w + L->fluxp[p];
          } else {
            *w = *w - fl;
            *w = *w - L->fluxp[p];
          }
        }
    }
  }
}

/* ------------------------------------- stretched z (GRID_VARIABLE_Z) --- */
/* CDU/V/W_variable_z = (CD?div_variable_z + CD?adv_variable_z) * 0.5 on a
   zeroed array, src/advection_variable_z.f90:15-321 (advection_method 2). */
#define UU(i, j, k) A3(o, o->U, i, j, k)
#define VV(i, j, k) A3(o, o->V, i, j, k)
#define WW(i, j, k) A3(o, o->W, i, j, k)
static double cd2vz_u(oracle *o, int i, int j, int k) {
  double Ax = 0.25 / o->dxmin, Ay = 0.25 / o->dymin;
  double d = -(((Ax * (UU(i + 1, j, k) + UU(i, j, k)) * (UU(i + 1, j, k) + UU(i, j, k)) -
                 Ax * (UU(i, j, k) + UU(i - 1, j, k)) * (UU(i, j, k) + UU(i - 1, j, k))) +
                (Ay * (UU(i, j + 1, k) + UU(i, j, k)) * (VV(i + 1, j, k) + VV(i, j, k)) -
                 Ay * (UU(i, j, k) + UU(i, j - 1, k)) * (VV(i + 1, j - 1, k) + VV(i, j - 1, k)))) +
               ((UU(i, j, k + 1) + UU(i, j, k)) * (WW(i + 1, j, k) + WW(i, j, k)) -
                (UU(i, j, k) + UU(i, j, k - 1)) * (WW(i + 1, j, k - 1) + WW(i, j, k - 1))) /
                   C1D(o->dzPr, k) / 4);
  Ax = 0.5 / o->dxmin;
  Ay = 0.125 / o->dymin;
  double Vadv = VV(i, j, k) + VV(i + 1, j, k) + VV(i, j - 1, k) + VV(i + 1, j - 1, k);
  double Wadv = (WW(i, j, k) + WW(i + 1, j, k) + WW(i, j, k - 1) + WW(i + 1, j, k - 1)) / 4;
  d = d - (Ax * (UU(i + 1, j, k) - UU(i - 1, j, k)) * UU(i, j, k) + Ay * (UU(i, j + 1, k) - UU(i, j - 1, k)) * Vadv +
           (UU(i, j, k + 1) - UU(i, j, k - 1)) / (C1D(o->zPr, k + 1) - C1D(o->zPr, k - 1)) * Wadv);
  return d * 0.5;
}
static double cd2vz_v(oracle *o, int i, int j, int k) {
  double Ax = 0.25 / o->dxmin, Ay = 0.25 / o->dymin;
  double d = -(((Ay * (VV(i, j + 1, k) + VV(i, j, k)) * (VV(i, j + 1, k) + VV(i, j, k)) -
                 Ay * (VV(i, j, k) + VV(i, j - 1, k)) * (VV(i, j, k) + VV(i, j - 1, k))) +
                (Ax * (VV(i + 1, j, k) + VV(i, j, k)) * (UU(i, j + 1, k) + UU(i, j, k)) -
                 Ax * (VV(i, j, k) + VV(i - 1, j, k)) * (UU(i - 1, j + 1, k) + UU(i - 1, j, k)))) +
               ((VV(i, j, k + 1) + VV(i, j, k)) * (WW(i, j + 1, k) + WW(i, j, k)) -
                (VV(i, j, k) + VV(i, j, k - 1)) * (WW(i, j + 1, k - 1) + WW(i, j, k - 1))) /
                   C1D(o->dzPr, k) / 4);
  Ax = 0.125 / o->dxmin;
  Ay = 0.5 / o->dymin;
  double Uadv = UU(i, j, k) + UU(i, j + 1, k) + UU(i - 1, j, k) + UU(i - 1, j + 1, k);
  double Wadv = (WW(i, j, k) + WW(i, j + 1, k) + WW(i, j, k - 1) + WW(i, j + 1, k - 1)) / 4;
  d = d - (Ax * (VV(i + 1, j, k) - VV(i - 1, j, k)) * Uadv + Ay * (VV(i, j + 1, k) - VV(i, j - 1, k)) * VV(i, j, k) +
           (VV(i, j, k + 1) - VV(i, j, k - 1)) / (C1D(o->zPr, k + 1) - C1D(o->zPr, k - 1)) * Wadv);
  return d * 0.5;
}
static double cd2vz_w(oracle *o, int i, int j, int k) {
  double Ax = 0.25 / o->dxmin, Ay = 0.25 / o->dymin;
  double d = -((((WW(i, j, k + 1) + WW(i, j, k)) * (WW(i, j, k + 1) + WW(i, j, k)) -
                 (WW(i, j, k) + WW(i, j, k - 1)) * (WW(i, j, k) + WW(i, j, k - 1))) /
                    C1D(o->dzW, k) / 4 +
                (Ay * (WW(i, j + 1, k) + WW(i, j, k)) * (VV(i, j, k + 1) + VV(i, j, k)) -
                 Ay * (WW(i, j, k) + WW(i, j - 1, k)) * (VV(i, j - 1, k + 1) + VV(i, j - 1, k)))) +
               (Ax * (WW(i + 1, j, k) + WW(i, j, k)) * (UU(i, j, k + 1) + UU(i, j, k)) -
                Ax * (WW(i, j, k) + WW(i - 1, j, k)) * (UU(i - 1, j, k + 1) + UU(i - 1, j, k))));
  Ax = 0.125 / o->dxmin;
  Ay = 0.125 / o->dymin;
  double Uadv = UU(i, j, k) + UU(i, j, k + 1) + UU(i - 1, j, k) + UU(i - 1, j, k + 1);
  double Vadv = VV(i, j, k) + VV(i, j, k + 1) + VV(i, j - 1, k) + VV(i, j - 1, k + 1);
  d = d - (Ax * (WW(i + 1, j, k) - WW(i - 1, j, k)) * Uadv + Ay * (WW(i, j + 1, k) - WW(i, j - 1, k)) * Vadv +
           (WW(i, j, k + 1) - WW(i, j, k - 1)) / (C1D(o->zW, k + 1) - C1D(o->zW, k - 1)) * WW(i, j, k));
  return d * 0.5;
}
#undef UU
#undef VV
#undef WW

/* z-direction face term of MomentumDiffusion_variable_z_nobranch_2ord for
   component c on the face above level f (momentum_diffusion.f90:298-420):
   U, V: 0.25 * nu4 * (F(f+1) - F(f)) / dzW(f); W: nu(f+1) * (W(f+1) - W(f)) /
   dzPr(f+1). */
static double zflux_vz(oracle *o, int c, int i, int j, int f) {
  const double *nu = o->Visc;
  if (c == 0)
    return 0.25 * (A3(o, nu, i + 1, j, f + 1) + A3(o, nu, i + 1, j, f) + A3(o, nu, i, j, f + 1) + A3(o, nu, i, j, f)) *
           (A3(o, o->U, i, j, f + 1) - A3(o, o->U, i, j, f)) / C1D(o->dzW, f);
  if (c == 1)
    return 0.25 * (A3(o, nu, i, j + 1, f + 1) + A3(o, nu, i, j + 1, f) + A3(o, nu, i, j, f + 1) + A3(o, nu, i, j, f)) *
           (A3(o, o->V, i, j, f + 1) - A3(o, o->V, i, j, f)) / C1D(o->dzW, f);
  return A3(o, nu, i, j, f + 1) * (A3(o, o->W, i, j, f + 1) - A3(o, o->W, i, j, f)) / C1D(o->dzPr, f + 1);
}

/* wall-model replacement term of a z face in the variable-z includes
   (wmfluxes-variable_z-nobranch-{U,V,W}-inc.f90): the face flux written as
   nu4 * dF / dzW(f) / 4 / dzPr(t) for U, V and nu * dW / dzPr(f+1) / dzW(t)
   for W, t the level of the cell written. */
static double zterm_vz(oracle *o, int c, int i, int j, int f, int t) {
  const double *nu = o->Visc;
  if (c == 0)
    return (A3(o, nu, i + 1, j, f + 1) + A3(o, nu, i + 1, j, f) + A3(o, nu, i, j, f + 1) + A3(o, nu, i, j, f)) *
           (A3(o, o->U, i, j, f + 1) - A3(o, o->U, i, j, f)) / C1D(o->dzW, f) / 4 / C1D(o->dzPr, t);
  if (c == 1)
    return (A3(o, nu, i, j + 1, f + 1) + A3(o, nu, i, j + 1, f) + A3(o, nu, i, j, f + 1) + A3(o, nu, i, j, f)) *
           (A3(o, o->V, i, j, f + 1) - A3(o, o->V, i, j, f)) / C1D(o->dzW, f) / 4 / C1D(o->dzPr, t);
  return A3(o, nu, i, j, f + 1) * (A3(o, o->W, i, j, f + 1) - A3(o, o->W, i, j, f)) / C1D(o->dzPr, f + 1) /
         C1D(o->dzW, t);
}

static void momentum_diffusion_vz(oracle *o) {
  double *S[3] = {o->Us, o->Vs, o->Ws};
  const int *N[3] = {o->Un, o->Vn, o->Wn};
  for (int c = 0; c < 3; c++) {
    const int *n = N[c];
    flux_fn F = FLUX[c];
    double *s = S[c];
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= n[2]; k++)
      for (int j = 1; j <= n[1]; j++)
        for (int i = 1; i <= n[0]; i++) {
          double a = A3(o, s, i, j, k);
          a = a + F(o, 0, i, j, k);
          a = a - F(o, 0, i - 1, j, k);
          a = a + F(o, 1, i, j, k);
          a = a - F(o, 1, i, j - 1, k);
          double tmp = zflux_vz(o, c, i, j, k);
          tmp = tmp - zflux_vz(o, c, i, j, k - 1);
          tmp = tmp / (c < 2 ? C1D(o->dzPr, k) : C1D(o->dzW, k));
          a = a + tmp;
          A3(o, s, i, j, k) = a;
        }
    if (!o->c.wall_model) continue;
    /* z lists only: stretched grids are accepted with periodic x and y and
       no solid bodies */
    for (int dir = 4; dir < 6; dir++) {
      wm_list *L = &o->wm[c * 6 + dir];
      int plus = dir % 2;
      for (int pass = 0; pass < 2; pass++)
        for (int p = 0; p < L->n; p++) {
          int q[3] = {L->ijk[3 * p], L->ijk[3 * p + 1], L->ijk[3 * p + 2]};
          int f = plus ? q[2] : q[2] - 1;
          int t = plus ? (pass == 0 ? q[2] + 1 : q[2]) : (pass == 0 ? q[2] : q[2] - 1);
          double fl = zterm_vz(o, c, q[0], q[1], f, t);
          double *w = &A3(o, s, q[0], q[1], t);
          if (pass == 0) {
            *w = *w + fl;
            *w = *w + L->fluxp[p];
          } else {
            *w = *w - fl;
            *w = *w - L->fluxm[p];
          }
        }
    }
  }
}

/* Convection (advection_method 4 on uniform grids, 2 on stretched z; explicit
   diffusion), src/dynamics.f90:230-499 */
/* BuoyancyForce, 2nd order (src/dynamics.f90:608-700), A = grav_acc /
   temperature_ref.  Buoyancy only: W += A/2 (T(k+1) + T(k)) - A T_ref.
   With moisture (apply_moist, src/dynamics.f90:695-710): the virtual
   temperature Tv = T (1 + 0.61 q) (theta_v, 766-770) is scattered to the two
   W faces of each cell in two passes over k (odd k first, then even), each
   W(k) += A/2 Tv(k) - A T_ref and W(k-1) += A/2 Tv(k) for k up to Wnz+1. */
static double theta_v(oracle *o, const double *T, const double *q, int i, int j, int k) {
  return A3(o, T, i, j, k) * (1.0 + 0.61 * A3(o, q, i, j, k));
}
static void buoyancy_force(oracle *o, double *Ws, const double *T) {
  const double A = o->c.grav_acc / o->c.temperature_ref, A2 = A / 2.0;
  if (o->c.enable_moisture) {
    const double *q = o->S[o->c.nscalars + 1];
    for (int start = 1; start <= 2; start++)
      for (int k = start; k <= o->Wn[2] + 1; k += 2)
        for (int j = 1; j <= o->Wn[1]; j++)
          for (int i = 1; i <= o->Wn[0]; i++) {
            const double tv = theta_v(o, T, q, i, j, k);
            A3(o, Ws, i, j, k) = A3(o, Ws, i, j, k) + A2 * tv - A * o->c.temperature_ref;
            A3(o, Ws, i, j, k - 1) = A3(o, Ws, i, j, k - 1) + A2 * tv;
          }
    return;
  }
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= o->Wn[2]; k++)
    for (int j = 1; j <= o->Wn[1]; j++)
      for (int i = 1; i <= o->Wn[0]; i++)
        A3(o, Ws, i, j, k) =
            A3(o, Ws, i, j, k) + A2 * (A3(o, T, i, j, k + 1) + A3(o, T, i, j, k)) - A * o->c.temperature_ref;
}

static void convection(oracle *o, int stage, double dt) {
  double *U2[3] = {o->U2, o->V2, o->W2};
  double *S[3] = {o->Us, o->Vs, o->Ws};
  const int *N[3] = {o->Un, o->Vn, o->Wn};
  for (int c = 0; c < 3; c++) {
    const int *n = N[c];
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= n[2]; k++)
      for (int j = 1; j <= n[1]; j++)
        for (int i = 1; i <= n[0]; i++) {
          if (stage > 0)
            A3(o, U2[c], i, j, k) = A3(o, S[c], i, j, k) * RK_rho[stage] * dt;
          else {
            A3(o, S[c], i, j, k) = 0;
            A3(o, U2[c], i, j, k) = 0;
          }
        }
  }
  if (o->variable_z) {
    /* CD?_variable_z: set(U2, 0), div, adv, multiply by 0.5 */
    for (int c = 0; c < 3; c++) {
      double *s = S[c];
      for (long p = 0; p < o->ntot; p++) s[p] = 0.0;
      const int *n = N[c];
#pragma omp parallel for collapse(2)
      for (int k = 1; k <= n[2]; k++)
        for (int j = 1; j <= n[1]; j++)
          for (int i = 1; i <= n[0]; i++)
            A3(o, s, i, j, k) = c == 0 ? cd2vz_u(o, i, j, k) : c == 1 ? cd2vz_v(o, i, j, k) : cd2vz_w(o, i, j, k);
    }
    momentum_diffusion_vz(o);
  } else {
    /* CD4div*: set(U2, 0) on the whole array, accumulate, multiply by -1/256 */
    for (int c = 0; c < 3; c++) {
      double *s = S[c];
      /* entries outside the interior end as 0 * (-1/256) = -0.0 */
      for (long p = 0; p < o->ntot; p++) s[p] = 0.0 * (-1.0 / 256);
      const int *n = N[c];
#pragma omp parallel for collapse(2)
      for (int k = 1; k <= n[2]; k++)
        for (int j = 1; j <= n[1]; j++)
          for (int i = 1; i <= n[0]; i++)
            A3(o, s, i, j, k) = c == 0 ? cd4_u(o, i, j, k) : c == 1 ? cd4_v(o, i, j, k) : cd4_w(o, i, j, k);
    }
    if (o->c.enable_buoyancy) buoyancy_force(o, o->Ws, o->S[o->c.nscalars]);
    momentum_diffusion(o);
  }
  for (int c = 0; c < 3; c++) {
    const int *n = N[c];
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= n[2]; k++)
      for (int j = 1; j <= n[1]; j++)
        for (int i = 1; i <= n[0]; i++)
          A3(o, U2[c], i, j, k) = A3(o, U2[c], i, j, k) + A3(o, S[c], i, j, k) * RK_beta[stage] * dt;
  }
}

/* OtherTerms with explicit diffusion: PressureGrad (2nd order) + add(U2,U),
   src/tsteps.f90:263-358, src/dynamics.f90:58-199 */
static void other_terms(oracle *o, double coef) {
  double A = -coef, Ax = -coef / o->dxmin, Ay = -coef / o->dymin, Az = -coef / o->dzmin;
  const double *Pr = o->Pr;
  if (o->c.pr_gradient_x != 0)
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= o->Un[2]; k++)
      for (int j = 1; j <= o->Un[1]; j++)
        for (int i = 1; i <= o->Un[0]; i++)
          A3(o, o->U2, i, j, k) = A3(o, o->U2, i, j, k) + A * o->c.pr_gradient_x;
  if (o->c.pr_gradient_y != 0)
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= o->Vn[2]; k++)
      for (int j = 1; j <= o->Vn[1]; j++)
        for (int i = 1; i <= o->Vn[0]; i++)
          A3(o, o->V2, i, j, k) = A3(o, o->V2, i, j, k) + A * o->c.pr_gradient_y;
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= o->Un[2]; k++)
    for (int j = 1; j <= o->Un[1]; j++)
      for (int i = 1; i <= o->Un[0]; i++)
        A3(o, o->U2, i, j, k) = A3(o, o->U2, i, j, k) + Ax * (A3(o, Pr, i + 1, j, k) - A3(o, Pr, i, j, k));
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= o->Vn[2]; k++)
    for (int j = 1; j <= o->Vn[1]; j++)
      for (int i = 1; i <= o->Vn[0]; i++)
        A3(o, o->V2, i, j, k) = A3(o, o->V2, i, j, k) + Ay * (A3(o, Pr, i, j + 1, k) - A3(o, Pr, i, j, k));
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= o->Wn[2]; k++)
    for (int j = 1; j <= o->Wn[1]; j++)
      for (int i = 1; i <= o->Wn[0]; i++) {
        if (o->variable_z)
          A3(o, o->W2, i, j, k) =
              A3(o, o->W2, i, j, k) - coef * (A3(o, Pr, i, j, k + 1) - A3(o, Pr, i, j, k)) / C1D(o->dzW, k);
        else
          A3(o, o->W2, i, j, k) = A3(o, o->W2, i, j, k) + Az * (A3(o, Pr, i, j, k + 1) - A3(o, Pr, i, j, k));
      }
#pragma omp parallel for
  for (long p = 0; p < o->ntot; p++) {
    o->U2[p] = o->U2[p] + o->U[p];
    o->V2[p] = o->V2[p] + o->V[p];
    o->W2[p] = o->W2[p] + o->W[p];
  }
}

/* ---------------------------------------------------------- scalars ----- */
static inline double flux_limiter(double r) {
  /* src/scalar_advection.f90:260-267 */
  double a = (1 + 2 * r) / 3;
  double b = 2.0 < a ? 2.0 : a;
  double c2 = 2 * r < b ? 2 * r : b;
  return 0.0 > c2 ? 0.0 : c2;
}
static inline double sgn1(double x) { return copysign(1.0, x); }

/* KappaScalar, src/scalar_advection.f90:102-269, restated per cell with the
   reference's accumulation order (x pass, y pass, z pass with odd/even
   ordering of the k faces). */
static void kappa_scalar(oracle *o, double *Sadv, const double *S) {
  const double eps = (double)1e-8f; /* real(knd), parameter :: eps = 1e-8 (default-real literal) */
  int nx = o->nx, ny = o->ny, nz = o->nz;
  const double Ax = 1 / o->dxmin, Ay = 1 / o->dymin, Az = 1 / o->dzmin;
  for (long p = 0; p < o->ntot; p++) Sadv[p] = 0;
  const double *U = o->U, *V = o->V, *W = o->W;
  /* x pass (faces i = 0..nx); a face writes only its own row */
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        double u = A3(o, U, i, j, k), sr, sl;
        if (u > 0) {
          sr = A3(o, S, i + 1, j, k) - A3(o, S, i, j, k);
          sl = A3(o, S, i, j, k) - A3(o, S, i - 1, j, k);
        } else {
          sr = A3(o, S, i, j, k) - A3(o, S, i + 1, j, k);
          sl = A3(o, S, i + 1, j, k) - A3(o, S, i + 2, j, k);
        }
        double slope = flux_limiter((sr + eps * sgn1(sl)) / (sl + eps * sgn1(sl)));
        if (A3(o, o->scflmask[0], i, j, k)) {
          double flux;
          if (u > 0)
            flux = u * (A3(o, S, i, j, k) + (A3(o, S, i, j, k) - A3(o, S, i - 1, j, k)) * slope / 2.0);
          else
            flux = u * (A3(o, S, i + 1, j, k) + (A3(o, S, i + 1, j, k) - A3(o, S, i + 2, j, k)) * slope / 2.0);
          A3(o, Sadv, i, j, k) = A3(o, Sadv, i, j, k) - Ax * flux;
          A3(o, Sadv, i + 1, j, k) = A3(o, Sadv, i + 1, j, k) + Ax * flux;
        }
      }
  /* y pass: a face writes only its own k plane */
#pragma omp parallel for
  for (int k = 1; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double v = A3(o, V, i, j, k), sr, sl;
        if (v > 0) {
          sr = A3(o, S, i, j + 1, k) - A3(o, S, i, j, k);
          sl = A3(o, S, i, j, k) - A3(o, S, i, j - 1, k);
        } else {
          sr = A3(o, S, i, j, k) - A3(o, S, i, j + 1, k);
          sl = A3(o, S, i, j + 1, k) - A3(o, S, i, j + 2, k);
        }
        double slope = flux_limiter((sr + eps * sgn1(sl)) / (sl + eps * sgn1(sl)));
        if (A3(o, o->scflmask[1], i, j, k)) {
          double flux;
          if (v > 0)
            flux = v * (A3(o, S, i, j, k) + (A3(o, S, i, j, k) - A3(o, S, i, j - 1, k)) * slope / 2.0);
          else
            flux = v * (A3(o, S, i, j + 1, k) + (A3(o, S, i, j + 1, k) - A3(o, S, i, j + 2, k)) * slope / 2.0);
          A3(o, Sadv, i, j, k) = A3(o, Sadv, i, j, k) - Ay * flux;
          A3(o, Sadv, i, j + 1, k) = A3(o, Sadv, i, j + 1, k) + Ay * flux;
        }
      }
  /* z pass: faces of one parity write disjoint level pairs (k, k+1) */
  for (int l = 0; l <= 1; l++)
#pragma omp parallel for
    for (int k = l; k <= nz; k += 2)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          double w = A3(o, W, i, j, k), sr, sl;
          if (w > 0) {
            sr = A3(o, S, i, j, k + 1) - A3(o, S, i, j, k);
            sl = A3(o, S, i, j, k) - A3(o, S, i, j, k - 1);
          } else {
            sr = A3(o, S, i, j, k) - A3(o, S, i, j, k + 1);
            sl = A3(o, S, i, j, k + 1) - A3(o, S, i, j, k + 2);
          }
          double slope = flux_limiter((sr + eps * sgn1(sl)) / (sl + eps * sgn1(sl)));
          if (A3(o, o->scflmask[2], i, j, k)) {
            double flux;
            if (w > 0)
              flux = w * (A3(o, S, i, j, k) + (A3(o, S, i, j, k) - A3(o, S, i, j, k - 1)) * slope / 2.0);
            else
              flux = w * (A3(o, S, i, j, k + 1) + (A3(o, S, i, j, k + 1) - A3(o, S, i, j, k + 2)) * slope / 2.0);
            A3(o, Sadv, i, j, k) = A3(o, Sadv, i, j, k) - Az * flux;
            A3(o, Sadv, i, j, k + 1) = A3(o, Sadv, i, j, k + 1) + Az * flux;
          }
        }
}

/* ScalarDiffusion_nobranch, src/scalar_diffusion.f90:404-507 */
static void scalar_diffusion(oracle *o, double *Sadv, const double *S) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  const double *T = o->TDiff;
  const double Ax = 1 / (2 * (o->dxmin * o->dxmin)), Ay = 1 / (2 * (o->dymin * o->dymin)),
               Az = 1 / (2 * (o->dzmin * o->dzmin));
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double a = A3(o, Sadv, i, j, k);
        a = a + (A3(o, T, i + 1, j, k) + A3(o, T, i, j, k)) * (A3(o, S, i + 1, j, k) - A3(o, S, i, j, k)) * Ax;
        a = a - (A3(o, T, i, j, k) + A3(o, T, i - 1, j, k)) * (A3(o, S, i, j, k) - A3(o, S, i - 1, j, k)) * Ax;
        a = a + (A3(o, T, i, j + 1, k) + A3(o, T, i, j, k)) * (A3(o, S, i, j + 1, k) - A3(o, S, i, j, k)) * Ay;
        a = a - (A3(o, T, i, j, k) + A3(o, T, i, j - 1, k)) * (A3(o, S, i, j, k) - A3(o, S, i, j - 1, k)) * Ay;
        a = a + (A3(o, T, i, j, k + 1) + A3(o, T, i, j, k)) * (A3(o, S, i, j, k + 1) - A3(o, S, i, j, k)) * Az;
        a = a - (A3(o, T, i, j, k) + A3(o, T, i, j, k - 1)) * (A3(o, S, i, j, k) - A3(o, S, i, j, k - 1)) * Az;
        A3(o, Sadv, i, j, k) = a;
      }
  const double Ad[3] = {Ax, Ay, Az};
  for (int d = 0; d < 3; d++)
    for (int pass = 0; pass < 2; pass++)
      for (int p = 0; p < o->scfl_n[d]; p++) {
        int q[3] = {o->scfl_ijk[d][3 * p], o->scfl_ijk[d][3 * p + 1], o->scfl_ijk[d][3 * p + 2]};
        int r[3] = {q[0], q[1], q[2]};
        r[d] += 1;
        double fl = (A3(o, T, r[0], r[1], r[2]) + A3(o, T, q[0], q[1], q[2])) *
                    (A3(o, S, r[0], r[1], r[2]) - A3(o, S, q[0], q[1], q[2])) * Ad[d];
        if (pass == 0)
          A3(o, Sadv, q[0], q[1], q[2]) = A3(o, Sadv, q[0], q[1], q[2]) - fl;
        else
          A3(o, Sadv, r[0], r[1], r[2]) = A3(o, Sadv, r[0], r[1], r[2]) + fl;
      }
}

/* ScalarRK3 stage for passive scalars, src/scalars.f90:88-299, 364-379 */
/* One scalar's RK3 stage (the `stage` procedure of ScalarRK3,
   src/scalars.f90:233-299): slot sc of kind 0 (passive), 1 (the temperature:
   its boundary conditions, no point sources) or 2 (the moisture: its
   boundary conditions, no point sources). */
static void bound_kind(oracle *o, double *S, int kind) {
  if (kind == 1)
    bound_temperature(o, S);
  else if (kind == 2)
    bound_moisture(o, S);
  else
    bound_scalar(o, S);
}
static void scalar_stage(oracle *o, int sc, int stage, double dt, int kind) {
  double *S = o->S[sc], *S2 = o->S2[sc], *Sadv = o->Sadv[sc];
  bound_kind(o, S, kind);
  if (stage > 0) {
    double a = RK_rho[stage] * dt;
#pragma omp parallel for
    for (long p = 0; p < o->ntot; p++) S2[p] = Sadv[p] * a;
  } else {
    for (long p = 0; p < o->ntot; p++) S2[p] = 0;
  }
  kappa_scalar(o, Sadv, S);
#pragma omp parallel for collapse(2)
  for (int k = 0; k <= o->nz + 1; k++)
    for (int j = 0; j <= o->ny + 1; j++)
      for (int i = 0; i <= o->nx + 1; i++)
        if (A3(o, o->Prtype, i, j, k) > 0) A3(o, Sadv, i, j, k) = 0;
  scalar_diffusion(o, Sadv, S);
  /* ScalarExtra / TemperatureExtra: volume (point) sources, then the WM
     divergence term (src/scalars.f90:301-331, 364-379) */
  for (int p = 0; p < o->src_n; p++) {
    if (kind || o->src_sc[p] != sc) continue;
    const int32_t *q = &o->src_ijk[3 * p];
    A3(o, Sadv, q[0], q[1], q[2]) = A3(o, Sadv, q[0], q[1], q[2]) + o->src_flux[p];
  }
  for (int p = 0; p < o->wmp.n; p++) {
    const int32_t *q = &o->wmp.ijk[3 * p];
    /* TemperatureExtra: the point's surface heat flux times its area factor
       first (scalars.f90:322-326) */
    if (kind == 1)
      A3(o, Sadv, q[0], q[1], q[2]) = A3(o, Sadv, q[0], q[1], q[2]) + o->wmp.tflux[p] * o->wmp_af[p];
    /* MoistureExtra: the moisture flux likewise (scalars.f90:350-354) */
    if (kind == 2)
      A3(o, Sadv, q[0], q[1], q[2]) = A3(o, Sadv, q[0], q[1], q[2]) + o->wmp.qflux[p] * o->wmp_af[p];
    A3(o, Sadv, q[0], q[1], q[2]) = A3(o, Sadv, q[0], q[1], q[2]) + o->wmp_div[p] * A3(o, S, q[0], q[1], q[2]);
  }
  double b = RK_beta[stage] * dt;
#pragma omp parallel for
  for (long p = 0; p < o->ntot; p++) S2[p] = S2[p] + Sadv[p] * b;
#pragma omp parallel for
  for (long p = 0; p < o->ntot; p++) S[p] = S[p] + S2[p];
  bound_kind(o, S, kind);
}

/* The temperature in solid cells after its stage (src/scalars.f90:131-207):
   over -1..n+2, a solid cell with fluid neighbours (within -1..n+2) takes
   their mean, summed in the order i-1, i+1, j-1, j+1, k-1, k+1; fluid cells
   are never written, so the loop order does not matter. */
static void temperature_solid_fix(oracle *o, double *T) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const int32_t *P = o->Prtype;
  for (int k = -1; k <= nz + 2; k++)
    for (int j = -1; j <= ny + 2; j++)
      for (int i = -1; i <= nx + 2; i++) {
        if (A3(o, P, i, j, k) <= 0) continue;
        double t = 0;
        int n = 0;
        if (i > -1 && A3(o, P, i - 1, j, k) <= 0) { t = t + A3(o, T, i - 1, j, k); n++; }
        if (i < nx + 2 && A3(o, P, i + 1, j, k) <= 0) { t = t + A3(o, T, i + 1, j, k); n++; }
        if (j > -1 && A3(o, P, i, j - 1, k) <= 0) { t = t + A3(o, T, i, j - 1, k); n++; }
        if (j < ny + 2 && A3(o, P, i, j + 1, k) <= 0) { t = t + A3(o, T, i, j + 1, k); n++; }
        if (k > -1 && A3(o, P, i, j, k - 1) <= 0) { t = t + A3(o, T, i, j, k - 1); n++; }
        if (k < nz + 2 && A3(o, P, i, j, k + 1) <= 0) { t = t + A3(o, T, i, j, k + 1); n++; }
        if (n > 0) A3(o, T, i, j, k) = t / n;
      }
}

/* The moisture in solid cells after its stage (src/scalars.f90:208):
   moisture_ref over the whole array -2..n+3. */
static void moisture_solid_fill(oracle *o, double *q) {
  for (int k = -2; k <= o->nz + 3; k++)
    for (int j = -2; j <= o->ny + 3; j++)
      for (int i = -2; i <= o->nx + 3; i++)
        if (A3(o, o->Prtype, i, j, k) > 0) A3(o, q, i, j, k) = o->c.moisture_ref;
}

/* ScalarRK3, src/scalars.f90:88-299: DivergenceWM, the temperature stage
   and its solid-cell values (buoyancy), the moisture stage and its solid
   cells, then the passive scalars. */
static void scalar_rk3(oracle *o, int stage, double dt) {
  const int nsc = o->c.nscalars, buoy = o->c.enable_buoyancy != 0;
  if (nsc == 0 && !buoy) return;
  wm_divergence(o);
  if (buoy) {
    scalar_stage(o, nsc, stage, dt, 1);
    temperature_solid_fix(o, o->S[nsc]);
  }
  if (buoy && o->c.enable_moisture) {
    scalar_stage(o, nsc + 1, stage, dt, 2);
    moisture_solid_fill(o, o->S[nsc + 1]);
  }
  for (int sc = 0; sc < nsc; sc++) scalar_stage(o, sc, stage, dt, 0);
}

/* ------------------------------------------------------------ FFT ------- */
/* The reference calls an external FFT Poisson library (PoisSolver,
   src/poisson.f90:76-210) that is not part of the reference tree.  This
   restatement solves the same discrete operator (Residue_variable_z,
   src/poisson.f90:23-60) directly: DFT in the periodic x and y directions,
   a tridiagonal system per mode in z.  The transform is an iterative
   radix-2 decimation-in-time FFT with twiddles exp(-2 pi i q/n) taken from a
   long-double table, and the tridiagonal systems are solved by Thomas
   elimination in the operation order of the GPU kernel; the singular horizontal-mean
   mode is integrated exactly with compensated sums.  The operation order is
   stated so that results can be compared bit for bit. */
typedef struct {
  double re, im;
} c2;

static c2 *twiddles(int n) {
  c2 *t = xcalloc(n / 2 + 1, sizeof(c2));
  for (int q = 0; q < n / 2; q++) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * q / n;
    t[q].re = (double)cosl(a);
    t[q].im = (double)-sinl(a);
  }
  return t;
}

static int ilog2i(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}
static int bitrev_i(int v, int logn) {
  int r = 0;
  for (int b = 0; b < logn; b++)
    if (v & (1 << b)) r |= 1 << (logn - 1 - b);
  return r;
}

/* In-place FFT of a strided line of length n (power of two). */
static void fft_line(c2 *x, int n, long stride, const c2 *tw, int inverse, c2 *work) {
  int logn = ilog2i(n);
  for (int m = 0; m < n; m++) work[bitrev_i(m, logn)] = x[m * stride];
  for (int len = 2; len <= n; len <<= 1) {
    int half = len / 2, tstep = n / len;
    for (int g0 = 0; g0 < n; g0 += len)
      for (int pos = 0; pos < half; pos++) {
        c2 w = tw[pos * tstep];
        if (inverse) w.im = -w.im;
        c2 a = work[g0 + pos], b = work[g0 + pos + half], c;
        c.re = b.re * w.re - b.im * w.im;
        c.im = b.re * w.im + b.im * w.re;
        work[g0 + pos].re = a.re + c.re;
        work[g0 + pos].im = a.im + c.im;
        work[g0 + pos + half].re = a.re - c.re;
        work[g0 + pos + half].im = a.im - c.im;
      }
  }
  for (int m = 0; m < n; m++) x[m * stride] = work[m];
}

/* Mixed-radix FFT for n = 2^a 3^b that is not a power of two: self-sorting
   (Stockham) stages, radix 3 first, then radix 2.  A stage with radix r and
   span ns (product of the earlier radices) takes element j + q n/r (q < r),
   multiplies it by tw[k q n/(ns r)] (k = j mod ns; conjugate for the
   inverse; no product when k q = 0), forms the r-point DFT and writes
   element (j / ns) ns r + k + q ns.  3-point DFT: s = b + c, d = b - c,
   y0 = a + s, m = a - s/2, t = (s3 d.im, -s3 d.re) (forward) or
   (-s3 d.im, s3 d.re) (inverse), y1 = m + t, y2 = m - t, s3 = sin(pi/3).
   tw holds exp(-2 pi i t/n) for t < n; work holds 2n values. */
static void fft_line_mr(c2 *x, int n, long stride, const c2 *tw, int inverse, c2 *work) {
  c2 *src = work, *dst = work + n;
  for (int m = 0; m < n; m++) src[m] = x[m * stride];
  const double s3 = 0.86602540378443864676;
  int rest = n, ns = 1;
  while (rest > 1) {
    const int r = rest % 3 == 0 ? 3 : 2, nb = n / r, tstride = n / (ns * r);
    for (int j = 0; j < nb; j++) {
      const int k = j % ns;
      c2 v[3];
      for (int q = 0; q < r; q++) {
        v[q] = src[j + q * nb];
        if (q > 0 && k > 0) {
          c2 w = tw[k * q * tstride], a = v[q];
          if (inverse) w.im = -w.im;
          v[q].re = a.re * w.re - a.im * w.im;
          v[q].im = a.re * w.im + a.im * w.re;
        }
      }
      const int base = (j / ns) * ns * r + k;
      if (r == 2) {
        dst[base].re = v[0].re + v[1].re;
        dst[base].im = v[0].im + v[1].im;
        dst[base + ns].re = v[0].re - v[1].re;
        dst[base + ns].im = v[0].im - v[1].im;
      } else {
        c2 sm = {v[1].re + v[2].re, v[1].im + v[2].im}, df = {v[1].re - v[2].re, v[1].im - v[2].im};
        c2 m = {v[0].re - 0.5 * sm.re, v[0].im - 0.5 * sm.im};
        c2 t = inverse ? (c2){-(s3 * df.im), s3 * df.re} : (c2){s3 * df.im, -(s3 * df.re)};
        dst[base].re = v[0].re + sm.re;
        dst[base].im = v[0].im + sm.im;
        dst[base + ns].re = m.re + t.re;
        dst[base + ns].im = m.im + t.im;
        dst[base + 2 * ns].re = m.re - t.re;
        dst[base + 2 * ns].im = m.im - t.im;
      }
    }
    c2 *tmp = src;
    src = dst;
    dst = tmp;
    ns *= r;
    rest /= r;
  }
  for (int m = 0; m < n; m++) x[m * stride] = src[m];
}

static c2 *twiddles_full(int n) {
  c2 *t = xcalloc(n, sizeof(c2));
  for (int q = 0; q < n; q++) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * q / n;
    t[q].re = (double)cosl(a);
    t[q].im = (double)-sinl(a);
  }
  return t;
}

static int is_pow2i(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* One transform of a line: radix 2 for powers of two, else mixed radix. */
static void fft_any(c2 *x, int n, long stride, const c2 *tw, int inverse, c2 *work) {
  if (is_pow2i(n))
    fft_line(x, n, stride, tw, inverse, work);
  else
    fft_line_mr(x, n, stride, tw, inverse, work);
}

/* Thomas elimination of one tridiagonal system a_k x_{k-1} + b_k x_k +
   c_k x_{k+1} = d_k (a_0 = c_{nz-1} = 0; strictly diagonally dominant for
   every mode but the horizontal mean): den_k = b_k - a_k c'_{k-1},
   c'_k = c_k / den_k, d'_k = (d_k - a_k d'_{k-1}) / den_k with each division
   taken as a multiplication by 1/den_k, then x_k = d'_k - c'_k x_{k+1}. */
static void thomas_solve(int nz, const double *a, const double *b, const double *c, c2 *d, double *cp) {
  double pc = 0, dr = 0, di = 0;
  for (int k = 0; k < nz; k++) {
    double inv = 1.0 / (b[k] - a[k] * pc);
    pc = c[k] * inv;
    dr = (d[k].re - a[k] * dr) * inv;
    di = (d[k].im - a[k] * di) * inv;
    cp[k] = pc;
    d[k].re = dr;
    d[k].im = di;
  }
  double xr = 0, xi = 0;
  for (int k = nz - 1; k >= 0; k--) {
    xr = d[k].re - cp[k] * xr;
    xi = d[k].im - cp[k] * xi;
    d[k].re = xr;
    d[k].im = xi;
  }
}

/* Kahan step: s += t with compensation cmp */
static void kahan(double *s, double *cmp, double t) {
  double y = t - *cmp, s2 = *s + y;
  *cmp = (s2 - *s) - y;
  *s = s2;
}

/* ------------------------------------------------ geometric multigrid --- */
/* The iterative solver of the pressure equation (poisson_solver = 3; the
   role of PoissMG, src/multigrid.f90:1047-1172, with MG_CGC's V-cycle
   994-1044, MG_GS 786-910, MG_res 912-977, Restrict 114-165, Prolongate
   58-111).  The reference's multigrid is vertex-centred with uniform
   coefficients; this one solves the cell-centred operator of
   Residue_variable_z (src/poisson.f90:23-60: periodic x/y, staggered
   Neumann walls in z, stretched z) that the direct solve inverts:
     level l: nx/2^l x ny/2^l x nz cells (coarsened in x and y only), ax =
       1/(dx 2^l)^2, ay = 1/(dy 2^l)^2, the fine z coefficients az(k), cz(k);
     smoother: z-line Gauss-Seidel, columns with (i + j) even, then odd; a
       column solves D_k P_k - az_k P_{k-1} - cz_k P_{k+1} = h_k - rhs_k with
       h_k = ax (P_w + P_e) + ay (P_s + P_n) and D_k = (((ax + ax) + (ay +
       ay)) + az_k) + cz_k by Thomas elimination (den_k = D_k - az_k e_{k-1},
       e_k = cz_k / den_k from tables; g_k = ((h_k - rhs_k) + az_k g_{k-1}) /
       den_k; P_k = g_k + e_k P_{k+1});
     residual r = rhs - s, s = ((P_w - P) + (P_e - P)) ax + ((P_s - P) + (P_n
       - P)) ay [+ az (P_b - P)] [+ cz (P_t - P)];
     restriction: coarse rhs = ((r1 + r2) + (r3 + r4)) * 0.25 over the 2x2
       children (i-pairs first), coarse P = 0 (MG_clear);
     prolongation (bilinear in x, y): fine P += 0.75 (0.75 c + 0.25 c_x) +
       0.25 (0.75 c_y + 0.25 c_xy), c_x / c_y the coarse neighbour on the
       child's side;
     cycle: 2 pre-smoothing sweeps, coarse correction, 2 post-smoothing
       sweeps; the coarsest level (halving stops where x or y would drop
       below 4 cells or become odd) is solved exactly (mg_coarse); cycles
       from P = 0 (PoissMG zeroes Phi) until the max-norm
       residual is below poisson_tol (the signed-max test of MG_res,
       src/multigrid.f90:975, taken on |r|) or mg_cycles_max cycles. */
typedef struct {
  int nx, ny;
  double ax, ay;
  double *phi, *rhs, *den, *e;
} mg_level;

static int mg_levels(int nx, int ny) {
  int L = 1;
  while (nx % 4 == 0 && ny % 4 == 0 && nx / 2 >= 4 && ny / 2 >= 4) {
    nx /= 2;
    ny /= 2;
    L++;
  }
  return L;
}

static void mg_sweep(const mg_level *v, int nz, const double *az, const double *cz, int color) {
  const int nx = v->nx, ny = v->ny;
  const long plane = (long)nx * ny;
  double *P = v->phi;
#pragma omp parallel for collapse(2)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      if (((i + j) & 1) != color) continue;
      const int iw = i > 1 ? i - 1 : nx, ie = i < nx ? i + 1 : 1, js = j > 1 ? j - 1 : ny, jn = j < ny ? j + 1 : 1;
      const long c = (i - 1) + (long)(j - 1) * nx;
      const long w = (iw - 1) + (long)(j - 1) * nx, e = (ie - 1) + (long)(j - 1) * nx;
      const long so = (i - 1) + (long)(js - 1) * nx, no = (i - 1) + (long)(jn - 1) * nx;
      double g = 0;
      for (int k = 1; k <= nz; k++) {
        const long o = (long)(k - 1) * plane;
        const double h = v->ax * (P[w + o] + P[e + o]) + v->ay * (P[so + o] + P[no + o]);
        g = ((h - v->rhs[c + o]) + az[k - 1] * g) / v->den[k - 1];
        P[c + o] = g;
      }
      for (int k = nz - 1; k >= 1; k--) {
        const long o = (long)(k - 1) * plane;
        P[c + o] = P[c + o] + v->e[k - 1] * P[c + o + plane];
      }
    }
}

static double mg_res_cell(const mg_level *v, int nz, const double *az, const double *cz, int i, int j, int k) {
  const int nx = v->nx, ny = v->ny;
  const long plane = (long)nx * ny, o = (long)(k - 1) * plane;
  const int iw = i > 1 ? i - 1 : nx, ie = i < nx ? i + 1 : 1, js = j > 1 ? j - 1 : ny, jn = j < ny ? j + 1 : 1;
  const double *P = v->phi;
  const double p = P[(i - 1) + (long)(j - 1) * nx + o];
  double s = ((P[(iw - 1) + (long)(j - 1) * nx + o] - p) + (P[(ie - 1) + (long)(j - 1) * nx + o] - p)) * v->ax +
             ((P[(i - 1) + (long)(js - 1) * nx + o] - p) + (P[(i - 1) + (long)(jn - 1) * nx + o] - p)) * v->ay;
  if (k > 1) s = s + az[k - 1] * (P[(i - 1) + (long)(j - 1) * nx + o - plane] - p);
  if (k < nz) s = s + cz[k - 1] * (P[(i - 1) + (long)(j - 1) * nx + o + plane] - p);
  return v->rhs[(i - 1) + (long)(j - 1) * nx + o] - s;
}

static double mg_residual_max(const mg_level *v, int nz, const double *az, const double *cz) {
  double m = 0;
#pragma omp parallel for collapse(2) reduction(max : m)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= v->ny; j++)
      for (int i = 1; i <= v->nx; i++) {
        const double r = fabs(mg_res_cell(v, nz, az, cz, i, j, k));
        m = r != r ? INFINITY : (r > m ? r : m);
      }
  return m;
}

static void mg_restrict(const mg_level *f, mg_level *c, int nz, const double *az, const double *cz) {
  const long cplane = (long)c->nx * c->ny;
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int J = 1; J <= c->ny; J++)
      for (int I = 1; I <= c->nx; I++) {
        const double r1 = mg_res_cell(f, nz, az, cz, 2 * I - 1, 2 * J - 1, k);
        const double r2 = mg_res_cell(f, nz, az, cz, 2 * I, 2 * J - 1, k);
        const double r3 = mg_res_cell(f, nz, az, cz, 2 * I - 1, 2 * J, k);
        const double r4 = mg_res_cell(f, nz, az, cz, 2 * I, 2 * J, k);
        const long q = (I - 1) + (long)(J - 1) * c->nx + (long)(k - 1) * cplane;
        c->rhs[q] = ((r1 + r2) + (r3 + r4)) * 0.25;
        c->phi[q] = 0;
      }
}

static void mg_prolong(const mg_level *c, mg_level *f, int nz) {
  const long cplane = (long)c->nx * c->ny, fplane = (long)f->nx * f->ny;
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= f->ny; j++)
      for (int i = 1; i <= f->nx; i++) {
        const int I = (i + 1) / 2, J = (j + 1) / 2;
        const int Ix = (i & 1) ? (I > 1 ? I - 1 : c->nx) : (I < c->nx ? I + 1 : 1);
        const int Jy = (j & 1) ? (J > 1 ? J - 1 : c->ny) : (J < c->ny ? J + 1 : 1);
        const double *E = c->phi + (long)(k - 1) * cplane;
        const double c00 = E[(I - 1) + (long)(J - 1) * c->nx], c10 = E[(Ix - 1) + (long)(J - 1) * c->nx];
        const double c01 = E[(I - 1) + (long)(Jy - 1) * c->nx], c11 = E[(Ix - 1) + (long)(Jy - 1) * c->nx];
        const double ex0 = 0.75 * c00 + 0.25 * c10, ex1 = 0.75 * c01 + 0.25 * c11;
        const long q = (i - 1) + (long)(j - 1) * f->nx + (long)(k - 1) * fplane;
        f->phi[q] = f->phi[q] + (0.75 * ex0 + 0.25 * ex1);
      }
}

/* Coarsest level: exact solve of L P = rhs (P = 0 on entry) by discrete
   Fourier transforms in x and y, evaluated by direct summation (a few
   columns), and a tridiagonal solve in z per mode (the direct solver's
   operator on the coarse grid):
     Y(m, j, k) = sum_i rhs(i, j, k) w_x(m (i-1) mod nx), i increasing,
     X(m, n, k) = sum_j Y(m, j, k) w_y(n (j-1) mod ny), j increasing,
   with w(q) = (cos 2pi q/N, -sin 2pi q/N) and complex products (a.re b.re -
   a.im b.im, a.re b.im + a.im b.re) accumulated as s = s + p;
     mode (m, n) != 0: Thomas elimination of a_k x_{k-1} + (lam - a_k - c_k)
       x_k + c_k x_{k+1} = X_k, lam = -4 sin^2(pi m/nx)/dx^2 - 4 sin^2(pi
       n/ny)/dy^2 (den_k = b_k - a_k cp_{k-1}, cp_k = c_k / den_k, dp_k = (X_k
       - a_k dp_{k-1}) / den_k, x_k = dp_k - cp_k x_{k+1});
     mode (0, 0): G_k = G_{k-1} + X_k dzPr(k) from G_0 = 0, x_{k+1} = x_k +
       G_k dzW(k) from x_1 = 0 (the singular Neumann problem);
   then the conjugate sums over n, then m, real part times 1/(nx ny). */
typedef struct {
  double re, im;
} mgc;
static mgc mgc_mul(mgc a, mgc b) {
  mgc r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
void oracle_mg_twiddles(int n, double *cs, double *sn) {
  for (int q = 0; q < n; q++) {
    long double t = 2.0L * 3.14159265358979323846264338327950288L * q / n;
    cs[q] = (double)cosl(t);
    sn[q] = -(double)sinl(t);
  }
}
static void mg_coarse(mg_level *v, int nz, const double *az, const double *cz, const double *dzPr,
                      const double *dzW) {
  const int nx = v->nx, ny = v->ny;
  const long plane = (long)nx * ny;
  double *cx = xcalloc(nx, sizeof(double)), *sx = xcalloc(nx, sizeof(double));
  double *cy = xcalloc(ny, sizeof(double)), *sy = xcalloc(ny, sizeof(double));
  oracle_mg_twiddles(nx, cx, sx);
  oracle_mg_twiddles(ny, cy, sy);
  mgc *Y = xcalloc(plane * nz, sizeof(mgc)), *X = xcalloc(plane * nz, sizeof(mgc));
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int m = 0; m < nx; m++) {
        mgc acc = {0, 0};
        for (int i = 0; i < nx; i++) {
          const mgc w = {cx[(m * i) % nx], sx[(m * i) % nx]};
          const mgc r = {v->rhs[i + (long)j * nx + (long)k * plane], 0};
          const mgc p = mgc_mul(r, w);
          acc.re = acc.re + p.re;
          acc.im = acc.im + p.im;
        }
        Y[m + (long)j * nx + (long)k * plane] = acc;
      }
  for (int k = 0; k < nz; k++)
    for (int n = 0; n < ny; n++)
      for (int m = 0; m < nx; m++) {
        mgc acc = {0, 0};
        for (int j = 0; j < ny; j++) {
          const mgc w = {cy[(n * j) % ny], sy[(n * j) % ny]};
          const mgc p = mgc_mul(Y[m + (long)j * nx + (long)k * plane], w);
          acc.re = acc.re + p.re;
          acc.im = acc.im + p.im;
        }
        X[m + (long)n * nx + (long)k * plane] = acc;
      }
  double *cp = xcalloc(nz, sizeof(double));
  mgc *dp = xcalloc(nz, sizeof(mgc));
  for (int n = 0; n < ny; n++)
    for (int m = 0; m < nx; m++) {
      mgc *col = X + m + (long)n * nx;
      if (m == 0 && n == 0) {
        double Gr = 0, Gi = 0, xr = 0, xi = 0;
        for (int k = 1; k <= nz; k++) {
          const mgc d = col[(long)(k - 1) * plane];
          col[(long)(k - 1) * plane].re = xr;
          col[(long)(k - 1) * plane].im = xi;
          if (k < nz) {
            Gr = Gr + d.re * dzPr[k];
            Gi = Gi + d.im * dzPr[k];
            xr = xr + Gr * dzW[k];
            xi = xi + Gi * dzW[k];
          }
        }
        continue;
      }
      const double sxm = sin(M_PI * m / nx), syn = sin(M_PI * n / ny);
      const double lam = -4 * sxm * sxm * v->ax - 4 * syn * syn * v->ay;
      double cprev = 0;
      mgc dprev = {0, 0};
      for (int k = 0; k < nz; k++) {
        const double b = (lam - az[k]) - cz[k];
        const double den = b - az[k] * cprev;
        cp[k] = cz[k] / den;
        const mgc d = col[(long)k * plane];
        dp[k].re = (d.re - az[k] * dprev.re) / den;
        dp[k].im = (d.im - az[k] * dprev.im) / den;
        cprev = cp[k];
        dprev = dp[k];
      }
      mgc x = dp[nz - 1];
      col[(long)(nz - 1) * plane] = x;
      for (int k = nz - 2; k >= 0; k--) {
        x.re = dp[k].re - cp[k] * x.re;
        x.im = dp[k].im - cp[k] * x.im;
        col[(long)k * plane] = x;
      }
    }
  /* inverse: conjugate twiddles over n, then over m */
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int m = 0; m < nx; m++) {
        mgc acc = {0, 0};
        for (int n = 0; n < ny; n++) {
          const mgc w = {cy[(n * j) % ny], -sy[(n * j) % ny]};
          const mgc p = mgc_mul(X[m + (long)n * nx + (long)k * plane], w);
          acc.re = acc.re + p.re;
          acc.im = acc.im + p.im;
        }
        Y[m + (long)j * nx + (long)k * plane] = acc;
      }
  const double scale = 1.0 / (double)plane;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double acc = 0;
        for (int m = 0; m < nx; m++) {
          const mgc w = {cx[(m * i) % nx], -sx[(m * i) % nx]};
          const mgc y = Y[m + (long)j * nx + (long)k * plane];
          acc = acc + (y.re * w.re - y.im * w.im);
        }
        v->phi[i + (long)j * nx + (long)k * plane] = acc * scale;
      }
  free(cx); free(sx); free(cy); free(sy); free(Y); free(X); free(cp); free(dp);
}

static void mg_vcycle(mg_level *lv, int l, int L, int nz, const double *az, const double *cz, const double *dzPr,
                      const double *dzW) {
  if (l == L - 1) {
    mg_coarse(&lv[l], nz, az, cz, dzPr, dzW);
    return;
  }
  for (int s = 0; s < 2; s++) {
    mg_sweep(&lv[l], nz, az, cz, 0);
    mg_sweep(&lv[l], nz, az, cz, 1);
  }
  mg_restrict(&lv[l], &lv[l + 1], nz, az, cz);
  mg_vcycle(lv, l + 1, L, nz, az, cz, dzPr, dzW);
  mg_prolong(&lv[l + 1], &lv[l], nz);
  for (int s = 0; s < 2; s++) {
    mg_sweep(&lv[l], nz, az, cz, 0);
    mg_sweep(&lv[l], nz, az, cz, 1);
  }
}

/* Line-solve tables of a level: den_k and e_k (k = 1..nz). */
void oracle_mg_tables(int nz, double ax, double ay, const double *az, const double *cz, double *den, double *e) {
  double eprev = 0;
  for (int k = 1; k <= nz; k++) {
    const double D = (((ax + ax) + (ay + ay)) + az[k - 1]) + cz[k - 1];
    den[k - 1] = D - az[k - 1] * eprev;
    e[k - 1] = cz[k - 1] / den[k - 1];
    eprev = e[k - 1];
  }
}

/* Multigrid solve; phi and rhs as poisson_direct.  Returns the cycles run
   (cycles_out) and the final max-norm residual (res_out). */
static void poisson_mg(oracle *o, const double *rhs, double *phi, int *cycles_out, double *res_out) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const int L = mg_levels(nx, ny);
  double *az = xcalloc(nz, sizeof(double)), *cz = xcalloc(nz, sizeof(double));
  for (int k = 1; k <= nz; k++) {
    az[k - 1] = k > 1 ? 1 / (C1D(o->dzW, k - 1) * C1D(o->dzPr, k)) : 0;
    cz[k - 1] = k < nz ? 1 / (C1D(o->dzW, k) * C1D(o->dzPr, k)) : 0;
  }
  mg_level *lv = xcalloc(L, sizeof(mg_level));
  for (int l = 0; l < L; l++) {
    lv[l].nx = nx >> l;
    lv[l].ny = ny >> l;
    const double dxl = o->dxmin * (double)(1 << l), dyl = o->dymin * (double)(1 << l);
    lv[l].ax = 1.0 / (dxl * dxl);
    lv[l].ay = 1.0 / (dyl * dyl);
    const long n = (long)lv[l].nx * lv[l].ny * nz;
    lv[l].phi = l == 0 ? phi : xcalloc(n, sizeof(double));
    lv[l].rhs = l == 0 ? (double *)rhs : xcalloc(n, sizeof(double));
    lv[l].den = xcalloc(nz, sizeof(double));
    lv[l].e = xcalloc(nz, sizeof(double));
    oracle_mg_tables(nz, lv[l].ax, lv[l].ay, az, cz, lv[l].den, lv[l].e);
  }
  const long n0 = (long)nx * ny * nz;
  for (long q = 0; q < n0; q++) phi[q] = 0;
  const int maxc = o->c.mg_cycles_max > 0 ? o->c.mg_cycles_max : 50;
  const double tol = o->c.poisson_tol > 0 ? o->c.poisson_tol : 1e-10;
  int cyc = 0;
  double R = mg_residual_max(&lv[0], nz, az, cz);
  double *dzPr = xcalloc(nz + 1, sizeof(double)), *dzW = xcalloc(nz + 1, sizeof(double));
  for (int k = 1; k <= nz; k++) {
    dzPr[k] = C1D(o->dzPr, k);
    dzW[k] = C1D(o->dzW, k);
  }
  while (cyc < maxc && !(R < tol)) {
    mg_vcycle(lv, 0, L, nz, az, cz, dzPr, dzW);
    cyc++;
    R = mg_residual_max(&lv[0], nz, az, cz);
  }
  if (cycles_out) *cycles_out = cyc;
  if (res_out) *res_out = R;
  for (int l = 0; l < L; l++) {
    if (l > 0) {
      free(lv[l].phi);
      free(lv[l].rhs);
    }
    free(lv[l].den);
    free(lv[l].e);
  }
  free(lv);
  free(az);
  free(cz);
  free(dzPr);
  free(dzW);
}

/* Direct solve; phi and rhs are interior arrays nx*ny*nz, x fastest.
   The singular mode pins Phi_hat(k=1) = 0. */
static void poisson_direct(oracle *o, const double *rhs, double *phi) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  {
    int a = nx, b = ny;
    while (a % 3 == 0) a /= 3;
    while (b % 3 == 0) b /= 3;
    if (!is_pow2i(a) || !is_pow2i(b)) {
      fprintf(stderr, "oracle: nx, ny of the form 2^a 3^b required by the FFT solver\n");
      abort();
    }
  }
  const int M = nx / 2 + 1;
  long plane = (long)M * ny;
  c2 *X = xcalloc(plane * nz, sizeof(c2));
  c2 *twx = is_pow2i(nx) ? twiddles(nx) : twiddles_full(nx), *twy = is_pow2i(ny) ? twiddles(ny) : twiddles_full(ny);
  int nmax = nx > ny ? nx : ny;
  nmax = nmax > nz ? nmax : nz;
  /* x transform of each real row, keep modes 0..nx/2 */
#pragma omp parallel
  {
    c2 *work = xcalloc(2 * (size_t)nmax, sizeof(c2)), *line = xcalloc(nx, sizeof(c2));
#pragma omp for collapse(2)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        for (int i = 0; i < nx; i++) {
          line[i].re = rhs[i + (long)j * nx + (long)k * nx * ny];
          line[i].im = 0;
        }
        fft_any(line, nx, 1, twx, 0, work);
        for (int m = 0; m < M; m++) X[(long)k * plane + (long)j * M + m] = line[m];
      }
#pragma omp for collapse(2)
    for (int k = 0; k < nz; k++)
      for (int m = 0; m < M; m++) fft_any(X + (long)k * plane + m, ny, M, twy, 0, work);
    free(work);
    free(line);
  }
  /* z systems */
  double *az = xcalloc(nz, sizeof(double)), *cz = xcalloc(nz, sizeof(double));
  for (int k = 1; k <= nz; k++) {
    az[k - 1] = k > 1 ? 1 / (C1D(o->dzW, k - 1) * C1D(o->dzPr, k)) : 0;
    cz[k - 1] = k < nz ? 1 / (C1D(o->dzW, k) * C1D(o->dzPr, k)) : 0;
  }
  double *lamx = xcalloc(M, sizeof(double)), *lamy = xcalloc(ny, sizeof(double));
  for (int m = 0; m < M; m++) {
    double sv = sin(M_PI * m / nx);
    lamx[m] = -4 * sv * sv / (o->dxmin * o->dxmin);
  }
  for (int j = 0; j < ny; j++) {
    double sv = sin(M_PI * j / ny);
    lamy[j] = -4 * sv * sv / (o->dymin * o->dymin);
  }
#pragma omp parallel
  {
    double *a = xcalloc(nz, sizeof(double)), *b = xcalloc(nz, sizeof(double)), *c = xcalloc(nz, sizeof(double));
    double *na = xcalloc(nz, sizeof(double));
    c2 *d = xcalloc(nz, sizeof(c2));
#pragma omp for collapse(2)
    for (int j = 0; j < ny; j++)
      for (int m = 0; m < M; m++) {
        for (int k = 0; k < nz; k++) d[k] = X[(long)k * plane + (long)j * M + m];
        if (m == 0 && j == 0) {
          /* horizontal-mean mode with x_1 = 0: G_{k-1} = G_k - d_k dzPr(k)
             from G_nz = 0, then x_{k+1} = x_k + G_k dzW(k); compensated */
          double Gr = 0, Gi = 0, cr = 0, ci = 0;
          for (int k = nz; k >= 2; k--) {
            double dz = C1D(o->dzPr, k);
            kahan(&Gr, &cr, -d[k - 1].re * dz);
            kahan(&Gi, &ci, -d[k - 1].im * dz);
            a[k - 2] = Gr;
            c[k - 2] = Gi;
          }
          double xr = 0, xi = 0;
          cr = ci = 0;
          d[0].re = 0;
          d[0].im = 0;
          for (int k = 1; k <= nz - 1; k++) {
            double dz = C1D(o->dzW, k);
            kahan(&xr, &cr, a[k - 1] * dz);
            kahan(&xi, &ci, c[k - 1] * dz);
            d[k].re = xr;
            d[k].im = xi;
          }
        } else {
          double lam = lamx[m] + lamy[j];
          for (int k = 0; k < nz; k++) {
            a[k] = az[k];
            c[k] = cz[k];
            b[k] = lam - a[k] - c[k];
          }
          thomas_solve(nz, a, b, c, d, na);
        }
        for (int k = 0; k < nz; k++) X[(long)k * plane + (long)j * M + m] = d[k];
      }
    free(a); free(b); free(c); free(na); free(d);
  }
  free(az); free(cz); free(lamx); free(lamy);
  /* inverse y, then inverse x from the half spectrum (Hermitian symmetry) */
  double scale = 1.0 / ((double)nx * ny);
#pragma omp parallel
  {
    c2 *work = xcalloc(2 * (size_t)nmax, sizeof(c2)), *line = xcalloc(nx, sizeof(c2));
#pragma omp for collapse(2)
    for (int k = 0; k < nz; k++)
      for (int m = 0; m < M; m++) fft_any(X + (long)k * plane + m, ny, M, twy, 1, work);
#pragma omp for collapse(2)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        const c2 *row = X + (long)k * plane + (long)j * M;
        for (int i = 0; i < nx; i++) {
          if (i < M) {
            line[i] = row[i];
          } else {
            line[i].re = row[nx - i].re;
            line[i].im = -row[nx - i].im;
          }
        }
        fft_any(line, nx, 1, twx, 1, work);
        for (int i = 0; i < nx; i++) phi[i + (long)j * nx + (long)k * nx * ny] = line[i].re * scale;
      }
    free(work);
    free(line);
  }
  free(X);
  free(twx);
  free(twy);
}

/* Residue_variable_z, src/poisson.f90:23-60 */
static double residue(oracle *o, const double *Phi, const double *RHS) {
  double Aw = 1 / (o->dxmin * o->dxmin), Ae = Aw, As = 1 / (o->dymin * o->dymin), An = As;
  double res = 0;
  for (int k = 1; k <= o->nz; k++) {
    double Ab = 1 / ((C1D(o->zPr, k) - C1D(o->zPr, k - 1)) * (C1D(o->zW, k) - C1D(o->zW, k - 1)));
    double At = 1 / ((C1D(o->zPr, k + 1) - C1D(o->zPr, k)) * (C1D(o->zW, k) - C1D(o->zW, k - 1)));
    double Ap = Aw + Ae + As + An + Ab + At;
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++) {
        double p = 0;
        p = p + A3(o, Phi, i - 1, j, k) * Aw;
        p = p + A3(o, Phi, i + 1, j, k) * Ae;
        p = p + A3(o, Phi, i, j - 1, k) * As;
        p = p + A3(o, Phi, i, j + 1, k) * An;
        p = p + A3(o, Phi, i, j, k - 1) * Ab;
        p = p + A3(o, Phi, i, j, k + 1) * At;
        p = p - A3(o, RHS, i, j, k);
        p = fabs(-p + Ap * A3(o, Phi, i, j, k));
        if (p > res) res = p;
      }
  }
  return res;
}

/* Sum in the device's order: blocks of 256 values, each the in-order sum of
   four 64-lane xor-butterfly sums; block sums added in order. */
static double tree_sum(const double *v, long n) {
  double total = 0;
  for (long b0 = 0; b0 < n; b0 += 256) {
    double blk = 0;
    for (int w = 0; w < 4; w++) {
      double lane[64];
      for (int l = 0; l < 64; l++) {
        long t = b0 + 64 * w + l;
        lane[l] = t < n ? v[t] : 0.0;
      }
      for (int off = 32; off > 0; off >>= 1) {
        double nl[64];
        for (int l = 0; l < 64; l++) nl[l] = lane[l] + lane[l ^ off];
        memcpy(lane, nl, sizeof lane);
      }
      blk += lane[0];
    }
    total += blk;
  }
  return total;
}

/* ------------------------------------------------ pressure correction --- */
/* PressureCorrection -> PrePoisson -> solver -> Bound_Phi -> PostPoisson,
   src/pressure.f90:169-266, 272-723, 728-991, 994-1229 */
static void pressure_correction(oracle *o, double *U, double *V, double *W, int use_q, double coef) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double dt2 = coef;
  double dt2_rec = 1.0 / dt2;
  bound_uvw(o, U, V, W);
  /* zero boundary-face velocity inside solids on open sides (pressure.f90:301-342) */
  if (o->c.btype[To] != NOSLIP && o->c.btype[To] != PERIODIC)
    for (int j = 1; j <= o->Wn[1]; j++)
      for (int i = 1; i <= o->Wn[0]; i++)
        if (A3(o, o->Wtype, i, j, nz) > 0) A3(o, W, i, j, nz) = 0;
  if (o->c.btype[Bo] != NOSLIP && o->c.btype[Bo] != PERIODIC)
    for (int j = 1; j <= o->Wn[1]; j++)
      for (int i = 1; i <= o->Wn[0]; i++)
        if (A3(o, o->Wtype, i, j, 0) > 0) A3(o, W, i, j, 0) = 0;
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double dz = o->variable_z ? C1D(o->dzPr, k) : o->dzmin;
        double r = (A3(o, U, i, j, k) - A3(o, U, i - 1, j, k)) / o->dxmin +
                   (A3(o, V, i, j, k) - A3(o, V, i, j - 1, k)) / o->dymin +
                   (A3(o, W, i, j, k) - A3(o, W, i, j, k - 1)) / dz;
        if (use_q) r = r - A3(o, o->Q, i, j, k);
        A3(o, o->RHS, i, j, k) = r * dt2_rec;
      }
  long plane = (long)nx * ny;
  double *rhs = xcalloc(plane * nz, sizeof(double)), *phi = xcalloc(plane * nz, sizeof(double));
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) rhs[(i - 1) + (j - 1) * nx + (k - 1) * plane] = A3(o, o->RHS, i, j, k);
  if (o->c.poisson_solver == 3)
    poisson_mg(o, rhs, phi, &o->last_mg_cycles, NULL);
  else
    poisson_direct(o, rhs, phi);
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) A3(o, o->Phi, i, j, k) = phi[(i - 1) + (j - 1) * nx + (k - 1) * plane];
  free(rhs);
  free(phi);
  bound_phi(o, o->Phi);
  o->last_res = residue(o, o->Phi, o->RHS);
  /* PostPoisson (2nd order) */
  double Au = dt2 / o->dxmin, Av = dt2 / o->dymin, Aw = dt2 / o->dzmin;
  const double *Phi = o->Phi;
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= o->Un[2]; k++)
    for (int j = 1; j <= o->Un[1]; j++)
      for (int i = 1; i <= o->Un[0]; i++)
        A3(o, U, i, j, k) = A3(o, U, i, j, k) - Au * (A3(o, Phi, i + 1, j, k) - A3(o, Phi, i, j, k));
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= o->Vn[2]; k++)
    for (int j = 1; j <= o->Vn[1]; j++)
      for (int i = 1; i <= o->Vn[0]; i++)
        A3(o, V, i, j, k) = A3(o, V, i, j, k) - Av * (A3(o, Phi, i, j + 1, k) - A3(o, Phi, i, j, k));
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= o->Wn[2]; k++)
    for (int j = 1; j <= o->Wn[1]; j++)
      for (int i = 1; i <= o->Wn[0]; i++) {
        if (o->variable_z)
          A3(o, W, i, j, k) =
              A3(o, W, i, j, k) - dt2 * (A3(o, Phi, i, j, k + 1) - A3(o, Phi, i, j, k)) / C1D(o->dzW, k);
        else
          A3(o, W, i, j, k) = A3(o, W, i, j, k) - Aw * (A3(o, Phi, i, j, k + 1) - A3(o, Phi, i, j, k));
      }
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) A3(o, o->Pr, i, j, k) = A3(o, o->Pr, i, j, k) + A3(o, Phi, i, j, k);
  /* FixPressureToReference: mean of the top layer (reference boundary To).
     The reference sums with an OpenMP reduction whose order is unspecified;
     the order used here is the device's: blocks of 256 consecutive cells
     (i fastest), 64-lane butterfly sums per wavefront, wavefront sums in
     order, block sums in order. */
  double *top = xcalloc((long)nx * ny, sizeof(double));
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) top[(i - 1) + (long)(j - 1) * nx] = A3(o, o->Pr, i, j, nz);
  double ref = tree_sum(top, (long)nx * ny);
  free(top);
  ref = ref / (double)(nx * ny);
  if (ref != 0)
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) A3(o, o->Pr, i, j, k) = A3(o, o->Pr, i, j, k) - ref;
  bound_pr(o, o->Pr);
  /* check_divergence (pressure.f90:884-989) computes on refreshed ghosts;
     the diagnostic works on copies so the state keeps the reference's
     (stale) ghosts, which the next SubgridModel call reads at init */
  double *Uc = xcalloc(o->ntot, sizeof(double)), *Vc = xcalloc(o->ntot, sizeof(double)),
         *Wc = xcalloc(o->ntot, sizeof(double));
  memcpy(Uc, U, o->ntot * sizeof(double));
  memcpy(Vc, V, o->ntot * sizeof(double));
  memcpy(Wc, W, o->ntot * sizeof(double));
  bound_uvw(o, Uc, Vc, Wc);
  double S = 0;
#pragma omp parallel for collapse(2) reduction(max : S)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double dz = o->variable_z ? C1D(o->dzPr, k) : o->dzmin;
        double p = (A3(o, Uc, i, j, k) - A3(o, Uc, i - 1, j, k)) / (o->dxmin) +
                   (A3(o, Vc, i, j, k) - A3(o, Vc, i, j - 1, k)) / (o->dymin) +
                   (A3(o, Wc, i, j, k) - A3(o, Wc, i, j, k - 1)) / (dz);
        if (use_q) p = p - A3(o, o->Q, i, j, k);
        if (fabs(p) > S) S = fabs(p);
      }
  free(Uc);
  free(Vc);
  free(Wc);
  o->last_div = S;
}

/* TimeStepLength with a constant dt and optional CFL check,
   src/dynamics.f90:511-600 */
static double cfl_max(oracle *o) {
  double m = 0;
  int nan = 0;
#pragma omp parallel for collapse(2) reduction(max : m) reduction(| : nan)
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++) {
        double a1 = fabs(A3(o, o->U, i, j, k)), a0 = fabs(A3(o, o->U, i - 1, j, k));
        double b1 = fabs(A3(o, o->V, i, j, k)), b0 = fabs(A3(o, o->V, i, j - 1, k));
        double c1 = fabs(A3(o, o->W, i, j, k)), c0 = fabs(A3(o, o->W, i, j, k - 1));
        double dzk = o->variable_z ? C1D(o->dzPr, k) : o->dzmin;
        double dxk = o->variable_z ? C1D(o->dxPr, i) : o->dxmin;
        double dyk = o->variable_z ? C1D(o->dyPr, j) : o->dymin;
        double p = (a1 > a0 ? a1 : a0) / dxk;
        p = p + (b1 > b0 ? b1 : b0) / dyk;
        p = p + (c1 > c0 ? c1 : c0) / dzk;
        if (p != p) nan = 1;
        if (p > m) m = p;
      }
  return nan ? INFINITY : m;
}

/* TMarchRK3's delta term of one stage (src/tsteps.f90:188-201): the mean of
   |U - U2| over each component's interior, added in the order U, V, W.  The
   reference sums sequentially; the sums here are tree_sum's (the device's
   order), over the interior in x-fastest order. */
static double delta_stage(oracle *o, double d) {
  const double *A[3] = {o->U, o->V, o->W}, *B[3] = {o->U2, o->V2, o->W2};
  const int *n[3] = {o->Un, o->Vn, o->Wn};
  for (int c = 0; c < 3; c++) {
    long N = (long)n[c][0] * n[c][1] * n[c][2];
    if (N <= 0) continue;
    double *v = xcalloc(N, sizeof(double));
    long t = 0;
    for (int k = 1; k <= n[c][2]; k++)
      for (int j = 1; j <= n[c][1]; j++)
        for (int i = 1; i <= n[c][0]; i++) v[t++] = fabs(A3(o, A[c], i, j, k) - A3(o, B[c], i, j, k));
    double sum = tree_sum(v, N);
    free(v);
    d = d + sum / (double)N;
  }
  return d;
}

/* ----------------------------------------------------- public driver --- */
int oracle_init(oracle *o) {
  /* InitialConditions tail, src/initial.f90:4284-4342 */
  bound_uvw(o, o->U, o->V, o->W);
  bound_pr(o, o->Pr);
  pressure_correction(o, o->U, o->V, o->W, 0, 1.0);
  sgs_smag(o);
  bound_visc(o, o->Visc);
  if (o->c.nscalars > 0 || o->c.enable_buoyancy) {
    for (int k = 1; k <= o->nz; k++)
      for (int j = 1; j <= o->ny; j++)
        for (int i = 1; i <= o->nx; i++)
          A3(o, o->TDiff, i, j, k) = (A3(o, o->Visc, i, j, k) - o->c.nu) / (double)0.6f + o->c.nu_scalar;
    bound_visc(o, o->TDiff);
  }
  if (o->c.enable_buoyancy) bound_temperature(o, o->S[o->c.nscalars]);
  if (o->c.enable_buoyancy && o->c.enable_moisture) bound_moisture(o, o->S[o->c.nscalars + 1]);
  for (int s = 0; s < o->c.nscalars; s++) bound_scalar(o, o->S[s]);
  wm_visc_points(o, 1);
  wm_uvw_fluxes(o);
  bound_visc(o, o->Visc);
  o->initialized = 1;
  return 0;
}

static void swap(double **a, double **b) {
  double *t = *a;
  *a = *b;
  *b = t;
}

int oracle_step(oracle *o, int32_t nsteps, elmm_diag *diag) {
  if (!o->initialized) {
    set_err("oracle_step before oracle_init");
    return ELMM_E_STATE;
  }
  double cfl = o->cfl_last;
  int done = 0;
  for (int s = 0; s < nsteps; s++) {
    if (!o->first_step_done) {
      /* first call of TMarchRK3, src/tsteps.f90:54-77 */
      memcpy(o->U2, o->U, o->ntot * sizeof(double));
      memcpy(o->V2, o->V, o->ntot * sizeof(double));
      memcpy(o->W2, o->W, o->ntot * sizeof(double));
      memcpy(o->Us, o->U, o->ntot * sizeof(double));
      memcpy(o->Vs, o->V, o->ntot * sizeof(double));
      memcpy(o->Ws, o->W, o->ntot * sizeof(double));
      bound_uvw(o, o->U, o->V, o->W);
      ib_momentum(o, o->U, o->V, o->W);
      o->first_step_done = 1;
    }
    /* TimeStepLength (src/dynamics.f90:511-549) */
    double dt = o->c.dt;
    int check_cfl = 0;
    if (o->c.variable_dt) {
      double m = cfl_max(o); /* +inf for a NaN velocity */
      if (!(m < INFINITY))
        dt = DBL_MIN;
      else if (m > 0)
        dt = fmin(o->c.cfl_target / m, o->c.dt_max);
      else
        dt = o->c.dt_max;
      if (o->c.end_time > 0 && dt + o->time > o->c.end_time) dt = o->c.end_time - o->time;
      cfl = m * dt;
    } else if (o->c.cfl_check) {
      double m = cfl_max(o);
      cfl = m * dt;
      check_cfl = 1;
    }
    double delta = 0;
    for (int st = 0; st < 3; st++) {
      double coef = 2 * RK_alpha[st] * dt;
      subgrid_stresses(o);
      convection(o, st, dt);
      scalar_rk3(o, st, dt);
      other_terms(o, coef);
      bound_uvw(o, o->U2, o->V2, o->W2);
      ib_momentum(o, o->U2, o->V2, o->W2);
      bound_uvw(o, o->U2, o->V2, o->W2);
      int use_q = o->c.ib_mass_sources != 0;
      if (use_q) ib_mass_sources(o, o->U2, o->V2, o->W2);
      pressure_correction(o, o->U2, o->V2, o->W2, use_q, coef);
      if (o->c.delta_diag) {
        delta = delta_stage(o, delta);
        if (st == 2) delta = delta / dt;
      }
      swap(&o->U, &o->U2);
      swap(&o->V, &o->V2);
      swap(&o->W, &o->W2);
      null_interior(o);
      bound_uvw(o, o->U, o->V, o->W);
      ib_momentum(o, o->U, o->V, o->W);
      bound_uvw(o, o->U, o->V, o->W);
    }
    /* the main loop's time update and divergence tests (src/main.f90:116-169) */
    o->steps_total++;
    done++;
    o->dt_last = dt;
    o->cfl_last = cfl;
    if (o->c.delta_diag) o->delta_last = delta;
    if (o->c.variable_dt) {
      o->time = o->time + dt;
      if (o->c.end_time > 0 && fabs(o->time - o->c.end_time) < 5 * (nextafter(fabs(o->c.end_time), INFINITY) -
                                                                    fabs(o->c.end_time)))
        o->time = o->c.end_time;
    } else {
      o->time = o->c.start_time + (double)o->steps_total * dt;
      if (o->c.end_time > 0 && fabs(o->time - o->c.end_time) < dt / 10) o->time = o->c.end_time;
    }
    if (o->steps_total >= 3 && o->c.variable_dt && dt < o->c.dt_min) {
      set_err("Solution diverged. 1");
      break;
    }
    if (o->steps_total >= 3 && check_cfl && o->c.cfl_max > 0 && !(cfl <= o->c.cfl_max)) {
      set_err("Solution diverged. 2");
      break;
    }
  }
  if (diag) {
    memset(diag, 0, sizeof *diag);
    diag->cfl = cfl;
    diag->poisson_res = o->last_res;
    diag->max_div = o->last_div;
    diag->refinements = o->c.poisson_solver == 3 ? o->last_mg_cycles : 0;
    diag->steps_done = done;
    diag->delta = o->delta_last;
    diag->dt = o->dt_last;
    diag->time = o->time;
  }
  return done < nsteps ? ELMM_E_DIVERGED : 0;
}

/* max |div U - Q| as in PostPoisson check_divergence (2nd order),
   src/pressure.f90:937-981 (without the Q term: the state after a step) */
double oracle_divergence(oracle *o) {
  double S = 0;
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++) {
        double dz = o->variable_z ? C1D(o->dzPr, k) : o->dzmin;
        double p = (A3(o, o->U, i, j, k) - A3(o, o->U, i - 1, j, k)) / (o->dxmin) +
                   (A3(o, o->V, i, j, k) - A3(o, o->V, i, j - 1, k)) / (o->dymin) +
                   (A3(o, o->W, i, j, k) - A3(o, o->W, i, j, k - 1)) / (dz);
        if (o->c.ib_mass_sources && o->first_step_done) p = p - A3(o, o->Q, i, j, k);
        if (fabs(p) > S) S = fabs(p);
      }
  return S;
}

/* ------------------------------------------------------ operator hooks --- */
void oracle_op_bound_uvw(oracle *o) { bound_uvw(o, o->U, o->V, o->W); }
void oracle_op_advection(oracle *o) {
  double *S[3] = {o->Us, o->Vs, o->Ws};
  const int *N[3] = {o->Un, o->Vn, o->Wn};
  for (int c = 0; c < 3; c++) {
    for (long p = 0; p < o->ntot; p++) S[c][p] = 0;
    for (int k = 1; k <= N[c][2]; k++)
      for (int j = 1; j <= N[c][1]; j++)
        for (int i = 1; i <= N[c][0]; i++) {
          double v;
          if (o->variable_z)
            v = c == 0 ? cd2vz_u(o, i, j, k) : c == 1 ? cd2vz_v(o, i, j, k) : cd2vz_w(o, i, j, k);
          else
            v = c == 0 ? cd4_u(o, i, j, k) : c == 1 ? cd4_v(o, i, j, k) : cd4_w(o, i, j, k);
          A3(o, S[c], i, j, k) = v;
        }
  }
}
void oracle_op_diffusion(oracle *o) {
  for (long p = 0; p < o->ntot; p++) o->Us[p] = o->Vs[p] = o->Ws[p] = 0;
  if (o->variable_z)
    momentum_diffusion_vz(o);
  else
    momentum_diffusion(o);
}
void oracle_op_wm_fluxes(oracle *o) { wm_uvw_fluxes(o); }
int32_t oracle_wm_list(oracle *o, int32_t list, double *fluxp, double *fluxm) {
  if (list < 0 || list >= 18) return -1;
  wm_list *L = &o->wm[list];
  for (int p = 0; p < L->n; p++) {
    if (fluxp) fluxp[p] = L->fluxp[p];
    if (fluxm) fluxm[p] = L->fluxm[p];
  }
  return L->n;
}
void oracle_op_kappa(oracle *o, int32_t sc) { kappa_scalar(o, o->Sadv[sc], o->S[sc]); }
void oracle_op_sgs(oracle *o) { subgrid_stresses(o); }
void oracle_op_ib_momentum(oracle *o) { ib_momentum(o, o->U, o->V, o->W); }
void oracle_op_mass_sources(oracle *o) { ib_mass_sources(o, o->U, o->V, o->W); }
void oracle_op_null(oracle *o) { null_interior(o); }
void oracle_op_ib_visc(oracle *o) {
  for (int p = 0; p < o->ib_n[3]; p++) {
    const int32_t *q = &o->ib_ijk[3][3 * p];
    A3(o, o->Visc, q[0], q[1], q[2]) = ib_interp(o, 3, p, o->Visc);
  }
}
int32_t oracle_wm_ustar(oracle *o, double *out) {
  if (out) memcpy(out, o->wmp.ustar, (size_t)o->wmp.n * sizeof(double));
  return o->wmp.n;
}
void oracle_op_wm_visc(oracle *o) { wm_visc_points(o, 1); }
void oracle_op_buoyancy(oracle *o) {
  if (o->c.enable_buoyancy) buoyancy_force(o, o->Ws, o->S[o->c.nscalars]);
}
void oracle_op_temp_fix(oracle *o) {
  if (o->c.enable_buoyancy) temperature_solid_fix(o, o->S[o->c.nscalars]);
}
void oracle_op_moist_fill(oracle *o) {
  if (o->c.enable_buoyancy && o->c.enable_moisture) moisture_solid_fill(o, o->S[o->c.nscalars + 1]);
}
void oracle_op_bound_moisture(oracle *o) {
  if (o->c.enable_buoyancy && o->c.enable_moisture) bound_moisture(o, o->S[o->c.nscalars + 1]);
}
void oracle_op_bound_temperature(oracle *o) {
  if (o->c.enable_buoyancy) bound_temperature(o, o->S[o->c.nscalars]);
}
int32_t oracle_op_wm_div(oracle *o, double *out) {
  wm_divergence(o);
  if (out) memcpy(out, o->wmp_div, (size_t)o->wmp.n * sizeof(double));
  return o->wmp.n;
}
void oracle_op_scalar_extra(oracle *o, int32_t sc) {
  double *S = o->S[sc], *Sadv = o->Sadv[sc];
  for (int p = 0; p < o->src_n; p++) {
    if (o->src_sc[p] != sc) continue;
    const int32_t *q = &o->src_ijk[3 * p];
    A3(o, Sadv, q[0], q[1], q[2]) = A3(o, Sadv, q[0], q[1], q[2]) + o->src_flux[p];
  }
  for (int p = 0; p < o->wmp.n; p++) {
    const int32_t *q = &o->wmp.ijk[3 * p];
    A3(o, Sadv, q[0], q[1], q[2]) = A3(o, Sadv, q[0], q[1], q[2]) + o->wmp_div[p] * A3(o, S, q[0], q[1], q[2]);
  }
}
void oracle_op_bound_pr(oracle *o) { bound_pr(o, o->Pr); }
void oracle_op_bound_scalar(oracle *o, int32_t sc) { bound_scalar(o, o->S[sc]); }
void oracle_op_bound_visc(oracle *o) { bound_visc(o, o->Visc); }
void oracle_op_poisson(oracle *o, const double *rhs, double *phi) { poisson_direct(o, rhs, phi); }
int32_t oracle_op_poisson_mg(oracle *o, const double *rhs, double *phi, double *res) {
  int cyc = 0;
  poisson_mg(o, rhs, phi, &cyc, res);
  return cyc;
}
double oracle_op_residual(oracle *o, const double *phi, const double *rhs) { return residue(o, phi, rhs); }
