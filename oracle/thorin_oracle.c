/*
 * thorin_oracle.c -- CPU oracle for the FARGO_THORIN hydro step (AlgoGas).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never linked into the product.
 *
 * A sequential C99 restatement of the reference's per-DT hydro step for the
 * in-scope options (single rank, so Zero_or_active=0, One_or_active=1,
 * Max_or_active=NRAD, MaxMO_or_active=NRAD-1; split.c:45-48 with
 * CPU_Number=1).  Every function cites the reference lines it follows.  The
 * arithmetic keeps the reference's operation order, so that, compiled with
 * gcc -std=c99 (no FMA contraction, as the reference's makefile:33 build),
 * the oracle reproduces the reference bit for bit.
 *
 * Parity pins:
 *   - bitwise against the reference itself, built from its sources by
 *     oracle/ref.mk and run on seeded disks (fixtures under tests/golden/,
 *     made by tests/golden/make_golden.py): every output field, the SOR
 *     iteration counts, PhysicalTime of every DT, the crash marks and the
 *     first crash's dump (tests/test_oracle_golden.py).  The golden runs
 *     cover WALL and damped boundaries (ZEROVRAD, INIT, G), the LP85, BL94,
 *     ZHU12 and SM03 opacities, stellar irradiation, Flock alpha with the
 *     viscosity cavity, ImposedDiskDrift, gas accretion, the two-fluid disk
 *     with pebble accretion, back-reaction, diffusion and accretion heating;
 *   - closed forms, for what the golden runs leave open: the opacity laws'
 *     published regime exponents and coefficients and their continuity at the
 *     regime borders, POWERLAW's scaling and its equality with LP85's ice
 *     branch (bitwise, and as a whole run against the reference's LP85 run on
 *     a disk colder than the ice line), the flux limiter's limits
 *     (tests/test_oracle_math.py).
 * The options the library refuses (OpenInner and the others that
 * thorin_ctx_create rejects) are not restated here.
 */
#define _GNU_SOURCE
#include "thorin_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323844          /* fondam.h:12 */
#define G_ 1.0                              /* fondam.h:11 */
#define GASCONST 1.0                        /* fondam.h:17 */
#define MOLWEIGHT 1.0                       /* fondam.h:18 */
#define CFLSECURITY 0.5                     /* SourceEuler.c:16 */
#define CVNR 1.41                           /* SourceEuler.c:19 */
#define SORMAXITERS 100000                  /* EnergySources.c:22 */
#define SOREPS 1.0e-6                       /* EnergySources.c:23 */
#define SORMIN 1.0e-15                      /* EnergySources.c:24 */
#define SORPRECISION 1.0e-15                /* EnergySources.c:25 */
#define OMEGAUPPER 1.9                      /* EnergySources.c:27 */

/* unit constants, same expressions as fondam.h:28-39 (folded by gcc) */
#define MSOL_SI 1.98855e30
#define GM_SI 1.32712440018e20
#define AU_SI 149597870700.0
#define R_STANDARD 8.3144598
#define MMW 2.4
#define R_SI (R_STANDARD/(MMW*0.001))
#define OPA2CU (MSOL_SI*1000.0/pow(AU_SI*100.0,2.0))
#define T2SI (GM_SI/R_SI/AU_SI)
#define TIME2SI (sqrt(pow(AU_SI,3.0)/GM_SI))
#define RHO2CGS (MSOL_SI*1000.0/pow(AU_SI*100.0,3.0))
#define STEFANBOLTZMANN (5.670367e-8*(1.0/MSOL_SI)*pow(1.0/TIME2SI,-3.0)*pow(1.0/T2SI,-4.0))
#define SQRT2PI_INV (1.0/sqrt(2.0*PI))

enum { F_DENS = THORIN_F_DENS };
#define NF THORIN_NFIELDS
/* private work fields beyond the public ids */
enum { W_RHOSTAR = NF, W_RHOINT, W_TEMPERINT, W_DRR, W_DRP, W_DPP, W_RADMOMP, W_RADMOMM,
       W_THETAMOMP, W_THETAMOMM, W_WORK, W_QRSTAR, W_VTHETARES, W_GRADTR, W_GRADTT,
       W_GRADTM, W_DIFFC, W_DIFFR, W_DIFFT, W_VOLDENS, W_A, W_B, W_ML, W_MLIP, W_MLIM,
       W_MLJP, W_MLJM, W_RHS, W_ABS, W_ORD, W_TEMPSHIFT, W_DQ,
       /* pebble fluid (Pebbles.c:37-56, SourceEuler.c:155-158) */
       W_PAGR, W_PAGT, W_PACCR, W_PACCT, W_GACCR, W_GACCT, W_DRAGR, W_DRAGT, W_MACC, NW };

struct oracle_ctx {
  thorin_params p;
  int nr, ns;
  int sm_rows, sm_cols;                 /* OpacityTable SM03: the gas table (Opacity2.c:22-23) */
  int heatsrc_max, heatsrc_index[8];    /* accretion heating (global.h:21-28) */
  double heatsrc[8];
  double *sm_pgrid, *sm_tgrid, *sm_ross;
  double *fld[NW];
  double *Radii, *Rinf, *Rsup, *Rmed, *Surf, *InvRmed, *InvSurf, *InvDiffRmed, *InvDiffRsup;
  double *InvRinf, *OmegaInv, *Rmed2, *GlobalRmed, *SigmaMed, *SigmaInf, *EnergyMed;
  double *VthetaMed, *WaveKiller, *globcsvec, *globtempervec, *Vmoy, *Vresidual, *VMed;
  int *Nshift;
  double *crash[4];                     /* the first crash's Rho, Vrad, Vtheta, Temperature (SourceEuler.c:276-282) */
  double sqrt_adiabind_inv, CV, invdtreb_sq;
  thorin_state st;
  int opa_filled;
  double indirect_x, indirect_y;
  /* pebbles */
  double *PebDensInit, *PebVradInit, *PebVthetaInit, *vt1D;
  double pebbulkdens, invdtpeb_sq;
};

static double *fieldalloc(int nr, int ns) {
  /* LowTasks.c:85: (Nr+3)*(Ns+1)+5 doubles, rings 0..Nr zeroed */
  return (double *)calloc((size_t)(nr + 3) * (ns + 1) + 5, sizeof(double));
}

/* ---------------- geometry: FillPolar1DArrays (SourceEuler.c:68-131) ---------------- */
static void o_fill_1d(oracle_ctx *o) {
  int nr = o->nr, ns = o->ns, i;
  double drrsep = (o->p.rmax - o->p.rmin) / (double)nr;
  FILE *in;
  char name[2 * THORIN_STRLEN];
  snprintf(name, sizeof name, "%s%s", o->p.outputdir, "radii.dat");
  in = fopen(name, "r");
  if (in == NULL) {
    if (o->p.log_grid)
      for (i = 0; i <= nr; i++) o->Radii[i] = o->p.rmin * exp((double)i / (double)nr * log(o->p.rmax / o->p.rmin));
    else
      for (i = 0; i <= nr; i++) o->Radii[i] = o->p.rmin + drrsep * (double)(i);
  } else {
    float t;
    for (i = 0; i <= nr; i++) { if (fscanf(in, "%f", &t) != 1) t = 0.f; o->Radii[i] = (double)t; }
    fclose(in);
  }
  for (i = 0; i < nr; i++) {
    o->GlobalRmed[i] = 2.0 / 3.0 * (o->Radii[i + 1] * o->Radii[i + 1] * o->Radii[i + 1] - o->Radii[i] * o->Radii[i] * o->Radii[i]);
    o->GlobalRmed[i] = o->GlobalRmed[i] / (o->Radii[i + 1] * o->Radii[i + 1] - o->Radii[i] * o->Radii[i]);
  }
  for (i = 0; i < nr; i++) {
    o->Rinf[i] = o->Radii[i];
    o->Rsup[i] = o->Radii[i + 1];
    o->Rmed[i] = 2.0 / 3.0 * (o->Rsup[i] * o->Rsup[i] * o->Rsup[i] - o->Rinf[i] * o->Rinf[i] * o->Rinf[i]);
    o->Rmed[i] = o->Rmed[i] / (o->Rsup[i] * o->Rsup[i] - o->Rinf[i] * o->Rinf[i]);
    o->Surf[i] = PI * (o->Rsup[i] * o->Rsup[i] - o->Rinf[i] * o->Rinf[i]) / (double)ns;
    o->InvRmed[i] = 1.0 / o->Rmed[i];
    o->InvSurf[i] = 1.0 / o->Surf[i];
    o->InvDiffRsup[i] = 1.0 / (o->Rsup[i] - o->Rinf[i]);
    o->InvRinf[i] = 1.0 / o->Rinf[i];
    o->OmegaInv[i] = pow(o->Rmed[i], 1.5);
    o->Rmed2[i] = o->Rmed[i] * o->Rmed[i];
  }
  o->Rinf[nr] = o->Radii[nr];
  for (i = 1; i < nr; i++) o->InvDiffRmed[i] = 1.0 / (o->Rmed[i] - o->Rmed[i - 1]);
}

/* ---------------- 1D profiles (Theo.c, Viscosity.c) ---------------- */
/* FViscosity, Viscosity.c:26-54 (single rank: GlobalRmed == Rmed) */
static double o_fviscosity(oracle_ctx *o, double rad) {
  const thorin_params *p = &o->p;
  double viscosity, rmin, rmax, scale, alpha;
  int i = 0;
  viscosity = p->viscosity;
  if (p->viscosity_alpha) {
    while (o->GlobalRmed[i] < rad) i++;
    viscosity = p->alphaviscosity * o->globcsvec[i] * o->globcsvec[i] * pow(rad, 1.5);
  }
  if (p->alpha_flock_on) {
    while (o->GlobalRmed[i] < rad) i++;
    alpha = (p->alphamri - p->alphadead) * (1.0 - tanh((p->tmri - o->globtempervec[i] * T2SI) / p->twidth)) / 2.0 + p->alphadead;
    viscosity = alpha * o->globcsvec[i] * o->globcsvec[i] * pow(rad, 1.5);
  }
  rmin = p->cavityradius - p->cavitywidth * p->aspectratio;
  rmax = p->cavityradius + p->cavitywidth * p->aspectratio;
  scale = 1.0 + (o->st.physical_time - o->st.physical_time_initial) * p->lambdadoubling;
  rmin *= scale;
  rmax *= scale;
  if (rad < rmin) viscosity *= p->cavityratio;
  if ((rad >= rmin) && (rad <= rmax)) viscosity *= exp((rmax - rad) / (rmax - rmin) * log(p->cavityratio));
  return viscosity;
}

/* AspectRatio, Viscosity.c:56-71 */
static double o_aspectratio(oracle_ctx *o, double rad) {
  const thorin_params *p = &o->p;
  double ar = p->aspectratio, rmin, rmax, scale;
  rmin = p->transitionradius - p->transitionwidth * p->aspectratio;
  rmax = p->transitionradius + p->transitionwidth * p->aspectratio;
  scale = 1.0 + (o->st.physical_time - o->st.physical_time_initial) * p->lambdadoubling;
  rmin *= scale;
  rmax *= scale;
  if (rad < rmin) ar *= p->transitionratio;
  if ((rad >= rmin) && (rad <= rmax)) ar *= exp((rmax - rad) / (rmax - rmin) * log(p->transitionratio));
  return ar;
}

/* mpi_make1Dprofile single-rank path (mpiTasks.c:12-36): sequential ring mean */
static void o_ring_mean(oracle_ctx *o, const double *f, double *out) {
  int i, j;
  for (i = 0; i < o->nr; i++) {
    double s = 0.;
    for (j = 0; j < o->ns; j++) s += f[i * o->ns + j];
    s /= (double)o->ns;
    out[i] = s;
  }
}

/* ---------------- EOS fields (SourceEuler.c:715-792) ---------------- */
static void o_sound_speed(oracle_ctx *o, const double *rho, const double *energy) {
  int i, j, l, ns = o->ns;
  double *cs = o->fld[THORIN_F_SOUNDSPEED];
  double g = o->p.adiabind;
  for (i = 0; i < o->nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      if (!o->p.energy_eq)
        cs[l] = o_aspectratio(o, o->Rmed[i]) * sqrt(G_ * 1.0 / o->Rmed[i]) * pow(o->Rmed[i], o->p.flaringindex);
      else
        cs[l] = sqrt(g * (g - 1.0) * fabs(energy[l]) / rho[l]);
    }
  o_ring_mean(o, cs, o->globcsvec);
}

static void o_pressure(oracle_ctx *o, const double *rho, const double *energy) {
  int i, j, l, ns = o->ns;
  double *pr = o->fld[THORIN_F_PRESSURE], *cs = o->fld[THORIN_F_SOUNDSPEED];
  for (i = 0; i < o->nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      if (!o->p.energy_eq) pr[l] = rho[l] * cs[l] * cs[l];
      else pr[l] = (o->p.adiabind - 1.0) * energy[l];
    }
}

static void o_temperature(oracle_ctx *o, const double *rho, const double *energy) {
  int i, j, l, ns = o->ns;
  double *t = o->fld[THORIN_F_TEMPER], *pr = o->fld[THORIN_F_PRESSURE];
  double tmp1, tmp2;
  for (i = 0; i < o->nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      tmp2 = 1.0 / (rho[l] * GASCONST);
      if (o->p.energy_eq) tmp1 = energy[l] * MOLWEIGHT * (o->p.adiabind - 1.0);
      else tmp1 = MOLWEIGHT * pr[l];
      t[l] = tmp1 * tmp2;
    }
  o_ring_mean(o, t, o->globtempervec);
}

/* ---------------- creation / initialisation ---------------- */
int oracle_create(const thorin_params *p, oracle_ctx **out) {
  oracle_ctx *o = (oracle_ctx *)calloc(1, sizeof(oracle_ctx));
  int k, n;
  if (!o) return -1;
  o->p = *p;
  o->nr = p->nrad;
  o->ns = p->nsec;
  if (o->nr < 4 || o->ns < 2) { free(o); return -2; }
  for (k = 0; k < NW; k++) o->fld[k] = fieldalloc(o->nr, o->ns);
  n = o->nr + 2;
#define A1(x) o->x = (double *)calloc((size_t)n + 1, sizeof(double))
  A1(Radii); A1(Rinf); A1(Rsup); A1(Rmed); A1(Surf); A1(InvRmed); A1(InvSurf); A1(InvDiffRmed);
  A1(InvDiffRsup); A1(InvRinf); A1(OmegaInv); A1(Rmed2); A1(GlobalRmed); A1(SigmaMed); A1(SigmaInf);
  A1(EnergyMed); A1(VthetaMed); A1(WaveKiller); A1(globcsvec); A1(globtempervec); A1(Vmoy);
  A1(VMed); A1(PebDensInit); A1(PebVradInit); A1(PebVthetaInit); A1(vt1D);
#undef A1
  o->pebbulkdens = p->pebblebulkdens / RHO2CGS;   /* Pebbles.c:55 */
  o->Vresidual = (double *)calloc((size_t)o->ns + 1, sizeof(double));
  o->Nshift = (int *)calloc((size_t)n, sizeof(int));
  o->sqrt_adiabind_inv = 1.0 / sqrt(p->adiabind);
  o->CV = GASCONST / (MOLWEIGHT * (p->adiabind - 1.0));
  o->st.dt_stellar = 1.0e38;
  o->st.omega_frame = p->omegaframe;
  o->st.sor_first = 1;
  o->st.mass_taper = 0.0;
  o_fill_1d(o);
  /* InitComputeAccel, SideEuler.c:101-118 */
  {
    int i, j, l;
    double *ab = o->fld[W_ABS], *od = o->fld[W_ORD];
    for (i = 0; i < o->nr; i++)
      for (j = 0; j < o->ns; j++) {
        l = j + i * o->ns;
        ab[l] = o->Rmed[i] * cos(2.0 * PI * (double)j / (double)o->ns);
        od[l] = o->Rmed[i] * sin(2.0 * PI * (double)j / (double)o->ns);
      }
  }
  /* SetWaveKillingZones, SideEuler.c:341-362 */
  if (p->damping_on || p->pebbles) {
    double DRin = p->rmin * p->dampingrminfrac, DRout = p->rmax * p->dampingrmaxfrac, damp;
    double tauin = p->dampingperiodfrac * 2.0 * PI * pow(p->rmin, 1.5);
    double tauout = p->dampingperiodfrac * 2.0 * PI * pow(DRout, 1.5);
    int i;
    for (i = 0; i < o->nr; i++) {
      o->WaveKiller[i] = 0.0;
      if (o->Rmed[i] < DRin) { damp = (DRin - o->Rmed[i]) / (DRin - p->rmin); o->WaveKiller[i] = damp * damp / tauin; }
      if (o->Rmed[i] > DRout) { damp = (o->Rmed[i] - DRout) / (p->rmax - DRout); o->WaveKiller[i] = damp / tauout; }
    }
  }
  /* InitOpacities (Opacity2.c:252-300): the gas table, Rosseland part */
  if (p->opacity_law == THORIN_OPA_SM03) {
    char name[2 * THORIN_STRLEN];
    FILE *in;
    int i, ok = 1;
    double *planck;
    snprintf(name, sizeof name, "%s%s", p->inputdir, "data_gasopacity_malygin.txt");
    in = fopen(name, "r");
    if (!in) { oracle_destroy(o); return -3; }
    ok = fscanf(in, "%i", &o->sm_rows) == 1 && fscanf(in, "%i", &o->sm_cols) == 1 && o->sm_rows > 1 && o->sm_cols > 2;
    if (ok) {
      o->sm_pgrid = (double *)malloc(sizeof(double) * o->sm_rows);
      o->sm_tgrid = (double *)malloc(sizeof(double) * o->sm_cols);
      o->sm_ross = (double *)malloc(sizeof(double) * o->sm_rows * o->sm_cols);
      planck = (double *)malloc(sizeof(double) * o->sm_rows * o->sm_cols);
      for (i = 0; ok && i < o->sm_rows; i++) ok = fscanf(in, "%lf", &o->sm_pgrid[i]) == 1;
      for (i = 0; ok && i < o->sm_cols; i++) ok = fscanf(in, "%lf", &o->sm_tgrid[i]) == 1;
      for (i = 0; ok && i < o->sm_rows * o->sm_cols; i++) ok = fscanf(in, "%lf", &planck[i]) == 1;
      for (i = 0; ok && i < o->sm_rows * o->sm_cols; i++) ok = fscanf(in, "%lf", &o->sm_ross[i]) == 1;
      free(planck);
    }
    fclose(in);
    if (!ok) { oracle_destroy(o); return -3; }
  }
  /* InitLabel, Init.c:56-70 */
  {
    int i, j;
    for (i = 0; i <= o->nr; i++)
      for (j = 0; j < o->ns; j++)
        o->fld[THORIN_F_LABEL][j + i * o->ns] = (o->Rmed[i] - p->rmin) / (p->rmax - p->rmin);
  }
  *out = o;
  return 0;
}

void oracle_destroy(oracle_ctx *o) {
  int k;
  if (!o) return;
  for (k = 0; k < NW; k++) free(o->fld[k]);
  for (k = 0; k < 4; k++) free(o->crash[k]);
  free(o->sm_pgrid); free(o->sm_tgrid); free(o->sm_ross);
  free(o->Radii); free(o->Rinf); free(o->Rsup); free(o->Rmed); free(o->Surf); free(o->InvRmed);
  free(o->InvSurf); free(o->InvDiffRmed); free(o->InvDiffRsup); free(o->InvRinf); free(o->OmegaInv);
  free(o->Rmed2); free(o->GlobalRmed); free(o->SigmaMed); free(o->SigmaInf); free(o->EnergyMed);
  free(o->VthetaMed); free(o->WaveKiller); free(o->globcsvec); free(o->globtempervec); free(o->Vmoy);
  free(o->VMed); free(o->Vresidual); free(o->Nshift);
  free(o->PebDensInit); free(o->PebVradInit); free(o->PebVthetaInit); free(o->vt1D);
  free(o);
}

/* InitializeFromFile route: Init.c:86-108 then cs, P, T (Init.c:131-133).
 * RefillSigma / RefillEnergy / FillVtheta follow Theo.c:33-57, 82-100, 132-150. */
int oracle_init_from_arrays(oracle_ctx *o, const double *dens, const double *vrad,
                            const double *vtheta, const double *temper) {
  int nr = o->nr, ns = o->ns, i, j, l;
  double *rho = o->fld[THORIN_F_DENS], *e = o->fld[THORIN_F_ENERGY];
  memcpy(rho, dens, sizeof(double) * nr * ns);
  memcpy(o->fld[THORIN_F_VRAD], vrad, sizeof(double) * nr * ns);
  memcpy(o->fld[THORIN_F_VTHETA], vtheta, sizeof(double) * nr * ns);
  if (o->p.energy_eq) {
    memcpy(e, temper, sizeof(double) * nr * ns);
    for (i = 0; i < nr; i++)
      for (j = 0; j < ns; j++) { l = j + i * ns; e[l] = e[l] * rho[l] / (o->p.adiabind - 1.0); }
  }
  for (i = 0; i < nr; i++) {
    double moy = 0.0;
    for (j = 0; j < ns; j++) moy += rho[j + i * ns];
    moy /= (double)ns;
    o->SigmaMed[i] = moy;
  }
  o->SigmaInf[0] = o->SigmaMed[0];
  for (i = 1; i < nr; i++)
    o->SigmaInf[i] = (o->SigmaMed[i - 1] * (o->Rmed[i] - o->Rinf[i]) + o->SigmaMed[i] * (o->Rinf[i] - o->Rmed[i - 1])) / (o->Rmed[i] - o->Rmed[i - 1]);
  if (o->p.energy_eq)
    for (i = 0; i < nr; i++) {
      double moy = 0.0;
      for (j = 0; j < ns; j++) moy += e[j + i * ns];
      moy /= (double)ns;
      o->EnergyMed[i] = moy;
    }
  if (o->p.damp_init)
    for (i = 0; i < nr; i++) {
      double moy = 0.0, vtcorr = o->Rmed[i] * o->st.omega_frame;
      for (j = 0; j < ns; j++) moy += o->fld[THORIN_F_VTHETA][j + i * ns] + vtcorr;
      moy /= (double)ns;
      o->VthetaMed[i] = moy;
    }
  o_sound_speed(o, rho, e);
  o_pressure(o, rho, e);
  o_temperature(o, rho, e);
  /* main.c:168-170 recomputes the sound speed once more */
  if (o->p.energy_eq) o_sound_speed(o, rho, e);
  return 0;
}

/* ---------------- ConditionCFL (SourceEuler.c:589-712) ---------------- */
static double o_min2(double a, double b) { return (b < a) ? b : a; }   /* SourceEuler.c:177-182 */
static double o_max2(double a, double b) { return (b > a) ? b : a; }   /* SourceEuler.c:184-189 */

static int o_condition_cfl_core(oracle_ctx *o, double deltaT, double *newdt_out) {
  int i, j, l, ns = o->ns, nr = o->nr, lip, ljp;
  double invdt1, invdt2, invdt3, invdt4, invdt5 = 1e-30, cs, newdt, dt;
  double dxrad, dxtheta, dvr, dvt, visc = 0.;
  const double *vt = o->fld[THORIN_F_VTHETA], *vr = o->fld[THORIN_F_VRAD];
  const double *soundspeed = o->fld[THORIN_F_SOUNDSPEED];
  newdt = 1e30;
  for (i = 0; i < nr; i++) {
    o->Vmoy[i] = 0.0;
    for (j = 0; j < ns; j++) o->Vmoy[i] += vt[j + i * ns];
    o->Vmoy[i] /= (double)ns;
  }
  for (i = 1; i < nr; i++) {
    dxrad = o->Rsup[i] - o->Rinf[i];
    dxtheta = o->Rmed[i] * 2.0 * PI / (double)ns;
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      o->Vresidual[j] = o->p.fast_transport ? vt[l] - o->Vmoy[i] : vt[l];
    }
    o->Vresidual[ns] = o->Vresidual[0];
    if (o->p.viscosity_alpha || (o->p.viscosity != 0.0)) visc = o_fviscosity(o, o->Rmed[i]);
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      lip = l + ns;
      ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      cs = soundspeed[l];
      invdt1 = cs / (o_min2(dxrad, dxtheta));
      invdt2 = fabs(vr[l]) / dxrad;
      invdt3 = fabs(o->Vresidual[j]) / dxtheta;
      dvr = vr[lip] - vr[l];
      dvt = vt[ljp] - vt[l];
      if (dvr >= 0.0) dvr = 1e-10; else dvr = -dvr;
      if (dvt >= 0.0) dvt = 1e-10; else dvt = -dvt;
      invdt4 = o_max2(dvr / dxrad, dvt / dxtheta);
      invdt4 *= 4.0 * CVNR * CVNR;
      if (o->p.viscosity_alpha || (o->p.viscosity != 0.0))
        invdt5 = visc * 4. / pow(o_min2(dxrad, dxtheta), 2);
      if (!o->p.pebbles)
        dt = CFLSECURITY / sqrt(invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4 + invdt5 * invdt5 + o->invdtreb_sq);
      else
        dt = CFLSECURITY / sqrt(invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4 + invdt5 * invdt5 +
                                o->invdtpeb_sq + o->invdtreb_sq);
      if (dt < newdt) newdt = dt;
    }
  }
  for (i = 0; i < nr - 1; i++) {
    dt = 2.0 * PI * CFLSECURITY / (double)ns / fabs(o->Vmoy[i] * o->InvRmed[i] - o->Vmoy[i + 1] * o->InvRmed[i + 1]);
    if (dt < newdt) newdt = dt;
  }
  if (newdt > o->st.dt_stellar) {
    double a = 1.0 * o->st.dt_stellar, b = o->p.minimumstep;
    double mx = a > b ? a : b;
    newdt = newdt < mx ? newdt : mx;
  }
  newdt = newdt < o->p.maximumstep ? newdt : o->p.maximumstep;
  if (o->st.cfl_count == 0) {
    if (o->p.firststep > 0.0) newdt = o->p.firststep;
    o->st.cfl_count = 1;
  }
  if (deltaT < newdt) newdt = deltaT;
  *newdt_out = newdt;
  /* a non-finite or absurd step count is refused (the reference would loop) */
  if (!(newdt > 0.0) || !(deltaT / newdt <= (double)THORIN_MAX_CFL_N)) return -1;
  return (int)(ceil(deltaT / newdt));
}

int oracle_condition_cfl(oracle_ctx *o, double deltaT, double invdtreb_sq, double *newdt) {
  o->invdtreb_sq = invdtreb_sq;
  return o_condition_cfl_core(o, deltaT, newdt);
}

/* MinStepForRebound, ReboundInterface.c:360-385 (pairs of planets only) */
static double o_minstep_rebound(const thorin_planets *sys) {
  int i, j;
  double dt = 1.e30, dtnew, r2, rv;
  for (i = 0; i < sys->nb; i++)
    for (j = i + 1; j < sys->nb; j++) {
      r2 = (sys->x[i] - sys->x[j]) * (sys->x[i] - sys->x[j]) + (sys->y[i] - sys->y[j]) * (sys->y[i] - sys->y[j]) +
           (sys->z[i] - sys->z[j]) * (sys->z[i] - sys->z[j]);
      rv = (sys->x[i] - sys->x[j]) * (sys->vx[i] - sys->vx[j]) + (sys->y[i] - sys->y[j]) * (sys->vy[i] - sys->vy[j]) +
           (sys->z[i] - sys->z[j]) * (sys->vz[i] - sys->vz[j]);
      dtnew = fabs(r2 / rv);
      if (dtnew < dt) dt = dtnew;
    }
  dt *= 0.25;
  dt *= dt;
  return 1.0 / dt;
}

/* ---------------- forces: ThicknessSmoothing + FillForcesArrays ---------------- */
/* ThicknessSmoothing, Force.c:75-102 */
static double o_thickness_smoothing(oracle_ctx *o, double x, double y) {
  int i, j;
  double r = sqrt(x * x + y * y), csavrg, H, sm = 0.0;
  const double *cs = o->fld[THORIN_F_SOUNDSPEED];
  if (r >= o->Rinf[0] && r <= o->Rsup[o->nr - 1]) {
    i = 0;
    while (o->Rsup[i] < r) i++;
    csavrg = 0.0;
    for (j = 0; j < o->ns; j++) csavrg += cs[j + i * o->ns];
    csavrg /= (double)o->ns;
    H = csavrg * o->OmegaInv[i] * o->sqrt_adiabind_inv;
    sm = o->p.thicknesssmoothing * H;
  }
  return sm;   /* MPI_MAX over one rank */
}

/* FillForcesArrays, Pframeforce.c:43-510 (gas part; the pebble branches are
 * added with the pebble fluid). */
static void o_fill_forces(oracle_ctx *o, thorin_planets *sys) {
  int i, j, l, nr = o->nr, ns = o->ns, k, npl = sys->nb, m, Nvert;
  double *agr = o->fld[THORIN_F_AGR], *agt = o->fld[THORIN_F_AGT];
  const double *rho = o->fld[THORIN_F_DENS], *cs = o->fld[THORIN_F_SOUNDSPEED];
  const double *abs_ = o->fld[W_ABS], *ord = o->fld[W_ORD];
  double xpl, ypl, zpl, mpl, dpl, invdpl3, smoothing;
  double d2, d, H, H2, zmax, deltaz;
  double rsm[THORIN_MAXPLANETS], rsm2[THORIN_MAXPLANETS], sigma, znode, znode2, denom, ax, ay, ar, at;
  double rst, rst2[THORIN_MAXPLANETS], rst3[THORIN_MAXPLANETS], rst4[THORIN_MAXPLANETS];
  double taper, omegainv, r2, Rsup_Rinf2, r2chk_1, r2chk_2, integstar, sum, tmp;
  double cs_avr, integstar_avr, sigma_avr, H_avr, x, y, mcell, axindir, ayindir;
  double indir_px, indir_py, axstar, aystar;
  int IntegrateWithPlanet[THORIN_MAXPLANETS], IntegrateLocally;
  double hillcut[THORIN_MAXPLANETS], rhill[THORIN_MAXPLANETS], smoothing2[THORIN_MAXPLANETS];
  double Hpl[THORIN_MAXPLANETS], s2[THORIN_MAXPLANETS], integpl[THORIN_MAXPLANETS];
  double rmorevert2[THORIN_MAXPLANETS], axpl[THORIN_MAXPLANETS], aypl[THORIN_MAXPLANETS];
  const double mt = o->st.mass_taper;
  /* pebbles (Pframeforce.c:56-57, 73-78, 150-214, 250-330, 342-441, 451-455, 504-508) */
  const int peb = o->p.pebbles, pgrav = o->p.pebble_gravity;
  double *pagr = o->fld[W_PAGR], *pagt = o->fld[W_PAGT];
  const double *prho = o->fld[THORIN_F_PEBBLEDENS], *tau = o->fld[THORIN_F_PEBBLESTOKES];
  const double palpha = o->p.pebblealpha;
  double tau_avr = 0.0, p_H_avr = 0.0, p_integstar_avr = 0.0, p_integstar = 0.0, pax = 0.0, pay = 0.0;
  double pH = 0.0, Hcorr = 0.0, pmcell = 0.0, p_integpl[THORIN_MAXPLANETS];
  for (i = 0; i < (nr + 1) * ns; i++) {
    agr[i] = 0.0; agt[i] = 0.0;
    if (peb) { pagr[i] = 0.0; pagt[i] = 0.0; }
  }
  axstar = 0.0;
  aystar = 0.0;
  for (k = 0; k < npl; k++) {
    axpl[k] = 0.0; aypl[k] = 0.0; hillcut[k] = 1.0;
    xpl = sys->x[k]; ypl = sys->y[k]; zpl = sys->z[k]; mpl = sys->mass[k];
    dpl = sqrt(xpl * xpl + ypl * ypl + zpl * zpl);
    rhill[k] = dpl * pow((1.0 / 3.0 * mpl), 1.0 / 3.0);
    smoothing = o_thickness_smoothing(o, xpl, ypl);
    Hpl[k] = smoothing / o->p.thicknesssmoothing;
    smoothing2[k] = smoothing * smoothing;
    rst = 0.5 * rhill[k];
    rst2[k] = rst * rst;
    rst3[k] = rst2[k] * rst;
    rst4[k] = rst3[k] * rst;
    rsm[k] = 0.05 * rhill[k];
    rsm2[k] = rsm[k] * rsm[k];
    rmorevert2[k] = 2.0 * rhill[k];
    rmorevert2[k] *= rmorevert2[k];
    if (o->p.indirect_term) {
      invdpl3 = 1.0 / (dpl * dpl * dpl);
      mpl *= mt;
      axstar += mpl * invdpl3 * xpl;
      aystar += mpl * invdpl3 * ypl;
    }
  }
  indir_px = -axstar;
  indir_py = -aystar;
  axstar = 0.0;
  aystar = 0.0;
  for (i = 0; i < nr; i++) {
    omegainv = o->OmegaInv[i];
    r2 = o->Rmed2[i];
    Rsup_Rinf2 = o->Rsup[i] - o->Rinf[i];
    Rsup_Rinf2 *= Rsup_Rinf2;
    Nvert = 10;
    cs_avr = 0.0;
    if (peb) tau_avr = 0.0;
    for (j = 0; j < ns; j++) {
      cs_avr += cs[j + i * ns];
      if (peb) tau_avr += tau[j + i * ns];
    }
    cs_avr /= (double)ns;
    H_avr = cs_avr * omegainv * o->sqrt_adiabind_inv;
    if (peb) {
      tau_avr /= (double)ns;
      p_H_avr = H_avr * sqrt(palpha / tau_avr);
      p_integstar_avr = 0.0;
    }
    H2 = H_avr * H_avr;
    zmax = 3.0 * H_avr;
    deltaz = zmax / (double)Nvert;
    sigma_avr = 0.0;
    integstar_avr = 0.0;
    for (m = 1; m <= Nvert; m++) {
      znode = ((double)m - 0.5) * deltaz;
      znode2 = znode * znode;
      sum = exp(-(0.5 * znode2 / H2));
      sigma_avr += sum;
      d2 = r2 + znode2;
      d = sqrt(d2);
      denom = d * d2;
      integstar_avr += sum / denom;
      if (peb) {
        znode2 *= palpha / tau_avr;
        d2 = r2 + znode2;
        d = sqrt(d2);
        denom = d * d2;
        p_integstar_avr += sum / denom;
      }
    }
    for (j = 0; j < ns; j++) {
      IntegrateLocally = 0;
      Nvert = 10;
      ax = 0.0;
      ay = 0.0;
      if (peb) { pax = 0.0; pay = 0.0; }
      l = j + i * ns;
      x = abs_[l];
      y = ord[l];
      mcell = rho[l] * o->Surf[i];
      H = cs[l] * omegainv * o->sqrt_adiabind_inv;
      H2 = H * H;
      if (fabs(H - H_avr) / H_avr > 0.05) IntegrateLocally = 1;
      if (peb) {
        pmcell = prho[l] * o->Surf[i];
        Hcorr = sqrt(palpha / tau[l]);
        pH = H * Hcorr;
        if (fabs(pH - p_H_avr) / p_H_avr > 0.05) IntegrateLocally = 1;
      }
      for (k = 0; k < npl; k++) {
        xpl = sys->x[k]; ypl = sys->y[k]; zpl = sys->z[k];
        s2[k] = (x - xpl) * (x - xpl) + (y - ypl) * (y - ypl);
        r2chk_1 = 10.0 * Hpl[k];
        r2chk_1 *= r2chk_1;
        r2chk_2 = 5.0 * rhill[k];
        r2chk_2 *= r2chk_2;
        if (mt == 1.0 && (s2[k] < r2chk_1 || s2[k] < r2chk_2)) {
          IntegrateLocally = 1;
          IntegrateWithPlanet[k] = 1;
          integpl[k] = 0.0;
          if (peb) p_integpl[k] = 0.0;
          d2 = s2[k] + zpl * zpl;
          if (o->p.exclude_hill) hillcut[k] = 1.0 / (exp(-(sqrt(d2) / rhill[k] - o->p.hillcut) / o->p.hillcut * 10.0) + 1.0);
        } else {
          IntegrateWithPlanet[k] = 0;
        }
        if (s2[k] < rmorevert2[k]) Nvert = 30;
      }
      if (IntegrateLocally) {
        zmax = 3.0 * H;
        deltaz = zmax / (double)Nvert;
        sigma = 0.0;
        integstar = 0.0;
        if (peb) p_integstar = 0.0;
        for (m = 1; m <= Nvert; m++) {
          znode = ((double)m - 0.5) * deltaz;
          znode2 = znode * znode;
          sum = exp(-(0.5 * znode2 / H2));
          sigma += sum;
          d2 = r2 + znode2;
          d = sqrt(d2);
          denom = d * d2;
          integstar += sum / denom;
          if (peb) {
            d2 = r2 + znode2 * Hcorr * Hcorr;
            d = sqrt(d2);
            denom = d * d2;
            p_integstar += sum / denom;
          }
          for (k = 0; k < npl; k++) {
            if (IntegrateWithPlanet[k]) {
              zpl = sys->z[k];
              d2 = s2[k] + (znode - zpl) * (znode - zpl);
              if (d2 < rst2[k]) {
                d = sqrt(d2);
                taper = -3.0 * d2 * d2 / rst4[k] + 4.0 * d * d2 / rst3[k];
              } else taper = 1.0;
              d2 += (rsm2[k] > Rsup_Rinf2 ? rsm2[k] : Rsup_Rinf2);
              d = sqrt(d2);
              denom = d * d2;
              integpl[k] += sum / denom * taper;
              if (zpl < -0.05 * H || zpl > 0.05 * H) {
                d2 = s2[k] + (-znode - zpl) * (-znode - zpl);
                if (d2 < rst2[k]) {
                  d = sqrt(d2);
                  taper = -3.0 * d2 * d2 / rst4[k] + 4.0 * d * d2 / rst3[k];
                } else taper = 1.0;
                d2 += (rsm2[k] > Rsup_Rinf2 ? rsm2[k] : Rsup_Rinf2);
                d = sqrt(d2);
                denom = d * d2;
                integpl[k] += sum / denom * taper;
              }
              if (peb) {
                d2 = s2[k] + (znode * Hcorr - zpl) * (znode * Hcorr - zpl);
                if (d2 < rst2[k]) {
                  d = sqrt(d2);
                  taper = -3.0 * d2 * d2 / rst4[k] + 4.0 * d * d2 / rst3[k];
                } else taper = 1.0;
                d2 += (rsm2[k] > Rsup_Rinf2 ? rsm2[k] : Rsup_Rinf2);
                d = sqrt(d2);
                denom = d * d2;
                p_integpl[k] += sum / denom * taper;
                if (zpl < -0.05 * pH || zpl > 0.05 * pH) {
                  d2 = s2[k] + (-znode * Hcorr - zpl) * (-znode * Hcorr - zpl);
                  if (d2 < rst2[k]) {
                    d = sqrt(d2);
                    taper = -3.0 * d2 * d2 / rst4[k] + 4.0 * d * d2 / rst3[k];
                  } else taper = 1.0;
                  d2 += (rsm2[k] > Rsup_Rinf2 ? rsm2[k] : Rsup_Rinf2);
                  d = sqrt(d2);
                  denom = d * d2;
                  p_integpl[k] += sum / denom * taper;
                }
              }
            }
          }
        }
      } else {
        sigma = sigma_avr;
        integstar = integstar_avr;
        if (peb) p_integstar = p_integstar_avr;
      }
      tmp = integstar / sigma;
      ar = -o->Rmed[i] * tmp;
      agr[l] += ar;
      axstar += mcell * tmp * x;
      aystar += mcell * tmp * y;
      if (peb) {
        tmp = p_integstar / sigma;
        ar = -o->Rmed[i] * tmp;
        pagr[l] += ar;
      }
      for (k = 0; k < npl; k++) {
        xpl = sys->x[k]; ypl = sys->y[k]; zpl = sys->z[k]; mpl = sys->mass[k];
        if (IntegrateWithPlanet[k]) {
          if (zpl < -0.05 * H || zpl > 0.05 * H) tmp = (x - xpl) * integpl[k] / (2.0 * sigma);
          else tmp = (x - xpl) * integpl[k] / sigma;
          ax += -mpl * tmp;
          if (o->p.exclude_hill) tmp *= hillcut[k];
          axpl[k] += mcell * tmp;
          if (zpl < -0.05 * H || zpl > 0.05 * H) tmp = (y - ypl) * integpl[k] / (2.0 * sigma);
          else tmp = (y - ypl) * integpl[k] / sigma;
          ay += -mpl * tmp;
          if (o->p.exclude_hill) tmp *= hillcut[k];
          aypl[k] += mcell * tmp;
          if (peb) {
            if (zpl < -0.05 * pH || zpl > 0.05 * pH) tmp = (x - xpl) * p_integpl[k] / (2.0 * sigma);
            else tmp = (x - xpl) * p_integpl[k] / sigma;
            pax += -mpl * tmp;
            if (pgrav) {
              if (o->p.exclude_hill) tmp *= hillcut[k];
              axpl[k] += pmcell * tmp;
            }
            if (zpl < -0.05 * pH || zpl > 0.05 * pH) tmp = (y - ypl) * p_integpl[k] / (2.0 * sigma);
            else tmp = (y - ypl) * p_integpl[k] / sigma;
            pay += -mpl * tmp;
            if (pgrav) {
              if (o->p.exclude_hill) tmp *= hillcut[k];
              aypl[k] += pmcell * tmp;
            }
          }
        } else {
          mpl *= mt;
          d2 = s2[k] + zpl * zpl + smoothing2[k];
          d = sqrt(d2);
          denom = d * d2;
          tmp = (x - xpl) / denom;
          ax += -mpl * tmp;
          if (o->p.exclude_hill) tmp *= hillcut[k];
          axpl[k] += mcell * tmp;
          tmp = (y - ypl) / denom;
          ay += -mpl * tmp;
          if (o->p.exclude_hill) tmp *= hillcut[k];
          aypl[k] += mcell * tmp;
          if (peb) {
            d2 += -smoothing2[k] + smoothing2[k] * Hcorr * Hcorr;
            d = sqrt(d2);
            denom = d * d2;
            tmp = (x - xpl) / denom;
            pax += -mpl * tmp;
            if (pgrav) {
              if (o->p.exclude_hill) tmp *= hillcut[k];
              axpl[k] += pmcell * tmp;
            }
            tmp = (y - ypl) / denom;
            pay += -mpl * tmp;
            if (pgrav) {
              if (o->p.exclude_hill) tmp *= hillcut[k];
              aypl[k] += pmcell * tmp;
            }
          }
        }
      }
      ar = (x * ax + y * ay) * o->InvRmed[i];
      at = (x * ay - y * ax) * o->InvRmed[i];
      agr[l] += ar;
      agt[l] += at;
      if (peb) {
        ar = (x * pax + y * pay) * o->InvRmed[i];
        at = (x * pay - y * pax) * o->InvRmed[i];
        pagr[l] += ar;
        pagt[l] += at;
      }
    }
  }
  for (k = 0; k < npl; k++) { sys->ax[k] = axpl[k]; sys->ay[k] = aypl[k]; }
  axindir = -axstar;
  ayindir = -aystar;
  if (o->p.indirect_term) { axindir += indir_px; ayindir += indir_py; }
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      x = abs_[l];
      y = ord[l];
      ar = (x * axindir + y * ayindir) * o->InvRmed[i];
      at = (x * ayindir - y * axindir) * o->InvRmed[i];
      agr[l] += ar;
      agt[l] += at;
      if (peb) { pagr[l] += ar; pagt[l] += at; }
    }
  /* IndirectTerm (Pframeforce.c:468-470), used by AdvanceSystemFromDisk */
  o->indirect_x = -axstar;
  o->indirect_y = -aystar;
}

/* DampingTW04, Pframeforce.c:571-603 */
static double o_damping_tw04(oracle_ctx *o, double m, double x, double y, double z, double vz) {
  double r, Omega, cs4inv, damp = 0.0, rhoavr = 0.0, csavr = 0.0;
  const double *rho = o->fld[THORIN_F_DENS], *cs = o->fld[THORIN_F_SOUNDSPEED];
  int i, j;
  r = sqrt(x * x + y * y);
  if ((o->p.verticaldamping > 0.0) && (r >= o->Rinf[0] && r <= o->Rsup[o->nr - 1])) {
    i = 0;
    while (o->Rsup[i] < r) i++;
    for (j = 0; j < o->ns; j++) { rhoavr += rho[j + i * o->ns]; csavr += cs[j + i * o->ns]; }
    csavr /= (double)o->ns;
    rhoavr /= (double)o->ns;
    cs4inv = pow(csavr, -4.0);
    Omega = pow(r, -1.5);
    damp = o->p.verticaldamping * m * rhoavr * Omega * cs4inv * (-2.176 * vz - 0.871 * z * Omega);
  }
  return damp;
}

/* AdvanceSystemFromDisk, Pframeforce.c:514-566 (torque log is host output) */
static void o_advance_from_disk(oracle_ctx *o, thorin_planets *sys, double dt) {
  int k;
  for (k = 0; k < sys->nb; k++) {
    if (sys->feeldisk[k]) {
      double damp = o_damping_tw04(o, sys->mass[k], sys->x[k], sys->y[k], sys->z[k], sys->vz[k]);
      sys->vx[k] += dt * sys->ax[k];
      sys->vy[k] += dt * sys->ay[k];
      sys->vz[k] += dt * (sys->az[k] + damp);
      sys->vx[k] += dt * o->indirect_x;
      sys->vy[k] += dt * o->indirect_y;
    }
  }
}

/* ---------------- boundaries (SideEuler.c:368-419, 490-503) ---------------- */
static void o_boundary(oracle_ctx *o, double step) {
  const thorin_params *p = &o->p;
  int i, j, l, ns = o->ns;
  double *vrad = o->fld[THORIN_F_VRAD], *vtheta = o->fld[THORIN_F_VTHETA];
  double *rho = o->fld[THORIN_F_DENS], *energy = o->fld[THORIN_F_ENERGY];
  double vrad0, vtheta0 = 0.0, rho0 = 0.0, energy0 = 0.0, DRin, DRout, lambda;
  if (!p->damping_on) return;
  DRin = p->rmin * p->dampingrminfrac;
  DRout = p->rmax * p->dampingrmaxfrac;
  for (i = 0; i < o->nr; i++) {
    if (o->Rmed[i] < DRin || o->Rmed[i] > DRout) {
      vrad0 = 0.0;
      if (p->damp_init) {
        rho0 = o->SigmaMed[i];
        energy0 = o->EnergyMed[i];
        vtheta0 = o->VthetaMed[i] - o->Rmed[i] * o->st.omega_frame;
      }
      if (p->damp_flux) {
        vrad0 = -p->gasflux / (2.0 * PI * o->Rinf[i] * o->SigmaInf[i]);
        rho0 = o->SigmaMed[i];
      }
      lambda = o->WaveKiller[i] * step;
      for (j = 0; j < ns; j++) {
        l = j + i * ns;
        vrad[l] = (vrad[l] + lambda * vrad0) / (1.0 + lambda);
        if (p->damp_init) {
          rho[l] = (rho[l] + lambda * rho0) / (1.0 + lambda);
          vtheta[l] = (vtheta[l] + lambda * vtheta0) / (1.0 + lambda);
          if (p->energy_eq) energy[l] = (energy[l] + lambda * energy0) / (1.0 + lambda);
        }
        if (p->damp_flux) rho[l] = (rho[l] + lambda * rho0) / (1.0 + lambda);
      }
    }
  }
}

static int o_detect_crash(oracle_ctx *o, const double *f) {   /* SourceEuler.c:48-66 */
  int l, n = o->nr * o->ns, b = 0;
  for (l = 0; l < n; l++) if (f[l] < 0.0) b = 1;
  return b;
}

/* ---------------- viscosity (Viscosity.c:85-260) ---------------- */
static void o_div_stress(oracle_ctx *o, const double *vr, const double *vt, const double *rho) {
  int i, j, l, nr = o->nr, ns = o->ns, lip, ljp, ljm, lim, ljmim;
  double *Drr = o->fld[W_DRR], *Drp = o->fld[W_DRP], *Dpp = o->fld[W_DPP];
  double *divergence = o->fld[THORIN_F_DIVV];
  double *Trr = o->fld[THORIN_F_TAURR], *Trp = o->fld[THORIN_F_TAURP], *Tpp = o->fld[THORIN_F_TAUPP];
  double dphi = 2.0 * M_PI / (double)ns, invdphi = 1.0 / dphi, onethird = 1.0 / 3.0, viscosity;
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      Drr[l] = (vr[lip] - vr[l]) * o->InvDiffRsup[i];
      Dpp[l] = (vt[ljp] - vt[l]) * invdphi * o->InvRmed[i] + 0.5 * (vr[lip] + vr[l]) * o->InvRmed[i];
      divergence[l] = (vr[lip] * o->Rsup[i] - vr[l] * o->Rinf[i]) * o->InvDiffRsup[i] * o->InvRmed[i];
      divergence[l] += (vt[ljp] - vt[l]) * invdphi * o->InvRmed[i];
    }
  for (i = 1; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      lim = l - ns;
      Drp[l] = 0.5 * (o->Rinf[i] * (vt[l] * o->InvRmed[i] - vt[lim] * o->InvRmed[i - 1]) * o->InvDiffRmed[i] +
                      (vr[l] - vr[ljm]) * invdphi * o->InvRinf[i]);
    }
  for (i = 0; i < nr; i++) {
    viscosity = o_fviscosity(o, o->Rmed[i]);
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      Trr[l] = 2.0 * rho[l] * viscosity * (Drr[l] - onethird * divergence[l]);
      Tpp[l] = 2.0 * rho[l] * viscosity * (Dpp[l] - onethird * divergence[l]);
    }
  }
  for (i = 1; i < nr; i++) {
    viscosity = o_fviscosity(o, o->Rinf[i]);
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      ljmim = ljm - ns;
      Trp[l] = 2.0 * 0.25 * (rho[l] + rho[lim] + rho[ljm] + rho[ljmim]) * viscosity * Drp[l];
    }
  }
}

static void o_visc_update(oracle_ctx *o, double *vr, double *vt, const double *rho, double DeltaT) {
  int i, j, l, nr = o->nr, ns = o->ns, lip, ljp, ljm, lim;
  const double *Trr = o->fld[THORIN_F_TAURR], *Trp = o->fld[THORIN_F_TAURP], *Tpp = o->fld[THORIN_F_TAUPP];
  double dphi = 2.0 * M_PI / (double)ns, invdphi = 1.0 / dphi;
  for (i = 1; i < nr - 1; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      vt[l] += DeltaT * o->InvRmed[i] * ((o->Rsup[i] * Trp[lip] - o->Rinf[i] * Trp[l]) * o->InvDiffRsup[i] +
                                         (Tpp[l] - Tpp[ljm]) * invdphi + 0.5 * (Trp[l] + Trp[lip])) /
               (0.5 * (rho[l] + rho[ljm]));
    }
  for (i = 1; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns; lim = l - ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      vr[l] += DeltaT * o->InvRinf[i] * ((o->Rmed[i] * Trr[l] - o->Rmed[i - 1] * Trr[lim]) * o->InvDiffRmed[i] +
                                         (Trp[ljp] - Trp[l]) * invdphi - 0.5 * (Tpp[l] + Tpp[lim])) /
               (0.5 * (rho[l] + rho[lim]));
    }
}

static void o_keplerian_edges(oracle_ctx *o, double *vt) {   /* Viscosity.c:221-260 */
  const thorin_params *p = &o->p;
  int j, nr = o->nr, ns = o->ns;
  double VKepIn, VKepOut;
  VKepIn = sqrt(G_ * 1.0 / o->Rmed[0]) *
           sqrt(1.0 - pow(o_aspectratio(o, o->Rmed[0]), 2.0) * pow(o->Rmed[0], 2.0 * p->flaringindex) *
                          (1. + p->sigmaslope - 2.0 * p->flaringindex));
  VKepOut = sqrt(G_ * 1.0 / o->Rmed[nr - 1]) *
            sqrt(1.0 - pow(o_aspectratio(o, o->Rmed[nr - 1]), 2.0) * pow(o->Rmed[nr - 1], 2.0 * p->flaringindex) *
                           (1. + p->sigmaslope - 2.0 * p->flaringindex));
  for (j = 0; j < ns; j++) vt[j] = VKepIn - o->Rmed[0] * o->st.omega_frame;
  for (j = 0; j < ns; j++) vt[j + (nr - 1) * ns] = VKepOut - o->Rmed[nr - 1] * o->st.omega_frame;
}

/* ---------------- SubStep1 / SubStep2 (SourceEuler.c:319-504) ---------------- */
static void o_substep1(oracle_ctx *o, double dt) {
  const thorin_params *p = &o->p;
  int i, j, l, lim, ljm, ljp, nr = o->nr, ns = o->ns;
  const double *rho = o->fld[THORIN_F_DENS], *vrad = o->fld[THORIN_F_VRAD], *vtheta = o->fld[THORIN_F_VTHETA];
  double *vradint = o->fld[THORIN_F_VRADINT], *vthetaint = o->fld[THORIN_F_VTHETAINT];
  const double *press = o->fld[THORIN_F_PRESSURE], *agr = o->fld[THORIN_F_AGR], *agt = o->fld[THORIN_F_AGT];
  double *accr = o->fld[W_GACCR], *acct = o->fld[W_GACCT];
  const double *dragr = o->fld[W_DRAGR], *dragt = o->fld[W_DRAGT];
  double gradp, vt2, dxtheta, tmp, invdxtheta, supp_torque;
  for (i = 1; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      gradp = (press[l] - press[lim]) * 2.0 / (rho[l] + rho[lim]) * o->InvDiffRmed[i];
      vt2 = vtheta[l] + vtheta[ljp] + vtheta[lim] + vtheta[ljp - ns];
      vt2 = vt2 / 4.0 + o->Rinf[i] * o->st.omega_frame;
      vt2 = vt2 * vt2;
      tmp = -gradp + 0.5 * (agr[l] + agr[lim]) + vt2 * o->InvRinf[i];
      vradint[l] = vrad[l] + dt * tmp;
      if (p->pebbles) {   /* SourceEuler.c:371-379 */
        accr[l] = tmp;
        if (p->back_reaction) {
          tmp = dragr[l] * 2.0 / (rho[l] + rho[lim]);
          vradint[l] += dt * tmp;
          accr[l] += tmp;
        }
      }
    }
  for (i = 0; i < nr; i++) {
    supp_torque = p->imposeddiskdrift * .5 * pow(o->Rmed[i], -2.5 + p->sigmaslope);
    dxtheta = 2.0 * PI / (double)ns * o->Rmed[i];
    invdxtheta = 1.0 / dxtheta;
    for (j = 0; j < ns; j++) {
      l = j + i * ns; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      gradp = (press[l] - press[ljm]) * 2.0 / (rho[l] + rho[ljm]) * invdxtheta;
      tmp = -gradp + 0.5 * (agt[l] + agt[ljm]);
      vthetaint[l] = vtheta[l] + dt * tmp;
      vthetaint[l] += dt * supp_torque;
      if (p->pebbles) {   /* SourceEuler.c:397-405 */
        acct[l] = tmp + supp_torque;
        if (p->back_reaction) {
          tmp = dragt[l] * 2.0 / (rho[l] + rho[ljm]);
          vthetaint[l] += dt * tmp;
          acct[l] += tmp;
        }
      }
    }
  }
  o_div_stress(o, vradint, vthetaint, rho);
  o_visc_update(o, vradint, vthetaint, rho, dt);
  if ((!p->damping_on) || (p->damping_on && p->damp_vrad)) o_keplerian_edges(o, vthetaint);
}

static void o_substep2(oracle_ctx *o, double dt) {
  int i, j, l, lim, lip, ljm, ljp, nr = o->nr, ns = o->ns;
  const double *rho = o->fld[THORIN_F_DENS], *vrad = o->fld[THORIN_F_VRADINT], *vtheta = o->fld[THORIN_F_VTHETAINT];
  double *qr = o->fld[W_RHOINT], *qt = o->fld[W_TEMPERINT];
  double *vradnew = o->fld[THORIN_F_VRADNEW], *vthetanew = o->fld[THORIN_F_VTHETANEW];
  const double *energy = o->fld[THORIN_F_ENERGY];
  double *energyint = o->fld[THORIN_F_ENERGYINT];
  double dxtheta, invdxtheta, dv;
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      dv = vrad[lip] - vrad[l];
      qr[l] = (dv < 0.0) ? CVNR * CVNR * rho[l] * dv * dv : 0.0;
      dv = vtheta[ljp] - vtheta[l];
      qt[l] = (dv < 0.0) ? CVNR * CVNR * rho[l] * dv * dv : 0.0;
    }
  for (i = 1; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns;
      vradnew[l] = vrad[l] - dt * 2.0 / (rho[l] + rho[lim]) * (qr[l] - qr[lim]) * o->InvDiffRmed[i];
    }
  for (i = 0; i < nr; i++) {
    dxtheta = 2.0 * PI / (double)ns * o->Rmed[i];
    invdxtheta = 1.0 / dxtheta;
    for (j = 0; j < ns; j++) {
      l = j + i * ns; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      vthetanew[l] = vtheta[l] - dt * 2.0 / (rho[l] + rho[ljm]) * (qt[l] - qt[ljm]) * invdxtheta;
    }
  }
  if (o->p.energy_eq)
    for (i = 0; i < nr; i++) {
      dxtheta = 2.0 * PI / (double)ns * o->Rmed[i];
      invdxtheta = 1.0 / dxtheta;
      for (j = 0; j < ns; j++) {
        l = j + i * ns; lip = l + ns; ljp = l + 1;
        if (j == ns - 1) ljp = i * ns;
        energyint[l] = energy[l] - dt * qr[l] * (vrad[lip] - vrad[l]) * o->InvDiffRsup[i] -
                       dt * qt[l] * (vtheta[ljp] - vtheta[l]) * invdxtheta;
      }
    }
}

/* ActualiseGas, SourceEuler.c:192-208: rings 0..nr inclusive */
static void o_actualise(oracle_ctx *o, double *dst, const double *src) {
  memcpy(dst, src, sizeof(double) * (size_t)(o->nr + 1) * o->ns);
}

/* ---------------- opacity (Opacity.c) ---------------- */
#define TEMP(i, j) pow(kappa[j] / kappa[i] * pow(rho, a[j] - a[i]), 1.0 / (b[i] - b[j]))
static double o_opacity_BL94(double rho, double T) {   /* Opacity.c:31-75 */
  static const double kappa[] = {2.0e-4, 2.0e16, 0.1, 2.0e81, 1.0e-8, 1.0e-36, 1.5e20, 0.348};
  static const double a[] = {0.0, 0.0, 0.0, 1.0, 2.0 / 3.0, 1.0 / 3.0, 1.0, 0.0};
  static const double b[] = {2.0, -7.0, 0.5, -24.0, 3.0, 10.0, -2.5, 0.0};
  double k_, k1, k2, k3, t1, t2;
  k_ = 1.0e-12;
  if (T <= TEMP(2, 3)) {
    k1 = kappa[0] * pow(T, b[0]);
    k2 = kappa[1] * pow(T, b[1]);
    k3 = kappa[2] * pow(T, b[2]);
    t1 = 1.0 / (k1 * k1) + 1.0 / (k2 * k2);
    t1 = 1.0 / (t1 * t1);
    t2 = pow(k3 / (1.0 + 1.0e22 * pow(T, -10.0)), 4.0);
    k_ = pow(t1 + t2, 0.25);
  } else if ((T > TEMP(2, 3)) && (T <= TEMP(3, 4))) k_ = kappa[3] * pow(rho, a[3]) * pow(T, b[3]);
  else if ((T > TEMP(3, 4)) && (T <= TEMP(4, 5))) k_ = kappa[4] * pow(rho, a[4]) * pow(T, b[4]);
  else if ((T > TEMP(4, 5)) && (T <= TEMP(5, 6))) k_ = kappa[5] * pow(rho, a[5]) * pow(T, b[5]);
  else if ((T > TEMP(5, 6)) && (T <= TEMP(6, 7))) k_ = kappa[6] * pow(rho, a[6]) * pow(T, b[6]);
  else k_ = kappa[7];
  return k_;
}
#undef TEMP

static double o_opacity_LP85(double rho, double T) {   /* Opacity.c:201-227 */
  double k;
  if (T < pow(2.0e-4 / 2.0e16, -1.0 / 9.0)) k = 2.0e-4 * T * T;
  else if (T < pow(2.0e16 / 5.0e-3, 1.0 / 8.0)) k = 2.0e16 * pow(T, -7.0);
  else if (T < pow(5.0e-3 / 2.0e34, -1.0 / 10.0) * pow(rho, 1.0 / 15.0)) k = 5.0e-3 * T;
  else if (T < pow(2.0e34 / 2.0e-8, 1.0 / 12.0)) k = 2.0e34 * pow(rho, 2.0 / 3.0) * pow(T, -9.0);
  else if (T < pow(2.0e-8 / 1.0e-36, 1.0 / 7.0) * pow(rho, 1.0 / 21.0)) k = 2.0e-8 * pow(rho, 2.0 / 3.0) * T * T * T;
  else if (T < pow(1.0e-36 / 1.5e20, -1.0 / 12.5) * pow(rho, 4.0 / 75.0)) k = 1.0e-36 * pow(rho, 1.0 / 3.0) * pow(T, 10.0);
  else k = 1.5e20 * rho * pow(T, -5.0 / 2.0);
  return k;
}

static double o_opacity_ZHU12(double rho, double T) {  /* Opacity.c:145-196 */
  double xlop, xlp, xlt, pre;
  pre = R_STANDARD * 1.0e7 / MMW * rho * T;
  xlp = log10(pre);
  xlt = log10(T);
  if (xlt < 2.23567 + 0.01899 * (xlp - 5.)) xlop = 1.5 * (xlt - 1.16331) - 0.736364;
  else if (xlt < 2.30713 + 0.01899 * (xlp - 5.)) xlop = -3.53154212 * xlt + 8.767726 - (7.24786 - 8.767726) * (xlp - 5.) / 16.;
  else if (xlt < (17.7 - 0.62 + 1.5 * 2.30713) / (1.5 + 5.832)) xlop = 1.5 * (xlt - 2.30713) + 0.62;
  else if (xlt < (5.9398 + 17.7) / (5.832 + 2.129)) xlop = -5.832 * xlt + 17.7;
  else if (xlt < (129.88071 + 5.9398 + (142.996475 - 129.88071) * 0.1 * (xlp + 4.)) / (2.129 + 42.98075)) xlop = 2.129 * xlt - 5.9398;
  else if (xlt < (129.88071 + 15.0125 + (142.996475 - 129.88071) * 0.1 * (xlp + 4.)) / (4.0625 + 42.98075))
    xlop = 129.88071 - 42.98075 * xlt + (142.996475 - 129.88071) * 0.1 * (xlp + 4.0);
  else if (xlt < 3.28 + xlp / 4. * 0.12) xlop = -15.0125 + 4.0625 * xlt;
  else if (xlt < 3.41 + 0.03328 * xlp / 4.) xlop = 58.9294 - 18.4808 * xlt + (61.6346 - 58.9294) * xlp / 4.;
  else if (xlt < 3.76 + (xlp - 4) / 2. * 0.03) xlop = -12.002 + 2.90477 * xlt + (xlp - 4) / 4. * (13.9953 - 12.002);
  else if (xlt < 4.07 + (xlp - 4) / 2. * 0.08) xlop = -39.4077 + 10.1935 * xlt + (xlp - 4) / 2. * (40.1719 - 39.4077);
  else if (xlt < 5.3715 + (xlp - 6) / 2. * 0.5594) xlop = 17.5935 - 3.3647 * xlt + (xlp - 6) / 2. * (17.5935 - 15.7376);
  else xlop = -0.48;
  if (xlop < 3.586 * xlt - 16.85 && xlt < 4.) xlop = 3.586 * xlt - 16.85;
  return pow(10., xlop);
}

/* OpacityTable SM03 (Opacity2.c:24-245): the Rosseland mean of Semenov et al.
 * (2003) dust and Malygin et al. (2014) gas.  The coefficient tables are the
 * published ones the reference holds (Opacity2.c:107-137). */
static double o_interp(const double *x, const double *y, int n, double x0) {   /* Opacity2.c:28-40 */
  int i = 1;
  if (n <= 1) return y[0];
  if (x0 < x[0]) return y[0];
  if (x0 > x[n - 1]) return y[n - 1];
  while ((x[i] < x0) && (i < n - 1)) i++;
  return y[i - 1] + (y[i] - y[i - 1]) * (x0 - x[i - 1]) / (x[i] - x[i - 1]);
}

static double o_opacity_malygin(const oracle_ctx *o, double rho, double T_gas) {   /* Opacity2.c:44-95, Rosseland */
  const int rows = o->sm_rows, cols = o->sm_cols;
  const double *pg = o->sm_pgrid, *tg = o->sm_tgrid, *a = o->sm_ross;
  double tP = 8.3144598 * 1.0e7 / 2.4 * rho * T_gas, tT = T_gas, fac, denom, mul1, mul2, mul3, mul4;
  int i, Tlow = 0, Thigh = 1, Plow, Phigh;
  if (tT < tg[0]) return 0.0;
  else if (tT > tg[cols - 1]) tT = tg[cols - 1];
  if (tP < pg[0]) tP = pg[0];
  else if (tP > pg[rows - 1]) tP = pg[rows - 1];
  for (i = 0; i < cols - 2; i++)
    if (tT >= tg[i]) { Tlow = i; Thigh = i + 1; }
  fac = ((double)rows) / log10(pg[rows - 1] / pg[0]);
  Plow = (int)(fac * log10(tP / pg[0]));
  if (Plow < 0) Plow = 0;
  if (Plow > rows - 2) Plow = rows - 2;
  Phigh = Plow + 1;
  denom = 1.0 / ((tg[Thigh] - tg[Tlow]) * (pg[Phigh] - pg[Plow]));
  mul1 = (tg[Thigh] - tT) * (pg[Phigh] - tP);
  mul2 = (tT - tg[Tlow]) * (pg[Phigh] - tP);
  mul3 = (tg[Thigh] - tT) * (tP - pg[Plow]);
  mul4 = (tT - tg[Tlow]) * (tP - pg[Plow]);
  return denom * (mul1 * a[Tlow + Plow * cols] + mul2 * a[Thigh + Plow * cols] + mul3 * a[Tlow + Phigh * cols] +
                  mul4 * a[Thigh + Phigh * cols]);
}

static double o_opacity_SM03(const oracle_ctx *o, double rho, double T_dust) {   /* Opacity2.c:99-245 */
  static const double eD[36] = {
    0.244153190819473e-03, -0.191588103585833e-03, 0.229132294521843e-03, 0.454088516995579e-06, -0.389280273285906e-08, 0.319006401785413e-11,
    0.292517586013911e+01, -0.801305197566973e-01, 0.923001367150898e-03, -0.349287851393541e-05, 0.587151269508960e-08, -0.371333580736933e-11,
    -0.125893536782728e+02, 0.160197849829283e+00, -0.582601311634824e-03, 0.110792133181219e-05, -0.105506373600323e-08, 0.404997080931974e-12,
    -0.192550086994197e+01, 0.354301116460647e-01, -0.127355043519226e-03, 0.233773506650911e-06, -0.207587683535083e-09, 0.728005983960845e-13,
    0.165244978116957e+01, -0.260479348963077e-02, 0.590717655258634e-05, -0.300202906476749e-08, 0.803836688553263e-12, -0.916709776405312e-16,
    0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  static const double logtt[8][4] = {
    {2.03742650, 2.92168648, 2.95520654, 2.96801571}, {2.07188201, 2.95808585, 2.99122608, 2.99869516},
    {2.11058971, 2.99738638, 3.02077549, 3.03181227}, {2.15533604, 3.04139269, 3.05269394, 3.06744284},
    {2.20139712, 3.08990511, 3.08707121, 3.10619090}, {2.25527251, 3.14457421, 3.12417806, 3.14860265},
    {2.31597035, 3.20736504, 3.16494737, 3.19589965}, {2.38738983, 3.28057837, 3.20978301, 3.24895362}};
  static const double logrr[8] = {-18.0, -16.0, -14.0, -12.0, -10.0, -8.0, -6.0, -4.0};
  double T_ev[4], temp[8], dT[5], T[5], tmax1, tmax2, tmin, T1, T2, TD, aKrL, aKrR, AA, BB, FF, aKext = 0.0;
  double logrho, kappa_gas;
  int i, j, KK, smooth;
  if (T_dust <= 0.9) return 0.9;
  logrho = log10(rho);
  for (i = 0; i < 4; i++) {
    for (j = 0; j < 8; j++) temp[j] = logtt[j][i];
    T_ev[i] = pow(10.0, o_interp(logrr, temp, 8, logrho));
  }
  T[0] = T_ev[0]; T[1] = 275.0; T[2] = 425.0; T[3] = 680.0;
  tmax1 = T_ev[1] > T_ev[2] ? T_ev[1] : T_ev[2];
  tmax2 = T_ev[2] > T_ev[3] ? T_ev[2] : T_ev[3];
  tmin = tmax1 < tmax2 ? tmax1 : tmax2;
  T[4] = tmin;
  dT[0] = 5.0; dT[1] = 5.0; dT[2] = 15.0; dT[3] = 5.0; dT[4] = 100.0;
  KK = 5;
  if (T_dust <= (T[0] + dT[0])) KK = 0;
  for (i = 1; i < 5; i++)
    if ((T_dust > (T[i - 1] + dT[i - 1])) && (T_dust <= (T[i] + dT[i]))) KK = i;
  if (KK != 5) {
    smooth = 0;
    /* the reference passes the double difference to the int abs() */
    for (i = 0; i < 5; i++)
      if (abs((int)(T_dust - T[i])) < dT[i]) smooth = 1;
    if (smooth == 1) {
      T1 = T[KK] - dT[KK];
      T2 = T[KK] + dT[KK];
      TD = T_dust - T[KK];
      j = (KK + 1) * 6;
      aKrL = ((((eD[j - 1] * T1 + eD[j - 2]) * T1 + eD[j - 3]) * T1 + eD[j - 4]) * T1 + eD[j - 5]) * T1 + eD[j - 6];
      if (KK == 4) aKrR = 0.0;
      else {
        j = (KK + 2) * 6;
        aKrR = ((((eD[j - 1] * T2 + eD[j - 2]) * T2 + eD[j - 3]) * T2 + eD[j - 4]) * T2 + eD[j - 5]) * T2 + eD[j - 6];
      }
      AA = 0.5e0 * (aKrL - aKrR);
      BB = 0.5e0 * (aKrL + aKrR);
      FF = PI / 2.0 / dT[KK];
      aKext = BB - AA * sin(FF * TD);
    } else {
      j = (KK + 1) * 6;
      aKext = ((((eD[j - 1] * T_dust + eD[j - 2]) * T_dust + eD[j - 3]) * T_dust + eD[j - 4]) * T_dust + eD[j - 5]) *
                  T_dust + eD[j - 6];
    }
  }
  kappa_gas = o_opacity_malygin(o, rho, T_dust);
  return aKext > kappa_gas ? aKext : kappa_gas;
}

static double o_opacity(oracle_ctx *o, double rho, double T) {
  switch (o->p.opacity_law) {
    case THORIN_OPA_SM03: return o_opacity_SM03(o, rho, T);
    case THORIN_OPA_LP85: return o_opacity_LP85(rho, T);
    case THORIN_OPA_ZHU12: return o_opacity_ZHU12(rho, T);
    case THORIN_OPA_CONST: return o->p.parametricopacity;
    case THORIN_OPA_POWERLAW: return o->p.powerlawopacity * T * T;
    default: return o_opacity_BL94(rho, T);
  }
}

/* exported for closed-form tests of the laws */
double oracle_opacity_law(int law, double rho_cgs, double T_K, double k0) {
  switch (law) {
    case THORIN_OPA_LP85: return o_opacity_LP85(rho_cgs, T_K);
    case THORIN_OPA_ZHU12: return o_opacity_ZHU12(rho_cgs, T_K);
    case THORIN_OPA_CONST: return k0;
    case THORIN_OPA_POWERLAW: return k0 * T_K * T_K;
    default: return o_opacity_BL94(rho_cgs, T_K);
  }
}

/* ---------------- implicit radiative diffusion (EnergySources.c) ---------------- */
static double o_flux_limiter(double s) {   /* EnergySources.c:684-696 */
  double tmp;
  if (s <= 2.0) { tmp = 3.0 + sqrt(9.0 + 10.0 * s * s); return 2.0 / tmp; }
  tmp = 10.0 * s + 9.0 + sqrt(81.0 + 180.0 * s);
  return 10.0 / tmp;
}
double oracle_flux_limiter(double s) { return o_flux_limiter(s); }

static double o_tau_eff(oracle_ctx *o, double tau) {   /* EnergySources.c:701-711 */
  if (o->p.stellar_irradiation) return 0.375 * tau + 0.5 + 0.25 / tau;
  return 0.375 * tau + 0.25 * sqrt(3.0) + 0.25 / tau;
}

/* Stellar irradiation of cell (i, j), 1 <= i <= nr-2: CalculateFlaring
 * (EnergySources.c:143-165: the sine of the grazing angle from the radial
 * derivative of the pressure scale height, 0 for surfaces tilted away) and
 * CalculateQirr (EnergySources.c:103-135: the heating it gives through the
 * effective optical depth).  The edge rings' flaring values of the reference
 * (EnergySources.c:166-183) feed no matrix row and are not formed. */
static double o_qirr(oracle_ctx *o, int i, int j, const double *cs, const double *opa, const double *rho) {
  const thorin_params *p = &o->p;
  const int ns = o->ns;
  const int l = j + i * ns, lip = l + ns, lim = l - ns;
  double Rstar, Teff4, Rstar2, Tirr4fac, csin, csout, Hin, Hout, dHdr, H, fl, tau, taueff, Tirr4, r2inv;
  Rstar = p->stellarradius * 6.957e8 / AU_SI;
  Teff4 = p->effectivetemperature / T2SI;
  Teff4 = pow(Teff4, 4.0);
  Rstar2 = 6.957e8 / AU_SI;
  Rstar2 = pow(p->stellarradius * Rstar2, 2.0);
  Tirr4fac = (1.0 - p->discalbedo) * Teff4 * Rstar2;
  csin = 0.5 * (cs[l] + cs[lim]);
  csout = 0.5 * (cs[l] + cs[lip]);
  Hin = csin * pow(o->Rinf[i], 1.5) * o->sqrt_adiabind_inv;
  Hout = csout * pow(o->Rsup[i], 1.5) * o->sqrt_adiabind_inv;
  dHdr = (Hout - Hin) * o->InvDiffRsup[i];
  H = cs[l] * o->OmegaInv[i] * o->sqrt_adiabind_inv;
  fl = atan(dHdr) - atan((H - 0.4 * Rstar) * o->InvRmed[i]);
  fl = (fl >= 0.0) ? sin(fl) : 0.0;
  if (!(fl > 0.0)) return 0.0;
  r2inv = 1.0 / (o->Rmed[i] * o->Rmed[i]);
  tau = p->opacitydrop * 0.5 * opa[l] * rho[l];
  taueff = o_tau_eff(o, tau);
  Tirr4 = Tirr4fac * r2inv * fl;
  return 2.0 * STEFANBOLTZMANN * Tirr4 / taueff;
}

/* SuccessiveOverrelaxation, EnergySources.c:406-511 (single rank) */
static int o_sor(oracle_ctx *o, double omega, int errcheck) {
  int n, i, j, l, nr = o->nr, ns = o->ns, ipass, jchess, lip, lim, ljp, ljm, changed;
  double resid, normres0 = 0.0, normres, dtemper;
  double *T = o->fld[THORIN_F_TEMPER];
  const double *Ml = o->fld[W_ML], *Mlip = o->fld[W_MLIP], *Mlim = o->fld[W_MLIM];
  const double *Mljp = o->fld[W_MLJP], *Mljm = o->fld[W_MLJM], *RHS = o->fld[W_RHS];
  (void)errcheck;
  for (i = 1; i < nr - 1; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns; lim = l - ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      resid = T[l] * Ml[l] + T[lip] * Mlip[l] + T[lim] * Mlim[l] + T[ljp] * Mljp[l] + T[ljm] * Mljm[l] - RHS[l];
      normres0 += fabs(resid);
    }
  for (n = 1; n <= SORMAXITERS; n++) {
    normres = 0.0;
    changed = 0;
    for (ipass = 1; ipass <= 2; ipass++) {
      jchess = (ipass == 1) ? 0 : 1;
      for (i = 1; i < nr - 1; i++) {
        for (j = jchess; j < ns; j += 2) {
          l = j + i * ns; lip = l + ns; lim = l - ns; ljp = l + 1;
          if (j == ns - 1) ljp = i * ns;
          ljm = l - 1;
          if (j == 0) ljm = i * ns + ns - 1;
          resid = T[l] * Ml[l] + T[lip] * Mlip[l] + T[lim] * Mlim[l] + T[ljp] * Mljp[l] + T[ljm] * Mljm[l] - RHS[l];
          normres += fabs(resid);
          dtemper = omega * resid / Ml[l];
          T[l] -= dtemper;
          if (dtemper > T[l] * SORPRECISION) changed++;
        }
        jchess = 1 - jchess;
      }
    }
    if (normres < SOREPS * normres0) return n;
    if (normres < SORMIN) return n;
    if (changed == 0) return n;
  }
  return n;
}

/* IterateRelaxationParameter, EnergySources.c:340-403 */
static int o_iterate_relaxation(oracle_ctx *o) {
  double omegamin = 1.0, omega = 1.0, omegamax = OMEGAUPPER;
  int nr = o->nr, ns = o->ns, n, nsplit = 9, Niter, count = 0;
  double *T = o->fld[THORIN_F_TEMPER];
  double *bck = (double *)malloc(sizeof(double) * (size_t)nr * ns);
  memcpy(bck, T, sizeof(double) * (size_t)nr * ns);
  o->st.sor_domega = (omegamax - omegamin) / nsplit;
  o->st.sor_niter_best = SORMAXITERS;
  o->st.sor_omega_best = 1.0;
  for (;;) {
    for (n = 0; n <= nsplit; n++) {
      Niter = o_sor(o, omega, 0);
      if (Niter == o->st.sor_niter_best) count++;
      if (Niter < o->st.sor_niter_best) { o->st.sor_niter_best = Niter; o->st.sor_omega_best = omega; }
      omega += o->st.sor_domega;
      memcpy(T, bck, sizeof(double) * (size_t)nr * ns);
    }
    if (o->st.sor_niter_best == SORMAXITERS) { free(bck); return -1; }
    if (count == nsplit + 1) { o->st.sor_domega = omegamax - omegamin; free(bck); return 0; }
    count = 0;
    omegamin = o->st.sor_omega_best - o->st.sor_domega;
    omegamax = o->st.sor_omega_best + o->st.sor_domega;
    if (omegamin < 1.0) omegamin = 1.0;
    if (omegamax > OMEGAUPPER) omegamax = OMEGAUPPER;
    o->st.sor_domega = (omegamax - omegamin) / nsplit;
    omega = omegamin;
  }
}

/* ImplicitRadiativeDiffusion, EnergySources.c:201-335, with the helpers it
 * calls (MidplaneVolumeDensity 658-680, OpacityProfile Opacity.c:231-265,
 * CalculateQminus 75-99, DiffusionCoefs 556-605, TemperatureGradient 609-653,
 * DiffusionTimestep 833-856). */
static int o_implicit_diffusion(oracle_ctx *o, double dt, double *omega_used) {
  const thorin_params *p = &o->p;
  int nr = o->nr, ns = o->ns, i, j, l, lip, lim, ljp, ljm, Niter;
  double *rho = o->fld[THORIN_F_DENS], *eint = o->fld[THORIN_F_ENERGYINT], *enew = o->fld[THORIN_F_ENERGYNEW];
  double *T = o->fld[THORIN_F_TEMPER], *cs = o->fld[THORIN_F_SOUNDSPEED];
  double *voldens = o->fld[W_VOLDENS], *opa = o->fld[THORIN_F_OPACITY], *qminus = o->fld[THORIN_F_QMINUS];
  double *gtr = o->fld[W_GRADTR], *gtt = o->fld[W_GRADTT], *gtm = o->fld[W_GRADTM];
  double *D = o->fld[W_DIFFC], *Dr = o->fld[W_DIFFR], *Dt = o->fld[W_DIFFT];
  double *A = o->fld[W_A], *B = o->fld[W_B], *Ml = o->fld[W_ML], *Mlip = o->fld[W_MLIP];
  double *Mlim = o->fld[W_MLIM], *Mljp = o->fld[W_MLJP], *Mljm = o->fld[W_MLJM], *RHS = o->fld[W_RHS];
  double *divv = o->fld[THORIN_F_DIVV], *qplus = o->fld[THORIN_F_QPLUS];
  double H, dxtheta, invdxtheta, invdxtheta2, afac, bfac, r, t, gradt, fac, s, dtmin, deltar2, dttmp;
  o_temperature(o, rho, eint);
  o_sound_speed(o, rho, eint);
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      H = cs[l] * o->OmegaInv[i] * o->sqrt_adiabind_inv;
      voldens[l] = rho[l] / H * SQRT2PI_INV;
    }
  if (p->opacity_law == THORIN_OPA_CONST || p->parametricopacity > 0.0) {
    if (!o->opa_filled) {
      double tmp = p->parametricopacity * OPA2CU;
      for (i = 0; i < nr; i++) for (j = 0; j < ns; j++) opa[j + i * ns] = tmp;
    }
  } else {
    for (i = 0; i < nr; i++)
      for (j = 0; j < ns; j++) {
        l = j + i * ns;
        opa[l] = o_opacity(o, voldens[l] * RHO2CGS, T[l] * T2SI) * OPA2CU;
      }
  }
  o->opa_filled = 1;
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      double tau, taueff;
      l = j + i * ns;
      tau = p->opacitydrop * 0.5 * opa[l] * rho[l];
      taueff = o_tau_eff(o, tau);
      qminus[l] = 2.0 * STEFANBOLTZMANN * pow(T[l], 3.0) / taueff;
    }
  /* TemperatureGradient */
  for (i = 1; i < nr; i++) {
    dxtheta = 2.0 * PI / (double)ns * o->Rmed[i];
    invdxtheta = 1.0 / dxtheta;
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns;
      gtr[l] = (T[l] - T[lim]) * o->InvDiffRmed[i];
      ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      gtt[l] = (T[l] - T[ljm]) * invdxtheta;
    }
  }
  for (i = 1; i < nr - 1; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      r = 0.5 * (gtr[lip] + gtr[l]);
      t = 0.5 * (gtt[ljp] + gtt[l]);
      gtm[l] = sqrt(r * r + t * t);
    }
  /* DiffusionCoefs */
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      if (i == 0) gradt = gtm[l + ns];
      else if (i == nr - 1) gradt = gtm[l - ns];
      else gradt = gtm[l];
      fac = 1.0 / (voldens[l] * opa[l]);
      s = 4.0 * gradt * fac / T[l];
      D[l] = o_flux_limiter(s) * 16.0 * STEFANBOLTZMANN * pow(T[l], 3.0) * fac;
    }
  for (i = 1; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      Dr[l] = 0.5 * (D[l] + D[lim]);
      Dt[l] = 0.5 * (D[l] + D[ljm]);
    }
  /* DiffusionTimestep */
  dtmin = 1.0e38;
  for (i = 0; i < nr; i++) {
    deltar2 = o->Rsup[i] - o->Rinf[i];
    deltar2 = deltar2 * deltar2 / 2.0;
    for (j = 0; j < ns; j++) {
      dttmp = deltar2 / D[j + i * ns];
      if (dttmp < dtmin) dtmin = dttmp;
    }
  }
  o->st.dt_stellar = dtmin;
  /* matrix */
  for (i = 1; i < nr; i++) {
    dxtheta = 2.0 * PI / (double)ns * o->Rmed[i];
    invdxtheta2 = 1.0 / dxtheta / dxtheta;
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      A[l] = Dr[l] * o->Rinf[i] * o->InvDiffRmed[i];
      H = cs[l] * o->OmegaInv[i] * o->sqrt_adiabind_inv;
      B[l] = Dt[l] * invdxtheta2 * 2.0 * H;
    }
  }
  for (i = 1; i < nr - 1; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      H = cs[l] * o->OmegaInv[i] * o->sqrt_adiabind_inv;
      bfac = dt / (rho[l] * o->CV);
      afac = bfac * o->InvDiffRsup[i] * 2.0 * H / o->Rmed[i];
      Ml[l] = 1.0 + divv[l] * (p->adiabind - 1.0) * dt + 4.0 * bfac * qminus[l] + afac * (A[lip] + A[l]) + bfac * (B[ljp] + B[l]);
      Mlip[l] = -afac * A[lip];
      Mlim[l] = -afac * A[l];
      Mljp[l] = -bfac * B[ljp];
      Mljm[l] = -bfac * B[l];
      RHS[l] = T[l] + bfac * (qplus[l] + 3.0 * qminus[l] * T[l]);
      if (p->stellar_irradiation) RHS[l] += bfac * o_qirr(o, i, j, cs, opa, rho);   /* EnergySources.c:290 */
      if (p->accret_heating || p->gas_accret_heating) {                        /* EnergySources.c:291-294 */
        int k;
        for (k = 0; k < o->heatsrc_max && k < 8; k++)
          if (l == o->heatsrc_index[k]) RHS[l] += bfac * o->heatsrc[k];
      }
    }
  if (o->st.sor_first) {
    o->st.sor_first = 0;
    if (o_iterate_relaxation(o) < 0) return -1;
    o->st.sor_omega = o->st.sor_omega_best;
    o->st.sor_niter_last = o->st.sor_niter_best;
  }
  o->st.sor_omega += o->st.sor_domega;
  if (o->st.sor_omega > OMEGAUPPER) { o->st.sor_omega = OMEGAUPPER; o->st.sor_domega = -o->st.sor_domega; }
  if (o->st.sor_omega < 1.0) { o->st.sor_omega = 1.0; o->st.sor_domega = -o->st.sor_domega; }
  *omega_used = o->st.sor_omega;
  Niter = o_sor(o, o->st.sor_omega, 1);
  if (Niter > o->st.sor_niter_last) o->st.sor_domega = -o->st.sor_domega;
  o->st.sor_niter_last = Niter;
  for (j = 0; j < ns; j++) T[j] = T[j + ns];
  for (j = 0; j < ns; j++) T[j + (nr - 1) * ns] = T[j + (nr - 2) * ns];
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) { l = j + i * ns; enew[l] = rho[l] * T[l] * o->CV; }
  return Niter;
}

/* SubStep3, SourceEuler.c:508-586 (ParametricCooling branch included) */
static int o_substep3(oracle_ctx *o, double dt, double *omega_used) {
  const thorin_params *p = &o->p;
  int nr = o->nr, ns = o->ns, i, j, l;
  const double *rho = o->fld[THORIN_F_DENS], *energy = o->fld[THORIN_F_ENERGYINT];
  double *enew = o->fld[THORIN_F_ENERGYNEW], *qplus = o->fld[THORIN_F_QPLUS];
  const double *Trr = o->fld[THORIN_F_TAURR], *Trp = o->fld[THORIN_F_TAURP], *Tpp = o->fld[THORIN_F_TAUPP];
  const double *divv = o->fld[THORIN_F_DIVV];
  double viscosity, r0 = o->Rmed[0], r1 = o->Rmed[1], r2 = o->Rmed[2], q1, q2;
  for (i = 1; i < nr; i++) {
    viscosity = o_fviscosity(o, o->Rmed[i]);
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      if (viscosity != 0.0) {
        qplus[l] = 0.5 / viscosity / rho[l] * (Trr[l] * Trr[l] + 2.0 * Trp[l] * Trp[l] + Tpp[l] * Tpp[l]);
        qplus[l] += (2.0 / 9.0) * viscosity * rho[l] * divv[l] * divv[l];
      } else qplus[l] = 0.0;
    }
  }
  viscosity = o_fviscosity(o, o->Rmed[0]);
  for (j = 0; j < ns; j++) {
    q1 = qplus[j + ns];
    q2 = qplus[j + 2 * ns];
    if (viscosity != 0.0) qplus[j] = q1 * exp(log(q1 / q2) * log(r0 / r1) / log(r1 / r2));
    else qplus[j] = 0.0;
  }
  if (p->parametric_cooling) return -2;   /* deprecated route; not in scope */
  return o_implicit_diffusion(o, dt, omega_used);
}

/* ---------------- transport (TransportEuler.c) ---------------- */
static void o_divise(oracle_ctx *o, const double *num, const double *den, double *res) {
  /* DivisePolarGrid, SideEuler.c:82-99 (with its Nrad/Nsec swap) */
  int i, j, l, ns = o->nr, nr = o->ns;
  for (i = 0; i <= nr; i++)
    for (j = 0; j < ns; j++) { l = j + ns * i; res[l] = num[l] / (den[l] + 1e-20); }
}

static void o_star_rad(oracle_ctx *o, const double *qb, const double *vr, double *qs, double dt) {
  /* ComputeStarRad, TransportEuler.c:232-273 */
  int i, j, l, lq, lip, lim, nr = o->nr, ns = o->ns;
  double *dq = o->fld[W_DQ], dqm, dqp;
  for (j = 0; j < ns; j++) {
    for (i = 0; i < nr; i++) {
      l = j + i * ns; lq = i + j * nr; lip = l + ns; lim = l - ns;
      if ((i == 0) || (i == nr - 1)) dq[lq] = 0.0;
      else {
        dqm = (qb[l] - qb[lim]) * o->InvDiffRmed[i];
        dqp = (qb[lip] - qb[l]) * o->InvDiffRmed[i + 1];
        if (dqp * dqm > 0.0) dq[lq] = 2.0 * dqp * dqm / (dqp + dqm);
        else dq[lq] = 0.0;
      }
    }
    for (i = 1; i < nr; i++) {
      l = j + i * ns; lq = i + j * nr; lim = l - ns;
      if (vr[l] > 0.0) qs[l] = qb[lim] + (o->Rmed[i] - o->Rmed[i - 1] - vr[l] * dt) * 0.5 * dq[lq - 1];
      else qs[l] = qb[l] - (o->Rmed[i + 1] - o->Rmed[i] + vr[l] * dt) * 0.5 * dq[lq];
    }
    qs[j] = qs[j + ns * nr] = 0.0;
  }
}

static void o_star_theta(oracle_ctx *o, const double *qb, const double *vt, double *qs, double dt) {
  /* ComputeStarTheta, TransportEuler.c:275-316 */
  int i, j, l, ljp, ljm, jm, nr = o->nr, ns = o->ns;
  double *dq = o->fld[W_DQ], dqm, dqp, dxtheta, ksi, invdxtheta;
  for (i = 0; i < nr; i++) {
    dxtheta = 2.0 * PI / (double)ns * o->Rmed[i];
    invdxtheta = 1.0 / dxtheta;
    for (j = 0; j < ns; j++) {
      l = j + i * ns; ljp = l + 1; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      if (j == ns - 1) ljp = i * ns;
      dqm = (qb[l] - qb[ljm]);
      dqp = (qb[ljp] - qb[l]);
      if ((dqp * dqm > 0.0) && (dqp + dqm != 0.0)) dq[l] = dqp * dqm / (dqp + dqm) * invdxtheta;
      else dq[l] = 0.0;
    }
    for (j = 0; j < ns; j++) {
      l = j + i * ns; jm = j - 1;
      if (j == 0) jm = ns - 1;
      ljm = jm + i * ns;
      ksi = vt[l] * dt;
      if (ksi > 0.0) qs[l] = qb[ljm] + (dxtheta - ksi) * dq[ljm];
      else qs[l] = qb[l] - (dxtheta + ksi) * dq[l];
    }
  }
}

static void o_vanleer_rad(oracle_ctx *o, const double *vr, double *qb, double dt) {
  /* VanLeerRadial, TransportEuler.c:384-415 */
  int i, j, nr = o->nr, ns = o->ns, l, lip;
  double *qrs = o->fld[W_QRSTAR], *rhos = o->fld[W_RHOSTAR], dtheta, varq;
  o_divise(o, qb, o->fld[W_RHOINT], o->fld[W_WORK]);
  o_star_rad(o, o->fld[W_WORK], vr, qrs, dt);
  dtheta = 2.0 * PI / (double)ns;
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns;
      varq = dt * dtheta * o->Rinf[i] * qrs[l] * rhos[l] * vr[l];
      varq -= dt * dtheta * o->Rsup[i] * qrs[lip] * rhos[lip] * vr[lip];
      qb[l] += varq * o->InvSurf[i];
    }
}

static void o_vanleer_theta(oracle_ctx *o, const double *vt, double *qb, double dt) {
  /* VanLeerTheta, TransportEuler.c:417-445 */
  int i, j, nr = o->nr, ns = o->ns, l, ljp;
  double *qrs = o->fld[W_QRSTAR], *rhos = o->fld[W_RHOSTAR], dxrad, varq, invsurf;
  o_divise(o, qb, o->fld[W_RHOINT], o->fld[W_WORK]);
  o_star_theta(o, o->fld[W_WORK], vt, qrs, dt);
  for (i = 0; i < nr; i++) {
    dxrad = (o->Rsup[i] - o->Rinf[i]) * dt;
    invsurf = 1.0 / o->Surf[i];
    for (j = 0; j < ns; j++) {
      l = j + i * ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      varq = dxrad * qrs[l] * rhos[l] * vt[l];
      varq -= dxrad * qrs[ljp] * rhos[ljp] * vt[ljp];
      qb[l] += varq * invsurf;
    }
  }
}

static void o_quantities_advection(oracle_ctx *o, double *rho, double *energy, const double *vt, double dt) {
  /* QuantitiesAdvection, TransportEuler.c:159-175 (energy == NULL:
   * QuantitiesAdvectionPebbles, 491-505) */
  o_star_theta(o, rho, vt, o->fld[W_RHOSTAR], dt);
  o_actualise(o, o->fld[W_RHOINT], rho);
  o_vanleer_theta(o, vt, o->fld[W_RADMOMP], dt);
  o_vanleer_theta(o, vt, o->fld[W_RADMOMM], dt);
  o_vanleer_theta(o, vt, o->fld[W_THETAMOMP], dt);
  o_vanleer_theta(o, vt, o->fld[W_THETAMOMM], dt);
  if (energy) o_vanleer_theta(o, vt, energy, dt);
  o_vanleer_theta(o, vt, rho, dt);
}

static void o_advect_shift(oracle_ctx *o, double *val) {
  /* AdvectSHIFT, TransportEuler.c:113-136 */
  int i, j, ji, nr = o->nr, ns = o->ns;
  double *tmp = o->fld[W_TEMPSHIFT];
  for (i = 0; i < nr; i++) {
    for (j = 0; j < ns; j++) {
      ji = j - o->Nshift[i];
      while (ji < 0) ji += ns;
      while (ji >= ns) ji -= ns;
      tmp[j + i * ns] = val[ji + i * ns];
    }
    for (j = 0; j < ns; j++) val[j + i * ns] = tmp[j + i * ns];
  }
}

static void o_transport(oracle_ctx *o, double *rho, double *vr, double *vt, double *energy, double dt) {
  /* Transport, TransportEuler.c:34-47, with ComputeLRMomenta 318-347,
   * OneWindRad 49-65, OneWindTheta 138-156, ComputeResiduals 70-111,
   * ComputeVelocities 350-381.  energy == NULL: TransportPebbles (449-505)
   * for the pebble fluid (no energy, same momenta and shift) */
  int i, j, l, lip, ljp, lim, ljm, nr = o->nr, ns = o->ns;
  double *rp = o->fld[W_RADMOMP], *rm = o->fld[W_RADMOMM], *tp = o->fld[W_THETAMOMP], *tm = o->fld[W_THETAMOMM];
  double *vtr = o->fld[W_VTHETARES];
  double of = o->st.omega_frame;
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lip = l + ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      rp[l] = rho[l] * vr[lip];
      rm[l] = rho[l] * vr[l];
      tp[l] = rho[l] * (vt[ljp] + o->Rmed[i] * of) * o->Rmed[i];
      tm[l] = rho[l] * (vt[l] + o->Rmed[i] * of) * o->Rmed[i];
    }
  /* OneWindRad */
  o_star_rad(o, rho, vr, o->fld[W_RHOSTAR], dt);
  o_actualise(o, o->fld[W_RHOINT], rho);
  o_vanleer_rad(o, vr, rp, dt);
  o_vanleer_rad(o, vr, rm, dt);
  o_vanleer_rad(o, vr, tp, dt);
  o_vanleer_rad(o, vr, tm, dt);
  if (energy) o_vanleer_rad(o, vr, energy, dt);
  o_vanleer_rad(o, vr, rho, dt);
  /* OneWindTheta: ComputeResiduals */
  {
    double invdt = 1.0 / dt, dpinvns = 2.0 * PI / (double)ns, moy, Ntilde, Nround, vtemp;
    for (i = 0; i < nr; i++) {
      moy = 0.0;
      for (j = 0; j < ns; j++) moy += vt[j + i * ns];
      o->VMed[i] = moy / (double)ns;
      Ntilde = o->VMed[i] * o->InvRmed[i] * dt * (double)ns / 2.0 / PI;
      Nround = floor(Ntilde + 0.5);
      o->Nshift[i] = (int)(long)Nround;
      if (o->p.fast_transport) {
        vtemp = (Ntilde - Nround) * o->Rmed[i] * invdt * dpinvns;
        for (j = 0; j < ns; j++) { l = j + i * ns; vtr[l] = vt[l] - o->VMed[i]; vt[l] = vtemp; }
      } else {
        for (j = 0; j < ns; j++) { l = j + i * ns; vtr[l] = vt[l]; vt[l] = 0.0; }
      }
    }
  }
  o_quantities_advection(o, rho, energy, vtr, dt);
  if (o->p.fast_transport) {
    o_quantities_advection(o, rho, energy, vt, dt);
    o_advect_shift(o, rp);
    o_advect_shift(o, rm);
    o_advect_shift(o, tp);
    o_advect_shift(o, tm);
    if (energy) o_advect_shift(o, energy);
    o_advect_shift(o, rho);
  }
  /* ComputeVelocities */
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      if (i == 0) vr[l] = 0.0;
      else vr[l] = (rp[lim] + rm[l]) / (rho[l] + rho[lim] + 1e-20);
      vt[l] = (tp[ljm] + tm[l]) / (rho[l] + rho[ljm] + 1e-15) / o->Rmed[i] - o->Rmed[i] * of;
    }
}

/* ---------------- pebble fluid (Pebbles.c) ---------------- */
#define TRAPEZMAX 16           /* Pebbles.c:21 */
#define TRAPEZEPS 1.0e-4       /* Pebbles.c:22 */
#define EVAPOREDSIGMA 1.0e-8   /* Pebbles.c:835 */
#define FLUX2CU (0.000003 / (2.0 * PI))   /* fondam.h:54 */

/* PebbleStokesNumbers, Pebbles.c:272-293 (Epstein drag from the pebble size) */
static void o_peb_stokes(oracle_ctx *o) {
  int i, j, l, nr = o->nr, ns = o->ns;
  const double *rho = o->fld[THORIN_F_DENS], *psize = o->fld[THORIN_F_PEBBLESIZE];
  double *tau = o->fld[THORIN_F_PEBBLESTOKES];
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      tau[l] = 2.0 * o->pebbulkdens * psize[l] / rho[l] * o->sqrt_adiabind_inv;
    }
}

/* CriticalCharTime, Pebbles.c:735-770 */
static void o_peb_cfl(oracle_ctx *o) {
  int i, j, l, ns = o->ns;
  const double *pvr = o->fld[THORIN_F_PEBBLEVRAD], *pvt = o->fld[THORIN_F_PEBBLEVTHETA];
  const double *vr = o->fld[THORIN_F_VRAD], *vt = o->fld[THORIN_F_VTHETA];
  double dxrad, dxtheta, pvtmed = 0.0, invdt1, invdt2, invdt3, invdt4, vres, temp;
  o->invdtpeb_sq = -1.e30;
  for (i = 1; i < o->nr; i++) {
    dxrad = o->Rsup[i] - o->Rinf[i];
    dxtheta = o->Rmed[i] * 2.0 * PI / (double)ns;
    if (o->p.fast_transport) {
      pvtmed = 0.0;
      for (j = 0; j < ns; j++) pvtmed += pvt[j + i * ns];
      pvtmed /= (double)ns;
    }
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      invdt1 = fabs(pvr[l]) / dxrad;
      vres = o->p.fast_transport ? pvt[l] - pvtmed : pvt[l];
      invdt2 = fabs(vres) / dxtheta;
      invdt3 = fabs(pvr[l] - vr[l]) / dxrad;
      invdt4 = fabs(pvt[l] - vt[l]) / dxtheta;
      temp = invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4;
      o->invdtpeb_sq = o_max2(o->invdtpeb_sq, temp);
    }
  }
}

/* Trapzd, Pebbles.c:295-320: refinement n of the trapezoidal rule for
 * exp(-z^2/(2 H2)) on [a,b]; *integral carries the previous refinement */
static double o_trapzd(int n, double a, double b, double H2, double *integral) {
  double iter, spacing, x, sum;
  int j, it;
  if (n == 1) {
    *integral = 0.5 * (b - a) * (exp(-0.5 * a * a / H2) + exp(-0.5 * b * b / H2));
    return *integral;
  }
  for (it = 1, j = 1; j < n - 1; j++) it <<= 1;
  iter = it;
  spacing = (b - a) / iter;
  x = a + 0.5 * spacing;
  sum = 0.0;
  for (j = 1; j <= it; j++) {
    sum += exp(-0.5 * x * x / H2);
    x += spacing;
  }
  *integral = 0.5 * (*integral + (b - a) * sum / iter);
  return *integral;
}

/* IntegrateColumnMass, Pebbles.c:325-342; returns -1 and sets *ok = 0 when
 * the refinement does not converge (the reference exits) */
static double o_column_mass(double a, double b, double Hpeb, int *ok) {
  int n;
  double integral, iold = 0.0, H2 = Hpeb * Hpeb, state = 0.0;
  for (n = 1; n <= TRAPEZMAX; n++) {
    integral = o_trapzd(n, a, b, H2, &state);
    if (n > 5)
      if (fabs(integral - iold) < TRAPEZEPS * fabs(iold) || (integral == 0.0 && iold == 0.0)) return integral;
    iold = integral;
  }
  *ok = 0;
  return 0.0;
}

/* GetGlobalIFrac, LowTasks.c:21-31 */
static double o_global_ifrac(oracle_ctx *o, double r) {
  int i = 0, nr = o->nr;
  if (r < o->GlobalRmed[0]) return 0.0;
  if (r > o->GlobalRmed[nr - 1]) return (double)nr - 1.0;
  while (o->GlobalRmed[i] <= r) i++;
  return (double)i + (r - o->GlobalRmed[i - 1]) / (o->GlobalRmed[i] - o->GlobalRmed[i - 1]) - 1.0;
}

/* Heating1, Planet.c:222-246 */
static void o_heating1(oracle_ctx *o, double Mplanet, double dMplanet, double Xplanet, double Yplanet, double dt,
                       int k) {
  const thorin_params *p = &o->p;
  int i, j, l;
  double Rperp, tdelay, taper, plrho, plradius, L, ang, dE;
  Rperp = sqrt(Xplanet * Xplanet + Yplanet * Yplanet);
  if (Rperp < o->Rinf[0] || Rperp > o->Rsup[o->nr - 1]) return;
  tdelay = 4.0 * p->dt * (double)p->heatingdelay;
  taper = o->st.physical_time / tdelay;
  if (taper > 1.0) taper = 1.0;
  plrho = p->planetarydensity / RHO2CGS;
  plradius = pow(3.0 * Mplanet / (4.0 * PI * plrho), 1.0 / 3.0);
  L = Mplanet * dMplanet / dt / plradius;
  ang = atan2(Yplanet, Xplanet);
  if (ang < 0.0) ang += 2.0 * PI;
  i = 0;
  while (o->Rsup[i] < Rperp) i++;
  j = floor((double)o->ns / 2.0 / PI * ang + 0.5);
  if (j == o->ns) j = 0;
  l = j + i * o->ns;
  o->heatsrc_index[k] = l;
  dE = L * taper / o->Surf[i];
  o->heatsrc[k] += dE;
}

/* qtrap/trapzd of exp(-z^2) (Planet.c:166-214, QTRAPEPS 1e-4, QTRAPMAX 16) */
static double o_gas_qtrap(double a, double b) {
  double s = 0.0, olds = 0.0, x, tnm, sum, del;
  int n, it, j;
  for (n = 1; n <= 16; n++) {
    if (n == 1) {
      s = 0.5 * (b - a) * (exp(-a * a) + exp(-b * b));
    } else {
      for (it = 1, j = 1; j < n - 1; j++) it <<= 1;
      tnm = it;
      del = (b - a) / tnm;
      x = a + 0.5 * del;
      for (sum = 0.0, j = 1; j <= it; j++, x += del) sum += exp(-x * x);
      s = 0.5 * (s + (b - a) * sum / tnm);
    }
    if (n > 5)
      if (fabs(s - olds) < 1.e-4 * fabs(olds) || (s == 0.0 && olds == 0.0)) return s;
    olds = s;
  }
  return s;
}

/* AccreteOntoPlanets, Planet.c:33-163 (single rank, sequential sums in the
 * reference's loop order; GasAccretHeating is a refused option) */
static void o_gas_accrete(oracle_ctx *o, thorin_planets *sys, double dt) {
  const thorin_params *p = &o->p;
  const int nr = o->nr, ns = o->ns;
  double *dens = o->fld[THORIN_F_DENS], *vrad = o->fld[THORIN_F_VRAD], *vtheta = o->fld[THORIN_F_VTHETA];
  const double *cs = o->fld[THORIN_F_SOUNDSPEED];
  const double *abs_ = o->fld[W_ABS], *ord = o->fld[W_ORD];
  int k, i, j, jf, l, lip, ljp, i_min, i_max, j_min, j_max;
  for (k = 0; k < sys->nb; k++) {
    double dMplanet, dPxPlanet, dPyPlanet, facc, facc1, facc2, frac1, frac2, Xplanet, Yplanet, Zplanet;
    double VXplanet, VYplanet, Mplanet, Rperp, Rplanet, RRoche, angle, PxPlanet, PyPlanet;
    if (!(sys->acc[k] > 1e-10)) continue;
    dMplanet = dPxPlanet = dPyPlanet = 0.0;
    facc = fabs(dt) * (sys->acc[k]);
    facc1 = 1.0 / 3.0 * facc;
    facc2 = 2.0 / 3.0 * facc;
    frac1 = 0.75;
    frac2 = 0.45;
    facc1 = facc1 < 1.0 ? facc1 : 1.0;
    facc2 = facc2 < 1.0 ? facc2 : 1.0;
    Xplanet = sys->x[k]; Yplanet = sys->y[k]; Zplanet = sys->z[k];
    VXplanet = sys->vx[k]; VYplanet = sys->vy[k]; Mplanet = sys->mass[k];
    Rperp = sqrt(Xplanet * Xplanet + Yplanet * Yplanet);
    Rplanet = sqrt(Xplanet * Xplanet + Yplanet * Yplanet + Zplanet * Zplanet);
    RRoche = pow((1.0 / 3.0 * Mplanet * o->st.mass_taper), (1.0 / 3.0)) * Rplanet;
    i_min = 0;
    i_max = nr - 1;
    while (i_min < nr && o->Rsup[i_min] < Rperp - RRoche) i_min++;
    while ((o->Rinf[i_max] > Rperp + RRoche) && (i_max > 0)) i_max--;
    angle = atan2(Yplanet, Xplanet);
    j_min = (int)((double)ns / 2.0 / PI * (angle - 2.0 * RRoche / Rperp));
    j_max = (int)((double)ns / 2.0 / PI * (angle + 2.0 * RRoche / Rperp));
    PxPlanet = Mplanet * VXplanet;
    PyPlanet = Mplanet * VYplanet;
    for (i = i_min; i <= i_max; i++)
      for (j = j_min; j <= j_max; j++) {
        double xc, yc, dx, dy, distance, vtcell, vrcell, vxcell, vycell, H, dz, z1, z2, factor, deltaM;
        jf = j;
        while (jf < 0) jf += ns;
        while (jf >= ns) jf -= ns;
        l = jf + i * ns;
        lip = l + ns;
        ljp = l + 1;
        if (jf == ns - 1) ljp = i * ns;
        xc = abs_[l];
        yc = ord[l];
        dx = Xplanet - xc;
        dy = Yplanet - yc;
        distance = sqrt(dx * dx + dy * dy);
        vtcell = 0.5 * (vtheta[l] + vtheta[ljp]) + o->Rmed[i] * o->st.omega_frame;
        vrcell = 0.5 * (vrad[l] + vrad[lip]);
        vxcell = (vrcell * xc - vtcell * yc) / o->Rmed[i];
        vycell = (vrcell * yc + vtcell * xc) / o->Rmed[i];
        H = cs[l] * o->OmegaInv[i] * o->sqrt_adiabind_inv;
        if (distance < frac1 * RRoche) {
          dz = sqrt(pow(frac1 * RRoche, 2) - pow(distance, 2));
          z1 = Zplanet - dz;
          z2 = Zplanet + dz;
          factor = o_gas_qtrap(z1 / (sqrt(2.0) * H), z2 / (sqrt(2.0) * H));
          factor *= 1.0 / (sqrt(2.0 * PI) * H);
          deltaM = facc1 * factor * dens[l] * o->Surf[i];
          dens[l] *= (1.0 - facc1 * factor);
          dPxPlanet += deltaM * vxcell;
          dPyPlanet += deltaM * vycell;
          dMplanet += deltaM;
        }
        if (distance < frac2 * RRoche) {
          dz = sqrt(pow(frac2 * RRoche, 2) - pow(distance, 2));
          z1 = Zplanet - dz;
          z2 = Zplanet + dz;
          factor = o_gas_qtrap(z1 / (sqrt(2.0) * H), z2 / (sqrt(2.0) * H));
          factor *= 1.0 / (sqrt(2.0 * PI) * H);
          deltaM = facc2 * factor * dens[l] * o->Surf[i];
          dens[l] *= (1.0 - facc2);
          dPxPlanet += deltaM * vxcell;
          dPyPlanet += deltaM * vycell;
          dMplanet += deltaM;
        }
      }
    if (p->gas_accret_heating && k < 8) o_heating1(o, Mplanet, dMplanet, Xplanet, Yplanet, dt, k);
    PxPlanet += dPxPlanet;
    PyPlanet += dPyPlanet;
    Mplanet += dMplanet;
    if (sys->feeldisk[k]) {
      sys->vx[k] = PxPlanet / Mplanet;
      sys->vy[k] = PyPlanet / Mplanet;
    }
    sys->mass[k] = Mplanet;
  }
  (void)p;
}

/* AccretePebblesOntoPlanets, Pebbles.c:346-538 (single rank; the
 * MPI_Allreduce sums are the local sums).  Returns -1 if a column
 * integration does not converge. */
static int o_peb_accrete(oracle_ctx *o, thorin_planets *sys, double dt) {
  int nr = o->nr, ns = o->ns, k, i, j, l, imin, imax, ncell, ok = 1;
  double *pdens = o->fld[THORIN_F_PEBBLEDENS], *macc = o->fld[W_MACC];
  const double *abs_ = o->fld[W_ABS], *ord = o->fld[W_ORD], *pvr = o->fld[THORIN_F_PEBBLEVRAD];
  const double *pvt = o->fld[THORIN_F_PEBBLEVTHETA], *tau = o->fld[THORIN_F_PEBBLESTOKES];
  const double *cs = o->fld[THORIN_F_SOUNDSPEED];
  const double of = o->st.omega_frame;
  double dMplanet, dPXplanet, dPYplanet, Xplanet, Yplanet, Zplanet, VXplanet, VYplanet, VZplanet;
  double Mplanet, Msolid, PXplanet, PYplanet, Rplanet, Rplanet3D, Omegaplanet, Rhill;
  double vtcorr, xcell, ycell, vrcell, vtcell, vxcell, vycell, flow, flowx, flowy, vrel;
  double ifrac, frac, Mt, Rbondi, tbondi, fac, Reff = 0.0, Reff2, R2, H, Hpeb, zmin, zmax, dM, mcell;
  double voldenspeb, tau_mean, dM_upper, Hpeb_mean, pdens_mean, veloc, dM_model, limiter;
  if (o->p.accret_heating) o->heatsrc_max = sys->nb;   /* Pebbles.c:379 */
  o_ring_mean(o, o->fld[THORIN_F_VTHETA], o->vt1D);
  for (k = 0; k < sys->nb; k++) {
    dMplanet = dPXplanet = dPYplanet = 0.0;
    Xplanet = sys->x[k]; Yplanet = sys->y[k]; Zplanet = sys->z[k];
    VXplanet = sys->vx[k]; VYplanet = sys->vy[k]; VZplanet = sys->vz[k];
    Mplanet = sys->mass[k];
    Msolid = sys->solid[k];
    PXplanet = VXplanet * Mplanet;
    PYplanet = VYplanet * Mplanet;
    Rplanet = Xplanet * Xplanet + Yplanet * Yplanet;
    Rplanet3D = sqrt(Rplanet + Zplanet * Zplanet);
    Rplanet = sqrt(Rplanet);
    Rhill = Rplanet3D * pow(1.0 / 3.0 * Mplanet, 1.0 / 3.0);
    Omegaplanet = pow(Rplanet, -1.5);
    ifrac = o_global_ifrac(o, Rplanet);
    frac = ifrac - floor(ifrac);
    i = (int)ifrac;
    flow = o->vt1D[i] * (1.0 - frac) + o->vt1D[i + 1] * frac + Rplanet * of;
    flowx = -flow * Yplanet / Rplanet;
    flowy = flow * Xplanet / Rplanet;
    vrel = sqrt((VXplanet - flowx) * (VXplanet - flowx) + (VYplanet - flowy) * (VYplanet - flowy) + VZplanet * VZplanet);
    Mt = sqrt(1.0 / 3.0) * pow(vrel, 3.0) / Omegaplanet;
    imax = nr - 1;
    imin = 0;
    while (o->Rinf[imax] > (Rplanet + Rhill) && imax >= imin) imax--;
    while (o->Rsup[imin] < (Rplanet - Rhill) && imin <= imax) imin++;
    /* step 1: mean Stokes number over the annulus of the Hill sphere */
    ncell = 0;
    tau_mean = 0.0;
    for (i = imin; i <= imax; i++)
      for (j = 0; j < ns; j++) { tau_mean += tau[j + i * ns]; ncell++; }
    tau_mean /= (double)ncell;
    /* step 2: effective accretion radius */
    if (Mplanet < Mt) {
      Rbondi = Mplanet / vrel / vrel;
      tbondi = Rbondi / vrel;
      fac = sqrt(tau_mean * pow(Rplanet, 1.5) / tbondi);
      Reff = Rbondi * fac;
    } else {
      fac = pow(tau_mean / 0.1, 1.0 / 3.0);
      Reff = Rhill * fac;
    }
    if (Reff > Rhill) Reff = Rhill;
    Reff2 = Reff * Reff;
    /* step 3: upper limit of the accreted mass, cell by cell */
    dM_upper = 0.0;
    Hpeb_mean = 0.0;
    pdens_mean = 0.0;
    ncell = 0;
    for (i = imin; i <= imax; i++) {
      vtcorr = o->Rmed[i] * of;
      (void)vtcorr;
      for (j = 0; j < ns; j++) {
        l = j + i * ns;
        xcell = abs_[l];
        ycell = ord[l];
        R2 = (xcell - Xplanet) * (xcell - Xplanet) + (ycell - Yplanet) * (ycell - Yplanet);
        if (R2 <= Reff2) {
          H = cs[l] * o->OmegaInv[i] * o->sqrt_adiabind_inv;
          Hpeb = H * sqrt(o->p.pebblealpha / tau[l]);
          zmin = Zplanet - sqrt(Reff2 - R2);
          zmax = Zplanet + sqrt(Reff2 - R2);
          mcell = o->Surf[i] * pdens[l];
          voldenspeb = pdens[l] / Hpeb * SQRT2PI_INV;
          dM = voldenspeb * o->Surf[i] * o_column_mass(zmin, zmax, Hpeb, &ok);
          if (!ok) return -1;
          if (dM > mcell) dM = 0.99999999 * mcell;
          macc[l] = dM;
          dM_upper += dM;
          Hpeb_mean += Hpeb;
          pdens_mean += pdens[l];
          ncell++;
        }
      }
    }
    if (ncell > 0) {
      Hpeb_mean /= (double)ncell;
      pdens_mean /= (double)ncell;
    }
    /* step 4: model accretion rate (2D/3D, Bondi/Hill) limits the upper bound */
    if (Mplanet < Mt) veloc = vrel;
    else veloc = Reff * Omegaplanet;
    if (Hpeb_mean < Reff) dM_model = 2.0 * Reff * veloc * pdens_mean;
    else dM_model = PI * Reff * Reff * veloc * pdens_mean / Hpeb_mean * SQRT2PI_INV;
    dM_model *= dt;
    if (dM_upper > dM_model) limiter = dM_model / dM_upper;
    else limiter = 1.0;
    /* step 5: remove the mass from the cells, collect mass and momentum */
    for (i = imin; i <= imax; i++) {
      vtcorr = o->Rmed[i] * of;
      for (j = 0; j < ns; j++) {
        l = j + i * ns;
        if (macc[l] > 0.0) {
          xcell = abs_[l];
          ycell = ord[l];
          vrcell = pvr[l];
          vtcell = pvt[l] + vtcorr;
          vxcell = (vrcell * xcell - vtcell * ycell) * o->InvRmed[i];
          vycell = (vrcell * ycell + vtcell * xcell) * o->InvRmed[i];
          dM = macc[l] * limiter;
          macc[l] = 0.0;
          mcell = o->Surf[i] * pdens[l];
          pdens[l] = (mcell - dM) / o->Surf[i];
          if (pdens[l] < 0.0) pdens[l] = 1.e-20;
          dMplanet += dM;
          dPXplanet += dM * vxcell;
          dPYplanet += dM * vycell;
        }
      }
    }
    /* optional step 6: accretion heating (Pebbles.c:525-528) */
    if (o->p.accret_heating && k < 8) o_heating1(o, Mplanet, dMplanet, Xplanet, Yplanet, dt, k);
    /* step 7: update the planet */
    PXplanet += dPXplanet;
    PYplanet += dPYplanet;
    Mplanet += dMplanet;
    Msolid += dMplanet;
    sys->mass[k] = Mplanet;
    sys->solid[k] = Msolid;
    sys->vx[k] = PXplanet / Mplanet;
    sys->vy[k] = PYplanet / Mplanet;
  }
  return 0;
}

/* EvaporatePebbles, Pebbles.c:837-866 */
static void o_peb_evaporate(oracle_ctx *o, double dt) {
  int i, j, l;
  double facc, evaptemp, *dens = o->fld[THORIN_F_PEBBLEDENS];
  const double *temp = o->fld[THORIN_F_TEMPER];
  if (o->p.evaporationrate <= 0.0) return;
  facc = o->p.evaporationrate * dt;
  if (facc > 1.0) facc = 1.0;
  evaptemp = o->p.evaporationtemperature / T2SI;
  for (i = 0; i < o->nr; i++)
    for (j = 0; j < o->ns; j++) {
      l = j + i * o->ns;
      if (temp[l] > evaptemp) {
        dens[l] *= (1.0 - facc);
        if (dens[l] < EVAPOREDSIGMA) dens[l] = EVAPOREDSIGMA;
      }
    }
}

/* DampPebbles, SideEuler.c:423-454 */
static void o_peb_damp(oracle_ctx *o, double dt) {
  int i, j, l, ns = o->ns;
  double DRin = o->p.rmin * o->p.dampingrminfrac, DRout = o->p.rmax * o->p.dampingrmaxfrac, lambda;
  double pdens0, pvr0, pvt0;
  double *pdens = o->fld[THORIN_F_PEBBLEDENS], *pvr = o->fld[THORIN_F_PEBBLEVRAD], *pvt = o->fld[THORIN_F_PEBBLEVTHETA];
  for (i = 0; i < o->nr; i++) {
    if (o->Rmed[i] < DRin || o->Rmed[i] > DRout) {
      lambda = o->WaveKiller[i] * dt;
      pdens0 = o->PebDensInit[i];
      pvr0 = o->PebVradInit[i];
      pvt0 = o->PebVthetaInit[i] - o->Rmed[i] * o->st.omega_frame;
      for (j = 0; j < ns; j++) {
        l = j + i * ns;
        pdens[l] = (pdens[l] + lambda * pdens0) / (1.0 + lambda);
        pvr[l] = (pvr[l] + lambda * pvr0) / (1.0 + lambda);
        pvt[l] = (pvt[l] + lambda * pvt0) / (1.0 + lambda);
      }
    }
  }
}

/* SourceTermsPebbles, Pebbles.c:549-612 */
static void o_peb_source_terms(oracle_ctx *o, double dt) {
  int i, j, l, lim, ljm, ljp, nr = o->nr, ns = o->ns;
  const double *pvr = o->fld[THORIN_F_PEBBLEVRAD], *pvt = o->fld[THORIN_F_PEBBLEVTHETA];
  const double *pdens = o->fld[THORIN_F_PEBBLEDENS], *tau = o->fld[THORIN_F_PEBBLESTOKES];
  const double *vr = o->fld[THORIN_F_VRAD], *vt = o->fld[THORIN_F_VTHETA];
  const double *pagr = o->fld[W_PAGR], *pagt = o->fld[W_PAGT];
  double *paccr = o->fld[W_PACCR], *pacct = o->fld[W_PACCT], *dragr = o->fld[W_DRAGR], *dragt = o->fld[W_DRAGT];
  double Omega, vt2, drag;
  o_peb_damp(o, dt);
  for (i = 1; i < nr; i++) {
    Omega = 1.0 / o->OmegaInv[i];
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      vt2 = pvt[l] + pvt[ljp] + pvt[lim] + pvt[ljp - ns];
      vt2 = vt2 / 4.0 + o->Rinf[i] * o->st.omega_frame;
      vt2 = vt2 * vt2;
      paccr[l] = vt2 * o->InvRinf[i] + 0.5 * (pagr[l] + pagr[lim]);
      if (o->p.back_reaction) {
        drag = Omega * (pvr[l] - vr[l]) / tau[l];
        dragr[l] = 0.5 * (pdens[l] + pdens[lim]) * drag;
      }
    }
  }
  for (i = 0; i < nr; i++) {
    Omega = 1.0 / o->OmegaInv[i];
    for (j = 0; j < ns; j++) {
      l = j + i * ns; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      pacct[l] = 0.5 * (pagt[l] + pagt[ljm]);
      if (o->p.back_reaction) {
        drag = Omega * (pvt[l] - vt[l]) / tau[l];
        dragt[l] = 0.5 * (pdens[l] + pdens[ljm]) * drag;
      }
    }
  }
}

/* SubStep1Pebbles, Pebbles.c:616-650 (semi-implicit drag step; vr, vt are
 * the gas velocities before SubStep1's update) */
static void o_peb_substep1(oracle_ctx *o, double dt) {
  int i, j, l, nr = o->nr, ns = o->ns;
  double *pvr = o->fld[THORIN_F_PEBBLEVRAD], *pvt = o->fld[THORIN_F_PEBBLEVTHETA];
  const double *tau = o->fld[THORIN_F_PEBBLESTOKES], *paccr = o->fld[W_PACCR], *pacct = o->fld[W_PACCT];
  const double *accr = o->fld[W_GACCR], *acct = o->fld[W_GACCT];
  const double *vr = o->fld[THORIN_F_VRAD], *vt = o->fld[THORIN_F_VTHETA];
  double Omega, efac;
  for (i = 0; i < nr; i++) {
    Omega = 1.0 / o->OmegaInv[i];
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      efac = exp(-dt * Omega / tau[l]);
      if (i > 0)
        pvr[l] = pvr[l] * efac + accr[l] * dt + (vr[l] + (paccr[l] - accr[l]) * tau[l] / Omega) * (1.0 - efac);
      pvt[l] = pvt[l] * efac + acct[l] * dt + (vt[l] + (pacct[l] - acct[l]) * tau[l] / Omega) * (1.0 - efac);
    }
  }
}

/* ParticleDiffusion, Pebbles.c:654-695 */
static void o_peb_diffusion(oracle_ctx *o) {
  int i, j, l, lim, ljm, nr = o->nr, ns = o->ns;
  double corr, dxtheta, invdxtheta;
  const double *rho = o->fld[THORIN_F_DENS], *pdens = o->fld[THORIN_F_PEBBLEDENS];
  double *pvr = o->fld[THORIN_F_PEBBLEVRAD], *pvt = o->fld[THORIN_F_PEBBLEVTHETA];
  for (i = 1; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns;
      corr = -o_fviscosity(o, o->Rinf[i]) / o->p.schmidtnumber;
      corr *= (rho[l] + rho[lim]) / (pdens[l] + pdens[lim]);
      corr *= (pdens[l] / rho[l] - pdens[lim] / rho[lim]) * o->InvDiffRmed[i];
      pvr[l] += corr;
    }
  for (i = 0; i < nr; i++) {
    dxtheta = 2.0 * PI / (double)ns * o->Rmed[i];
    invdxtheta = 1.0 / dxtheta;
    for (j = 0; j < ns; j++) {
      l = j + i * ns; ljm = l - 1;
      if (j == 0) ljm = i * ns + ns - 1;
      corr = -o_fviscosity(o, o->Rmed[i]) / o->p.schmidtnumber;
      corr *= (rho[l] + rho[ljm]) / (pdens[l] + pdens[ljm]);
      corr *= (pdens[l] / rho[l] - pdens[ljm] / rho[ljm]) * invdxtheta;
      pvt[l] += corr;
    }
  }
}

/* EvolvePebbleDisk, Pebbles.c:699-708 (TransportPebbles: TransportEuler.c:449-505) */
static void o_peb_evolve(oracle_ctx *o, double dt) {
  o_peb_damp(o, dt);
  o_transport(o, o->fld[THORIN_F_PEBBLEDENS], o->fld[THORIN_F_PEBBLEVRAD], o->fld[THORIN_F_PEBBLEVTHETA], NULL, dt);
  o_peb_damp(o, dt);
}

/* EtaPressureSupport (Pebbles.c:236-270) and InitPebblesViaFlux
 * (Pebbles.c:79-202): the coagulation-drift equilibrium pebble disk of a
 * fresh run (EquilPebbleDisk, Pebbles.c:59-64) */
static void o_peb_equilibrium(oracle_ctx *o) {
  const thorin_params *p = &o->p;
  int i, j, l, lim, ljp, nr = o->nr, ns = o->ns;
  double *etafc = o->fld[W_WORK], *etacc = o->fld[THORIN_F_ETA];
  const double *vt = o->fld[THORIN_F_VTHETA], *rho = o->fld[THORIN_F_DENS], *vr = o->fld[THORIN_F_VRAD];
  const double *cs = o->fld[THORIN_F_SOUNDSPEED];
  double *pdens = o->fld[THORIN_F_PEBBLEDENS], *tau = o->fld[THORIN_F_PEBBLESTOKES];
  double *psize = o->fld[THORIN_F_PEBBLESIZE], *pvr = o->fld[THORIN_F_PEBBLEVRAD];
  double *pvt = o->fld[THORIN_F_PEBBLEVTHETA];
  const double of = o->st.omega_frame, pebbleflux = p->pebbleflux * FLUX2CU, Ndf = 0.5;
  double vtcorr, vkinv, vtcc;
  for (i = 0; i < nr; i++) {
    vtcorr = o->Rmed[i] * of;
    vkinv = sqrt(o->Rmed[i]);
    for (j = 0; j < ns; j++) {
      l = j + i * ns; ljp = l + 1;
      if (j == ns - 1) ljp = i * ns;
      vtcc = 0.5 * (vt[l] + vt[ljp]);
      etacc[l] = 1.0 - (vtcc + vtcorr) * vkinv;
    }
  }
  for (i = 1; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns; lim = l - ns;
      etafc[l] = 0.5 * (etacc[l] + etacc[lim]);
    }
  for (j = 0; j < ns; j++) etafc[j] = etafc[j + ns];
  for (i = 0; i < nr; i++) {
    double rho1 = 0.0, vr1 = 0.0, etafc1 = 0.0, etacc1 = 0.0, cs1 = 0.0;
    double tau1, psize1, taudf, adf, taufrag, afrag, vkepl, fac, etaterm, pvr1, pvrtemp, pvt1, pdens1;
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      rho1 += rho[l]; vr1 += vr[l]; etafc1 += etafc[l]; etacc1 += etacc[l]; cs1 += cs[l];
    }
    rho1 /= (double)ns; vr1 /= (double)ns; etafc1 /= (double)ns; etacc1 /= (double)ns; cs1 /= (double)ns;
    tau1 = sqrt(3.0) * p->pebblecoagulation * pebbleflux / (32.0 * PI * pow(o->Rmed[i], -1.5) * rho1);
    tau1 = sqrt(tau1) / (o->Rmed[i] * etacc1);
    psize1 = fabs(tau1) * sqrt(p->adiabind) * rho1 / (2.0 * o->pebbulkdens);
    taudf = p->fragmentthreshold * sqrt(o->Rmed[i]) / (2.0 * etacc1 * (1.0 - Ndf));
    adf = fabs(taudf) * sqrt(p->adiabind) * rho1 / (2.0 * o->pebbulkdens);
    taufrag = p->fragmentfactor / (3.0 * p->fragmentalpha) * pow(p->fragmentthreshold / cs1, 2);
    afrag = fabs(taufrag) * sqrt(p->adiabind) * rho1 / (2.0 * o->pebbulkdens);
    if ((psize1 < adf) && (psize1 < afrag)) {
    } else if ((taudf > 0.0) && (adf < psize1) && (adf < afrag)) {
      tau1 = taudf; psize1 = adf;
    } else {
      tau1 = taufrag; psize1 = afrag;
    }
    vtcorr = o->Rmed[i] * of;
    vkepl = pow(o->Rmed[i], -0.5);
    fac = 1.0 / (tau1 * tau1 + 1.0);
    etaterm = etafc1 * vkepl;
    pvr1 = -2.0 * tau1 * fac * (etaterm - 0.5 / tau1 * vr1);
    etaterm = etacc1 * vkepl;
    pvrtemp = -2.0 * tau1 * fac * (etaterm - 0.5 / tau1 * vr1);
    pvt1 = -fac * (etaterm - 0.5 * tau1 * vr1);
    pvt1 += (vkepl - vtcorr);
    pdens1 = pebbleflux / (2.0 * PI * o->Rmed[i] * fabs(pvrtemp));
    if (pdens1 / rho1 > p->pebbletogasmax) pdens1 = rho1 * p->pebbletogasmax;
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      psize[l] = psize1; tau[l] = tau1; pdens[l] = pdens1; pvr[l] = pvr1; pvt[l] = pvt1;
    }
    o->PebDensInit[i] = pdens1;
    o->PebVradInit[i] = pvr1;
    o->PebVthetaInit[i] = pvt1 + o->Rmed[i] * of;
  }
}

/* BckpFieldsForBC, Pebbles.c:206-232 (damping targets of a restarted run) */
static void o_peb_backup(oracle_ctx *o) {
  int i, j, l, ns = o->ns;
  const double *pdens = o->fld[THORIN_F_PEBBLEDENS], *pvr = o->fld[THORIN_F_PEBBLEVRAD];
  const double *pvt = o->fld[THORIN_F_PEBBLEVTHETA];
  for (i = 0; i < o->nr; i++) {
    o->PebDensInit[i] = 0.0; o->PebVradInit[i] = 0.0; o->PebVthetaInit[i] = 0.0;
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      o->PebDensInit[i] += pdens[l]; o->PebVradInit[i] += pvr[l]; o->PebVthetaInit[i] += pvt[l];
    }
    o->PebDensInit[i] /= (double)ns;
    o->PebVradInit[i] /= (double)ns;
    o->PebVthetaInit[i] /= (double)ns;
    o->PebVthetaInit[i] += o->Rmed[i] * o->st.omega_frame;
  }
}

/* Pebble disk set-up (Init.c:138-144): mode 0 = equilibrium disk of a fresh
 * run from the current gas state; mode 1 = restart from the given pebble
 * density, velocities and size (RestartPebbleDisk, Pebbles.c:67-75). */
int oracle_init_pebbles(oracle_ctx *o, int mode, const double *pdens, const double *pvrad,
                        const double *pvtheta, const double *psize) {
  size_t n = (size_t)o->nr * o->ns;
  if (!o->p.pebbles) return -1;
  if (mode == 0) { o_peb_equilibrium(o); return 0; }
  if (mode != 1 || !pdens || !pvrad || !pvtheta || !psize) return -1;
  memcpy(o->fld[THORIN_F_PEBBLEDENS], pdens, n * sizeof(double));
  memcpy(o->fld[THORIN_F_PEBBLEVRAD], pvrad, n * sizeof(double));
  memcpy(o->fld[THORIN_F_PEBBLEVTHETA], pvtheta, n * sizeof(double));
  memcpy(o->fld[THORIN_F_PEBBLESIZE], psize, n * sizeof(double));
  o_peb_backup(o);
  return 0;
}

/* ---------------- AlgoGas (SourceEuler.c:210-317) ---------------- */
int oracle_algogas(oracle_ctx *o, thorin_planets *sys, thorin_nbody_fn nbody, void *user,
                   thorin_step_log *log) {
  const thorin_params *p = &o->p;
  double dt, dtemp = 0.0, DT = p->dt, newdt, omega_used = 0.0;
  int N, crashed, niter, rc = 0;
  if (log) memset(log, 0, sizeof(*log));
  dt = DT / 1.0;
  while (dtemp < 0.999999999 * DT) {
    double mt = o->st.physical_time / (p->masstaper * 2.0 * M_PI);
    o->st.mass_taper = (mt > 1.0 ? 1.0 : pow(sin(mt * M_PI / 2.0), 2.0));
    if (p->pebbles) o_peb_stokes(o);
    o->invdtreb_sq = o_minstep_rebound(sys);
    if (p->pebbles) o_peb_cfl(o);
    N = o_condition_cfl_core(o, DT - dtemp, &newdt);
    if (N < 0) return -5;
    dt = (DT - dtemp) / (double)N;
    {   /* InitHeating, Planet.c:251-258 */
      int k;
      for (k = 0; k < 8; k++) { o->heatsrc[k] = 0.0; o->heatsrc_index[k] = -1; }
    }
    o_gas_accrete(o, sys, dt);
    if (p->pebbles) {
      if (o_peb_accrete(o, sys, dt)) return -4;
      o_peb_evaporate(o, dt);
    }
    dtemp += dt;
    o_fill_forces(o, sys);
    o_advance_from_disk(o, sys, dt);
    if (nbody) nbody(user, sys, o->st.physical_time, dt, o->st.omega_frame * dt);
    o_boundary(o, dt);
    crashed = o_detect_crash(o, o->fld[THORIN_F_DENS]);
    crashed = o_detect_crash(o, o->fld[THORIN_F_ENERGY]);
    if (p->pebbles) crashed = o_detect_crash(o, o->fld[THORIN_F_PEBBLEDENS]);
    if (crashed) {
      if (o->st.already_crashed == 0) {   /* WriteDiskPolar (..., 999), SourceEuler.c:276-282 */
        const int src[4] = {THORIN_F_DENS, THORIN_F_VRAD, THORIN_F_VTHETA, THORIN_F_TEMPER};
        size_t n = (size_t)(o->nr + 1) * o->ns;
        int k;
        for (k = 0; k < 4; k++) {
          if (!o->crash[k]) o->crash[k] = (double *)malloc(n * sizeof(double));
          memcpy(o->crash[k], o->fld[src[k]], n * sizeof(double));
        }
        if (log) { log->crash_dumped = 1; log->crash_time = o->st.physical_time; }
      }
      o->st.already_crashed++;
    }
    o_pressure(o, o->fld[THORIN_F_DENS], o->fld[THORIN_F_ENERGY]);
    if (p->pebbles) o_peb_source_terms(o, dt);
    o_substep1(o, dt);
    if (p->pebbles) {
      o_peb_substep1(o, dt);
      if (p->diffusive_particles) o_peb_diffusion(o);
    }
    o_substep2(o, dt);
    o_actualise(o, o->fld[THORIN_F_VRAD], o->fld[THORIN_F_VRADNEW]);
    o_actualise(o, o->fld[THORIN_F_VTHETA], o->fld[THORIN_F_VTHETANEW]);
    o_boundary(o, dt);
    niter = 0;
    if (p->energy_eq) {
      o_div_stress(o, o->fld[THORIN_F_VRAD], o->fld[THORIN_F_VTHETA], o->fld[THORIN_F_DENS]);
      niter = o_substep3(o, dt, &omega_used);
      if (niter < 0) rc = -3;
      o_actualise(o, o->fld[THORIN_F_ENERGY], o->fld[THORIN_F_ENERGYNEW]);
    }
    o_transport(o, o->fld[THORIN_F_DENS], o->fld[THORIN_F_VRAD], o->fld[THORIN_F_VTHETA],
                p->energy_eq ? o->fld[THORIN_F_ENERGY] : NULL, dt);
    o_boundary(o, dt);
    o_temperature(o, o->fld[THORIN_F_DENS], o->fld[THORIN_F_ENERGY]);
    if (p->energy_eq) o_sound_speed(o, o->fld[THORIN_F_DENS], o->fld[THORIN_F_ENERGY]);
    if (p->pebbles) o_peb_evolve(o, dt);
    o->st.physical_time += dt;
    if (log && log->nsub < THORIN_MAXSUB) {
      log->cfl_n[log->nsub] = N;
      log->dt[log->nsub] = dt;
      log->sor_niter[log->nsub] = niter;
      log->sor_omega[log->nsub] = omega_used;
      log->crashed[log->nsub] = crashed;
      log->crashed_any |= crashed;
      log->nsub++;
    }
    if (rc) return rc;
  }
  return 0;
}

/* ---------------- accessors ---------------- */
int oracle_get_field(oracle_ctx *o, int field, double *out, size_t count) {
  size_t n = (size_t)(o->nr + 1) * o->ns;
  if (field >= THORIN_X_CRASH_DENS && field <= THORIN_X_CRASH_TEMPER) {
    if (!o->crash[field - THORIN_X_CRASH_DENS]) return -1;
    if (count > n) count = n;
    memcpy(out, o->crash[field - THORIN_X_CRASH_DENS], count * sizeof(double));
    return 0;
  }
  if (field < 0 || field >= NW) return -1;
  if (count > n) count = n;
  memcpy(out, o->fld[field], count * sizeof(double));
  return 0;
}

int oracle_set_field(oracle_ctx *o, int field, const double *in, size_t count) {
  size_t n = (size_t)(o->nr + 1) * o->ns;
  if (field < 0 || field >= NW) return -1;
  if (count > n) count = n;
  memcpy(o->fld[field], in, count * sizeof(double));
  return 0;
}

int oracle_get_state(oracle_ctx *o, thorin_state *s) { *s = o->st; return 0; }
int oracle_set_state(oracle_ctx *o, const thorin_state *s) { o->st = *s; return 0; }

int oracle_get_1d(oracle_ctx *o, int which, double *out, size_t count) {
  size_t k, n = (size_t)o->nr + 1;
  const double *src = NULL;
  if (count > n) count = n;
  switch (which) {
    case 0: src = o->Radii; break;
    case 1: src = o->Rinf; break;
    case 2: src = o->Rsup; break;
    case 3: src = o->Rmed; break;
    case 4: src = o->Surf; break;
    case 5: src = o->InvDiffRmed; break;
    case 6: src = o->SigmaMed; break;
    case 7: src = o->SigmaInf; break;
    case 8: src = o->EnergyMed; break;
    case 9: src = o->WaveKiller; break;
    case 10: src = o->globcsvec; break;
    case 11: src = o->globtempervec; break;
    case 12: src = o->VthetaMed; break;
    case 13: for (k = 0; k < count; k++) out[k] = (double)o->Nshift[k]; return 0;
    case 14: src = o->VMed; break;
    case 15: src = o->PebDensInit; break;
    case 16: src = o->PebVradInit; break;
    case 17: src = o->PebVthetaInit; break;
    default: return -1;
  }
  memcpy(out, src, count * sizeof(double));
  return 0;
}

/* GasTotalMass, SideEuler.c:26-51 */
double oracle_gas_total_mass(oracle_ctx *o) {
  int i, j;
  double total = 0.0;
  const double *d = o->fld[THORIN_F_DENS];
  for (i = 0; i < o->nr; i++)
    for (j = 0; j < o->ns; j++) total += o->Surf[i] * d[j + i * o->ns];
  return total;
}
