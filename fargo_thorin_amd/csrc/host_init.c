/*
 * host_init.c -- context creation and initialisation on the host.
 *
 * Geometry: FillPolar1DArrays (SourceEuler.c:68-131).  Profiles: Theo.c.
 * Initial state: InitEuler (SourceEuler.c:133-175), InitGasDensityEnergy and
 * InitGasVelocity (Pframeforce.c:620-751), InitLabel (Init.c:56-70), the
 * InitializeFromFile and restart routes of Initialization (Init.c:72-145),
 * InitComputeAccel (SideEuler.c:101-118), SetWaveKillingZones
 * (SideEuler.c:341-362).  Compiled with gcc -std=c99 like the reference, so
 * every initial array is bit-identical to the reference's.
 */
#define _GNU_SOURCE
#include "thorin_internal.h"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* one error text per thread: the ensemble steps each disk on a thread of
 * its own, and copies a failing disk's text to the caller's thread */
static __thread char g_err[512];
void thorin_set_error(const char *msg) {
  strncpy(g_err, msg, sizeof(g_err) - 1);
  g_err[sizeof(g_err) - 1] = 0;
}
void thorin_set_errorf(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
const char *thorin_last_error(void) { return g_err; }

static double *alloc1(int n) { return (double *)calloc((size_t)n, sizeof(double)); }

/* FViscosity (Viscosity.c:26-54), host copy used at initialisation */
double thorin_fviscosity_host(const thorin_ctx *c, double rad) {
  const thorin_params *p = &c->p;
  double viscosity, rmin, rmax, scale, alpha;
  int i = 0;
  viscosity = p->viscosity;
  if (p->viscosity_alpha) {
    while (c->GlobalRmed[i] < rad) i++;
    viscosity = p->alphaviscosity * c->globcsvec[i] * c->globcsvec[i] * pow(rad, 1.5);
  }
  if (p->alpha_flock_on) {
    while (c->GlobalRmed[i] < rad) i++;
    alpha = (p->alphamri - p->alphadead) * (1.0 - tanh((p->tmri - c->globtempervec[i] * TH_T2SI) / p->twidth)) / 2.0 + p->alphadead;
    viscosity = alpha * c->globcsvec[i] * c->globcsvec[i] * pow(rad, 1.5);
  }
  rmin = p->cavityradius - p->cavitywidth * p->aspectratio;
  rmax = p->cavityradius + p->cavitywidth * p->aspectratio;
  scale = 1.0 + (c->st.physical_time - c->st.physical_time_initial) * p->lambdadoubling;
  rmin *= scale;
  rmax *= scale;
  if (rad < rmin) viscosity *= p->cavityratio;
  if ((rad >= rmin) && (rad <= rmax)) viscosity *= exp((rmax - rad) / (rmax - rmin) * log(p->cavityratio));
  return viscosity;
}

/* AspectRatio (Viscosity.c:56-71) */
double thorin_aspectratio_host(const thorin_ctx *c, double rad) {
  const thorin_params *p = &c->p;
  double ar = p->aspectratio, rmin, rmax, scale;
  rmin = p->transitionradius - p->transitionwidth * p->aspectratio;
  rmax = p->transitionradius + p->transitionwidth * p->aspectratio;
  scale = 1.0 + (c->st.physical_time - c->st.physical_time_initial) * p->lambdadoubling;
  rmin *= scale;
  rmax *= scale;
  if (rad < rmin) ar *= p->transitionratio;
  if ((rad >= rmin) && (rad <= rmax)) ar *= exp((rmax - rad) / (rmax - rmin) * log(p->transitionratio));
  return ar;
}

/* The cavity part of FViscosity is a per-ring multiplier (1.0 outside the
 * cavity, so that multiplying by it is exact); it only changes with time when
 * LambdaDoubling != 0, in which case the device path refreshes it per substep. */
void thorin_cavity_factors(thorin_ctx *c) {
  const thorin_params *p = &c->p;
  double rmin = p->cavityradius - p->cavitywidth * p->aspectratio;
  double rmax = p->cavityradius + p->cavitywidth * p->aspectratio;
  double scale = 1.0 + (c->st.physical_time - c->st.physical_time_initial) * p->lambdadoubling;
  int i, w;
  rmin *= scale;
  rmax *= scale;
  for (w = 0; w < 2; w++) {
    double *dst = w ? c->cav_inf : c->cav_med;
    const double *r = w ? c->Rinf : c->Rmed;
    for (i = 0; i <= c->nr; i++) {
      double rad = r[i], f = 1.0;
      if (rad < rmin) f = p->cavityratio;
      if ((rad >= rmin) && (rad <= rmax)) f = exp((rmax - rad) / (rmax - rmin) * log(p->cavityratio));
      dst[i] = f;
    }
  }
}

/* FillPolar1DArrays, SourceEuler.c:68-131 */
static void fill_1d(thorin_ctx *c) {
  int nr = c->nr, ns = c->ns, i;
  double drrsep = (c->p.rmax - c->p.rmin) / (double)nr;
  char name[2 * THORIN_STRLEN];
  FILE *in;
  snprintf(name, sizeof name, "%s%s", c->p.outputdir, "radii.dat");
  in = fopen(name, "r");
  if (in == NULL) {
    if (c->p.log_grid)
      for (i = 0; i <= nr; i++) c->Radii[i] = c->p.rmin * exp((double)i / (double)nr * log(c->p.rmax / c->p.rmin));
    else
      for (i = 0; i <= nr; i++) c->Radii[i] = c->p.rmin + drrsep * (double)(i);
  } else {
    float t;
    for (i = 0; i <= nr; i++) { if (fscanf(in, "%f", &t) != 1) t = 0.f; c->Radii[i] = (double)t; }
    fclose(in);
  }
  for (i = 0; i < nr; i++) {
    c->GlobalRmed[i] = 2.0 / 3.0 * (c->Radii[i + 1] * c->Radii[i + 1] * c->Radii[i + 1] - c->Radii[i] * c->Radii[i] * c->Radii[i]);
    c->GlobalRmed[i] = c->GlobalRmed[i] / (c->Radii[i + 1] * c->Radii[i + 1] - c->Radii[i] * c->Radii[i]);
  }
  for (i = 0; i < nr; i++) {
    c->Rinf[i] = c->Radii[i];
    c->Rsup[i] = c->Radii[i + 1];
    c->Rmed[i] = 2.0 / 3.0 * (c->Rsup[i] * c->Rsup[i] * c->Rsup[i] - c->Rinf[i] * c->Rinf[i] * c->Rinf[i]);
    c->Rmed[i] = c->Rmed[i] / (c->Rsup[i] * c->Rsup[i] - c->Rinf[i] * c->Rinf[i]);
    c->Surf[i] = TH_PI * (c->Rsup[i] * c->Rsup[i] - c->Rinf[i] * c->Rinf[i]) / (double)ns;
    c->InvRmed[i] = 1.0 / c->Rmed[i];
    c->InvSurf[i] = 1.0 / c->Surf[i];
    c->InvDiffRsup[i] = 1.0 / (c->Rsup[i] - c->Rinf[i]);
    c->InvRinf[i] = 1.0 / c->Rinf[i];
    c->OmegaInv[i] = pow(c->Rmed[i], 1.5);
    c->Rmed2[i] = c->Rmed[i] * c->Rmed[i];
  }
  c->Rinf[nr] = c->Radii[nr];
  for (i = 1; i < nr; i++) c->InvDiffRmed[i] = 1.0 / (c->Rmed[i] - c->Rmed[i - 1]);
  /* per-ring factors for the kernels */
  for (i = 0; i <= nr; i++) {
    c->pow_rmed15[i] = pow(c->Rmed[i], 1.5);
    c->pow_rinf15[i] = pow(c->Rinf[i], 1.5);
  }
  for (i = 0; i < nr; i++) {
    double dxrad = c->Rsup[i] - c->Rinf[i], dxtheta = c->Rmed[i] * 2.0 * TH_PI / (double)ns;
    double m = (dxtheta < dxrad) ? dxtheta : dxrad;   /* min2, SourceEuler.c:177-182 */
    c->cfl_pow2[i] = pow(m, 2);
    c->supp_torque[i] = c->p.imposeddiskdrift * .5 * pow(c->Rmed[i], -2.5 + c->p.sigmaslope);
  }
}

/* Theo.c:15-23 */
static double sigma_profile(const thorin_ctx *c, double r) {
  double cavity = 1.0;
  if (r < c->p.cavityradius) cavity = 1.0 / c->p.cavityratio;
  return cavity * 1.0 * c->p.sigma0 * pow(r, -c->p.sigmaslope);   /* ScalingFactor = 1 */
}
/* Theo.c:60-74 */
static double energy_profile(const thorin_ctx *c, double r) {
  double fac = 1.0 / (1.0 * (c->p.adiabind - 1.0));
  double power = -c->p.sigmaslope - 1.0 + 2.0 * c->p.flaringindex;
  return 1.0 * c->p.sigma0 * pow(c->p.aspectratio, 2.0) * pow(r, power) * fac;
}

static void ring_mean(const thorin_ctx *c, const double *f, double *out) {   /* mpiTasks.c:12-36 */
  int i, j;
  for (i = 0; i < c->nr; i++) {
    double s = 0.;
    for (j = 0; j < c->ns; j++) s += f[i * c->ns + j];
    s /= (double)c->ns;
    out[i] = s;
  }
}

/* ComputeSoundSpeed / ComputePressureField / ComputeTemperatureField,
 * SourceEuler.c:715-792, on the host staging arrays */
static void host_eos(thorin_ctx *c) {
  int i, j, l, ns = c->ns;
  double g = c->p.adiabind, tmp1, tmp2;
  for (i = 0; i < c->nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      if (!c->p.energy_eq)
        c->h_cs[l] = thorin_aspectratio_host(c, c->Rmed[i]) * sqrt(1.0 * 1.0 / c->Rmed[i]) * pow(c->Rmed[i], c->p.flaringindex);
      else
        c->h_cs[l] = sqrt(g * (g - 1.0) * fabs(c->h_energy[l]) / c->h_dens[l]);
    }
  ring_mean(c, c->h_cs, c->globcsvec);
  for (i = 0; i < c->nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      if (!c->p.energy_eq) c->h_press[l] = c->h_dens[l] * c->h_cs[l] * c->h_cs[l];
      else c->h_press[l] = (g - 1.0) * c->h_energy[l];
    }
  for (i = 0; i < c->nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      tmp2 = 1.0 / (c->h_dens[l] * 1.0);
      if (c->p.energy_eq) tmp1 = c->h_energy[l] * 1.0 * (g - 1.0);
      else tmp1 = 1.0 * c->h_press[l];
      c->h_temper[l] = tmp1 * tmp2;
    }
  ring_mean(c, c->h_temper, c->globtempervec);
}

static int ctx_create(const thorin_params *p, thorin_ctx **out, int with_device);

int thorin_ctx_create(const thorin_params *p, thorin_ctx **out) { return ctx_create(p, out, 1); }

/* Host-only context: geometry and initial state without device buffers
 * (for inspection of the initialisation on a machine without a GPU); every
 * device entry point rejects such a context. */
int thorin_ctx_create_host(const thorin_params *p, thorin_ctx **out) { return ctx_create(p, out, 0); }

static int ctx_create(const thorin_params *p, thorin_ctx **out, int with_device) {
  thorin_ctx *c;
  int n, i, j;
  if (!p || !out) { thorin_set_error("null argument"); return -1; }
  if (p->nrad < 8 || p->nsec < 4 || (p->nsec & 1)) {
    thorin_set_error("NRAD must be >= 8 and NSEC even and >= 4 (red-black SOR needs an even NSEC)");
    return -1;
  }
  if (p->nrad + 2 > 16384) { thorin_set_error("NRAD exceeds MAX1D (types.h:65)"); return -1; }
  if (p->open_inner || p->open_outer || p->non_reflecting || p->outer_source_mass) {
    thorin_set_error("OPEN / NONREFLECTING boundaries and OuterSourceMass are not supported by this build");
    return -1;
  }
  if (p->corotating) { thorin_set_error("Frame COROTATING/GUIDING-CENTER is not supported by this build"); return -1; }
  if (p->parametric_cooling) { thorin_set_error("CoolingTime > 0 (parametric cooling) is not supported by this build"); return -1; }
  if (p->advect_label) { thorin_set_error("LabelAdvection is not supported by this build"); return -1; }
  if (!p->energy_eq) { thorin_set_error("EnergyEquation NO (isothermal) is not supported by this build"); return -1; }
  if (p->prescribed_accretion || p->aerodynamic_drag || p->torque_density) {
    thorin_set_error("prescribed accretion / aerodynamic drag / torque density are not supported by this build");
    return -1;
  }
  c = (thorin_ctx *)calloc(1, sizeof(thorin_ctx));
  if (!c) { thorin_set_error("out of host memory"); return -1; }
  c->p = *p;
  c->nr = p->nrad;
  c->ns = p->nsec;
  c->nfield = (size_t)(c->nr + 1) * c->ns;
  n = c->nr + 3;
  c->Radii = alloc1(n); c->Rinf = alloc1(n); c->Rsup = alloc1(n); c->Rmed = alloc1(n); c->Surf = alloc1(n);
  c->InvRmed = alloc1(n); c->InvSurf = alloc1(n); c->InvDiffRmed = alloc1(n); c->InvDiffRsup = alloc1(n);
  c->InvRinf = alloc1(n); c->OmegaInv = alloc1(n); c->Rmed2 = alloc1(n); c->GlobalRmed = alloc1(n);
  c->SigmaMed = alloc1(n); c->SigmaInf = alloc1(n); c->EnergyMed = alloc1(n); c->VthetaMed = alloc1(n);
  c->WaveKiller = alloc1(n); c->globcsvec = alloc1(n); c->globtempervec = alloc1(n);
  c->pow_rmed15 = alloc1(n); c->pow_rinf15 = alloc1(n); c->cav_med = alloc1(n); c->cav_inf = alloc1(n);
  c->cfl_pow2 = alloc1(n); c->supp_torque = alloc1(n);
  c->h_dens = alloc1((int)c->nfield); c->h_vrad = alloc1((int)c->nfield); c->h_vtheta = alloc1((int)c->nfield);
  c->h_energy = alloc1((int)c->nfield); c->h_label = alloc1((int)c->nfield); c->h_cs = alloc1((int)c->nfield);
  c->h_press = alloc1((int)c->nfield); c->h_temper = alloc1((int)c->nfield);
  c->h_abs = alloc1((int)c->nfield); c->h_ord = alloc1((int)c->nfield);
  c->h_trig = alloc1(2 * c->ns);
  if (!c->h_dens || !c->h_vrad || !c->h_vtheta || !c->h_energy || !c->h_label || !c->h_cs || !c->h_press ||
      !c->h_temper || !c->h_abs || !c->h_ord || !c->h_trig) {
    thorin_set_error("out of host memory");
    thorin_ctx_destroy(c);
    return -1;
  }
  c->st.dt_stellar = 1.0e38;
  c->st.omega_frame = p->omegaframe;
  c->st.sor_first = 1;
  fill_1d(c);
  thorin_cavity_factors(c);
  /* constants */
  c->k.opa2cu = TH_OPA2CU;
  c->k.t2si = TH_T2SI;
  c->k.rho2cgs = TH_RHO2CGS;
  c->k.stefan = TH_STEFANBOLTZMANN;
  c->k.sqrt2pi_inv = TH_SQRT2PI_INV;
  c->k.sqrt3 = sqrt(3.0);
  c->k.sqrt_adiabind_inv = 1.0 / sqrt(p->adiabind);
  c->k.cv = 1.0 / (1.0 * (p->adiabind - 1.0));
  c->k.lr01 = log(c->Rmed[0] / c->Rmed[1]);
  c->k.lr12 = log(c->Rmed[1] / c->Rmed[2]);
  /* FViscosity's radial search (Viscosity.c:35) must land on the ring itself,
   * which the kernels assume when they index the ring profiles directly */
  for (i = 0; i < c->nr; i++) {
    int k1 = 0, k2 = 0;
    while (c->GlobalRmed[k1] < c->Rmed[i]) k1++;
    if (i > 0) while (c->GlobalRmed[k2] < c->Rinf[i]) k2++;
    if (k1 != i || (i > 0 && k2 != i)) {
      thorin_set_error("radial grid: GlobalRmed search does not return the ring index");
      thorin_ctx_destroy(c);
      return -1;
    }
  }
  /* InitComputeAccel, SideEuler.c:101-118; the device forms Rmed[i] times
     these per-sector values, the same products */
  for (j = 0; j < c->ns; j++) {
    c->h_trig[j] = cos(2.0 * TH_PI * (double)j / (double)c->ns);
    c->h_trig[c->ns + j] = sin(2.0 * TH_PI * (double)j / (double)c->ns);
  }
  for (i = 0; i < c->nr; i++)
    for (j = 0; j < c->ns; j++) {
      c->h_abs[j + i * c->ns] = c->Rmed[i] * cos(2.0 * TH_PI * (double)j / (double)c->ns);
      c->h_ord[j + i * c->ns] = c->Rmed[i] * sin(2.0 * TH_PI * (double)j / (double)c->ns);
    }
  /* SetWaveKillingZones, SideEuler.c:341-362 */
  if (p->damping_on || p->pebbles) {
    double DRin = p->rmin * p->dampingrminfrac, DRout = p->rmax * p->dampingrmaxfrac, damp;
    double tauin = p->dampingperiodfrac * 2.0 * TH_PI * pow(p->rmin, 1.5);
    double tauout = p->dampingperiodfrac * 2.0 * TH_PI * pow(DRout, 1.5);
    for (i = 0; i < c->nr; i++) {
      c->WaveKiller[i] = 0.0;
      if (c->Rmed[i] < DRin) { damp = (DRin - c->Rmed[i]) / (DRin - p->rmin); c->WaveKiller[i] = damp * damp / tauin; }
      if (c->Rmed[i] > DRout) { damp = (c->Rmed[i] - DRout) / (p->rmax - DRout); c->WaveKiller[i] = damp / tauout; }
    }
  }
  /* InitOpacities (Opacity2.c:252-300): the gas table of OpacityTable SM03,
   * INPUTDIR/data_gasopacity_malygin.txt; only its Rosseland part is used */
  if (p->opacity_law == THORIN_OPA_SM03) {
    char name[2 * THORIN_STRLEN];
    FILE *in;
    int ok, k;
    double v;
    snprintf(name, sizeof name, "%s%s", p->inputdir, "data_gasopacity_malygin.txt");
    in = fopen(name, "r");
    if (!in) {
      char m[2 * THORIN_STRLEN + 64];
      snprintf(m, sizeof m, "OpacityTable SM03: opacity file '%s' not found", name);
      thorin_set_error(m);
      thorin_ctx_destroy(c);
      return -1;
    }
    ok = fscanf(in, "%i", &c->sm_rows) == 1 && fscanf(in, "%i", &c->sm_cols) == 1 && c->sm_rows > 1 &&
         c->sm_cols > 2 && c->sm_rows <= 4096 && c->sm_cols <= 4096;
    if (ok) {
      c->sm_pgrid = alloc1(c->sm_rows);
      c->sm_tgrid = alloc1(c->sm_cols);
      c->sm_ross = alloc1(c->sm_rows * c->sm_cols);
      ok = c->sm_pgrid && c->sm_tgrid && c->sm_ross;
      for (k = 0; ok && k < c->sm_rows; k++) ok = fscanf(in, "%lf", &c->sm_pgrid[k]) == 1;
      for (k = 0; ok && k < c->sm_cols; k++) ok = fscanf(in, "%lf", &c->sm_tgrid[k]) == 1;
      for (k = 0; ok && k < c->sm_rows * c->sm_cols; k++) ok = fscanf(in, "%lf", &v) == 1;   /* Planck */
      for (k = 0; ok && k < c->sm_rows * c->sm_cols; k++) ok = fscanf(in, "%lf", &c->sm_ross[k]) == 1;
    }
    fclose(in);
    if (!ok) {
      thorin_set_error("OpacityTable SM03: the opacity file is short or malformed");
      thorin_ctx_destroy(c);
      return -1;
    }
  }
  if (with_device && thorin_dev_create(c) != 0) { thorin_ctx_destroy(c); return -1; }
  *out = c;
  return 0;
}

int thorin_ctx_destroy(thorin_ctx *c) {
  if (!c) return 0;
  thorin_dev_destroy(c);
  free(c->Radii); free(c->Rinf); free(c->Rsup); free(c->Rmed); free(c->Surf); free(c->InvRmed);
  free(c->InvSurf); free(c->InvDiffRmed); free(c->InvDiffRsup); free(c->InvRinf); free(c->OmegaInv);
  free(c->Rmed2); free(c->GlobalRmed); free(c->SigmaMed); free(c->SigmaInf); free(c->EnergyMed);
  free(c->VthetaMed); free(c->WaveKiller); free(c->globcsvec); free(c->globtempervec);
  free(c->h_pdens); free(c->h_pvr); free(c->h_pvt); free(c->h_psize); free(c->h_tau); free(c->h_eta);
  free(c->PebDensInit); free(c->PebVradInit); free(c->PebVthetaInit);
  free(c->pow_rmed15); free(c->pow_rinf15); free(c->cav_med); free(c->cav_inf); free(c->cfl_pow2);
  free(c->supp_torque);
  free(c->h_dens); free(c->h_vrad); free(c->h_vtheta); free(c->h_energy); free(c->h_label);
  free(c->h_cs); free(c->h_press); free(c->h_temper); free(c->h_abs); free(c->h_ord); free(c->h_trig);
  free(c->sm_pgrid); free(c->sm_tgrid); free(c->sm_ross);
  free(c);
  return 0;
}

/* InitGasVelocity, Pframeforce.c:644-751 (CentrifugalBalance is a CLI
 * option of the reference driver, -b, not exposed here) */
static void init_gas_velocity(thorin_ctx *c) {
  int i, j, l, nr = c->nr, ns = c->ns;
  double r, rg, omega, ri, viscosity, rhoinf;
  double *vr = c->h_vrad, *vt = c->h_vtheta, *cs = c->h_cs;
  for (i = 0; i <= nr; i++) {
    if (i == nr) { r = c->Rmed[nr - 1]; ri = c->Rinf[nr - 1]; }
    else { r = c->Rmed[i]; ri = c->Rinf[i]; }
    viscosity = thorin_fviscosity_host(c, r);
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      rg = r;
      omega = sqrt(1.0 * 1.0 / rg / rg / rg);
      vt[l] = omega * r * sqrt(1.0 - cs[l] * cs[l] * r / c->p.adiabind);
      vt[l] -= c->st.omega_frame * r;
      vr[l] = c->p.imposeddiskdrift * c->p.sigma0 / c->SigmaInf[i] / ri;
      if (c->p.damp_flux) {
        if (i == 0) rhoinf = c->SigmaMed[i] - 0.5 * (c->SigmaMed[i + 1] - c->SigmaMed[i]);
        else rhoinf = 0.5 * (c->SigmaMed[i - 1] + c->SigmaMed[i]);
        vr[l] -= c->p.gasflux / (2.0 * TH_PI * ri * rhoinf);
      } else if (c->p.viscosity_alpha) {
        vr[l] -= 3.0 * viscosity / r * (-c->p.sigmaslope + 2.0 * c->p.flaringindex + 1.0);
      } else {
        vr[l] -= 3.0 * viscosity / r * (-c->p.sigmaslope + .5);
      }
    }
  }
  if (!c->p.damp_flux)
    for (j = 0; j < ns; j++) vr[j] = vr[j + ns * nr] = 0.0;
}

/* RefillSigma / RefillEnergy / FillVtheta (Theo.c:33-57, 82-100, 132-150) */
static void refill_profiles(thorin_ctx *c) {
  int i, j, nr = c->nr, ns = c->ns;
  for (i = 0; i < nr; i++) {
    double moy = 0.0;
    for (j = 0; j < ns; j++) moy += c->h_dens[j + i * ns];
    moy /= (double)ns;
    c->SigmaMed[i] = moy;
  }
  c->SigmaInf[0] = c->SigmaMed[0];
  for (i = 1; i < nr; i++)
    c->SigmaInf[i] = (c->SigmaMed[i - 1] * (c->Rmed[i] - c->Rinf[i]) + c->SigmaMed[i] * (c->Rinf[i] - c->Rmed[i - 1])) /
                     (c->Rmed[i] - c->Rmed[i - 1]);
  if (c->p.energy_eq)
    for (i = 0; i < nr; i++) {
      double moy = 0.0;
      for (j = 0; j < ns; j++) moy += c->h_energy[j + i * ns];
      moy /= (double)ns;
      c->EnergyMed[i] = moy;
    }
}

static void fill_vtheta_med(thorin_ctx *c) {
  int i, j, ns = c->ns;
  for (i = 0; i < c->nr; i++) {
    double moy = 0.0, vtcorr = c->Rmed[i] * c->st.omega_frame;
    for (j = 0; j < ns; j++) moy += c->h_vtheta[j + i * ns] + vtcorr;
    moy /= (double)ns;
    c->VthetaMed[i] = moy;
  }
}

/* Initialization, Init.c:72-145.  Modes 0 and 1 make the initial state
 * (InitEuler, or the InitializeFromFile arrays with RefillSigma, RefillEnergy
 * and FillVtheta).  Mode 2 is the restart route: the reference forms the
 * damping profiles from that initial state first and only then reads the
 * restart files into the fields (Init.c:106-126), so mode 2 keeps the
 * profiles of the last mode 0/1 state (making the analytic one first when
 * there is none; an analytic state is refilled as the reference refills
 * it) and replaces the fields. */
static void init_fields_from(thorin_ctx *c, const double *dens, const double *vrad, const double *vtheta,
                             const double *temper) {
  int i, j, l, nr = c->nr, ns = c->ns;
  size_t ncell = (size_t)nr * ns;
  memcpy(c->h_dens, dens, sizeof(double) * ncell);
  memcpy(c->h_vrad, vrad, sizeof(double) * ncell);
  memcpy(c->h_vtheta, vtheta, sizeof(double) * ncell);
  memcpy(c->h_energy, temper, sizeof(double) * ncell);
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      c->h_energy[l] = c->h_energy[l] * c->h_dens[l] / (c->p.adiabind - 1.0);
    }
}

int thorin_init_state(thorin_ctx *c, int mode, const double *dens, const double *vrad,
                      const double *vtheta, const double *temper) {
  int i, j, l, nr = c->nr, ns = c->ns;
  if (mode < 0 || mode > 2) { thorin_set_error("thorin_init_state: mode must be 0, 1 or 2"); return -1; }
  if (mode != 0 && (!dens || !vrad || !vtheta || !temper)) {
    thorin_set_error("thorin_init_state: mode 1/2 needs dens, vrad, vtheta and temper arrays");
    return -1;
  }
  if (mode == 2) {
    if (c->t0_mode != 2) {   /* RefillSigma/RefillEnergy/FillVtheta of the analytic state */
      if (thorin_init_state(c, 0, NULL, NULL, NULL, NULL)) return -1;
      refill_profiles(c);
      if (c->p.damp_init) fill_vtheta_med(c);
    }
    init_fields_from(c, dens, vrad, vtheta, temper);
    host_eos(c);
    if (c->dev && thorin_dev_upload_init(c) != 0) return -1;
    return 0;
  }
  memset(c->h_dens, 0, sizeof(double) * c->nfield);
  memset(c->h_vrad, 0, sizeof(double) * c->nfield);
  memset(c->h_vtheta, 0, sizeof(double) * c->nfield);
  memset(c->h_energy, 0, sizeof(double) * c->nfield);
  memset(c->h_cs, 0, sizeof(double) * c->nfield);
  memset(c->h_press, 0, sizeof(double) * c->nfield);
  memset(c->h_temper, 0, sizeof(double) * c->nfield);
  /* InitEuler, SourceEuler.c:133-175 */
  for (i = 0; i < nr; i++) {
    c->SigmaMed[i] = sigma_profile(c, c->Rmed[i]);
    c->SigmaInf[i] = sigma_profile(c, c->Rinf[i]);
  }
  for (i = 0; i < nr; i++) c->EnergyMed[i] = energy_profile(c, c->Rmed[i]);
  for (i = 0; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      c->h_dens[l] = c->SigmaMed[i];
      c->h_energy[l] = c->EnergyMed[i];
    }
  host_eos(c);
  init_gas_velocity(c);
  if (c->p.damp_init) fill_vtheta_med(c);
  /* InitLabel, Init.c:56-70 */
  for (i = 0; i <= nr; i++)
    for (j = 0; j < ns; j++) c->h_label[j + i * ns] = (c->Rmed[i] - c->p.rmin) / (c->p.rmax - c->p.rmin);
  if (mode == 1) {
    init_fields_from(c, dens, vrad, vtheta, temper);
    refill_profiles(c);
    if (c->p.damp_init) fill_vtheta_med(c);
    host_eos(c);
  }
  c->t0_mode = mode + 1;
  /* main.c:168-170: cs once more after InitComputeAccel (same values) */
  if (c->dev && thorin_dev_upload_init(c) != 0) return -1;
  return 0;
}

/* ---------------- pebble fluid set-up (Init.c:138-144) ---------------- */
#define TH_FLUX2CU (0.000003 / (2.0 * TH_PI))   /* fondam.h:54 */

/* EtaPressureSupport (Pebbles.c:236-270) + InitPebblesViaFlux
 * (Pebbles.c:79-202): coagulation-drift equilibrium of a fresh run, from the
 * initial gas state */
static void peb_equilibrium(thorin_ctx *c) {
  const thorin_params *p = &c->p;
  int i, j, l, lim, ljp, nr = c->nr, ns = c->ns;
  double *etacc = c->h_eta, *etafc = (double *)calloc(c->nfield, sizeof(double));
  const double *vt = c->h_vtheta, *rho = c->h_dens, *vr = c->h_vrad, *cs = c->h_cs;
  const double of = c->st.omega_frame, pebbleflux = p->pebbleflux * TH_FLUX2CU, Ndf = 0.5;
  const double pebbulkdens = p->pebblebulkdens / TH_RHO2CGS;
  for (i = 0; i < nr; i++) {
    double vtcorr = c->Rmed[i] * of, vkinv = sqrt(c->Rmed[i]);
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      ljp = (j == ns - 1) ? i * ns : l + 1;
      etacc[l] = 1.0 - (0.5 * (vt[l] + vt[ljp]) + vtcorr) * vkinv;
    }
  }
  for (i = 1; i < nr; i++)
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      lim = l - ns;
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
    /* drift-limited, differential-drift and fragmentation-limited sizes */
    tau1 = sqrt(3.0) * p->pebblecoagulation * pebbleflux / (32.0 * TH_PI * pow(c->Rmed[i], -1.5) * rho1);
    tau1 = sqrt(tau1) / (c->Rmed[i] * etacc1);
    psize1 = fabs(tau1) * sqrt(p->adiabind) * rho1 / (2.0 * pebbulkdens);
    taudf = p->fragmentthreshold * sqrt(c->Rmed[i]) / (2.0 * etacc1 * (1.0 - Ndf));
    adf = fabs(taudf) * sqrt(p->adiabind) * rho1 / (2.0 * pebbulkdens);
    taufrag = p->fragmentfactor / (3.0 * p->fragmentalpha) * pow(p->fragmentthreshold / cs1, 2);
    afrag = fabs(taufrag) * sqrt(p->adiabind) * rho1 / (2.0 * pebbulkdens);
    if (!((psize1 < adf) && (psize1 < afrag))) {
      if ((taudf > 0.0) && (adf < psize1) && (adf < afrag)) { tau1 = taudf; psize1 = adf; }
      else { tau1 = taufrag; psize1 = afrag; }
    }
    /* steady drift velocities; the density from the imposed mass flux */
    vkepl = pow(c->Rmed[i], -0.5);
    fac = 1.0 / (tau1 * tau1 + 1.0);
    etaterm = etafc1 * vkepl;
    pvr1 = -2.0 * tau1 * fac * (etaterm - 0.5 / tau1 * vr1);
    etaterm = etacc1 * vkepl;
    pvrtemp = -2.0 * tau1 * fac * (etaterm - 0.5 / tau1 * vr1);
    pvt1 = -fac * (etaterm - 0.5 * tau1 * vr1);
    pvt1 += (vkepl - c->Rmed[i] * of);
    pdens1 = pebbleflux / (2.0 * TH_PI * c->Rmed[i] * fabs(pvrtemp));
    if (pdens1 / rho1 > p->pebbletogasmax) pdens1 = rho1 * p->pebbletogasmax;
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      c->h_psize[l] = psize1; c->h_tau[l] = tau1; c->h_pdens[l] = pdens1; c->h_pvr[l] = pvr1; c->h_pvt[l] = pvt1;
    }
    c->PebDensInit[i] = pdens1;
    c->PebVradInit[i] = pvr1;
    c->PebVthetaInit[i] = pvt1 + c->Rmed[i] * of;
  }
  free(etafc);
}

/* BckpFieldsForBC (Pebbles.c:206-232): ordered ring means as damping targets */
static void peb_backup(thorin_ctx *c) {
  int i, j, l, ns = c->ns;
  for (i = 0; i < c->nr; i++) {
    c->PebDensInit[i] = 0.0; c->PebVradInit[i] = 0.0; c->PebVthetaInit[i] = 0.0;
    for (j = 0; j < ns; j++) {
      l = j + i * ns;
      c->PebDensInit[i] += c->h_pdens[l]; c->PebVradInit[i] += c->h_pvr[l]; c->PebVthetaInit[i] += c->h_pvt[l];
    }
    c->PebDensInit[i] /= (double)ns;
    c->PebVradInit[i] /= (double)ns;
    c->PebVthetaInit[i] /= (double)ns;
    c->PebVthetaInit[i] += c->Rmed[i] * c->st.omega_frame;
  }
}

int thorin_init_pebbles(thorin_ctx *c, int mode, const double *pdens, const double *pvrad,
                        const double *pvtheta, const double *psize) {
  size_t ncell;
  if (!c) { thorin_set_error("null context"); return -1; }
  if (!c->p.pebbles) { thorin_set_error("thorin_init_pebbles: PebbleAccretion is NO in the parameters"); return -1; }
  if (mode != 0 && mode != 1) { thorin_set_error("thorin_init_pebbles: mode must be 0 or 1"); return -1; }
  if (mode == 1 && (!pdens || !pvrad || !pvtheta || !psize)) {
    thorin_set_error("thorin_init_pebbles: mode 1 needs pdens, pvrad, pvtheta and psize arrays");
    return -1;
  }
  ncell = (size_t)c->nr * c->ns;
  if (!c->h_pdens) {
    c->h_pdens = alloc1((int)c->nfield); c->h_pvr = alloc1((int)c->nfield); c->h_pvt = alloc1((int)c->nfield);
    c->h_psize = alloc1((int)c->nfield); c->h_tau = alloc1((int)c->nfield); c->h_eta = alloc1((int)c->nfield);
    c->PebDensInit = alloc1(c->nr + 3); c->PebVradInit = alloc1(c->nr + 3); c->PebVthetaInit = alloc1(c->nr + 3);
    if (!c->h_pdens || !c->h_pvr || !c->h_pvt || !c->h_psize || !c->h_tau || !c->h_eta || !c->PebDensInit ||
        !c->PebVradInit || !c->PebVthetaInit) {
      thorin_set_error("out of host memory");
      return -1;
    }
  }
  if (mode == 0) {
    peb_equilibrium(c);
  } else {
    memcpy(c->h_pdens, pdens, sizeof(double) * ncell);
    memcpy(c->h_pvr, pvrad, sizeof(double) * ncell);
    memcpy(c->h_pvt, pvtheta, sizeof(double) * ncell);
    memcpy(c->h_psize, psize, sizeof(double) * ncell);
    memset(c->h_tau, 0, sizeof(double) * c->nfield);
    peb_backup(c);
  }
  c->peb_ready = 1;
  if (c->dev && thorin_dev_upload_pebbles(c) != 0) return -1;
  return 0;
}

/* initial host copy of a field (valid after thorin_init_state) */
int thorin_host_field(thorin_ctx *c, int field, double *out, size_t count) {
  const double *src = NULL;
  if (count > c->nfield) count = c->nfield;
  switch (field) {
    case THORIN_F_DENS: src = c->h_dens; break;
    case THORIN_F_VRAD: src = c->h_vrad; break;
    case THORIN_F_VTHETA: src = c->h_vtheta; break;
    case THORIN_F_ENERGY: src = c->h_energy; break;
    case THORIN_F_LABEL: src = c->h_label; break;
    case THORIN_F_TEMPER: src = c->h_temper; break;
    case THORIN_F_PRESSURE: src = c->h_press; break;
    case THORIN_F_SOUNDSPEED: src = c->h_cs; break;
    case THORIN_F_PEBBLEDENS: src = c->h_pdens; break;
    case THORIN_F_PEBBLEVRAD: src = c->h_pvr; break;
    case THORIN_F_PEBBLEVTHETA: src = c->h_pvt; break;
    case THORIN_F_PEBBLESIZE: src = c->h_psize; break;
    case THORIN_F_PEBBLESTOKES: src = c->h_tau; break;
    case THORIN_F_ETA: src = c->h_eta; break;
    default: thorin_set_error("no host copy of this field"); return -1;
  }
  if (!src) { thorin_set_error("no host copy of this field (pebbles not set up)"); return -1; }
  memcpy(out, src, sizeof(double) * count);
  return 0;
}

int thorin_get_geometry(thorin_ctx *c, thorin_geometry *g) {
  g->nrad = c->nr;
  g->nsec = c->ns;
  g->radii = c->Radii; g->rinf = c->Rinf; g->rsup = c->Rsup; g->rmed = c->Rmed; g->surf = c->Surf;
  g->inv_rmed = c->InvRmed; g->inv_surf = c->InvSurf; g->inv_diff_rmed = c->InvDiffRmed;
  g->inv_diff_rsup = c->InvDiffRsup; g->inv_rinf = c->InvRinf; g->omega_inv = c->OmegaInv;
  g->rmed2 = c->Rmed2; g->sigma_med = c->SigmaMed; g->sigma_inf = c->SigmaInf;
  g->energy_med = c->EnergyMed; g->vtheta_med = c->VthetaMed; g->wave_killer = c->WaveKiller;
  g->globcsvec = c->globcsvec; g->globtempervec = c->globtempervec;
  return 0;
}

int thorin_get_state(thorin_ctx *c, thorin_state *s) { *s = c->st; return 0; }
int thorin_set_state(thorin_ctx *c, const thorin_state *s) { c->st = *s; return 0; }
