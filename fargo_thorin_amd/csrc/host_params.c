/*
 * host_params.c -- in.par reader with the reference's semantics.
 *
 * Registry and defaults: var.c:17-140.  Parsing: Interpret.c:65-148 (first
 * token is the key, case-insensitive; the value starts after any of
 * "\t :=>_"; '#' starts a comment; a value that does not parse as a number
 * keeps the previous line's number, exactly as the reference's sscanf into
 * the same variable does).  Derived flags: Interpret.c:149-265.
 */
#define _GNU_SOURCE
#include "thorin_internal.h"
#include <ctype.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { P_INT = 0, P_REAL = 1, P_STRING = 2 };

typedef struct {
  const char *name;
  int type;
  size_t off;
  size_t size;        /* capacity for strings */
  int necessary;
  const char *deflt;
} pdesc;

#define R(key, fld, nec, d) {key, P_REAL, offsetof(thorin_params, fld), 0, nec, d}
#define I(key, fld, nec, d) {key, P_INT, offsetof(thorin_params, fld), 0, nec, d}
#define S(key, fld, nec, d) {key, P_STRING, offsetof(thorin_params, fld), sizeof(((thorin_params *)0)->fld), nec, d}

/* var.c:19-139, same order */
static const pdesc registry[] = {
  R("DT", dt, 1, "1."), R("SIGMA0", sigma0, 1, "173."), I("NINTERM", ninterm, 1, "10."),
  I("NTOT", ntot, 1, "1501."), S("OUTPUTDIR", outputdir, 1, "out"), S("INPUTDIR", inputdir, 0, "in"),
  S("INNERBOUNDARY", innerboundary, 0, "WALL"), S("OUTERBOUNDARY", outerboundary, 0, "WALL"),
  S("LABELADVECTION", labeladvection, 0, "NO"), S("TRANSPORT", transport, 0, "FAST"),
  S("PLANETCONFIG", planetconfig, 0, "Systems/SolarSystem.cfg"), R("MASSTAPER", masstaper, 0, "0.0000001"),
  S("RADIALSPACING", gridspacing, 0, "ARITHMETIC"), I("NRAD", nrad, 1, "64.0"), I("NSEC", nsec, 1, "64.0"),
  R("RMIN", rmin, 1, "1.0"), R("RMAX", rmax, 1, "1.0"), R("THICKNESSSMOOTHING", thicknesssmoothing, 0, "0.0"),
  R("ROCHESMOOTHING", rochesmoothing, 0, "0.0"), R("ASPECTRATIO", aspectratio, 1, "0.05"),
  R("VISCOSITY", viscosity, 0, "0.0"), R("ALPHAVISCOSITY", alphaviscosity, 0, "0.0"),
  S("ALPHAFLOCK", alphaflock, 0, "NO"), R("TMRI", tmri, 0, "1000.0"), R("TWIDTH", twidth, 0, "25.0"),
  R("ALPHAMRI", alphamri, 0, "0.0"), R("ALPHADEAD", alphadead, 0, "0.0"), R("SIGMASLOPE", sigmaslope, 1, "0.0"),
  R("RELEASERADIUS", releaseradius, 0, "0.0"), R("RELEASEDATE", releasedate, 0, "0.0"),
  R("OMEGAFRAME", omegaframe, 0, "0.0"), S("DISK", disk, 0, "YES"), S("FRAME", frame, 0, "FIXED"),
  S("OUTERSOURCEMASS", outersourcemass, 0, "NO"), S("WRITEDENSITY", writedensity, 0, "YES"),
  S("WRITEVELOCITY", writevelocity, 0, "YES"), S("INDIRECTTERM", indirectterm, 0, "YES"),
  S("EXCLUDEHILL", excludehill, 0, "NO"), R("IMPOSEDDISKDRIFT", imposeddiskdrift, 0, "0.0"),
  R("FLARINGINDEX", flaringindex, 0, "0.0"), R("ECCENTRICITY", eccentricity, 0, "0.0"),
  R("CAVITYRADIUS", cavityradius, 0, "0.0"), R("CAVITYRATIO", cavityratio, 0, "1.0"),
  R("CAVITYWIDTH", cavitywidth, 0, "1.0"), R("TRANSITIONRADIUS", transitionradius, 0, "0.0"),
  R("TRANSITIONRATIO", transitionratio, 0, "1.0"), R("TRANSITIONWIDTH", transitionwidth, 0, "1.0"),
  R("LAMBDADOUBLING", lambdadoubling, 0, "0.0"),
  S("ENERGYEQUATION", energyequation, 0, "NO"), S("WRITETEMPERATURE", writetemperature, 0, "NO"),
  S("WRITEENERGY", writeenergy, 0, "NO"), S("WRITEDIVV", writedivv, 0, "NO"), S("WRITEQPLUS", writeqplus, 0, "NO"),
  S("WRITEQBALANCE", writeqbalance, 0, "NO"), S("WRITEOPACITY", writeopacity, 0, "NO"),
  R("ADIABIND", adiabind, 0, "1.4"), R("COOLINGTIME", coolingtime, 0, "-1.0"),
  S("STELLARIRRADIATION", stellarirradiation, 0, "NO"), R("OPACITYDROP", opacitydrop, 0, "0.6"),
  R("EFFECTIVETEMPERATURE", effectivetemperature, 0, "5656.0"), R("STELLARRADIUS", stellarradius, 0, "3.0"),
  R("DISCALBEDO", discalbedo, 0, "0.5"), R("PARAMETRICOPACITY", parametricopacity, 0, "0.0"),
  S("INITIALIZEFROMFILE", initializefromfile, 0, "NO"), S("DENSINFILE", densinfile, 0, "in/gasdens.cfg"),
  S("VRADINFILE", vradinfile, 0, "in/gasvrad.cfg"), S("VTHETAINFILE", vthetainfile, 0, "in/gasvtheta.cfg"),
  S("TEMPERINFILE", temperinfile, 0, "in/gastemper.cfg"),
  S("DAMPTOWARDS", damptowards, 0, "ZEROVRAD"), R("DAMPINGRMINFRAC", dampingrminfrac, 0, "1.25"),
  R("DAMPINGRMAXFRAC", dampingrmaxfrac, 0, "0.84"), R("DAMPINGPERIODFRAC", dampingperiodfrac, 0, "1.0"),
  I("NOUTELEMENTS", noutelements, 0, "1"), R("PLANETARYDENSITY", planetarydensity, 0, "1.0"),
  S("RESOLVECOLLISIONS", resolvecollisions, 0, "NO"), I("TARGETNPL", targetnpl, 0, "-1.0"),
  R("IAS15PRECISSION", ias15precission, 0, "1.e-9"), R("IAS15MINDT", ias15mindt, 0, "0.0"),
  S("DISCARDPLANETS", discardplanets, 0, "NO"),
  S("WRITETORQUEFILES", writetorquefiles, 0, "YES"), R("HILLCUT", hillcut, 0, "0.8"),
  R("VERTICALDAMPING", verticaldamping, 0, "0.1"), S("PLANETSFEELDISK", planetsfeeldisk, 0, "NO"),
  R("ACCRETIONRATE", accretionrate, 0, "0.0"), S("GASACCRETIONHEATING", gasaccretionheating, 0, "NO"),
  S("PEBBLEACCRETION", pebbleaccretion, 0, "NO"), S("BACKREACTION", backreaction, 0, "NO"),
  S("ACCRETIONALHEATING", accretionalheating, 0, "NO"), S("WRITEETA", writeeta, 0, "NO"),
  R("PEBBLEFLUX", pebbleflux, 0, "2.0e-4"), R("PEBBLEALPHA", pebblealpha, 0, "1.0e-4"),
  R("PEBBLECOAGULATION", pebblecoagulation, 0, "0.5"), R("PEBBLEBULKDENS", pebblebulkdens, 0, "1.0"),
  R("SCHMIDTNUMBER", schmidtnumber, 0, "1.0"), S("PARTICLEDIFFUSION", particlediffusion, 0, "NO"),
  I("HEATINGDELAY", heatingdelay, 0, "100"),
  R("PARAMETRICACCRETION", parametricaccretion, 0, "0.0"), S("TORQUEMAPINFILE", torquemapinfile, 0, "NO"),
  I("GETTORQUEFORPLANET", gettorqueforplanet, 0, "-1"), R("PEBBLETOGASMAX", pebbletogasmax, 0, "0.01"),
  R("FRAGMENTFACTOR", fragmentfactor, 0, "1.0"), R("FRAGMENTTHRESHOLD", fragmentthreshold, 0, "1.0"),
  R("FRAGMENTALPHA", fragmentalpha, 0, "1.0e-3"), S("OPACITYTABLE", opacitytable, 0, "BL94"),
  R("GASFLUX", gasflux, 0, "0.0"), S("DAMPING", damping, 0, "NO"),
  R("EVAPORATIONTEMPERATURE", evaporationtemperature, 0, "1000.0"),
  R("EVAPORATIONRATE", evaporationrate, 0, "1.0e-3"), R("FIRSTSTEP", firststep, 0, "-1.0"),
  R("MINIMUMSTEP", minimumstep, 0, "1.0e-3"), R("MAXIMUMSTEP", maximumstep, 0, "1.0e6"),
  S("AERODYNAMICDRAG", aerodynamicdrag, 0, "NO"), S("ELEMENTSWITHDISK", elementswithdisk, 0, "NO"),
  S("PEBBLEGRAVITY", pebblegravity, 0, "NO"),
  /* extension of this build: coefficient of the POWERLAW opacity; 0 = off */
  R("POWERLAWOPACITY", powerlawopacity, 0, "0.0"),
};
#define NREG ((int)(sizeof(registry) / sizeof(registry[0])))

/* var(), Interpret.c:34-62: only non-mandatory keys receive their default */
int thorin_params_defaults(thorin_params *p) {
  int k;
  memset(p, 0, sizeof(*p));
  for (k = 0; k < NREG; k++) {
    const pdesc *d = &registry[k];
    double temp = 0.0;
    char *base = (char *)p + d->off;
    sscanf(d->deflt, "%lf", &temp);
    if (d->necessary) continue;
    if (d->type == P_INT) *(int *)base = (int)(double)temp;
    else if (d->type == P_REAL) *(double *)base = temp;
    else { strncpy(base, d->deflt, d->size - 1); base[d->size - 1] = 0; }
  }
  return 0;
}

static int first_is(const char *s, char lo) { return s[0] == lo || s[0] == (char)toupper((unsigned char)lo); }

/* Interpret.c:149-265 */
int thorin_params_derive(thorin_params *p) {
  size_t n;
  p->advect_label = first_is(p->labeladvection, 'y');
  p->outer_source_mass = first_is(p->outersourcemass, 'y');
  p->fast_transport = !first_is(p->transport, 's');
  p->open_inner = first_is(p->innerboundary, 'o');
  p->non_reflecting = first_is(p->innerboundary, 'n');
  p->open_outer = first_is(p->outerboundary, 'o');
  p->damping_on = first_is(p->innerboundary, 'd') || first_is(p->outerboundary, 'd') || first_is(p->damping, 'y');
  p->damp_vrad = first_is(p->damptowards, 'z');
  p->damp_init = first_is(p->damptowards, 'i');
  p->damp_flux = first_is(p->damptowards, 'g');
  p->log_grid = first_is(p->gridspacing, 'l');
  p->is_disk = !first_is(p->disk, 'n');
  p->corotating = first_is(p->frame, 'c') || first_is(p->frame, 'g');
  p->guiding_center = first_is(p->frame, 'g');
  p->write_velocity = !first_is(p->writevelocity, 'n');
  p->write_density = !first_is(p->writedensity, 'n');
  p->indirect_term = !first_is(p->indirectterm, 'n');
  p->exclude_hill = first_is(p->excludehill, 'y');
  if (p->alphaviscosity != 0.0 && p->viscosity != 0.0) {
    thorin_set_error("You cannot use at the same time VISCOSITY and ALPHAVISCOSITY.");
    return -1;
  }
  p->viscosity_alpha = (p->alphaviscosity != 0.0);
  p->alpha_flock_on = first_is(p->alphaflock, 'y');
  if (p->thicknesssmoothing != 0.0 && p->rochesmoothing != 0.0) {
    thorin_set_error("You cannot use at the same time `ThicknessSmoothing' and `RocheSmoothing'.");
    return -1;
  }
  if (p->thicknesssmoothing <= 0.0 && p->rochesmoothing <= 0.0) {
    thorin_set_error("A non-vanishing potential smoothing length is required.");
    return -1;
  }
  p->roche_smoothing = (p->rochesmoothing != 0.0);
  n = strlen(p->outputdir);
  if (n == 0 || p->outputdir[n - 1] != '/') strncat(p->outputdir, "/", sizeof(p->outputdir) - n - 1);
  n = strlen(p->inputdir);
  if (n == 0 || p->inputdir[n - 1] != '/') strncat(p->inputdir, "/", sizeof(p->inputdir) - n - 1);
  p->energy_eq = first_is(p->energyequation, 'y');
  p->write_temperature = p->energy_eq || first_is(p->writetemperature, 'y');
  p->write_energy = first_is(p->writeenergy, 'y');
  p->write_divergence = first_is(p->writedivv, 'y');
  p->write_qplus = first_is(p->writeqplus, 'y');
  p->write_qbalance = first_is(p->writeqbalance, 'y');
  p->write_opacity = first_is(p->writeopacity, 'y');
  p->parametric_cooling = (p->coolingtime > 0.0);
  p->stellar_irradiation = first_is(p->stellarirradiation, 'y');
  p->init_from_file = first_is(p->initializefromfile, 'y');
  p->write_torque = first_is(p->writetorquefiles, 'y');
  p->write_torque_map = first_is(p->torquemapinfile, 'y');
  p->collisions = first_is(p->resolvecollisions, 'y');
  p->monitor_npl = (p->targetnpl >= 0);
  p->feel_disk = first_is(p->planetsfeeldisk, 'y');
  p->gas_accret_heating = first_is(p->gasaccretionheating, 'y');
  p->pebbles = first_is(p->pebbleaccretion, 'y');
  p->back_reaction = first_is(p->backreaction, 'y');
  /* Interpret.c:241 tests `*PARTICLEDIFFUSION = 'Y'` (an assignment), so
   * particle diffusion is always on and the string's first letter becomes 'Y' */
  if (p->particlediffusion[0] != 'y') p->particlediffusion[0] = 'Y';
  p->diffusive_particles = 1;
  p->accret_heating = first_is(p->accretionalheating, 'y');
  p->write_eta = first_is(p->writeeta, 'y');
  p->prescribed_accretion = (p->parametricaccretion > 0.0);
  p->torque_density = (p->gettorqueforplanet >= 0);
  p->discard_planets = first_is(p->discardplanets, 'y');
  p->aerodynamic_drag = first_is(p->aerodynamicdrag, 'y');
  p->elements_with_disk = first_is(p->elementswithdisk, 'y');
  p->pebble_gravity = first_is(p->pebblegravity, 'y');
  if (p->parametricopacity > 0.0) p->opacity_law = THORIN_OPA_CONST;
  else if (p->powerlawopacity > 0.0) p->opacity_law = THORIN_OPA_POWERLAW;
  else if (p->opacitytable[0] == 'Z') p->opacity_law = THORIN_OPA_ZHU12;
  else if (p->opacitytable[0] == 'L') p->opacity_law = THORIN_OPA_LP85;
  else if (p->opacitytable[0] == 'S') p->opacity_law = THORIN_OPA_SM03;
  else p->opacity_law = THORIN_OPA_BL94;
  return 0;
}

/* Programmatic access for bindings: one "KEY value" line of an in.par file.
 * thorin_params_derive() must be called after the last set. */
thorin_params *thorin_params_new(void) {
  thorin_params *p = (thorin_params *)malloc(sizeof(thorin_params));
  if (p) thorin_params_defaults(p);
  return p;
}

void thorin_params_free(thorin_params *p) { free(p); }

int thorin_params_set(thorin_params *p, const char *key, const char *value) {
  char nm[300], sv[290];
  double temp = 0.0;
  size_t i;
  int k;
  snprintf(nm, sizeof nm, "%s", key);
  for (i = 0; i < strlen(nm); i++) nm[i] = (char)toupper((unsigned char)nm[i]);
  sscanf(value, "%lf", &temp);
  sv[0] = 0;
  sscanf(value, "%289s", sv);
  for (k = 0; k < NREG; k++) {
    const pdesc *d = &registry[k];
    char *base = (char *)p + d->off;
    if (strcmp(nm, d->name) != 0) continue;
    if (d->type == P_INT) *(int *)base = (int)temp;
    else if (d->type == P_REAL) *(double *)base = temp;
    else { strncpy(base, sv, d->size - 1); base[d->size - 1] = 0; }
    return 0;
  }
  thorin_set_error("unknown parameter key");
  return -1;
}

/* value of a key as text (numbers with %.17g) */
int thorin_params_get(const thorin_params *p, const char *key, char *out, int outlen) {
  char nm[300];
  size_t i;
  int k;
  snprintf(nm, sizeof nm, "%s", key);
  for (i = 0; i < strlen(nm); i++) nm[i] = (char)toupper((unsigned char)nm[i]);
  for (k = 0; k < NREG; k++) {
    const pdesc *d = &registry[k];
    const char *base = (const char *)p + d->off;
    if (strcmp(nm, d->name) != 0) continue;
    if (d->type == P_INT) snprintf(out, (size_t)outlen, "%d", *(const int *)base);
    else if (d->type == P_REAL) snprintf(out, (size_t)outlen, "%.17g", *(const double *)base);
    else snprintf(out, (size_t)outlen, "%s", base);
    return 0;
  }
  thorin_set_error("unknown parameter key");
  return -1;
}

/* derived flag by name (Interpret.c:149-265), for bindings and tests */
int thorin_params_flag(const thorin_params *p, const char *name) {
#define FL(n) if (strcmp(name, #n) == 0) return p->n
  FL(advect_label); FL(fast_transport); FL(open_inner); FL(open_outer); FL(non_reflecting);
  FL(damping_on); FL(damp_vrad); FL(damp_init); FL(damp_flux); FL(log_grid); FL(is_disk);
  FL(corotating); FL(indirect_term); FL(exclude_hill); FL(viscosity_alpha); FL(alpha_flock_on);
  FL(energy_eq); FL(write_temperature); FL(write_energy); FL(parametric_cooling);
  FL(stellar_irradiation); FL(init_from_file); FL(write_torque); FL(feel_disk); FL(pebbles);
  FL(back_reaction); FL(diffusive_particles); FL(opacity_law); FL(write_density); FL(write_velocity);
#undef FL
  return -1;
}

/* ReadVariables, Interpret.c:65-148 */
int thorin_params_read(const char *path, thorin_params *p) {
  char nm[300], s[350], stringval[290];
  char *s1;
  double temp = 0.0;
  int k, success;
  unsigned char seen[sizeof(registry) / sizeof(registry[0])];
  FILE *in;
  thorin_params_defaults(p);
  memset(seen, 0, sizeof seen);
  stringval[0] = 0;
  in = fopen(path, "r");
  if (!in) {
    thorin_set_error("Unable to read parameter file.");
    return -1;
  }
  while (fgets(s, 349, in) != NULL) {
    success = sscanf(s, "%299s ", nm);
    if ((nm[0] != '#') && (success == 1)) {
      size_t i;
      s1 = s + strlen(nm);
      sscanf(s1 + strspn(s1, "\t :=>_"), "%lf", &temp);
      sscanf(s1 + strspn(s1, "\t :=>_"), "%289s ", stringval);
      for (i = 0; i < strlen(nm); i++) nm[i] = (char)toupper((unsigned char)nm[i]);
      for (k = 0; k < NREG; k++) {
        const pdesc *d = &registry[k];
        char *base = (char *)p + d->off;
        if (strcmp(nm, d->name) != 0) continue;
        seen[k] = 1;
        if (d->type == P_INT) *(int *)base = (int)temp;
        else if (d->type == P_REAL) *(double *)base = temp;
        else { strncpy(base, stringval, d->size - 1); base[d->size - 1] = 0; }
      }
    }
  }
  fclose(in);
  for (k = 0; k < NREG; k++)
    if (registry[k].necessary && !seen[k]) {
      static char msg[160];
      snprintf(msg, sizeof msg, "Fatal error : undefined mandatory variable %s", registry[k].name);
      thorin_set_error(msg);
      return -1;
    }
  return thorin_params_derive(p);
}
