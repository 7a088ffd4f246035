/*
 * thorin_main.c -- host driver of the GPU hydro step, mirroring the
 * reference's main loop (src_main/main.c:24-230): read in.par, set up the
 * planets and the disk, then call AlgoGas once per DT and write the
 * gas*<n>.dat outputs every NINTERM steps (Output.c:153-237).
 *
 *   thorin_amd [-v] [-0] [-t] [-s N] [-o outputdir] in.par
 *
 * -t prints the wall time of the AlgoGas calls (CLOCK_MONOTONIC around each
 * call, the first three excluded) when the run ends.
 *
 * Output files and formats follow the reference: raw native-endian float64
 * gas<name><n>.dat (no header, ring-major), dims.dat, used_rad.dat,
 * omegaframe.dat, tqwk<k>.dat and nbody.orbits.dat.  The N-body state for
 * restarts is written as nbody.state<n>.dat (text) because REBOUND's binary
 * snapshot is not part of this build.
 */
#define _GNU_SOURCE
#include "../../include/thorin_amd.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <time.h>

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void die(const char *what) {
  fprintf(stderr, "thorin_amd: %s: %s\n", what, thorin_last_error());
  exit(1);
}

static void mkdir_p(const char *dir) {
  char tmp[THORIN_STRLEN];
  size_t n;
  char *q;
  snprintf(tmp, sizeof tmp, "%s", dir);
  n = strlen(tmp);
  if (n && tmp[n - 1] == '/') tmp[n - 1] = 0;
  for (q = tmp + 1; *q; q++)
    if (*q == '/') { *q = 0; mkdir(tmp, 0755); *q = '/'; }
  mkdir(tmp, 0755);
}

static void write_polar(const thorin_params *p, thorin_ctx *c, int field, const char *name, int n, double *buf) {
  char fn[2 * THORIN_STRLEN];
  FILE *f;
  size_t cnt = (size_t)p->nrad * p->nsec;
  if (thorin_download_field(c, field, buf, cnt)) die("download");
  snprintf(fn, sizeof fn, "%sgas%s%d.dat", p->outputdir, name, n);
  f = fopen(fn, "w");
  if (!f) { fprintf(stderr, "cannot open %s\n", fn); exit(1); }
  fwrite(buf, sizeof(double), cnt, f);
  fclose(f);
  printf("Writing 'gas%s%d.dat'...0/done\n", name, n);
}

/* SendOutput, Output.c:202-237 */
static void send_output(const thorin_params *p, thorin_ctx *c, int n, double *buf, double *buf2) {
  int ln = p->write_density ? n : 0;
  printf("\n*** OUTPUT %d ***\n", n);
  write_polar(p, c, THORIN_F_DENS, "dens", ln, buf);
  ln = p->write_velocity ? n : 0;
  write_polar(p, c, THORIN_F_VRAD, "vrad", ln, buf);
  write_polar(p, c, THORIN_F_VTHETA, "vtheta", ln, buf);
  if (p->write_energy) write_polar(p, c, THORIN_F_ENERGY, "energy", n, buf);
  if (p->write_temperature) write_polar(p, c, THORIN_F_TEMPER, "temper", n, buf);
  if (p->write_divergence) write_polar(p, c, THORIN_F_DIVV, "divv", n, buf);
  if (p->write_qplus) write_polar(p, c, THORIN_F_QPLUS, "qplus", n, buf);
  if (p->write_qbalance) {
    char fn[2 * THORIN_STRLEN];
    size_t k, cnt = (size_t)p->nrad * p->nsec;
    FILE *f;
    if (thorin_download_field(c, THORIN_F_QPLUS, buf, cnt) || thorin_download_field(c, THORIN_F_QMINUS, buf2, cnt)) die("download");
    for (k = 0; k < cnt; k++) buf[k] = buf[k] - buf2[k];   /* ActualizeQbalance, Output.c:240-256 */
    snprintf(fn, sizeof fn, "%sgasqbalance%d.dat", p->outputdir, n);
    f = fopen(fn, "w");
    if (f) { fwrite(buf, sizeof(double), cnt, f); fclose(f); }
  }
  if (p->write_opacity) write_polar(p, c, THORIN_F_OPACITY, "opacity", n, buf);
  if (p->pebbles) {   /* WritePebbles, Pebbles.c:772-778 */
    write_polar(p, c, THORIN_F_PEBBLEDENS, "pebbledens", n, buf);
    write_polar(p, c, THORIN_F_PEBBLEVRAD, "pebblevrad", n, buf);
    write_polar(p, c, THORIN_F_PEBBLEVTHETA, "pebblevtheta", n, buf);
    if (p->write_eta) write_polar(p, c, THORIN_F_ETA, "eta", n, buf);
  }
}

/* ReadfromFile, Init.c:23-54: dst holds the t=0 field; a missing file
 * leaves it (with the reference's warning), a short one replaces only what
 * it holds. */
static int read_binary(const char *fn, double *dst, size_t cnt) {
  FILE *f = fopen(fn, "r");
  size_t got;
  if (!f) { fprintf(stderr, "WARNING ! Can't read %s. Restarting with t=0 settings.\n", fn); return -1; }
  got = fread(dst, sizeof(double), cnt, f);
  fclose(f);
  if (got != cnt) fprintf(stderr, "WARNING ! %s holds %zu of %zu values.\n", fn, got, cnt);
  return got == cnt ? 0 : -1;
}

/* the first crash's fields (SourceEuler.c:276-282: WriteDiskPolar(..., 999)) */
static void write_crash_dump(const thorin_params *p, thorin_ctx *c, double *buf) {
  static const int ids[4] = {THORIN_X_CRASH_DENS, THORIN_X_CRASH_VRAD, THORIN_X_CRASH_VTHETA, THORIN_X_CRASH_TEMPER};
  static const char *names[4] = {"dens", "vrad", "vtheta", "temper"};
  char fn[2 * THORIN_STRLEN];
  size_t cnt = (size_t)p->nrad * p->nsec;
  int k;
  FILE *f;
  for (k = 0; k < 4; k++) {
    if (thorin_download_field(c, ids[k], buf, cnt)) die("download");
    snprintf(fn, sizeof fn, "%sgas%s999.dat", p->outputdir, names[k]);
    f = fopen(fn, "w");
    if (!f) { fprintf(stderr, "cannot open %s\n", fn); exit(1); }
    fwrite(buf, sizeof(double), cnt, f);
    fclose(f);
    printf("Writing 'gas%s999.dat'...0/done\n", names[k]);
  }
}

/* ReadfromAsciiFile, Init.c:148-199 */
static void read_ascii(const char *fn, double *dst, size_t cnt) {
  FILE *f = fopen(fn, "r");
  size_t k;
  double tmp;
  if (!f) { fprintf(stderr, "\nError! Can't find the initialization file %s \n", fn); exit(1); }
  for (k = 0; k < cnt; k++)
    if (fscanf(f, "%lf", &dst[k]) == EOF) {
      fprintf(stderr, "\nError! Number of values in the initialization file %s\nis smaller than the size of hydrodynamic arrays.\n", fn);
      exit(1);
    }
  if (fscanf(f, "%lf", &tmp) != EOF) {
    fprintf(stderr, "\nError! Number of values in the initialization file %s\nis larger than the size of hydrodynamic arrays.\n", fn);
    exit(1);
  }
  fclose(f);
}

static void write_nbody_state(const thorin_params *p, const thorin_planets *s, double t, int n) {
  char fn[2 * THORIN_STRLEN];
  FILE *f;
  int k;
  snprintf(fn, sizeof fn, "%snbody.state%d.dat", p->outputdir, n);
  f = fopen(fn, "w");
  if (!f) return;
  fprintf(f, "%d %.17g\n", s->nb, t);
  for (k = 0; k < s->nb; k++)
    fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", s->mass[k], s->x[k], s->y[k], s->z[k], s->vx[k],
            s->vy[k], s->vz[k], s->gas[k], s->solid[k]);
  fclose(f);
}

/* nbody.orbits.dat (ReboundInterface.c:197-200, 244-247, 405-441): per
 * planet its index, the time, a, e, inc, Omega, omega, f about the primary
 * (of mass 1, plus the initial disk's mass inside the planet's distance from
 * the star's axis with ElementsWithDisk), mass, gas and solid mass times
 * MassTaper, the planet's radius at PlanetaryDensity, x, y, z, vx, vy, vz.
 * sigma_med is NULL for the rows written before the disk exists. */
#define ORB_RHO2CGS (1.98855e30 * 1000.0 / pow(149597870700.0 * 100.0, 3.0))   /* fondam.h:27,30,38 */
static void write_orbits(FILE *f, const thorin_params *p, const thorin_geometry *g, const double *sigma_med,
                         const thorin_planets *s, double t, double taper) {
  const double rdenominv = 1.0 / (4.0 * M_PI * (p->planetarydensity / ORB_RHO2CGS));
  int k, i;
  if (!f) return;
  for (k = 0; k < s->nb; k++) {
    double el[6], mp = 1.0;
    if (p->elements_with_disk && sigma_med) {
      /* GetInteriorDiskMass (ReboundInterface.c:387-402): whole rings below
       * the cylindrical radius, the ring that holds it in proportion */
      const double rp = sqrt(s->x[k] * s->x[k] + s->y[k] * s->y[k]);
      double md = 0.0;
      for (i = 0; i < p->nrad && rp >= g->rsup[i]; i++) md += p->nsec * g->surf[i] * sigma_med[i];
      if (i < p->nrad && rp > g->rinf[i] && rp < g->rsup[i])
        md += (rp - g->rinf[i]) / (g->rsup[i] - g->rinf[i]) * p->nsec * g->surf[i] * sigma_med[i];
      mp += md;
    }
    if (thorin_orbit_elements(s, k, mp, el)) {
      for (i = 0; i < 6; i++) el[i] = NAN;
    }
    fprintf(f, "%d\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g\t%.12g"
               "\t%.12g\t%.12g\n",
            k + 1, t, el[0], el[1], el[2], el[3], el[4], el[5], s->mass[k] * taper, s->gas[k] * taper,
            s->solid[k] * taper, pow(3.0 * s->mass[k] * rdenominv, 1.0 / 3.0), s->x[k], s->y[k], s->z[k], s->vx[k],
            s->vy[k], s->vz[k]);
  }
  fflush(f);
}

static int read_nbody_state(const thorin_params *p, thorin_planets *s, double *t, int n) {
  char fn[2 * THORIN_STRLEN];
  FILE *f;
  int k, nb;
  snprintf(fn, sizeof fn, "%snbody.state%d.dat", p->outputdir, n);
  f = fopen(fn, "r");
  if (!f) return -1;
  if (fscanf(f, "%d %lf", &nb, t) != 2 || nb != s->nb) { fclose(f); return -1; }
  for (k = 0; k < nb; k++)
    if (fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", &s->mass[k], &s->x[k], &s->y[k], &s->z[k], &s->vx[k], &s->vy[k],
               &s->vz[k], &s->gas[k], &s->solid[k]) != 9) { fclose(f); return -1; }
  fclose(f);
  return 0;
}

int main(int argc, char **argv) {
  thorin_params p;
  thorin_planets sys;
  thorin_ctx *c = NULL;
  thorin_state st;
  thorin_step_log *lg;
  const char *parfile = NULL, *outdir = NULL;
  int i, k, restart = 0, nrestart = 0, only_init = 0, verbose = 0, begin_i = 0, timing = 0, ntimed = 0, ncalls = 0;
  FILE *orbits;
  thorin_geometry geo;
  double t_orbits = 0.0;
  double t_calls = 0.0;
  size_t cnt;
  double *buf, *buf2;
  char fn[2 * THORIN_STRLEN];
  FILE *f;
  for (i = 1; i < argc; i++) {
    if (argv[i][0] == '-') {
      if (strchr(argv[i], 'v')) verbose = 1;
      if (strchr(argv[i], 't')) timing = 1;
      if (strchr(argv[i], '0')) only_init = 1;
      if (strchr(argv[i], 's')) { restart = 1; if (++i < argc) nrestart = atoi(argv[i]); }
      else if (strchr(argv[i], 'o')) { if (++i < argc) outdir = argv[i]; }
    } else parfile = argv[i];
  }
  if (!parfile) { fprintf(stderr, "Usage: %s [-v] [-0] [-s N] [-o dir] in.par\n", argv[0]); return 1; }
  if (thorin_params_read(parfile, &p)) die("parameters");
  if (outdir) {
    snprintf(p.outputdir, sizeof p.outputdir, "%s", outdir);
    if (thorin_params_derive(&p)) die("parameters");
  }
  mkdir_p(p.outputdir);
  if (thorin_planets_from_config(&p, &sys)) die("planets");
  printf("%d planet(s) found.\n", sys.nb);
  if (thorin_ctx_create(&p, &c)) die("context");
  cnt = (size_t)p.nrad * p.nsec;
  buf = (double *)malloc(sizeof(double) * (cnt + p.nsec));
  buf2 = (double *)malloc(sizeof(double) * (cnt + p.nsec));
  if (restart) {
    /* Init.c:81-126: the initial state first (its ring profiles are the
     * damping targets), then the restart files over its fields.  A missing
     * gastemper file keeps the t=0 temperature (the reference would take
     * its t=0 energy for a temperature there). */
    double *d = malloc(sizeof(double) * cnt), *vr = malloc(sizeof(double) * cnt), *vt = malloc(sizeof(double) * cnt),
           *t = malloc(sizeof(double) * cnt), tt = 0.0;
    if (p.init_from_file) {
      read_ascii(p.densinfile, d, cnt);
      read_ascii(p.vradinfile, vr, cnt);
      read_ascii(p.vthetainfile, vt, cnt);
      read_ascii(p.temperinfile, t, cnt);
      if (thorin_init_state(c, 1, d, vr, vt, t)) die("init");
    } else if (thorin_init_state(c, 0, NULL, NULL, NULL, NULL)) die("init");
    if (thorin_download_field(c, THORIN_F_DENS, d, cnt) || thorin_download_field(c, THORIN_F_VRAD, vr, cnt) ||
        thorin_download_field(c, THORIN_F_VTHETA, vt, cnt) || thorin_download_field(c, THORIN_F_TEMPER, t, cnt))
      die("download");
    snprintf(fn, sizeof fn, "%sgasdens%d.dat", p.outputdir, nrestart); read_binary(fn, d, cnt);
    snprintf(fn, sizeof fn, "%sgasvrad%d.dat", p.outputdir, nrestart); read_binary(fn, vr, cnt);
    snprintf(fn, sizeof fn, "%sgasvtheta%d.dat", p.outputdir, nrestart); read_binary(fn, vt, cnt);
    snprintf(fn, sizeof fn, "%sgastemper%d.dat", p.outputdir, nrestart); read_binary(fn, t, cnt);
    if (thorin_init_state(c, 2, d, vr, vt, t)) die("restart");
    if (read_nbody_state(&p, &sys, &tt, nrestart) == 0) {
      thorin_get_state(c, &st);
      st.physical_time = st.physical_time_initial = tt;
      thorin_set_state(c, &st);
    }
    t_orbits = tt;
    begin_i = nrestart * p.ninterm;
    free(d); free(vr); free(vt); free(t);
  } else if (p.init_from_file) {
    double *d = malloc(sizeof(double) * cnt), *vr = malloc(sizeof(double) * cnt), *vt = malloc(sizeof(double) * cnt),
           *t = malloc(sizeof(double) * cnt);
    fprintf(stderr, "Initializing from files...\n");
    read_ascii(p.densinfile, d, cnt);
    read_ascii(p.vradinfile, vr, cnt);
    read_ascii(p.vthetainfile, vt, cnt);
    read_ascii(p.temperinfile, t, cnt);
    if (thorin_init_state(c, 1, d, vr, vt, t)) die("init");
    free(d); free(vr); free(vt); free(t);
  } else {
    if (thorin_init_state(c, 0, NULL, NULL, NULL, NULL)) die("init");
  }
  /* pebble disk (Init.c:138-144): RestartPebbleDisk or EquilPebbleDisk */
  if (p.pebbles) {
    if (restart) {
      /* RestartPebbleDisk (Pebbles.c:67-75) reads into the zeroed pebble grids */
      double *pd = calloc(cnt, sizeof(double)), *pvr = calloc(cnt, sizeof(double)),
             *pvt = calloc(cnt, sizeof(double)), *ps = calloc(cnt, sizeof(double));
      snprintf(fn, sizeof fn, "%sgaspebbledens%d.dat", p.outputdir, nrestart); read_binary(fn, pd, cnt);
      snprintf(fn, sizeof fn, "%sgaspebblesize%d.dat", p.outputdir, 0); read_binary(fn, ps, cnt);
      snprintf(fn, sizeof fn, "%sgaspebblevrad%d.dat", p.outputdir, nrestart); read_binary(fn, pvr, cnt);
      snprintf(fn, sizeof fn, "%sgaspebblevtheta%d.dat", p.outputdir, nrestart); read_binary(fn, pvt, cnt);
      if (thorin_init_pebbles(c, 1, pd, pvr, pvt, ps)) die("pebble restart");
      free(pd); free(pvr); free(pvt); free(ps);
    } else {
      if (thorin_init_pebbles(c, 0, NULL, NULL, NULL, NULL)) die("pebble disk");
      /* InitPebblesViaFlux writes the Stokes numbers and sizes of output 0 */
      write_polar(&p, c, THORIN_F_PEBBLESTOKES, "pebblestokes", 0, buf);
      write_polar(&p, c, THORIN_F_PEBBLESIZE, "pebblesize", 0, buf);
    }
  }
  if (verbose) printf("NRAD %d NSEC %d DT %g NTOT %d NINTERM %d\n", p.nrad, p.nsec, p.dt, p.ntot, p.ninterm);
  /* used_rad.dat (SourceEuler.c:119-129) and dims.dat (Output.c:191-200) */
  {
    thorin_geometry g;
    thorin_get_geometry(c, &g);
    snprintf(fn, sizeof fn, "%sused_rad.dat", p.outputdir);
    f = fopen(fn, "w");
    if (f) { for (i = 0; i <= p.nrad; i++) fprintf(f, "%.18g\n", g.radii[i]); fclose(f); }
    snprintf(fn, sizeof fn, "%sdims.dat", p.outputdir);
    f = fopen(fn, "w");
    if (f) { fprintf(f, "%d\t%d\t\t%d\t%d\t%f\t%d\t%d\t%d\n", 0, 0, 0, 0, p.rmax, p.ntot / p.ninterm, p.nrad, p.nsec); fclose(f); }
  }
  /* the row of the set-up, MassTaper taken as 1 (ReboundInterface.c:199, 246) */
  snprintf(fn, sizeof fn, "%snbody.orbits.dat", p.outputdir);
  orbits = fopen(fn, restart ? "a" : "w");
  if (!orbits) fprintf(stderr, "WARNING ! Can't write %s\n", fn);
  thorin_get_geometry(c, &geo);
  write_orbits(orbits, &p, &geo, NULL, &sys, t_orbits, 1.0);
  lg = (thorin_step_log *)calloc(1, sizeof(thorin_step_log));
  for (i = begin_i; i <= p.ntot; i++) {
    int ts = i / p.ninterm;
    if (p.ninterm * ts == i) {
      thorin_get_state(c, &st);
      send_output(&p, c, ts, buf, buf2);
      write_nbody_state(&p, &sys, st.physical_time, ts);
      snprintf(fn, sizeof fn, "%somegaframe.dat", p.outputdir);
      f = fopen(fn, "a");
      if (f) { fprintf(f, "%d\t%#.18g\n", ts, st.omega_frame); fclose(f); }
      if (only_init) break;
    }
    {
      const double t0 = now_s();
      k = thorin_algogas(c, &sys, thorin_nbody_step, NULL, lg);
      if (++ncalls > 3) { t_calls += now_s() - t0; ntimed++; }
    }
    if (k < 0) die("AlgoGas");
    if (lg->crash_dumped) {   /* SourceEuler.c:275-284 */
      printf("\nCrash! at time %.12g\n", lg->crash_time);
      write_crash_dump(&p, c, buf);
    }
    for (k = 0; k < lg->nsub; k++) printf("%c", lg->crashed[k] ? 'c' : '.');
    printf("\n");
    fflush(stdout);
    if (p.noutelements > 0 && (i + 1) % p.noutelements == 0) {   /* main.c:202-207 */
      thorin_get_state(c, &st);
      write_orbits(orbits, &p, &geo, geo.sigma_med, &sys, st.physical_time, st.mass_taper);
    }
    if (p.write_torque)
      for (k = 0; k < lg->tq_rows; k++) {
        snprintf(fn, sizeof fn, "%stqwk%d.dat", p.outputdir, k);
        f = fopen(fn, "a");
        if (f) {
          fprintf(f, "%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\n", lg->tq[k][0], lg->tq[k][1], lg->tq[k][2],
                  lg->tq[k][3], lg->tq[k][4], lg->tq[k][5], lg->tq[k][6]);
          fclose(f);
        }
      }
  }
  if (timing && ntimed > 0)
    printf("timing: %d AlgoGas calls after 3: %.4f ms per call (%.1f calls/s)\n", ntimed, 1e3 * t_calls / ntimed,
           ntimed / t_calls);
  if (orbits) fclose(orbits);
  thorin_ctx_destroy(c);
  free(buf);
  free(buf2);
  free(lg);
  return 0;
}
