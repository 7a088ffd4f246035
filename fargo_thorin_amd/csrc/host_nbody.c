/*
 * host_nbody.c -- the planetary system on the host.
 *
 * The reference integrates planets with a vendored REBOUND (IAS15) called
 * from AlgoGas (ReboundInterface.c:262-355).  REBOUND is not part of this
 * build: the host keeps its own integrator behind the same hook
 * (thorin_nbody_fn).  The planet file is read as in SetupReboundSimulation
 * (ReboundInterface.c:130-208): a line starting with a letter holds a name
 * and nine floats (mass, gas, solid, a, e, inc, Omega, omega, f) parsed as
 * float, angles in degrees.
 *
 * Integrator: heliocentric equations of motion of N planets around a unit
 * mass primary, advanced over each substep with a Gragg-Bulirsch-Stoer step
 * (modified midpoint + polynomial extrapolation) to round-off accuracy.  For
 * the reference's test setups (one planet on a circular orbit) this agrees
 * with IAS15 to ~1e-15 relative, which changes the disk forces far below the
 * hydro parity tolerance.
 */
#define _GNU_SOURCE
#include "thorin_internal.h"
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* orbital elements -> heliocentric state (the standard two-body relations) */
static void elements_to_state(double mu, double a, double e, double inc, double Om, double om, double f,
                              double *x, double *y, double *z, double *vx, double *vy, double *vz) {
  double r = a * (1.0 - e * e) / (1.0 + e * cos(f));
  double v0 = sqrt(mu / a / (1.0 - e * e));
  double cO = cos(Om), sO = sin(Om), co = cos(om), so = sin(om), cf = cos(f), sf = sin(f);
  double ci = cos(inc), si = sin(inc);
  *x = r * (cO * (co * cf - so * sf) - sO * (so * cf + co * sf) * ci);
  *y = r * (sO * (co * cf - so * sf) + cO * (so * cf + co * sf) * ci);
  *z = r * (so * cf + co * sf) * si;
  *vx = v0 * ((e + cf) * (-ci * co * sO - cO * so) - sf * (co * cO - ci * so * sO));
  *vy = v0 * ((e + cf) * (ci * co * cO - sO * so) - sf * (co * sO + ci * so * cO));
  *vz = v0 * ((e + cf) * co * si - sf * si * so);
}

/* the angle in [0, pi] whose cosine is num/den, negated when side < 0 (an
 * angle measured from the x axis takes the sign of its y component); 0 or
 * pi where rounding puts the cosine outside [-1, 1], and 0 for den = 0 */
static double angle_of(double num, double den, double side) {
  const double c = num / den;
  if (!(c > -1.0 && c < 1.0)) return c <= -1.0 ? TH_PI : 0.0;
  return side < 0.0 ? -acos(c) : acos(c);
}

/* heliocentric state -> osculating elements about a primary of mass mp (G =
 * 1), in the conventions of the elements OutputElements writes
 * (ReboundInterface.c:405-441, through the vendored REBOUND's orbit routine):
 * a from the vis-viva energy, e from the Laplace-Runge-Lenz vector, the
 * angles from the x axis.  Within 1e-8 rad of the reference plane the node
 * is undefined: Omega = 0 and omega, f come from the longitudes of the
 * pericentre and of the planet. */
int thorin_orbit_elements(const thorin_planets *sys, int k, double mp, double el[6]) {
  if (!sys || !el || k < 0 || k >= sys->nb || !(mp > 0.0)) {
    thorin_set_error("thorin_orbit_elements: bad planet index or primary mass");
    return -1;
  }
  const double mu = sys->mass[k] + mp;
  const double x = sys->x[k], y = sys->y[k], z = sys->z[k];
  const double vx = sys->vx[k], vy = sys->vy[k], vz = sys->vz[k];
  const double r = sqrt(x * x + y * y + z * z);
  if (!(r > 0.0)) {
    thorin_set_error("thorin_orbit_elements: planet at the primary");
    return -1;
  }
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double a = -mu / (v2 - 2.0 * mu / r);
  const double hx = y * vz - z * vy, hy = z * vx - x * vz, hz = x * vy - y * vx;
  const double h = sqrt(hx * hx + hy * hy + hz * hz);
  const double rdot = (x * vx + y * vy + z * vz) / r;
  /* e vector = ((v^2 - mu/r) r_vec - (r rdot) v_vec) / mu */
  const double w = v2 - mu / r, rr = r * rdot;
  const double ex = (w * x - rr * vx) / mu, ey = (w * y - rr * vy) / mu, ez = (w * z - rr * vz) / mu;
  const double e = sqrt(ex * ex + ey * ey + ez * ez);
  const double inc = angle_of(hz, h, 1.0);
  const double nx = -hy, ny = hx, nn = sqrt(nx * nx + ny * ny);   /* z x h: the ascending node */
  const double Om = angle_of(nx, nn, ny);
  double om, f;
  if (inc < 1e-8 || inc > TH_PI - 1e-8) {
    const double varpi = angle_of(ex, e, ey), lam = angle_of(x, r, y);
    if (inc < TH_PI / 2.0) {
      om = varpi - Om;
      f = lam - varpi;
    } else {
      om = Om - varpi;
      f = varpi - lam;
    }
  } else {
    const double u = angle_of(nx * x + ny * y, nn * r, z);        /* argument of latitude */
    om = angle_of(nx * ex + ny * ey, nn * e, ez);
    f = u - om;
  }
  el[0] = a; el[1] = e; el[2] = inc; el[3] = Om; el[4] = om; el[5] = f;
  return 0;
}

int thorin_planets_from_config(const thorin_params *p, thorin_planets *sys) {
  FILE *in = fopen(p->planetconfig, "r");
  char s[512], nm[512];
  double rad = TH_PI / 180.0;
  memset(sys, 0, sizeof(*sys));
  if (!in) { thorin_set_error("can't open PlanetConfig file"); return -1; }
  while (fgets(s, 510, in) != NULL) {
    float mass, gas, solid, a, e, inc, Om, om, f;
    char *s1;
    int k;
    if (sscanf(s, "%511s ", nm) != 1) continue;
    if (!isalpha((unsigned char)s[0])) continue;
    s1 = s + strlen(nm);
    if (sscanf(s1 + strspn(s1, "\t :=>_"), "%f %f %f %f %f %f %f %f %f", &mass, &gas, &solid, &a, &e, &inc, &Om, &om, &f) != 9) {
      fclose(in);
      thorin_set_error("error reading PlanetConfig file");
      return -1;
    }
    if (sys->nb >= THORIN_MAXPLANETS) { fclose(in); thorin_set_error("too many planets"); return -1; }
    k = sys->nb++;
    sys->mass[k] = (double)mass;
    sys->gas[k] = gas;
    sys->solid[k] = solid;
    elements_to_state(1.0 * (1.0 + (double)mass), (double)a, (double)e, (double)inc * rad, (double)Om * rad,
                      (double)om * rad, (double)f * rad, &sys->x[k], &sys->y[k], &sys->z[k], &sys->vx[k],
                      &sys->vy[k], &sys->vz[k]);
    sys->acc[k] = p->accretionrate;
    sys->feeldisk[k] = p->feel_disk;
  }
  fclose(in);
  return 0;
}

/* heliocentric accelerations of the planets (primary mass 1, G = 1) */
static void nbody_accel(int n, const double *m, const double *q, double *a) {
  int i, j;
  for (i = 0; i < n; i++) {
    double x = q[6 * i], y = q[6 * i + 1], z = q[6 * i + 2];
    double r2 = x * x + y * y + z * z, r = sqrt(r2), f = -(1.0 + m[i]) / (r2 * r);
    a[3 * i] = f * x;
    a[3 * i + 1] = f * y;
    a[3 * i + 2] = f * z;
  }
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++) {
      double dx, dy, dz, d2, d, rj2, rj, fd, fi;
      if (i == j) continue;
      dx = q[6 * j] - q[6 * i];
      dy = q[6 * j + 1] - q[6 * i + 1];
      dz = q[6 * j + 2] - q[6 * i + 2];
      d2 = dx * dx + dy * dy + dz * dz;
      d = sqrt(d2);
      rj2 = q[6 * j] * q[6 * j] + q[6 * j + 1] * q[6 * j + 1] + q[6 * j + 2] * q[6 * j + 2];
      rj = sqrt(rj2);
      fd = m[j] / (d2 * d);
      fi = m[j] / (rj2 * rj);
      a[3 * i] += fd * dx - fi * q[6 * j];
      a[3 * i + 1] += fd * dy - fi * q[6 * j + 1];
      a[3 * i + 2] += fd * dz - fi * q[6 * j + 2];
    }
}

static void deriv(int n, const double *m, const double *q, double *dq) {
  double a[3 * THORIN_MAXPLANETS];
  int i;
  nbody_accel(n, m, q, a);
  for (i = 0; i < n; i++) {
    dq[6 * i] = q[6 * i + 3];
    dq[6 * i + 1] = q[6 * i + 4];
    dq[6 * i + 2] = q[6 * i + 5];
    dq[6 * i + 3] = a[3 * i];
    dq[6 * i + 4] = a[3 * i + 1];
    dq[6 * i + 5] = a[3 * i + 2];
  }
}

/* modified midpoint over H with nsub substeps (Gragg) */
static void midpoint(int n, const double *m, const double *q0, double H, int nsub, double *out) {
  int dim = 6 * n, k, s;
  double h = H / nsub;
  double *zm = (double *)malloc(sizeof(double) * dim), *zn = (double *)malloc(sizeof(double) * dim);
  double *d = (double *)malloc(sizeof(double) * dim);
  deriv(n, m, q0, d);
  for (k = 0; k < dim; k++) { zm[k] = q0[k]; zn[k] = q0[k] + h * d[k]; }
  for (s = 1; s < nsub; s++) {
    deriv(n, m, zn, d);
    for (k = 0; k < dim; k++) {
      double tmp = zm[k] + 2.0 * h * d[k];
      zm[k] = zn[k];
      zn[k] = tmp;
    }
  }
  deriv(n, m, zn, d);
  for (k = 0; k < dim; k++) out[k] = 0.5 * (zm[k] + zn[k] + h * d[k]);
  free(zm); free(zn); free(d);
}

/* one Bulirsch-Stoer step: extrapolate the midpoint results to h -> 0 */
static int bs_step(int n, const double *m, double *q, double H) {
  enum { KMAX = 10 };
  static const int seq[KMAX] = {2, 4, 6, 8, 10, 12, 14, 16, 18, 20};
  int dim = 6 * n, k, j, c;
  double *T[KMAX];
  double err = 0.0, scale = 0.0;
  for (k = 0; k < KMAX; k++) T[k] = (double *)malloc(sizeof(double) * dim);
  for (k = 0; k < KMAX; k++) {
    midpoint(n, m, q, H, seq[k], T[k]);
    /* Neville-Aitken in h^2, in place: T[k-j] becomes the order-j estimate */
    for (j = k - 1; j >= 0; j--) {
      double r = (double)seq[k] / (double)seq[j];
      double f = 1.0 / (r * r - 1.0);
      for (c = 0; c < dim; c++) T[j][c] = T[j + 1][c] + (T[j + 1][c] - T[j][c]) * f;
    }
    if (k >= 3) {
      err = 0.0;
      scale = 0.0;
      for (c = 0; c < dim; c++) {
        double dlt = fabs(T[0][c] - T[1][c]);
        if (dlt > err) err = dlt;
        if (fabs(T[0][c]) > scale) scale = fabs(T[0][c]);
      }
      if (err <= 1e-15 * (scale > 0 ? scale : 1.0)) break;
    }
  }
  if (k == KMAX) k = KMAX - 1;
  for (c = 0; c < dim; c++) q[c] = T[0][c];
  for (j = 0; j < KMAX; j++) free(T[j]);
  return (err <= 1e-13 * (scale > 0 ? scale : 1.0)) ? 0 : 1;
}

/* AdvanceSystemRebound + RotatePsys + SynchronizeFargoRebound
 * (SourceEuler.c:260-269).  The substep is split until each Bulirsch-Stoer
 * step converges (close encounters); if 2^11 pieces still do not converge,
 * the planets are left as they were and 1 is returned (thorin_algogas then
 * fails the DT with "N-body step failed"). */
int thorin_nbody_step(void *user, thorin_planets *sys, double t, double dt, double angle) {
  int n = sys->nb, i, pieces = 1, tries, bad = 1;
  double q[6 * THORIN_MAXPLANETS], q0[6 * THORIN_MAXPLANETS], sint, cost;
  (void)user;
  (void)t;
  if (n <= 0) return 0;
  for (i = 0; i < n; i++) {
    q0[6 * i] = sys->x[i]; q0[6 * i + 1] = sys->y[i]; q0[6 * i + 2] = sys->z[i];
    q0[6 * i + 3] = sys->vx[i]; q0[6 * i + 4] = sys->vy[i]; q0[6 * i + 5] = sys->vz[i];
  }
  for (tries = 0; tries < 12 && bad; tries++, pieces *= 2) {
    int k;
    bad = 0;
    memcpy(q, q0, sizeof(double) * 6 * n);
    for (k = 0; k < pieces; k++) bad |= bs_step(n, sys->mass, q, dt / pieces);
  }
  if (bad) {
    thorin_set_error("thorin_nbody_step: the Bulirsch-Stoer step did not converge in 2048 pieces of the substep");
    return 1;
  }
  sint = sin(angle);
  cost = cos(angle);
  for (i = 0; i < n; i++) {
    double xt = q[6 * i], yt = q[6 * i + 1];
    sys->x[i] = xt * cost + yt * sint;
    sys->y[i] = -xt * sint + yt * cost;
    sys->z[i] = q[6 * i + 2];
    xt = q[6 * i + 3];
    yt = q[6 * i + 4];
    sys->vx[i] = xt * cost + yt * sint;
    sys->vy[i] = -xt * sint + yt * cost;
    sys->vz[i] = q[6 * i + 5];
  }
  return 0;
}
