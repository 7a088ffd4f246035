/*
 * algogas_shim.c -- the reference's AlgoGas on the GPU library.
 *
 * Compiled with the reference's own headers (src_main/fargo.h) and linked
 * with its objects, where this AlgoGas replaces SourceEuler.c's
 * (SourceEuler.c:210-317, prototype proto.h:74, called once per DT by
 * main.c:201).  The reference's driver stays as it is: parameters, set-up,
 * REBOUND, outputs and restarts are its own; every DT of gas (and pebbles)
 * runs in libthorin_amd.so, whose fields stay resident in HBM.
 *
 *   first call   the library's context from the parameters the reference
 *                read (one line per var.c key below), the initial state made
 *                as Initialization() makes it (Init.c:72-145), then the
 *                reference's own fields uploaded over it, so that both start
 *                from the same bits
 *   every call   the planets (mass, gas, solid, position, velocity, accretion
 *                rate, FeelDisk) go in; the library calls back once per
 *                substep for the N-body part, which runs the reference's own
 *                AdvanceSystemRebound, RotatePsys and SynchronizeFargoRebound
 *                (SourceEuler.c:260-269) on the library's copy of the planets
 *                (its accretion and disk kicks included); afterwards the
 *                planets, PhysicalTime, MassTaper and OmegaFrame come back
 *   outputs      the fields the next SendOutput writes (Output.c:202-237) are
 *                downloaded after the DT that precedes it (main.c:176-179);
 *                the torque rows of UpdateLog (Force.c:27-70), the first
 *                crash's dump (SourceEuler.c:275-284) and the substep marks
 *                are written here as the reference writes them
 *
 * Build: oracle/ref.mk (target thorin_gpu) compiles this file and links it
 * with the reference's objects, SourceEuler.o's AlgoGas made weak.
 */
#include "fargo.h"
#include "thorin_amd.h"

extern int begin_i;             /* main.c:15 */
extern boolean Restart;         /* main.c:14 */
extern int NbRestart;
extern boolean Write_Density, Write_Velocity;

static thorin_ctx *Ctx;
static PolarGrid *Eta;          /* WriteEta's field (EtaCellCentered is Pebbles.c's own) */
static int EtaPending = -1;      /* the output whose gaseta file holds stale data */
static thorin_planets Tp;       /* the library's copy of the planets */
static thorin_step_log Log;
static struct reb_simulation *Rsim;
static long Ncalls;

static void shim_fail(const char *what) {
  masterprint("\nGPU hydro step: %s: %s\n", what, thorin_last_error());
  prs_exit(1);
}

static void shim_str(char *dst, size_t cap, const char *src, const char *key) {
  size_t n = strlen(src);
  if (n >= cap) {
    masterprint("\nGPU hydro step: parameter %s is longer than %d characters\n", key, (int)cap - 1);
    prs_exit(1);
  }
  memcpy(dst, src, n + 1);
}

/* the parameters the reference read (var.c:19-139, Interpret.c:65-266) */
#define SHIM_STR(fld, g) shim_str(p->fld, sizeof p->fld, g, #g)
static void params_from_reference(thorin_params *p) {
  thorin_params_defaults(p);
  p->dt = (double)DT;
  p->sigma0 = (double)SIGMA0;
  p->ninterm = (int)NINTERM;
  p->ntot = (int)NTOT;
  SHIM_STR(outputdir, OUTPUTDIR);
  SHIM_STR(inputdir, INPUTDIR);
  SHIM_STR(innerboundary, INNERBOUNDARY);
  SHIM_STR(outerboundary, OUTERBOUNDARY);
  SHIM_STR(labeladvection, ADVLABEL);
  SHIM_STR(transport, TRANSPORT);
  SHIM_STR(planetconfig, PLANETCONFIG);
  p->masstaper = (double)MASSTAPER;
  SHIM_STR(gridspacing, GRIDSPACING);
  p->nrad = (int)NRAD;
  p->nsec = (int)NSEC;
  p->rmin = (double)RMIN;
  p->rmax = (double)RMAX;
  p->thicknesssmoothing = (double)THICKNESSSMOOTHING;
  p->rochesmoothing = (double)ROCHESMOOTHING;
  p->aspectratio = (double)ASPECTRATIO;
  p->viscosity = (double)VISCOSITY;
  p->alphaviscosity = (double)ALPHAVISCOSITY;
  SHIM_STR(alphaflock, ALPHAFLOCK);
  p->tmri = (double)TMRI;
  p->twidth = (double)TWIDTH;
  p->alphamri = (double)ALPHAMRI;
  p->alphadead = (double)ALPHADEAD;
  p->sigmaslope = (double)SIGMASLOPE;
  p->releaseradius = (double)RELEASERADIUS;
  p->releasedate = (double)RELEASEDATE;
  p->omegaframe = (double)OMEGAFRAME;
  SHIM_STR(disk, DISK);
  SHIM_STR(frame, FRAME);
  SHIM_STR(outersourcemass, OUTERSOURCEMASS);
  SHIM_STR(writedensity, WRITEDENSITY);
  SHIM_STR(writevelocity, WRITEVELOCITY);
  SHIM_STR(indirectterm, INDIRECTTERM);
  SHIM_STR(excludehill, EXCLUDEHILL);
  p->imposeddiskdrift = (double)IMPOSEDDISKDRIFT;
  p->flaringindex = (double)FLARINGINDEX;
  p->eccentricity = (double)ECCENTRICITY;
  p->cavityradius = (double)CAVITYRADIUS;
  p->cavityratio = (double)CAVITYRATIO;
  p->cavitywidth = (double)CAVITYWIDTH;
  p->transitionradius = (double)TRANSITIONRADIUS;
  p->transitionratio = (double)TRANSITIONRATIO;
  p->transitionwidth = (double)TRANSITIONWIDTH;
  p->lambdadoubling = (double)LAMBDADOUBLING;
  SHIM_STR(energyequation, ENERGYEQUATION);
  SHIM_STR(writetemperature, WRITETEMPERATURE);
  SHIM_STR(writeenergy, WRITEENERGY);
  SHIM_STR(writedivv, WRITEDIVV);
  SHIM_STR(writeqplus, WRITEQPLUS);
  SHIM_STR(writeqbalance, WRITEQBALANCE);
  SHIM_STR(writeopacity, WRITEOPACITY);
  p->adiabind = (double)ADIABIND;
  p->coolingtime = (double)COOLINGTIME;
  SHIM_STR(stellarirradiation, STELLARIRRADIATION);
  p->opacitydrop = (double)OPACITYDROP;
  p->effectivetemperature = (double)EFFECTIVETEMPERATURE;
  p->stellarradius = (double)STELLARRADIUS;
  p->discalbedo = (double)DISCALBEDO;
  p->parametricopacity = (double)PARAMETRICOPACITY;
  SHIM_STR(initializefromfile, INITIALIZEFROMFILE);
  SHIM_STR(densinfile, DENSINFILE);
  SHIM_STR(vradinfile, VRADINFILE);
  SHIM_STR(vthetainfile, VTHETAINFILE);
  SHIM_STR(temperinfile, TEMPERINFILE);
  SHIM_STR(damptowards, DAMPTOWARDS);
  p->dampingrminfrac = (double)DAMPINGRMINFRAC;
  p->dampingrmaxfrac = (double)DAMPINGRMAXFRAC;
  p->dampingperiodfrac = (double)DAMPINGPERIODFRAC;
  p->noutelements = (int)NOUTELEMENTS;
  p->planetarydensity = (double)PLANETARYDENSITY;
  SHIM_STR(resolvecollisions, RESOLVECOLLISIONS);
  p->targetnpl = (int)TARGETNPL;
  p->ias15precission = (double)IAS15PRECISSION;
  p->ias15mindt = (double)IAS15MINDT;
  SHIM_STR(discardplanets, DISCARDPLANETS);
  SHIM_STR(writetorquefiles, WRITETORQUEFILES);
  p->hillcut = (double)HILLCUT;
  p->verticaldamping = (double)VERTICALDAMPING;
  SHIM_STR(planetsfeeldisk, PLANETSFEELDISK);
  p->accretionrate = (double)ACCRETIONRATE;
  SHIM_STR(gasaccretionheating, GASACCRETIONHEATING);
  SHIM_STR(pebbleaccretion, PEBBLEACCRETION);
  SHIM_STR(backreaction, BACKREACTION);
  SHIM_STR(accretionalheating, ACCRETIONALHEATING);
  SHIM_STR(writeeta, WRITEETA);
  p->pebbleflux = (double)PEBBLEFLUX;
  p->pebblealpha = (double)PEBBLEALPHA;
  p->pebblecoagulation = (double)PEBBLECOAGULATION;
  p->pebblebulkdens = (double)PEBBLEBULKDENS;
  p->schmidtnumber = (double)SCHMIDTNUMBER;
  SHIM_STR(particlediffusion, PARTICLEDIFFUSION);
  p->heatingdelay = (int)HEATINGDELAY;
  p->parametricaccretion = (double)PARAMETRICACCRETION;
  SHIM_STR(torquemapinfile, TORQUEMAPINFILE);
  p->gettorqueforplanet = (int)GETTORQUEFORPLANET;
  p->pebbletogasmax = (double)PEBBLETOGASMAX;
  p->fragmentfactor = (double)FRAGMENTFACTOR;
  p->fragmentthreshold = (double)FRAGMENTTHRESHOLD;
  p->fragmentalpha = (double)FRAGMENTALPHA;
  SHIM_STR(opacitytable, OPACITYTABLE);
  p->gasflux = (double)GASFLUX;
  SHIM_STR(damping, DAMPING);
  p->evaporationtemperature = (double)EVAPORATIONTEMPERATURE;
  p->evaporationrate = (double)EVAPORATIONRATE;
  p->firststep = (double)FIRSTSTEP;
  p->minimumstep = (double)MINIMUMSTEP;
  p->maximumstep = (double)MAXIMUMSTEP;
  SHIM_STR(aerodynamicdrag, AERODYNAMICDRAG);
  SHIM_STR(elementswithdisk, ELEMENTSWITHDISK);
  SHIM_STR(pebblegravity, PEBBLEGRAVITY);
  thorin_params_derive(p);
}
#undef SHIM_STR

static void planets_in(const PlanetarySystem *sys, thorin_planets *t) {
  int k;
  t->nb = sys->nb;
  for (k = 0; k < sys->nb; k++) {
    t->mass[k] = sys->mass[k];
    t->gas[k] = sys->gas[k];
    t->solid[k] = sys->solid[k];
    t->x[k] = sys->x[k];
    t->y[k] = sys->y[k];
    t->z[k] = sys->z[k];
    t->vx[k] = sys->vx[k];
    t->vy[k] = sys->vy[k];
    t->vz[k] = sys->vz[k];
    t->ax[k] = sys->ax[k];
    t->ay[k] = sys->ay[k];
    t->az[k] = sys->az[k];
    t->acc[k] = sys->acc[k];
    t->feeldisk[k] = sys->FeelDisk[k] ? 1 : 0;
  }
}

static void planets_out(const thorin_planets *t, PlanetarySystem *sys) {
  int k;
  sys->nb = t->nb;
  for (k = 0; k < t->nb; k++) {
    sys->mass[k] = t->mass[k];
    sys->gas[k] = t->gas[k];
    sys->solid[k] = t->solid[k];
    sys->x[k] = t->x[k];
    sys->y[k] = t->y[k];
    sys->z[k] = t->z[k];
    sys->vx[k] = t->vx[k];
    sys->vy[k] = t->vy[k];
    sys->vz[k] = t->vz[k];
    sys->ax[k] = t->ax[k];
    sys->ay[k] = t->ay[k];
    sys->az[k] = t->az[k];
  }
}

/* the N-body part of a substep (SourceEuler.c:260-269), REBOUND's own: the
 * library's planets (accreted mass, disk kicks) go into sys, AdvanceSystemRebound
 * pushes them into the simulation, and its result comes back */
static int shim_nbody(void *user, thorin_planets *t, double time, double dt, double angle) {
  PlanetarySystem *sys = (PlanetarySystem *)user;
  int k;
  planets_out(t, sys);
  PhysicalTime = time;
  AdvanceSystemRebound(sys, Rsim, dt);
  RotatePsys(Rsim, angle);
  SynchronizeFargoRebound(sys, Rsim);
  t->nb = sys->nb;
  for (k = 0; k < sys->nb; k++) {
    t->x[k] = sys->x[k];
    t->y[k] = sys->y[k];
    t->z[k] = sys->z[k];
    t->vx[k] = sys->vx[k];
    t->vy[k] = sys->vy[k];
    t->vz[k] = sys->vz[k];
    t->mass[k] = sys->mass[k];
  }
  return 0;
}

static void download(int field, PolarGrid *g) {
  if (thorin_download_field(Ctx, field, g->Field, (size_t)NRAD * NSEC)) shim_fail("download");
}

static void upload(int field, PolarGrid *g, size_t n) {
  if (thorin_upload_field(Ctx, field, g->Field, n)) shim_fail("upload");
}

/* the library's context and its initial state (Init.c:72-145) */
static void shim_init(PolarGrid *Rho, PolarGrid *Vrad, PolarGrid *Vtheta, PolarGrid *Energy) {
  thorin_params p;
  thorin_state st;
  const size_t n = (size_t)NRAD * NSEC;
  params_from_reference(&p);
  if (thorin_ctx_create(&p, &Ctx)) shim_fail("context");
  /* the initial state whose ring profiles the damping targets are: InitEuler
   * or the InitializeFromFile arrays, then the restart files over it */
  if (InitFromFile) {
    PolarGrid *d = CreatePolarGrid(NRAD, NSEC, "dens0"), *vr = CreatePolarGrid(NRAD, NSEC, "vrad0");
    PolarGrid *vt = CreatePolarGrid(NRAD, NSEC, "vtheta0"), *t = CreatePolarGrid(NRAD, NSEC, "temper0");
    ReadfromAsciiFile(d, DENSINFILE);
    ReadfromAsciiFile(vr, VRADINFILE);
    ReadfromAsciiFile(vt, VTHETAINFILE);
    ReadfromAsciiFile(t, TEMPERINFILE);
    if (thorin_init_state(Ctx, 1, d->Field, vr->Field, vt->Field, t->Field)) shim_fail("initial state");
  } else if (thorin_init_state(Ctx, 0, NULL, NULL, NULL, NULL)) {
    shim_fail("initial state");
  }
  if (Restart &&
      thorin_init_state(Ctx, 2, Rho->Field, Vrad->Field, Vtheta->Field, Temperature->Field))
    shim_fail("restart state");
  /* the reference's own fields (its restart read them; its energy is the
   * conserved variable) */
  upload(THORIN_F_DENS, Rho, n);
  upload(THORIN_F_VRAD, Vrad, n + NSEC);
  upload(THORIN_F_VTHETA, Vtheta, n);
  upload(THORIN_F_ENERGY, Energy, n);
  if (Pebbles) {             /* Init.c:138-144 */
    if (Restart) {           /* RestartPebbleDisk (Pebbles.c:67-75) */
      PolarGrid *ps = CreatePolarGrid(NRAD, NSEC, "pebblesize0");
      ReadfromFile(ps, "gaspebblesize", 0);
      if (thorin_init_pebbles(Ctx, 1, PebbleDens->Field, PebbleVrad->Field, PebbleVtheta->Field, ps->Field))
        shim_fail("pebble restart");
    } else if (thorin_init_pebbles(Ctx, 0, NULL, NULL, NULL, NULL)) {   /* EquilPebbleDisk */
      shim_fail("pebble disk");
    }
  }
  if (thorin_get_state(Ctx, &st)) shim_fail("state");
  st.physical_time = PhysicalTime;
  st.physical_time_initial = PhysicalTimeInitial;
  st.omega_frame = OmegaFrame;
  if (thorin_set_state(Ctx, &st)) shim_fail("state");
}

/* the fields SendOutput writes (Output.c:202-237, WritePebbles); output <
 * 0: the four the main loop's MonitorIntegral sums read (main.c:216-220) */
static void download_for_output(PolarGrid *Rho, PolarGrid *Vrad, PolarGrid *Vtheta, PolarGrid *Energy,
                                int output) {
  download(THORIN_F_DENS, Rho);
  download(THORIN_F_VRAD, Vrad);
  download(THORIN_F_VTHETA, Vtheta);
  download(THORIN_F_ENERGY, Energy);
  if (output < 0) return;
  if (Write_Temperature || Write_Qbalance) download(THORIN_F_TEMPER, Temperature);
  if (Write_Divergence) download(THORIN_F_DIVV, DivergenceVelocity);
  if (Write_Qplus || Write_Qbalance) download(THORIN_F_QPLUS, Qplus);
  if (Write_Qbalance) download(THORIN_F_QMINUS, Qminus);
  if (Write_Opacity) download(THORIN_F_OPACITY, Opacity);
  if (Pebbles) {
    download(THORIN_F_PEBBLEDENS, PebbleDens);
    download(THORIN_F_PEBBLEVRAD, PebbleVrad);
    download(THORIN_F_PEBBLEVTHETA, PebbleVtheta);
    if (Write_Eta) {
      if (!Eta) Eta = CreatePolarGrid(NRAD, NSEC, "eta");
      download(THORIN_F_ETA, Eta);
      EtaPending = output;
    }
  }
}

void AlgoGas(PolarGrid *Rho, PolarGrid *Vrad, PolarGrid *Vtheta, PolarGrid *Energy, PolarGrid *Label,
             PlanetarySystem *sys, struct reb_simulation *rsim) {
  thorin_state st;
  int k, rc;
  char fn[1024];
  FILE *f;
  (void)Label;
  Rsim = rsim;
  /* WritePebbles wrote its own (never updated) EtaCellCentered; the eta of
   * that output goes over it */
  if (EtaPending >= 0) {
    WriteDiskPolar(Eta, EtaPending);
    EtaPending = -1;
  }
  if (!Ctx) shim_init(Rho, Vrad, Vtheta, Energy);
  planets_in(sys, &Tp);
  rc = thorin_algogas(Ctx, &Tp, shim_nbody, sys, &Log);
  if (rc < 0) shim_fail("AlgoGas");
  planets_out(&Tp, sys);
  if (thorin_get_state(Ctx, &st)) shim_fail("state");
  PhysicalTime = st.physical_time;
  MassTaper = st.mass_taper;
  OmegaFrame = st.omega_frame;
  /* UpdateLog (Force.c:27-70), first substep of the DT */
  for (k = 0; k < Log.tq_rows && WriteTorque; k++) {
    sprintf(fn, "%stqwk%d.dat", OUTPUTDIR, k);
    f = fopenp(fn, "a");
    fprintf(f, "%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\t%#.18g\n", Log.tq[k][0], Log.tq[k][1],
            Log.tq[k][2], Log.tq[k][3], Log.tq[k][4], Log.tq[k][5], Log.tq[k][6]);
    fclose(f);
  }
  /* per substep: the SOR iterations under -d (EnergySources.c:314-315), the
   * crash mark, and before the first crash's mark its message and dump
   * (SourceEuler.c:272-287); the dump holds the fields that substep found */
  for (k = 0; k < Log.nsub; k++) {
    if (debug == YES && EnergyEq) printf("SOR: Niter = %d\n", Log.sor_niter[k]);
    if (Log.crashed[k] && Log.crash_dumped) {
      Log.crash_dumped = 0;
      fprintf(stdout, "\nCrash! at time %.12g\n", Log.crash_time);
      download(THORIN_X_CRASH_DENS, Rho);
      download(THORIN_X_CRASH_VRAD, Vrad);
      download(THORIN_X_CRASH_VTHETA, Vtheta);
      download(THORIN_X_CRASH_TEMPER, Temperature);
      WriteDiskPolar(Rho, 999);
      WriteDiskPolar(Vrad, 999);
      WriteDiskPolar(Vtheta, 999);
      WriteDiskPolar(Temperature, 999);
    }
    masterprint(Log.crashed[k] ? "c" : ".");
    fflush(stdout);
  }
  masterprint("\n");
  /* the main loop's next index; SendOutput runs before its AlgoGas */
  Ncalls++;
  if ((begin_i + Ncalls) % NINTERM == 0)
    download_for_output(Rho, Vrad, Vtheta, Energy, (int)((begin_i + Ncalls) / NINTERM));
  else if (MonitorIntegral == YES)
    download_for_output(Rho, Vrad, Vtheta, Energy, -1);
}
