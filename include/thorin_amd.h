/*
 * thorin_amd.h -- C ABI of the MI355X-native FARGO_THORIN hydro step.
 *
 * The library (fargo_thorin_amd/libthorin_amd.so) keeps every PolarGrid field
 * resident in HBM and advances it by one DT per thorin_algogas() call.  All
 * signatures use plain C types: no torch, no HIP types.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   thorin_algogas        <- AlgoGas()                 src_main/SourceEuler.c:210-317,
 *                                                       prototype src_main/proto.h:74,
 *                                                       called once per DT by main.c:201
 *   thorin_ctx_create     <- InitEuler()+FillPolar1DArrays()+InitComputeAccel()
 *                                                       SourceEuler.c:68-175, SideEuler.c:101-118
 *   thorin_init_state     <- Initialization()          Init.c:72-145 (analytic, InitFromFile and
 *                                                       restart routes)
 *   thorin_init_pebbles   <- EquilPebbleDisk()/RestartPebbleDisk()  Init.c:138-144,
 *                                                       Pebbles.c:59-75, 79-232
 *   thorin_upload_field /
 *   thorin_download_field <- the host PolarGrid ->Field arrays (types.h:37-44) that
 *                            WriteDiskPolar() (Output.c:153-189) and ReadfromFile()
 *                            (Init.c:23-54) read and write
 *   thorin_params_read    <- ReadVariables()/InitVariables()  Interpret.c:65-266, var.c:17-140
 *
 * Layout: a field is (nrad+1) rings of nsec doubles, cell l = j + i*nsec
 * (SourceEuler.c:359); ring nrad is the ghost ring that vrad and the star
 * values use (TransportEuler.c:271, Pframeforce.c:749).
 *
 * Error convention: every entry point returns 0 on success, a negative value
 * on a HIP/allocation/argument error and a positive value for a physics
 * warning (crash detected, SOR not converged); thorin_last_error() describes
 * the last failure.  The library never calls exit().
 */
#ifndef THORIN_AMD_H
#define THORIN_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define THORIN_MAXPLANETS 120   /* types.h:67 MAXPLANETS */
#define THORIN_STRLEN     512

/* Opacity laws (Interpret.c:250-265, Opacity.c).  POWERLAW (kappa = k0*T^2 in
 * cgs, k0 = PowerLawOpacity) has no counterpart in the reference. */
enum { THORIN_OPA_BL94 = 0, THORIN_OPA_LP85 = 1, THORIN_OPA_ZHU12 = 2,
       THORIN_OPA_CONST = 3, THORIN_OPA_POWERLAW = 4, THORIN_OPA_SM03 = 5 };

/* Every in.par key of var.c:19-139, plus the flags ReadVariables derives
 * from them (Interpret.c:149-265). */
typedef struct thorin_params {
  /* var.c numeric keys */
  double dt, sigma0; int ninterm, ntot;
  double masstaper; int nrad, nsec; double rmin, rmax;
  double thicknesssmoothing, rochesmoothing, aspectratio, viscosity, alphaviscosity;
  double tmri, twidth, alphamri, alphadead, sigmaslope, releaseradius, releasedate;
  double omegaframe, imposeddiskdrift, flaringindex, eccentricity;
  double cavityradius, cavityratio, cavitywidth;
  double transitionradius, transitionratio, transitionwidth, lambdadoubling;
  double adiabind, coolingtime, opacitydrop, effectivetemperature, stellarradius;
  double discalbedo, parametricopacity;
  double dampingrminfrac, dampingrmaxfrac, dampingperiodfrac;
  int noutelements; double planetarydensity; int targetnpl;
  double ias15precission, ias15mindt;
  double hillcut, verticaldamping, accretionrate;
  double pebbleflux, pebblealpha, pebblecoagulation, pebblebulkdens, schmidtnumber;
  int heatingdelay;
  double parametricaccretion; int gettorqueforplanet;
  double pebbletogasmax, fragmentfactor, fragmentthreshold, fragmentalpha, gasflux;
  double evaporationtemperature, evaporationrate, firststep, minimumstep, maximumstep;
  double powerlawopacity;          /* POWERLAW opacity coefficient (cgs); not in var.c */
  /* var.c string keys (verbatim) */
  char outputdir[THORIN_STRLEN], inputdir[THORIN_STRLEN];
  char innerboundary[64], outerboundary[64], labeladvection[64], transport[64];
  char planetconfig[THORIN_STRLEN], gridspacing[64], alphaflock[64], disk[64], frame[64];
  char outersourcemass[64], writedensity[64], writevelocity[64], indirectterm[64];
  char excludehill[64], energyequation[64], writetemperature[64], writeenergy[64];
  char writedivv[64], writeqplus[64], writeqbalance[64], writeopacity[64];
  char stellarirradiation[64], initializefromfile[64];
  char densinfile[THORIN_STRLEN], vradinfile[THORIN_STRLEN];
  char vthetainfile[THORIN_STRLEN], temperinfile[THORIN_STRLEN];
  char damptowards[64], resolvecollisions[64], discardplanets[64], writetorquefiles[64];
  char planetsfeeldisk[64], gasaccretionheating[64], pebbleaccretion[64], backreaction[64];
  char accretionalheating[64], writeeta[64], particlediffusion[64], torquemapinfile[64];
  char opacitytable[64], damping[64], aerodynamicdrag[64], elementswithdisk[64];
  char pebblegravity[64];
  /* derived flags (Interpret.c:149-265) */
  int advect_label, outer_source_mass, fast_transport, open_inner, open_outer, non_reflecting;
  int damping_on, damp_vrad, damp_init, damp_flux, log_grid, is_disk, corotating, guiding_center;
  int write_velocity, write_density, indirect_term, exclude_hill, viscosity_alpha, alpha_flock_on;
  int roche_smoothing, energy_eq, write_temperature, write_energy, write_divergence, write_qplus;
  int write_qbalance, write_opacity, parametric_cooling, stellar_irradiation, init_from_file;
  int write_torque, write_torque_map, collisions, monitor_npl, feel_disk, gas_accret_heating;
  int pebbles, back_reaction, diffusive_particles, accret_heating, write_eta;
  int prescribed_accretion, torque_density, discard_planets, aerodynamic_drag;
  int elements_with_disk, pebble_gravity, opacity_law;
} thorin_params;

/* Planetary system as AlgoGas sees it (types.h:96-117), heliocentric. */
typedef struct thorin_planets {
  int nb;
  double mass[THORIN_MAXPLANETS], gas[THORIN_MAXPLANETS], solid[THORIN_MAXPLANETS];
  double x[THORIN_MAXPLANETS], y[THORIN_MAXPLANETS], z[THORIN_MAXPLANETS];
  double vx[THORIN_MAXPLANETS], vy[THORIN_MAXPLANETS], vz[THORIN_MAXPLANETS];
  double ax[THORIN_MAXPLANETS], ay[THORIN_MAXPLANETS], az[THORIN_MAXPLANETS];
  double acc[THORIN_MAXPLANETS];
  int feeldisk[THORIN_MAXPLANETS];
} thorin_planets;

/* Host N-body step that AlgoGas calls between the force evaluation and the
 * boundary conditions: AdvanceSystemRebound + RotatePsys +
 * SynchronizeFargoRebound (SourceEuler.c:260-269, ReboundInterface.c:262-355,
 * Psys.c:336-355).  t is PhysicalTime at the start of the substep. */
typedef int (*thorin_nbody_fn)(void *user, thorin_planets *sys, double t, double dt,
                               double rotation_angle);

/* Scalars AlgoGas keeps between calls (the reference's globals and
 * function-local statics: global.h:14-20, SourceEuler.c:601,
 * EnergySources.c:39-41,210-212). */
typedef struct thorin_state {
  double physical_time;      /* PhysicalTime */
  double physical_time_initial;
  double omega_frame;        /* OmegaFrame */
  double mass_taper;         /* MassTaper */
  double dt_stellar;         /* dt_stellar (EnergySources.c:855), 1e38 initially */
  double sor_omega, sor_domega; /* omega, domega */
  int sor_first, sor_niter_last, sor_niter_best; double sor_omega_best;
  int cfl_count;             /* ConditionCFL's static count */
  int already_crashed;       /* AlreadyCrashed */
} thorin_state;

/* What one thorin_algogas() call did; the log mirrors what the reference
 * prints ('.'/'c' per substep, SourceEuler.c:285-287; SOR Niter under -d,
 * EnergySources.c:314-315). */
#define THORIN_MAXSUB 4096
typedef struct thorin_step_log {
  int nsub;                      /* substeps taken in this DT */
  int crashed_any;               /* any substep reported a crash */
  int cfl_n[THORIN_MAXSUB];      /* GasTimeStepsCFL of each substep */
  double dt[THORIN_MAXSUB];      /* dt of each substep */
  int sor_niter[THORIN_MAXSUB];  /* SOR iterations of each substep (0 without energy eq.) */
  double sor_omega[THORIN_MAXSUB];
  int crashed[THORIN_MAXSUB];
  double lost_mass;              /* LostMass accumulated by VanLeerRadial (OpenInner) */
  /* the first crash of the run (AlreadyCrashed == 0, SourceEuler.c:275-284)
   * happened in this call: crash_time is PhysicalTime at the start of that
   * substep ("Crash! at time"), and THORIN_X_CRASH_* hold dens, vrad, vtheta
   * and the temperature as DetectCrash saw them (the gas*999.dat dumps) */
  int crash_dumped;
  double crash_time;
  /* torque log of the first substep, one row per planet (UpdateLog,
   * Force.c:27-70): PhysicalTime, tq, tq_norm, x, y, ax, ay */
  int tq_rows;
  double tq[8][7];
} thorin_step_log;

/* Field identifiers for upload/download.  Names are the reference's
 * CreatePolarGrid names ("gas"+name is the output file prefix). */
enum {
  THORIN_F_DENS = 0, THORIN_F_VRAD, THORIN_F_VTHETA, THORIN_F_ENERGY, THORIN_F_LABEL,
  THORIN_F_TEMPER, THORIN_F_PRESSURE, THORIN_F_SOUNDSPEED, THORIN_F_QPLUS, THORIN_F_QMINUS,
  THORIN_F_OPACITY, THORIN_F_DIVV, THORIN_F_TAURR, THORIN_F_TAURP, THORIN_F_TAUPP,
  THORIN_F_AGR, THORIN_F_AGT, THORIN_F_ENERGYINT, THORIN_F_ENERGYNEW,
  THORIN_F_VRADINT, THORIN_F_VTHETAINT, THORIN_F_VRADNEW, THORIN_F_VTHETANEW,
  THORIN_F_QBALANCE,
  THORIN_F_PEBBLEDENS, THORIN_F_PEBBLEVRAD, THORIN_F_PEBBLEVTHETA, THORIN_F_PEBBLESIZE,
  THORIN_F_PEBBLESTOKES, THORIN_F_ETA, THORIN_F_TORQUE,
  THORIN_NFIELDS
};

/* Read-only diagnostics for thorin_download_field: the energy equation's
 * five-point system (EnergySources.c:256-297, rings 1..nrad-2) as the last
 * thorin_rad_system assembled it, and the diffusion coefficient.  A gas
 * step forms the rows inside its SOR launches and stores only the
 * right-hand side: after thorin_algogas, THORIN_X_RHS and THORIN_X_DIFF are
 * its last solve's, the five diagonals are left as they were. */
enum { THORIN_X_ML = 64, THORIN_X_MLIP, THORIN_X_MLIM, THORIN_X_MLJP, THORIN_X_MLJM, THORIN_X_RHS,
       THORIN_X_DIFF,
       /* the fields of the run's first crash, taken where DetectCrash tests
        * them (after the boundary conditions, before SubStep1), as
        * WriteDiskPolar(..., 999) writes them (SourceEuler.c:279-282) */
       THORIN_X_CRASH_DENS, THORIN_X_CRASH_VRAD, THORIN_X_CRASH_VTHETA, THORIN_X_CRASH_TEMPER };

/* thorin_algogas refuses a substep whose CFL step count N is not finite or
 * exceeds this (the reference would loop on it); it returns < 0 with the
 * state of the substep's start. */
#define THORIN_MAX_CFL_N 10000000

/* 1D geometry (global.h:11-24) filled exactly as FillPolar1DArrays
 * (SourceEuler.c:68-131) and Theo.c.  Arrays have nrad+1 entries. */
typedef struct thorin_geometry {
  int nrad, nsec;
  const double *radii, *rinf, *rsup, *rmed, *surf, *inv_rmed, *inv_surf;
  const double *inv_diff_rmed, *inv_diff_rsup, *inv_rinf, *omega_inv, *rmed2;
  const double *sigma_med, *sigma_inf, *energy_med, *vtheta_med, *wave_killer;
  const double *globcsvec, *globtempervec;
} thorin_geometry;

typedef struct thorin_ctx thorin_ctx;

/* ---- parameters (Interpret.c:65-266) ---- */
int  thorin_params_defaults(thorin_params *p);
int  thorin_params_read(const char *path, thorin_params *p);
int  thorin_params_derive(thorin_params *p);
/* bindings: heap-allocated defaults, one "KEY value" in.par line at a time,
 * read back a key, read a derived flag */
thorin_params *thorin_params_new(void);
void thorin_params_free(thorin_params *p);
int  thorin_params_set(thorin_params *p, const char *key, const char *value);
int  thorin_params_get(const thorin_params *p, const char *key, char *out, int outlen);
int  thorin_params_flag(const thorin_params *p, const char *name);

/* ---- context ---- */
int  thorin_ctx_create(const thorin_params *p, thorin_ctx **out);
/* host-only context (no device buffers): geometry and initial state only;
 * thorin_host_field() returns the initial dens/vrad/vtheta/energy/label/
 * temper/pressure/soundspeed as Initialization() leaves them */
int  thorin_ctx_create_host(const thorin_params *p, thorin_ctx **out);
int  thorin_host_field(thorin_ctx *ctx, int field, double *out, size_t count);
int  thorin_ctx_destroy(thorin_ctx *ctx);
int  thorin_get_geometry(thorin_ctx *ctx, thorin_geometry *g);
int  thorin_get_state(thorin_ctx *ctx, thorin_state *s);
int  thorin_set_state(thorin_ctx *ctx, const thorin_state *s);

/* Initialization (Init.c:72-145).  mode 0: analytic InitEuler profiles;
 * mode 1: the four host arrays (dens, vrad, vtheta, temper; nrad*nsec each,
 * ring-major) replace the fields as the InitializeFromFile route does
 * (energy = T*rho/(gamma-1), then RefillSigma/RefillEnergy/FillVtheta and
 * cs, P, T); mode 2: restart, the same arrays are the gas*N.dat contents. */
int  thorin_init_state(thorin_ctx *ctx, int mode, const double *dens, const double *vrad,
                       const double *vtheta, const double *temper);

/* Pebble disk (Init.c:138-144), after thorin_init_state when PebbleAccretion
 * is YES.  mode 0: the coagulation-drift equilibrium of a fresh run from the
 * initial gas state (EquilPebbleDisk: EtaPressureSupport + InitPebblesViaFlux,
 * Pebbles.c:59-64, 79-202, 236-270), which also fills PEBBLESIZE,
 * PEBBLESTOKES and ETA; mode 1: restart, the four host arrays (nrad*nsec
 * each: pebble density, vrad, vtheta, size) replace the pebble fields
 * (RestartPebbleDisk, Pebbles.c:67-75) and their ordered ring means become
 * the damping targets (BckpFieldsForBC, Pebbles.c:206-232). */
int  thorin_init_pebbles(thorin_ctx *ctx, int mode, const double *pdens, const double *pvrad,
                         const double *pvtheta, const double *psize);

/* ---- fields ---- */
int  thorin_upload_field(thorin_ctx *ctx, int field, const double *host, size_t count);
int  thorin_download_field(thorin_ctx *ctx, int field, double *host, size_t count);

/* ---- the hot path: one DT of AlgoGas ---- */
int  thorin_algogas(thorin_ctx *ctx, thorin_planets *sys, thorin_nbody_fn nbody,
                    void *nbody_user, thorin_step_log *log);

/* ---- an ensemble of disks advanced together.  No counterpart entry point in
 * the reference, which runs one disk per process; each disk follows
 * exactly the AlgoGas of thorin_algogas (SourceEuler.c:210-317).
 * The disks' host logic runs on one thread each; the GPU work of all disks
 * is merged so that each kernel of a substep is ONE launch over every disk's
 * grid (the disks may differ in size, parameters and substep count).  Each
 * disk's results are bit-identical to thorin_algogas on that disk alone.
 * The contexts stay usable on their own between ensemble calls.
 * thorin_ensemble_algogas advances every disk by one DT: sys, nbody_user,
 * logs and rc_out hold one entry per disk (nbody_user, logs and rc_out may
 * be NULL); nbody is called from the disk's thread.  Returns -1 if any
 * disk failed, else the largest per-disk return code. */
#define THORIN_MAXDISKS 64
typedef struct thorin_ensemble thorin_ensemble;
int  thorin_ensemble_create(thorin_ctx *const *disks, int n, thorin_ensemble **out);
int  thorin_ensemble_algogas(thorin_ensemble *e, thorin_planets *const *sys, thorin_nbody_fn nbody,
                             void *const *nbody_user, thorin_step_log *logs, int *rc_out);
/* Each merged substep (the launches between two host syncs) is captured as
 * a HIP graph and replayed while it repeats (on by default). */
int  thorin_ensemble_set_graphs(thorin_ensemble *e, int on);
/* launches issued so far, the per-disk kernel operations they carried, and
 * the graphs captured and replayed (any pointer may be NULL) */
int  thorin_ensemble_stats(const thorin_ensemble *e, long long *launches, long long *kernel_ops,
                           long long *graph_captures, long long *graph_replays);
void thorin_ensemble_destroy(thorin_ensemble *e);

/* ---- autotuner.  No counterpart in the reference.  Launch shapes that
 * leave every result bit-identical: SOR iterations per launch (1, 2 or 3,
 * with a halo of twice that many cells in LDS and four temperature buffers)
 * and rings per tile
 * of the radial transport (8, 16, 32).  thorin_autotune times each on the
 * context's grid with HIP events (on a synthetic system; the temperature is
 * restored, the energy system is assembled again by the next substep) and
 * keeps the fastest: for the SOR, the shape that finishes a solve of the
 * expected length first (the most iterations of the last solves, 3 before
 * any; a solve runs whole launches); thorin_ensemble_autotune picks one shape
 * for all disks (least time summed over them).  Call them between steps. */
typedef struct thorin_tuning {
  int sor_iters_per_launch;       /* 1, 2 or 3 (default 3) */
  int rad_rows;                   /* 8, 16 or 32 (default 8) */
  double sor_us_per_iter[3];      /* measured: per SOR iteration, 1 / 2 / 3 per launch */
  double rad_us[3];               /* measured: one radial transport launch, 8 / 16 / 32 rings */
  int fused_chain;                /* 1 (default): SubStep1, ViscousTerms and SubStep2 (with the damping
                                     boundary around them) as one kernel where the options allow it (the
                                     gas alone, a damping boundary on vrad only); 0: one kernel per step.
                                     Same bits either way; the autotuner leaves it as it is */
  int sor_assembly;               /* 1 (default): the gas step's energy system is formed inside its SOR
                                     launches; 0: a separate matrix pass assembles it and the SOR loads
                                     it.  Same bits either way; the autotuner leaves it as it is */
} thorin_tuning;
int  thorin_autotune(thorin_ctx *ctx, thorin_tuning *out);
int  thorin_set_tuning(thorin_ctx *ctx, const thorin_tuning *t);
int  thorin_get_tuning(thorin_ctx *ctx, thorin_tuning *out);
int  thorin_ensemble_autotune(thorin_ensemble *e, thorin_tuning *out);

/* Disk quantities the host needs once per DT for the torque log
 * (UpdateLog, Force.c:27-70): ring means of dens and soundspeed
 * (nrad doubles each) and the last disk->planet accelerations in sys. */
int  thorin_ring_means(thorin_ctx *ctx, int field, double *out);

/* Ring profiles of the last substep (nrad doubles): 0 globcsvec,
 * 1 globtempervec (mpiTasks.c:12-62), 2 Vmoy (SourceEuler.c:609-618),
 * 3 VMed and 4 Nshift (TransportEuler.c:85-95). */
int  thorin_ring_profile(thorin_ctx *ctx, int which, double *out);

/* Total gas mass (GasTotalMass, SideEuler.c:26-51), reduced on the GPU. */
int  thorin_gas_total_mass(thorin_ctx *ctx, double *out);

/* Kernel timing with HIP events recorded on the context's stream around the
 * kernels whose id bit is set in mask (ids: thorin_prof_kernel_name).
 * Enabling resets the totals; read returns total milliseconds and launches. */
int  thorin_prof_enable(thorin_ctx *ctx, int on, unsigned long long mask);
int  thorin_prof_read(thorin_ctx *ctx, int kernel_id, double *total_ms, long long *count);
const char *thorin_prof_kernel_name(int kernel_id);

/* ---- the energy equation's implicit solve on its own (EnergySources.c:
 * 200-335), for solver studies.  No counterpart entry point in the reference
 * (there the solve is internal to ImplicitRadiativeDiffusion).
 * thorin_rad_system assembles the system of the current fields for a step dt
 * (SubStep3's T, cs, opacity, Q+, Q-, diffusion coefficients and the
 * five-point matrix) and keeps T as the initial guess; thorin_rad_solve
 * solves it from that guess, leaving T in THORIN_F_TEMPER.
 *   method 0: FP64 red-black SOR of the reference (stop: the in-sweep
 *             sum|A.T - RHS| below tol times its initial value);
 *   method 1: FP32 red-black SOR correction sweeps (inner_sweeps per round,
 *             from e = 0) inside FP64 iterative refinement; stop when the
 *             FP64 sum|A.T - RHS| is below tol times its initial value. */
typedef struct thorin_cnn thorin_cnn;
typedef struct thorin_solve_opts {
  int method, inner_sweeps, max_iters;
  double tol, omega;
  thorin_cnn *warm;          /* learned warm start (method 1 only), NULL: none */
  int sweep_block;           /* method 1: FP32 sweeps per launch, 1..4 (0: 4); more than one
                                runs them on an LDS tile with a halo (same iterates) */
} thorin_solve_opts;
typedef struct thorin_solve_result {
  int sweeps, rounds, converged;
  double ms;                 /* GPU time of the solve (HIP events on the stream) */
  double resid0, resid;      /* FP64 sum|A.T - RHS| of T0 and after */
  int warm_used;             /* the warm start lowered the residual and was kept */
  double resid_warm;         /* FP64 sum|A.T - RHS| of the network's guess */
} thorin_solve_result;
int  thorin_rad_system(thorin_ctx *ctx, double dt, double *dt_used);   /* dt <= 0: the hydro CFL step */
int  thorin_rad_solve(thorin_ctx *ctx, const thorin_solve_opts *o, thorin_solve_result *r);

/* ---- the same solve with no allocation.  No counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * methods, stop tests and bits as thorin_rad_solve, without a warm start (the
 * network allocates its activations) and with kernel timing off.  Where
 * thorin_rad_solve allocates its FP32 planes (matrix diagonals, residual,
 * correction double buffer: 8 floats per cell) on first use, this one takes
 * them from one library-wide device array of 8 * (THORIN_NOALLOC_MAX_CELLS + 64)
 * floats (51 MB), which each GPU the library runs on holds until the process
 * ends; grids with (NRAD+1)*NSEC above THORIN_NOALLOC_MAX_CELLS (1024x1536
 * fits) are refused.  Every call on a GPU shares that GPU's array and clears
 * it first: calls that overlap on one GPU (two host threads, two contexts)
 * race. */
#define THORIN_NOALLOC_MAX_CELLS 1600000
int  thorin_rad_solve_noalloc(thorin_ctx *ctx, const thorin_solve_opts *o, thorin_solve_result *r);

/* ---- learned warm start of the energy solve.  No counterpart in the
 * reference.  A 4-level, 32-channel U-Net (BF16 weights and activations,
 * FP32 accumulation on MFMA, FP32 master weights and Adam) maps the features
 * of an assembled system (the Jacobi correction of T0 scaled by its RMS s,
 * the off-diagonals over the diagonal, log of the diagonal) to
 * (T* - T0)/(T0 s); thorin_rad_solve with opts.warm starts from
 * T0 + prediction when that lowers the residual.  Grids must have NRAD and
 * NSEC divisible by 8.  Parameters: per convolution W[9][cin][cout] then
 * b[cout], FP32; convolutions 16->32, 7 x 32->32, 32->16. */
int    thorin_cnn_create(unsigned long long seed, thorin_cnn **out);
int    thorin_cnn_destroy(thorin_cnn *n);
size_t thorin_cnn_param_count(const thorin_cnn *n);
int    thorin_cnn_get_params(thorin_cnn *n, float *out, size_t count);
int    thorin_cnn_set_params(thorin_cnn *n, const float *in, size_t count);
/* a training pair: features of the system thorin_rad_system assembled and
 * the target from the T that the last thorin_rad_solve left */
int    thorin_cnn_add_pair(thorin_cnn *n, thorin_ctx *ctx);
int    thorin_cnn_num_pairs(const thorin_cnn *n);
/* iters Adam steps over the stored pairs in turn (on the GPU); mean loss of
 * the first and the last pass over the pairs */
int    thorin_cnn_train(thorin_cnn *n, int iters, float lr, double *loss_first, double *loss_last);
/* forward + backward on host arrays (x: H*W*16 features, t: H*W target):
 * loss (mean square over rings 1..H-2), parameter gradient, prediction */
int    thorin_cnn_grad(thorin_cnn *n, int H, int W, const float *x, const float *t, double *loss, float *grad,
                       float *y);
/* the same with the residual loss when aux != NULL: aux holds 5 planes of
 * H*W floats (f0, wip, wim, wjp, wjm) and the loss is the mean square over
 * rings 1..H-2 of rho = -f0 + y + wip y[i+1] + wim y[i-1] + wjp y[j+1] + wjm y[j-1]
 * (y = 0 on rings 0 and H-1); t may be NULL */
int    thorin_cnn_grad2(thorin_cnn *n, int H, int W, const float *x, const float *t, const float *aux,
                        double *loss, float *grad, float *y);
/* training objective: 0 = mean square error of the normalised correction
 * (default); 1 = mean square of the normalised residual of the guess */
int    thorin_cnn_set_loss(thorin_cnn *n, int mode);
/* the network's grid: the system's NRAD x NSEC coarsened f = 1 (default), 2
 * or 4 times along both axes.  With f > 1 the features are averaged over
 * f x f cells, the output is interpolated bilinearly back to the system's
 * grid (cell centres; clamped across the rings, periodic along the sectors)
 * and the residual loss is taken there.  NRAD/f and NSEC/f must be divisible
 * by 8.  Set before the first pair.  In thorin_cnn_grad / _grad2, H and W
 * are then the system's grid: x and t hold (H/f)(W/f) pixels, aux 5 H W, and
 * the prediction comes back on the network's grid. */
int    thorin_cnn_set_coarsen(thorin_cnn *n, int f);
int    thorin_cnn_coarsen(const thorin_cnn *n);

/* ---- the device math library, for checking it.  No counterpart in the
 * reference (whose host builds call glibc's).  out[i] = f(x[i]), or
 * f(x[i], y[i]) for pow and atan2, evaluated on the GPU by the device
 * functions the kernels call, compiled with the kernels' flags; host arrays
 * of n <= 2^26 values. */
enum { THORIN_MATH_POW = 0, THORIN_MATH_EXP, THORIN_MATH_LOG, THORIN_MATH_LOG10, THORIN_MATH_TANH,
       THORIN_MATH_SIN, THORIN_MATH_ATAN2, THORIN_MATH_ASIN, THORIN_MATH_ATAN, THORIN_MATH_SQRT,
       THORIN_MATH_N };
int  thorin_math_eval(int fn, const double *x, const double *y, double *out, size_t n);
/* The context's opacity law evaluated on the GPU as the energy kernels
 * evaluate it (OpacityProfile, Opacity.c:231-262, called from EnergySources.c:237): kappa in
 * cm^2/g from the volume density in g/cm^3 and T in K; host arrays of
 * n <= 2^26 values.  No counterpart entry point in the reference. */
int  thorin_opacity_eval(thorin_ctx *ctx, const double *rho_cgs, const double *T_K, double *kappa_cgs, size_t n);

/* Synchronise the device and return the library's last error text. */
int  thorin_sync(thorin_ctx *ctx);
const char *thorin_last_error(void);

/* ---- host N-body (REBOUND stays host-side; this is the library's own
 * integrator used when no REBOUND is linked) ---- */
int  thorin_planets_from_config(const thorin_params *p, thorin_planets *sys);
int  thorin_nbody_step(void *user, thorin_planets *sys, double t, double dt,
                       double rotation_angle);
/* Osculating elements of planet k about a primary of mass primary_mass (1,
 * or 1 plus the disk mass inside the planet's orbit with ElementsWithDisk):
 * el = a, e, inc, Omega, omega, f (radians), as OutputElements writes them
 * to nbody.orbits.dat (ReboundInterface.c:405-441). */
int  thorin_orbit_elements(const thorin_planets *sys, int k, double primary_mass, double el[6]);

#ifdef __cplusplus
}
#endif
#endif
