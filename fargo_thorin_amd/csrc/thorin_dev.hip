// thorin_dev.hip -- MI355X (gfx950) kernels and the device driver of AlgoGas.
//
// One DT of the reference's AlgoGas (src_main/SourceEuler.c:210-317) runs as
// a sequence of HIP kernels on one stream; every PolarGrid lives in HBM for
// the whole run.  All arithmetic is IEEE FP64 compiled with
// -ffp-contract=off and keeps the reference's operation order, so every
// stencil reproduces the reference bit for bit; only OCML's exp/log/pow/tanh
// differ from glibc by at most a few ulp.  Ring means that feed integer
// decisions (Vmoy, VMed -> Nshift, globcsvec, globtempervec) are summed in
// the reference's sequential order (mpiTasks.c:26-32, TransportEuler.c:86-91,
// SourceEuler.c:609-618); every other reduction is a fixed-order block tree
// and hence run-to-run deterministic.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "thorin_internal.h"

#define HCHECK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      char m_[256];                                                                      \
      snprintf(m_, sizeof m_, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, \
               __LINE__);                                                                \
      thorin_set_error(m_);                                                              \
      return -1;                                                                         \
    }                                                                                    \
  } while (0)

#define DPI 3.14159265358979323844
#define CFLSECURITY 0.5
#define CVNR 1.41
#define SORMAXITERS 100000
#define SOREPS 1.0e-6
#define SORMIN 1.0e-15
#define SORPRECISION 1.0e-15
#define OMEGAUPPER 1.9

static const int TB = 256;          // threads per block of the cell kernels
static const int RB = 256;          // threads per block of the ring kernels
static const int MAXPL = THORIN_MAXPLANETS;

// ---------------------------------------------------------------------------
// device buffers
enum {
  B_DENS = THORIN_F_DENS, B_VRAD = THORIN_F_VRAD, B_VTHETA = THORIN_F_VTHETA,
  B_ENERGY = THORIN_F_ENERGY, B_LABEL = THORIN_F_LABEL, B_TEMPER = THORIN_F_TEMPER,
  B_PRESSURE = THORIN_F_PRESSURE, B_CS = THORIN_F_SOUNDSPEED, B_QPLUS = THORIN_F_QPLUS,
  B_QMINUS = THORIN_F_QMINUS, B_OPACITY = THORIN_F_OPACITY, B_DIVV = THORIN_F_DIVV,
  B_TAURR = THORIN_F_TAURR, B_TAURP = THORIN_F_TAURP, B_TAUPP = THORIN_F_TAUPP,
  B_AGR = THORIN_F_AGR, B_AGT = THORIN_F_AGT, B_EINT = THORIN_F_ENERGYINT,
  B_ENEW = THORIN_F_ENERGYNEW, B_VRADINT = THORIN_F_VRADINT, B_VTHETAINT = THORIN_F_VTHETAINT,
  B_VRADNEW = THORIN_F_VRADNEW, B_VTHETANEW = THORIN_F_VTHETANEW,
  B_DIFF = THORIN_NFIELDS, B_ML, B_MLIP, B_MLIM, B_MLJP, B_MLJM, B_RHS, B_TBACK, B_TEMPER2, B_TEMPER3, B_TEMPER4,
  B_RP, B_RM, B_TP, B_TM, B_XRP, B_XRM, B_XTP, B_XTM, B_XE, B_XRHO,
  // pebble fluid work fields (Pebbles.c:37-56, SourceEuler.c:155-158)
  B_PAGR, B_PAGT, B_PACCR, B_PACCT, B_GACCR, B_GACCT, B_DRAGR, B_DRAGT, B_MACC,
  // the first crash's dens, vrad, vtheta, temperature (SourceEuler.c:279-282)
  B_CRASH_DENS, B_CRASH_VRAD, B_CRASH_VTHETA, B_CRASH_TEMPER,
  NBUF
};
enum {
  B_PDENS = THORIN_F_PEBBLEDENS, B_PVR = THORIN_F_PEBBLEVRAD, B_PVT = THORIN_F_PEBBLEVTHETA,
  B_PSIZE = THORIN_F_PEBBLESIZE, B_TAU = THORIN_F_PEBBLESTOKES
};
#define PEB_MAXPL 8             // planets that accrete pebbles (as the force kernel)

struct DevCtl {                 // device-side control block, mirrored to pinned host memory
  double newdt_raw;             // ConditionCFL minimum before the clamps
  double dt_stellar;            // DiffusionTimestep minimum
  int crash_e, crash_rho;
  int sor_done, sor_niter, sor_iter;
  int sor_cur;                  // T buffer holding the latest iterate (0: B_TEMPER, 1: B_TEMPER2, 2: B_TEMPER3,
                                // 3: B_TEMPER4)
  double sor_normres0, sor_normres;
  double axstar, aystar;
  double axpl[MAXPL], aypl[MAXPL];
  double tq_rho[24], tq_cs[24]; // ring means of rho and cs: the torque log's rings next to the logged
                                // planets (2 each), then the ring of each planet that DampingTW04 damps
  int crash_peb;                // DetectCrashPebbles of this substep
  int acc_fail;                 // IntegrateColumnMass did not converge
  double pacc[PEB_MAXPL][3];    // pebble accretion per planet: dM, dPX, dPY
  double gacc[8][3];            // gas accretion per accreting planet, in order: dM, dPX, dPY
};

static_assert(sizeof(DevCtl) % 8 == 0, "k_cfl_final mirrors DevCtl in 8-byte words");

struct Mat32 {                  // FP32 matrix diagonals, residual and correction double buffer
  float *ml, *mip, *mim, *mjp, *mjm, *r;
  float *e0, *e1;
};

struct TqRings {                // rings of the leading block's ring means (-1: none)
  int r[24];
};

struct PlanetK {                // per-planet constants of FillForcesArrays (Pframeforce.c:93-121)
  double x, y, z, m, rhill, rst2, rst3, rst4, rsm2, rmorevert2;
  int iring;                    // ring of ThicknessSmoothing, -1 outside the grid
  int pad;
};

struct BOff {                   // first workgroup of this disk in a batched launch (0 otherwise),
  int blk0;                     // for the kernels without a Geo argument
};

struct PlanetsK {               // the planets of the force kernel, passed by value (no copy)
  PlanetK p[8];
};

struct Geo {                    // passed by value to every kernel
  int nr, ns;
  int blk0;                     // first workgroup of this disk in a batched launch (0 otherwise)
  int ncb;                      // TB-wide column blocks per ring
  const double *Rinf, *Rsup, *Rmed, *Surf, *InvRmed, *InvSurf, *InvDiffRmed, *InvDiffRsup, *InvRinf;
  const double *OmegaInv, *SigmaMed, *SigmaInf, *EnergyMed, *VthetaMed, *WaveKiller;
  const double *pow_rmed15, *pow_rinf15, *cav_med, *cav_inf, *cfl_pow2, *supp_torque;
  // ring constants of the azimuthal spacing, formed on the host by the
  // expressions the kernels would evaluate per cell:
  //   InvDxTheta[i] = 1 / (2 pi / NSEC * Rmed[i])      InvDxTheta2[i] = 1 / dxtheta / dxtheta
  //   DxThetaA[i] = Rmed[i] * 2 * pi / NSEC             invdphi = 1 / (2 pi / NSEC)
  const double *InvDxTheta, *InvDxTheta2, *DxThetaA;
  double invdphi;
  const int *dampring;
  // cos and sin of the sector angles 2 pi j / NSEC (host libm): a cell's
  // x = Rmed[i] cos_j and y = Rmed[i] sin_j are the reference's
  // CellAbscissa / CellOrdinate (SideEuler.c:101-118) bit for bit
  const double *cosj, *sinj;
  double *gcs, *gtemp, *Vmoy, *VMed, *ringmax;
  double *sig_avr, *ist_avr;    // ring-averaged vertical integrals of FillForcesArrays
  int *Nshift;
  double adiabind, sqrt_adiabind_inv, cv, viscosity, alphavisc, alphamri, alphadead, tmri, twidth;
  double t2si, opa2cu, rho2cgs, stefan, sqrt2pi_inv, sqrt3, opacitydrop, parametricopacity;
  double powerlawopacity, gasflux, omega_frame, lr01, lr12, thicksmooth;
  int visc_alpha, alpha_flock, fast_transport, damp_init, damp_flux, opacity_law, stellar_irr;
  double rstar, tirr4fac;       // stellar irradiation: R_star in AU and (1-A) Teff^4 R_star^2 in code units
  // OpacityTable SM03: the gas table (rows pressures x cols temperatures)
  const double *sm_pg, *sm_tg, *sm_ross;
  int sm_rows, sm_cols;
  double sm_fac;                // rows / log10(p_max / p_min), as the reference forms it per call
  int indirect_term, exclude_hill;
  double hillcut;
  // LP85 thresholds folded on the host (Opacity.c:205-220)
  double lp_t1, lp_t2, lp_t3a, lp_t4, lp_t5a, lp_t6a;
  // pebble fluid
  int pebbles, back_reaction, pebble_gravity;
  double palpha, pbulk, schmidt;
  const double *peb_dens0, *peb_vr0, *peb_vt0;   // PebDensInit, PebVradInit, PebVthetaInit
  const int *pdampring;                          // rings DampPebbles acts on
  double *tau_avr, *phavr, *pist_avr;            // ring means of FillForcesArrays' pebble part
  double *pebmax;                                // CriticalCharTime ring maxima
  double *pvmed;                                 // ordered ring mean of the pebble vtheta
};

// FillForcesArrays' launch: what it depends on besides the fields (the
// planets, MassTaper, the torque log's rings) and what the host needs of it
// afterwards.  A gas-only substep enqueues the next substep's force kernels
// behind its own end (the fields they read are final there), so that they
// run while the host makes its part of the next substep; the next substep
// uses them if its inputs are those (same kernels, same inputs: same bits).
struct ForceSet {
  int valid;
  int nb, want_tq;
  double mt, t;
  double x[8], y[8], z[8], m[8];
  int feel[8];
  int ntq, dslot[8];
  double tq_t[8], tq_x[8], tq_y[8], tq_m[8], tq_frac[8];
  double indir_px, indir_py;
};

// accretion heating sources of one substep (heatsrc, heatsrc_index,
// heatsrc_max of global.h:21-28): k_matrix adds bfac*q[k] to the RHS of cell
// idx[k] for k < n, in k order (EnergySources.c:291-294)
struct HeatK {
  int n, pad;
  int idx[8];
  double q[8];
};

struct thorin_dev {
  hipStream_t s;
  int nr, ns;
  size_t nf;
  double *b[NBUF];
  double *d1base;
  double *trig;                 // cos, sin of the sector angles (2*NSEC)
  double *sm;                   // OpacityTable SM03: pressure grid, temperature grid, Rosseland table
  Geo g;
  DevCtl *ctl;                  // device
  DevCtl *hctl;                 // pinned host mirror

  double *red;                  // block partial sums (forces)
  double *sorp;                 // SOR partial sums (2 passes)
  int *sorc;                    // SOR partial counts
  double *pstellar;             // DiffusionTimestep block minima
  double *mresp;                // tile partials of the initial SOR residual (and k_mp_resid's block partials)
  double *ringtmp;              // per-ring scratch
  int nblk_cells, nblk_sor;
  int opa_filled;
  int sys_form;                 // 1: the SOR's first launch forms the energy system's rows (gas step);
                                // 0: B_ML .. B_RHS hold them (thorin_rad_system, the autotuner)
  double sys_dt;                // the system's dt and accretion heating sources (sys_form)
  HeatK sys_heat;
  int cfl_ready;                // hctl->newdt_raw holds the CFL minimum of the current fields
  int niter_hist[4], nsolve;    // SOR iteration counts of the last four solves
  int rs_rows;                  // rings per workgroup of k_ring_means (LDS holds them whole)
  int sor_nit;                  // SOR iterations per launch (1, 2 or 3; thorin_tune)
  int visc_chain;               // SubStep1 .. SubStep2 as one kernel where the options allow it
  int sor_assembly;             // the gas step's energy system formed inside its SOR launches
  int sor_queue_min;            // ensemble: SOR iterations to queue at least
  ForceSet spec;                // force kernels enqueued ahead for the next substep
  int heat_max;                 // heatsrc_max: set by the pebble accretion when AccretionalHeating is on
  unsigned *hflag;              // pinned host word: k_cfl_final writes the substep's sequence number
                                // there after the control block's mirror (the host polls it)
  unsigned flag_seq;            // the last sequence number given to k_cfl_final
  hipEvent_t ev_t0, ev_t1;      // the energy solve's timing (thorin_rad_solve*)
  double *h_part;               // pinned: block partials of the FP64 residual (nblk_cells)
  int device;                   // the HIP device of the context's stream
  int damp_nlo, damp_hi0;       // k_boundary's ring bands: 0 .. damp_nlo-1 and damp_hi0 .. NRAD-1
  int rad_rows;                 // rings per tile of the radial transport (8, 16, 32; thorin_tune)
  int rad_ready;                // thorin_rad_system has assembled the energy system
  struct Mat32 m32;             // FP32 copies for the mixed-precision solve (allocated on first use)
  double cfl_invdtreb;          // MinStepForRebound term it was formed with
  double *acc_stage;            // pinned staging of the pebble accretion's host part
  size_t acc_stage_n;
  // kernel timing with HIP events (thorin_prof_*)
  int prof_on;
  unsigned long long prof_mask;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> *prof_pend;
  std::vector<hipEvent_t> *prof_pool;
  double prof_ms[64];
  long long prof_cnt[64];
  struct Rec *rec;              // ensemble mode: operations are recorded, not issued (thorin_ensemble_*)
  int tt_npt_min;               // smallest register count per thread of k_transport_theta (ensemble)
  int ec_min;                   // smallest cells per thread of the end-of-substep kernels (ensemble)
};

// kernels that the event timer knows, by launch name
static const char *const KNAMES[] = {
    "k_eos_cfl", "k_cfl_final", "k_forces", "k_forces_reduce", "k_indirect", "k_boundary", "k_substep1",
    "k_divstress", "k_visc_update", "k_substep2",
    "k_matrix", "k_sor_iter", "k_sor_check", "k_transport_rad",
    "k_transport_theta", "k_velocities", "k_substep3", "k_peb_cfl", "k_peb_damp", "k_peb_source",
    "k_peb_substep1", "k_peb_accrete", "k_peb_evaporate", "k_ring_means", "k_eos_cfl_vel", "k_visc_chain", NULL};

// a template instance ("k_name<3>") counts as its kernel
static int kid_of(const char *name) {
  if (name[0] == '(') name++;                 // "(k_name<a, b>)": a macro argument with a comma
  const size_t n = strcspn(name, "<");
  for (int i = 0; KNAMES[i]; i++)
    if (strlen(KNAMES[i]) == n && !strncmp(KNAMES[i], name, n)) return i;
  return -1;
}

// ---------------------------------------------------------------------------
// small device helpers

// FViscosity (Viscosity.c:26-54) at Rmed[i] (pw = pow_rmed15, cav = cav_med)
// or Rinf[i] (pow_rinf15, cav_inf); the radial search of the reference lands
// on ring i, whose ring means are gcs_i and gtemp_i.
__device__ __forceinline__ double fvisc_val(const Geo &g, double gcs_i, double gtemp_i, double pw, double cav) {
  double v = g.viscosity;
  if (g.visc_alpha) v = g.alphavisc * gcs_i * gcs_i * pw;
  if (g.alpha_flock) {
    const double a = (g.alphamri - g.alphadead) * (1.0 - tanh((g.tmri - gtemp_i * g.t2si) / g.twidth)) / 2.0 + g.alphadead;
    v = a * gcs_i * gcs_i * pw;
  }
  return v * cav;
}
__device__ __forceinline__ double fvisc_med(const Geo &g, int i) {
  return fvisc_val(g, g.gcs[i], g.gtemp[i], g.pow_rmed15[i], g.cav_med[i]);
}
__device__ __forceinline__ double fvisc_inf(const Geo &g, int i) {
  return fvisc_val(g, g.gcs[i], g.gtemp[i], g.pow_rinf15[i], g.cav_inf[i]);
}

// x^3 rounded once (double-double product): glibc's pow(x,3.0) is correctly
// rounded in all but pathological cases, so this matches it where OCML's
// pow would differ by an ulp.
__device__ __forceinline__ double cube_cr(double x) {
  double p = x * x;
  double e1 = __builtin_fma(x, x, -p);
  double q = p * x;
  double e2 = __builtin_fma(p, x, -q);
  return q + (e2 + e1 * x);
}

__device__ __forceinline__ double opacity_bl94(double rho, double T) {
  // Opacity.c:31-75
  const double kappa[8] = {2.0e-4, 2.0e16, 0.1, 2.0e81, 1.0e-8, 1.0e-36, 1.5e20, 0.348};
  const double a[8] = {0.0, 0.0, 0.0, 1.0, 2.0 / 3.0, 1.0 / 3.0, 1.0, 0.0};
  const double b[8] = {2.0, -7.0, 0.5, -24.0, 3.0, 10.0, -2.5, 0.0};
#define TEMP(i, j) pow(kappa[j] / kappa[i] * pow(rho, a[j] - a[i]), 1.0 / (b[i] - b[j]))
  double k_ = 1.0e-12;
  if (T <= TEMP(2, 3)) {
    double k1 = kappa[0] * pow(T, b[0]);
    double k2 = kappa[1] * pow(T, b[1]);
    double k3 = kappa[2] * pow(T, b[2]);
    double t1 = 1.0 / (k1 * k1) + 1.0 / (k2 * k2);
    t1 = 1.0 / (t1 * t1);
    double t2 = pow(k3 / (1.0 + 1.0e22 * pow(T, -10.0)), 4.0);
    k_ = pow(t1 + t2, 0.25);
  } else if ((T > TEMP(2, 3)) && (T <= TEMP(3, 4))) k_ = kappa[3] * pow(rho, a[3]) * pow(T, b[3]);
  else if ((T > TEMP(3, 4)) && (T <= TEMP(4, 5))) k_ = kappa[4] * pow(rho, a[4]) * pow(T, b[4]);
  else if ((T > TEMP(4, 5)) && (T <= TEMP(5, 6))) k_ = kappa[5] * pow(rho, a[5]) * pow(T, b[5]);
  else if ((T > TEMP(5, 6)) && (T <= TEMP(6, 7))) k_ = kappa[6] * pow(rho, a[6]) * pow(T, b[6]);
  else k_ = kappa[7];
#undef TEMP
  return k_;
}

// T < k * pow(rho, e), decided from a single-precision estimate of the right
// side where T is clear of it by more than 1e-5 (relative), and from the
// double expression itself otherwise.  The estimate's error is below 1e-6
// for 1e-30 < rho < 1e30 (float rounding of rho, log2f and exp2f within an
// ulp each, e within an ulp), so every decision is the double expression's.
__device__ __forceinline__ bool lt_kpow(double T, double k, double rho, double e, float ef) {
  const float rf = (float)rho;
  if (rf > 1e-30f && rf < 1e30f) {
    const double est = k * (double)exp2f(log2f(rf) * ef);
    if (T < est * (1.0 - 1e-5)) return true;
    if (T > est * (1.0 + 1e-5)) return false;
  }
  return T < k * pow(rho, e);
}

__device__ __forceinline__ double opacity_lp85(const Geo &g, double rho, double T) {
  // Opacity.c:201-227; the constant thresholds come folded from the host
  if (T < g.lp_t1) return 2.0e-4 * T * T;
  if (T < g.lp_t2) return 2.0e16 * pow(T, -7.0);
  if (lt_kpow(T, g.lp_t3a, rho, 1.0 / 15.0, 1.0f / 15.0f)) return 5.0e-3 * T;
  if (T < g.lp_t4) return 2.0e34 * pow(rho, 2.0 / 3.0) * pow(T, -9.0);
  if (lt_kpow(T, g.lp_t5a, rho, 1.0 / 21.0, 1.0f / 21.0f)) return 2.0e-8 * pow(rho, 2.0 / 3.0) * T * T * T;
  if (lt_kpow(T, g.lp_t6a, rho, 4.0 / 75.0, 4.0f / 75.0f)) return 1.0e-36 * pow(rho, 1.0 / 3.0) * pow(T, 10.0);
  return 1.5e20 * rho * pow(T, -5.0 / 2.0);
}

__device__ __forceinline__ double opacity_zhu12(double rho, double T) {
  // Opacity.c:145-196
  double xlop, xlp, xlt, pre;
  pre = 8.3144598 * 1.0e7 / 2.4 * rho * T;
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

// OpacityTable SM03 (Opacity2.c:24-245): the Rosseland mean of Semenov et
// al. (2003) dust and Malygin et al. (2014) gas.  The dust coefficients and
// evaporation temperatures are the published tables the reference holds
// (Opacity2.c:107-137).
// interpolate (Opacity2.c:28-40) of x0 in x[0..n-1] with y = column c of
// the row-major table tab of ncol columns
__device__ __forceinline__ double sm_interp_col(const double *x, const double *tab, int ncol, int c, int n,
                                                double x0) {
  if (n <= 1) return tab[c];
  if (x0 < x[0]) return tab[c];
  if (x0 > x[n - 1]) return tab[(n - 1) * ncol + c];
  int i = 1;
  while ((x[i] < x0) && (i < n - 1)) i++;
  const double y0 = tab[(i - 1) * ncol + c], y1 = tab[i * ncol + c];
  return y0 + (y1 - y0) * (x0 - x[i - 1]) / (x[i] - x[i - 1]);
}

// opacity_malygin (Opacity2.c:44-95), Rosseland table: bilinear in the
// pressure-temperature grid.  The temperature cell is the last i <= cols-3
// with T >= tgrid[i] (the reference's scan; a binary search on the ascending
// grid finds the same i); the pressure cell comes from the log-uniform
// index estimate, clamped
__device__ __forceinline__ double sm_gas(const Geo &g, double rho, double T_gas) {
  const int rows = g.sm_rows, cols = g.sm_cols;
  const double *pg = g.sm_pg, *tg = g.sm_tg, *a = g.sm_ross;
  double tP = 8.3144598 * 1.0e7 / 2.4 * rho * T_gas;
  double tT = T_gas;
  if (tT < tg[0]) return 0.0;
  else if (tT > tg[cols - 1]) tT = tg[cols - 1];
  if (tP < pg[0]) tP = pg[0];
  else if (tP > pg[rows - 1]) tP = pg[rows - 1];
  int lo = 0, hi = cols - 3;               // last i in [0, cols-3] with tg[i] <= tT (tg[0] <= tT holds)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tT >= tg[mid]) lo = mid;
    else hi = mid - 1;
  }
  const int Tlow = lo, Thigh = lo + 1;
  int Plow = (int)(g.sm_fac * log10(tP / pg[0]));
  Plow = Plow < 0 ? 0 : (Plow > rows - 2 ? rows - 2 : Plow);
  const int Phigh = Plow + 1;
  const double denom = 1.0 / ((tg[Thigh] - tg[Tlow]) * (pg[Phigh] - pg[Plow]));
  const double mul1 = (tg[Thigh] - tT) * (pg[Phigh] - tP);
  const double mul2 = (tT - tg[Tlow]) * (pg[Phigh] - tP);
  const double mul3 = (tg[Thigh] - tT) * (tP - pg[Plow]);
  const double mul4 = (tT - tg[Tlow]) * (tP - pg[Plow]);
  return denom * (mul1 * a[Tlow + Plow * cols] + mul2 * a[Thigh + Plow * cols] + mul3 * a[Tlow + Phigh * cols] +
                  mul4 * a[Thigh + Phigh * cols]);
}

__device__ __forceinline__ double opacity_sm03(const Geo &g, double rho, double T_dust) {   // Opacity2.c:99-245
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
  if (T_dust <= 0.9) return 0.9;
  const double logrho = log10(rho);
  // the evaporation temperatures: column c of logtt interpolated in log rho
  // (the arrays below are indexed by constants or selected, so they stay in
  // registers)
  double T_ev[4];
#pragma unroll
  for (int c = 0; c < 4; c++) T_ev[c] = pow(10.0, sm_interp_col(logrr, &logtt[0][0], 4, c, 8, logrho));
  const double tmax1 = T_ev[1] > T_ev[2] ? T_ev[1] : T_ev[2];
  const double tmax2 = T_ev[2] > T_ev[3] ? T_ev[2] : T_ev[3];
  const double T4 = tmax1 < tmax2 ? tmax1 : tmax2;
  // T[k] and dT[k] of the reference's five regime borders
  auto Tk = [&](int k) { return k == 0 ? T_ev[0] : (k == 1 ? 275.0 : (k == 2 ? 425.0 : (k == 3 ? 680.0 : T4))); };
  auto dTk = [](int k) { return k == 2 ? 15.0 : (k == 4 ? 100.0 : 5.0); };
  int KK = 5;
  if (T_dust <= (Tk(0) + dTk(0))) KK = 0;
#pragma unroll
  for (int i = 1; i < 5; i++)
    if ((T_dust > (Tk(i - 1) + dTk(i - 1))) && (T_dust <= (Tk(i) + dTk(i)))) KK = i;
  double aKext = 0.0;
  if (KK != 5) {
    int smooth = 0;
#pragma unroll
    for (int i = 0; i < 5; i++)       // the reference passes the double difference to the int abs()
      if (abs((int)(T_dust - Tk(i))) < dTk(i)) smooth = 1;
    const double TK = Tk(KK), dTK = dTk(KK);
    if (smooth == 1) {
      const double T1 = TK - dTK, T2 = TK + dTK, TD = T_dust - TK;
      int j = (KK + 1) * 6;
      const double aKrL = ((((eD[j - 1] * T1 + eD[j - 2]) * T1 + eD[j - 3]) * T1 + eD[j - 4]) * T1 + eD[j - 5]) * T1 +
                          eD[j - 6];
      double aKrR = 0.0;
      if (KK != 4) {
        j = (KK + 2) * 6;
        aKrR = ((((eD[j - 1] * T2 + eD[j - 2]) * T2 + eD[j - 3]) * T2 + eD[j - 4]) * T2 + eD[j - 5]) * T2 + eD[j - 6];
      }
      const double AA = 0.5e0 * (aKrL - aKrR), BB = 0.5e0 * (aKrL + aKrR);
      const double FF = DPI / 2.0 / dTK;
      aKext = BB - AA * sin(FF * TD);
    } else {
      const int j = (KK + 1) * 6;
      aKext = ((((eD[j - 1] * T_dust + eD[j - 2]) * T_dust + eD[j - 3]) * T_dust + eD[j - 4]) * T_dust + eD[j - 5]) *
                  T_dust + eD[j - 6];
    }
  }
  const double kappa_gas = sm_gas(g, rho, T_dust);
  return aKext > kappa_gas ? aKext : kappa_gas;
}

// SM: the SM03 table law compiled in (its table lookups need more scalar
// registers than SubStep3 has to spare; the other laws run an instance
// without it)
template <bool SM = true>
__device__ __forceinline__ double opacity_law(const Geo &g, double rho, double T) {
  switch (g.opacity_law) {
    case THORIN_OPA_SM03: return SM ? opacity_sm03(g, rho, T) : 0.0;
    case THORIN_OPA_LP85: return opacity_lp85(g, rho, T);
    case THORIN_OPA_ZHU12: return opacity_zhu12(rho, T);
    case THORIN_OPA_CONST: return g.parametricopacity;
    case THORIN_OPA_POWERLAW: return g.powerlawopacity * T * T;
    default: return opacity_bl94(rho, T);
  }
}

// fixed-order block reduction (sum), blockDim.x == n threads, power of two
// Block reductions of N threads (a multiple of the 64-lane wave): a
// butterfly inside each wave (every lane ends with the same value: IEEE
// addition is commutative), then the N/64 wave results in wave order.  sh
// needs N/64 doubles; every thread gets the result.
template <int N>
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (N == 64) return v;
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int k = 1; k < N / 64; k++) r += sh[k];
  __syncthreads();
  return r;
}

// M block sums at once, each added in the order block_sum forms it (the
// butterfly in each wave, then the waves in order): one barrier pair for all
template <int N, int M>
__device__ __forceinline__ void block_sums(double (&v)[M], double *sh) {
#pragma unroll
  for (int m = 0; m < M; m++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[m] += __shfl_xor(v[m], o, 64);
  if (N == 64) return;
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int m = 0; m < M; m++) sh[m * (N / 64) + w] = v[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; m++) {
    double r = sh[m * (N / 64)];
#pragma unroll
    for (int k = 1; k < N / 64; k++) r += sh[m * (N / 64) + k];
    v[m] = r;
  }
  __syncthreads();
}

template <int N>
__device__ __forceinline__ double block_min(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  if (N == 64) return v;
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int k = 1; k < N / 64; k++) r = sh[k] < r ? sh[k] : r;
  __syncthreads();
  return r;
}

template <int N>
__device__ __forceinline__ double block_max(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  if (N == 64) return v;
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int k = 1; k < N / 64; k++) r = sh[k] > r ? sh[k] : r;
  __syncthreads();
  return r;
}

// Cell kernels run on a 1D grid of TB-wide column blocks over rings
// i0 .. i0+nring-1.  Workgroups are dispatched round-robin over the 8 XCDs
// (block b to XCD b % 8), each with its own L2; numbering the logical blocks
// so that XCD x gets the contiguous band x*per .. x*per+per-1 keeps the
// neighbouring rings of a radial stencil in the same L2.  L is the logical
// block (ring-major); false for the padding blocks of the last band.
__device__ __forceinline__ bool xcd_cell(const Geo &g, int i0, int nring, int &i, int &j, int &L, int boff = 0) {
  const int total = g.ncb * nring, per = (total + 7) >> 3;
  const int b = (int)blockIdx.x - g.blk0 - boff;   // boff: leading blocks with another job (a multiple of 8)
  L = (b & 7) * per + (b >> 3);
  if (L >= total) return false;
  const int q = L / g.ncb;
  i = i0 + q;
  j = (L - q * g.ncb) * TB + threadIdx.x;
  return true;
}

// Ordered sum of one ring read by one lane straight from global memory (the
// association of ordered_ring_sum): the next 16 values are loaded before the
// 16 dependent adds of the current ones.
__device__ __forceinline__ double ordered_row_sum_global(const double *__restrict__ row, int ns) {
  double s = 0.0;
  double a[16], b[16];
  int j = 0;
  if (ns >= 16) {
#pragma unroll
    for (int k = 0; k < 16; k++) a[k] = row[k];
    for (j = 16; j + 16 <= ns; j += 16) {
#pragma unroll
      for (int k = 0; k < 16; k++) b[k] = row[j + k];
#pragma unroll
      for (int k = 0; k < 16; k++) s += a[k];
#pragma unroll
      for (int k = 0; k < 16; k++) a[k] = b[k];
    }
#pragma unroll
    for (int k = 0; k < 16; k++) s += a[k];
  }
  for (; j < ns; j++) s += row[j];
  return s;
}

// The same sequential sum formed by one wave: chunks of RS_CHUNK values are
// loaded coalesced (the next chunk's loads in flight while the current one
// is added) and staged through the wave's LDS slice st (RS_CHUNK doubles),
// from which lane 0 adds them in index order.  Every lane of the wave calls
// it; the result is valid in lane 0.  One lane walking the ring through
// global memory waits a memory latency per 16 values; here the wait is one
// latency per chunk, hidden behind the additions.
static const int RS_CH = 8;
static const int RS_CHUNK = 64 * RS_CH;
__device__ __forceinline__ double wave_ordered_row_sum(const double *__restrict__ row, int ns, double *st) {
  const int lane = threadIdx.x & 63;
  double v[RS_CH];
#pragma unroll
  for (int k = 0; k < RS_CH; k++) {
    const int j = k * 64 + lane;
    v[k] = j < ns ? row[j] : 0.0;
  }
  double s = 0.0;
  for (int c0 = 0; c0 < ns; c0 += RS_CHUNK) {
#pragma unroll
    for (int k = 0; k < RS_CH; k++) st[k * 64 + lane] = v[k];
#pragma unroll
    for (int k = 0; k < RS_CH; k++) {
      const int j = c0 + RS_CHUNK + k * 64 + lane;
      v[k] = j < ns ? row[j] : 0.0;
    }
    // LDS operations of one wave complete in order: lane 0 reads what the
    // other lanes wrote above
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0) {
      if (ns - c0 >= RS_CHUNK) {
        // a full chunk: the next 8 values are read while 8 are added
        double a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = st[k];
        for (int q = 8; q < RS_CHUNK; q += 8) {
#pragma unroll
          for (int k = 0; k < 8; k++) b[k] = st[q + k];
#pragma unroll
          for (int k = 0; k < 8; k++) s += a[k];
#pragma unroll
          for (int k = 0; k < 8; k++) a[k] = b[k];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) s += a[k];
      } else {
        for (int q = 0; q < ns - c0; q++) s += st[q];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  return s;
}

// plain round-robin numbering (logical block = launch block) for kernels
// whose cost varies strongly from ring to ring (the force kernel near a
// planet), so that the expensive rings spread over all XCDs
__device__ __forceinline__ bool rr_cell(const Geo &g, int i0, int nring, int &i, int &j, int &L, int boff = 0) {
  L = (int)blockIdx.x - g.blk0 - boff;
  if (L >= g.ncb * nring) return false;
  const int q = L / g.ncb;
  i = i0 + q;
  j = (L - q * g.ncb) * TB + threadIdx.x;
  return true;
}

// the same banding for a grid of ntx x nty tiles (tile x fastest)
__device__ __forceinline__ bool xcd_tile(int ntx, int nty, int &tx, int &ty, int &L, int blk0 = 0) {
  const int total = ntx * nty, per = (total + 7) >> 3;
  const int b = (int)blockIdx.x - blk0;
  L = (b & 7) * per + (b >> 3);
  if (L >= total) return false;
  ty = L / ntx;
  tx = L - ty * ntx;
  return true;
}

// Per-thread part of a block reduction over p[tid], p[tid+NT], ... (that
// order): eight loads are issued before their adds, so the loop is not one
// memory latency per element.
template <int NT>
__device__ __forceinline__ double strided_sum(const double *__restrict__ p, int n) {
  double s = 0.0;
  int b = threadIdx.x;
  for (; b + 7 * NT < n; b += 8 * NT) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = p[b + u * NT];
#pragma unroll
    for (int u = 0; u < 8; u++) s += v[u];
  }
  for (; b < n; b += NT) s += p[b];
  return s;
}

template <int NT>
__device__ __forceinline__ double strided_min(const double *__restrict__ p, int n, double m) {
  // eight loads in flight per round, the last round's guarded (a minimum
  // does not depend on the order)
  for (int b = threadIdx.x; b < n; b += 8 * NT) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = (b + u * NT < n) ? p[b + u * NT] : m;
#pragma unroll
    for (int u = 0; u < 8; u++) m = v[u] < m ? v[u] : m;
  }
  return m;
}

// sequential ring sum in the reference's order (one thread walks the ring).
// The walk goes in batches of 16 LDS values held in two register buffers:
// the reads of batch b+1 are issued before the 16 dependent adds of batch b,
// so each batch's reads have the previous batch's adds to arrive in and the
// chain runs at the FP64 add latency.
__device__ __forceinline__ void ring_ld16(double *r, const double *sh) {
#pragma unroll
  for (int k = 0; k < 16; k++) r[k] = sh[k];
}
__device__ __forceinline__ void ring_add16(double &s, const double *r) {
#pragma unroll
  for (int k = 0; k < 16; k++) s += r[k];
}
__device__ __forceinline__ double ordered_ring_sum(const double *sh, int ns) {
  double s = 0.0;
  const int nb = ns >> 4;                      // full batches of 16
  double a[16], b[16];
  if (nb > 0) ring_ld16(a, sh);
  int q = 0;
  for (; q + 1 < nb; q += 2) {
    ring_ld16(b, sh + (q + 1) * 16);
    ring_add16(s, a);
    if (q + 2 < nb) ring_ld16(a, sh + (q + 2) * 16);
    ring_add16(s, b);
  }
  if (q < nb) ring_add16(s, a);                // an odd batch count: the last one
  for (int j = nb << 4; j < ns; j++) s += sh[j];
  return s;
}

// ---------------------------------------------------------------------------
// End of a substep, one workgroup per ring:
//  * ComputeTemperatureField + ComputeSoundSpeed (SourceEuler.c:765-792,
//    715-737) from rho and the energy, with the ordered ring means
//    globcsvec (and globtempervec when the Flock alpha needs it);
//  * the per-cell part of the next substep's ConditionCFL
//    (SourceEuler.c:589-672: ordered Vmoy, the five inverse time limits).
//    The reference takes the minimum over cells of
//    CFLSECURITY/sqrt(S + invdtpeb_sq + invdtreb_sq) with S the sum of the
//    five squares; every step of that expression is monotone in S (IEEE
//    rounding of +, sqrt and / is), so the minimum is the expression at the
//    largest S.  The block maximum of S goes to ringmax[i] and k_cfl_final
//    applies the rest once;
//  * the ring-averaged vertical integrals of the next FillForcesArrays
//    (Pframeforce.c:150-189), which only depend on globcsvec, and their
//    pebble counterparts from the Stokes-number ring means of k_peb_cfl.
// The ring sums are sequential in j; lanes 0, 1, 2 of the first wave walk
// them side by side while the other waves wait.  The CFL terms that need no
// ring mean are formed before the sums (EC cells per thread) and the sum
// under the square root is completed afterwards in the reference's order.
// With do_eos = 0 the temperature and sound speed already in memory are used.
//
// VEL (the end of a gas-only substep): ComputeVelocities
// (TransportEuler.c:350-381) from the transport's momenta and the damping of
// ApplyBoundaryCondition (SideEuler.c:380-414, ZEROVRAD or no damping: vrad
// only, pointwise, with the ring's lambda) are done here too, and vr, vtheta
// written: the workgroup of ring i forms vtheta of ring i (and of its next
// sector) and vr of rings i and i+1 with the same expressions as the owner
// of each cell, so every value equals what the separate kernels give.
#define EC_MAX 12              // cells per thread: NSEC <= RB * EC_MAX
// instances of EC cells per thread (3, 6, 12): the per-cell CFL terms live
// in registers, so the smallest instance with NSEC <= RB * EC runs
// The kernel runs at four waves per SIMD (at most 128 registers): with four
// workgroups of four waves per CU, 1024 rings are all resident at once, and
// the sequential ring sums of every ring run in one overlapped round rather
// than two.
#define EC_PAD 4
struct VelSrc {                 // VEL: the momenta of the transport and the damping
  const double *rp, *rm, *tp, *tm;
  double step;                  // dt (lambda = WaveKiller * step)
  int damp;                     // ZEROVRAD damping on
};

template <bool VEL, int EC>
__device__ __forceinline__ void b_eos_cfl_t(const Geo &g, const double *__restrict__ rho, const double *__restrict__ e,
                                            double *__restrict__ T, double *__restrict__ cs, double *__restrict__ vr,
                                            double *__restrict__ vt, int do_eos, const int *__restrict__ guard,
                                            const VelSrc &vs) {
  extern __shared__ double sh[];       // (2 or 3)*ns: vt, cs [, T]
  __shared__ double red[RB / 64];
  __shared__ double vmoy, gcs_i, gtemp_i;
  if (guard && *guard == 0) return;
  // rings in XCD bands (XCD x = block % 8 takes rings x*per .. x*per+per-1):
  // the neighbouring rings this ring reads are then mostly in the same L2
  const int bb = (int)blockIdx.x - g.blk0, per = (g.nr + 7) >> 3;
  const int i = (bb & 7) * per + (bb >> 3);
  if (i >= g.nr) return;
  const int ns = g.ns, nr = g.nr, tid = threadIdx.x;
  const bool want_t = g.alpha_flock && do_eos;
  // rows EC_PAD doubles longer than NSEC: lanes 0, 1, 2 read distinct banks
  double *s_vt = sh, *s_cs = sh + (ns + EC_PAD), *s_t = sh + 2 * (ns + EC_PAD);
  const double gam = g.adiabind;
  const size_t base = (size_t)i * ns;
  const bool cfl = (i >= 1 && i < nr);
  const double dxrad = g.Rsup[i] - g.Rinf[i];
  const double dxtheta = g.DxThetaA[i];
  const double mn = (dxtheta < dxrad) ? dxtheta : dxrad;
  // pre: the CFL terms of the sound speed and of vr; adr: the radial
  // shear term; the azimuthal shear needs vtheta of the next sector, taken
  // from LDS after the barrier
  double pre[EC], adr[EC];
#pragma unroll
  for (int k = 0; k < EC; k++) {
    const int j = tid + k * RB;
    pre[k] = 0.0;
    adr[k] = 0.0;
    if (j < ns) {
      const size_t l = base + j;
      double vtl, vrl = 0.0, vrp = 0.0;
      if (VEL) {
        const double rl = rho[l];
        const size_t ljm = (j == 0) ? base + ns - 1 : l - 1;
        vtl = (vs.tp[ljm] + vs.tm[l]) / (rl + rho[ljm] + 1e-15) / g.Rmed[i] - g.Rmed[i] * g.omega_frame;
        if (i > 0) vrl = (vs.rp[l - ns] + vs.rm[l]) / (rl + rho[l - ns] + 1e-20);
        if (vs.damp && g.dampring[i]) {
          const double lam = g.WaveKiller[i] * vs.step;
          vrl = (vrl + lam * 0.0) / (1.0 + lam);
        }
        if (i + 1 < nr) {
          vrp = (vs.rp[l] + vs.rm[l + ns]) / (rho[l + ns] + rl + 1e-20);
          if (vs.damp && g.dampring[i + 1]) {
            const double lam = g.WaveKiller[i + 1] * vs.step;
            vrp = (vrp + lam * 0.0) / (1.0 + lam);
          }
        } else {
          vrp = vr[l + ns];        // the ghost ring (no kernel of the substep writes it)
        }
        vt[l] = vtl;
        vr[l] = vrl;
      } else {
        vtl = vt[l];
      }
      s_vt[j] = vtl;
      double c;
      if (do_eos) {
        const double tmp2 = 1.0 / (rho[l] * 1.0);
        const double tmp1 = e[l] * 1.0 * (gam - 1.0);
        const double t = tmp1 * tmp2;
        c = sqrt(gam * (gam - 1.0) * fabs(e[l]) / rho[l]);
        T[l] = t;
        cs[l] = c;
        if (want_t) s_t[j] = t;
      } else {
        c = cs[l];
      }
      s_cs[j] = c;
      if (cfl) {
        const double invdt1 = c / mn;
        if (!VEL) {
          vrl = vr[l];
          vrp = vr[l + ns];
        }
        const double invdt2 = fabs(vrl) / dxrad;
        double dvr = vrp - vrl;
        if (dvr >= 0.0) dvr = 1e-10; else dvr = -dvr;
        adr[k] = dvr / dxrad;
        pre[k] = invdt1 * invdt1 + invdt2 * invdt2;
      }
    }
  }
  __syncthreads();
  double msum = 0.0;
  if (tid < 3) {
    const double *src = (tid == 0) ? s_vt : ((tid == 1) ? s_cs : s_t);
    if (tid < 2 || want_t) {
      const double m = ordered_ring_sum(src, ns) / (double)ns;
      msum = m;
      if (tid == 0) {
        vmoy = m;
        g.Vmoy[i] = m;
      } else if (tid == 1) {
        gcs_i = m;
        g.gcs[i] = m;
      } else {
        gtemp_i = m;
        g.gtemp[i] = m;
      }
    }
    if (tid == 2 && !want_t) gtemp_i = g.gtemp[i];
  }
  if (tid < 64) {
    // ring-averaged Gaussian vertical integrals of FillForcesArrays from
    // globcsvec: node q on lane q - 1, the terms then added in q order
    const double m = __shfl(msum, 1, 64);
    const double H_avr = m * g.OmegaInv[i] * g.sqrt_adiabind_inv;
    const double H2 = H_avr * H_avr;
    const double r2 = g.Rmed[i] * g.Rmed[i];
    const double deltaz = 3.0 * H_avr / (double)10;
    // pebbles: H_peb = H sqrt(alpha/tau) with the ring-mean Stokes number
    // (Pframeforce.c:160-164, 180-186)
    const double ta = g.pebbles ? g.tau_avr[i] : 1.0;
    double t_sig = 0.0, t_ist = 0.0, t_pist = 0.0;
    if (tid < 10) {
      const int q = tid + 1;
      const double znode = ((double)q - 0.5) * deltaz;
      const double znode2 = znode * znode;
      const double sum = exp(-(0.5 * znode2 / H2));
      t_sig = sum;
      const double d2 = r2 + znode2;
      const double dd = sqrt(d2);
      const double denom = dd * d2;
      t_ist = sum / denom;
      if (g.pebbles) {
        const double pz2 = znode2 * (g.palpha / ta);
        const double pd2 = r2 + pz2;
        const double pd = sqrt(pd2);
        t_pist = sum / (pd * pd2);
      }
    }
    double sig = 0.0, ist = 0.0, pist = 0.0;
#pragma unroll
    for (int q = 0; q < 10; q++) {
      sig += __shfl(t_sig, q, 64);
      ist += __shfl(t_ist, q, 64);
      pist += __shfl(t_pist, q, 64);
    }
    if (tid == 1) {
      g.sig_avr[i] = sig;
      g.ist_avr[i] = ist;
      if (g.pebbles) {
        g.phavr[i] = H_avr * sqrt(g.palpha / ta);
        g.pist_avr[i] = pist;
      }
    }
  }
  __syncthreads();
  double m = -1.0;
  if (cfl) {
    const int viscous = g.visc_alpha || (g.viscosity != 0.0);
    const double visc = viscous ? fvisc_val(g, gcs_i, gtemp_i, g.pow_rmed15[i], g.cav_med[i]) : 0.0;
    const double invdt5 = viscous ? visc * 4. / g.cfl_pow2[i] : 1e-30;
    const double vm = vmoy;
#pragma unroll
    for (int k = 0; k < EC; k++) {
      const int j = tid + k * RB;
      if (j < ns) {
        const double vres = g.fast_transport ? s_vt[j] - vm : s_vt[j];
        const double invdt3 = fabs(vres) / dxtheta;
        // vtheta of the next sector: the value its owner formed (the same
        // expression) or read
        const int jp = (j == ns - 1) ? 0 : j + 1;
        double dvt = s_vt[jp] - s_vt[j];
        if (dvt >= 0.0) dvt = 1e-10; else dvt = -dvt;
        const double a = adr[k], b = dvt / dxtheta;
        double invdt4 = (b > a) ? b : a;
        invdt4 *= 4.0 * CVNR * CVNR;
        const double S = pre[k] + invdt3 * invdt3 + invdt4 * invdt4 + invdt5 * invdt5;
        if (S > m) m = S;
      }
    }
  }
  m = block_max<RB>(m, red);
  if (tid == 0) g.ringmax[i] = m;
}
template <int EC>
__device__ __forceinline__ void b_eos_cfl(const Geo &g, const double *__restrict__ rho, const double *__restrict__ e,
                                          double *__restrict__ T, double *__restrict__ cs, const double *__restrict__ vr,
                                          const double *__restrict__ vt, int do_eos, const int *__restrict__ guard) {
  const VelSrc none = {nullptr, nullptr, nullptr, nullptr, 0.0, 0};
  b_eos_cfl_t<false, EC>(g, rho, e, T, cs, const_cast<double *>(vr), const_cast<double *>(vt), do_eos, guard, none);
}
template <int EC>
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(4)))
k_eos_cfl(Geo g, const double *__restrict__ rho, const double *__restrict__ e, double *__restrict__ T,
          double *__restrict__ cs, const double *__restrict__ vr, const double *__restrict__ vt, int do_eos,
          const int *__restrict__ guard) {
  b_eos_cfl<EC>(g, rho, e, T, cs, vr, vt, do_eos, guard);
}
template <int EC>
__device__ __forceinline__ void b_eos_cfl_vel(const Geo &g, const double *__restrict__ rho,
                                              const double *__restrict__ e, double *__restrict__ T,
                                              double *__restrict__ cs, double *__restrict__ vr, double *__restrict__ vt,
                                              const int *__restrict__ guard, const VelSrc &vs) {
  b_eos_cfl_t<true, EC>(g, rho, e, T, cs, vr, vt, 1, guard, vs);
}
template <int EC>
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(4)))
k_eos_cfl_vel(Geo g, const double *__restrict__ rho, const double *__restrict__ e, double *__restrict__ T,
              double *__restrict__ cs, double *__restrict__ vr, double *__restrict__ vt,
              const int *__restrict__ guard, VelSrc vs) {
  b_eos_cfl_vel<EC>(g, rho, e, T, cs, vr, vt, guard, vs);
}

// ConditionCFL's minimum (SourceEuler.c:620-676): the cell term at the
// largest S of rings 1..NRAD-1, with the pebble limit invdtpeb_sq
// (CriticalCharTime, the maximum of k_peb_cfl's ring maxima) and the N-body
// limit invdtreb_sq added in the reference's order, then the shear limit
// between rings.  The minimum of the DiffusionTimestep partials of this
// substep's k_substep3 (nstellar of them; 0: none formed) goes to
// ctl->dt_stellar.
// one workgroup of CF_T threads: the partials and ring values are few per
// thread, so their loads take about one memory latency
#define CF_T 1024
__device__ __forceinline__ void cfl_final_body(const Geo &g, DevCtl *ctl, const double *__restrict__ pstellar,
                                               int nstellar, double invdtreb_sq, DevCtl *__restrict__ hout) {
  // the four extrema (largest S and pebble term, smallest stellar partial and
  // shear limit) are independent: each thread forms its share of all four,
  // then one reduction; min and max are exact, so the order does not matter
  __shared__ double red[4][CF_T / 64];
  double ms = 1.0e38;
  double S = -1.0, P = -1.e30, msh = 1e30;
  const int tid = threadIdx.x, nr = g.nr;
  auto shear = [&](double va, double vb, int i) {
    const double dt = 2.0 * DPI * CFLSECURITY / (double)g.ns / fabs(va * g.InvRmed[i] - vb * g.InvRmed[i + 1]);
    if (dt < msh) msh = dt;
  };
  if (nr <= CF_T) {
    // one ring per thread: the ring values are loaded with the first batch
    // of stellar partials, one memory latency for all
    const double v = (tid >= 1 && tid < nr) ? g.ringmax[tid] : -1.0;
    const double u = (g.pebbles && tid >= 1 && tid < nr) ? g.pebmax[tid] : -1.e30;
    const double va = (tid < nr - 1) ? g.Vmoy[tid] : 0.0, vb = (tid < nr - 1) ? g.Vmoy[tid + 1] : 0.0;
    if (nstellar > 0) ms = strided_min<CF_T>(pstellar, nstellar, ms);
    S = v > S ? v : S;
    P = u > P ? u : P;
    if (tid < nr - 1) shear(va, vb, tid);
  } else {
    if (nstellar > 0) ms = strided_min<CF_T>(pstellar, nstellar, ms);
    for (int i = 1 + tid; i < nr; i += CF_T) {
      const double v = g.ringmax[i];
      S = v > S ? v : S;
      if (g.pebbles) { const double u = g.pebmax[i]; P = u > P ? u : P; }
    }
    for (int i = tid; i < nr - 1; i += CF_T) shear(g.Vmoy[i], g.Vmoy[i + 1], i);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double u = __shfl_xor(S, o, 64);
    S = u > S ? u : S;
    u = __shfl_xor(P, o, 64);
    P = u > P ? u : P;
    u = __shfl_xor(ms, o, 64);
    ms = u < ms ? u : ms;
    u = __shfl_xor(msh, o, 64);
    msh = u < msh ? u : msh;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = S;
    red[1][w] = P;
    red[2][w] = ms;
    red[3][w] = msh;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    // the waves' extrema across the first wave's lanes (the other lanes hold
    // the values of a thread without data)
    const int k = threadIdx.x;
    S = k < CF_T / 64 ? red[0][k] : -1.0;
    P = k < CF_T / 64 ? red[1][k] : -1.e30;
    ms = k < CF_T / 64 ? red[2][k] : 1.0e38;
    msh = k < CF_T / 64 ? red[3][k] : 1e30;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double u = __shfl_xor(S, o, 64);
      S = u > S ? u : S;
      u = __shfl_xor(P, o, 64);
      P = u > P ? u : P;
      u = __shfl_xor(ms, o, 64);
      ms = u < ms ? u : ms;
      u = __shfl_xor(msh, o, 64);
      msh = u < msh ? u : msh;
    }
  }
  if (threadIdx.x == 0) {
    if (nstellar > 0) ctl->dt_stellar = ms;
    double m = 1e30;
    if (g.pebbles) {
      const double dt = CFLSECURITY / sqrt(S + P + invdtreb_sq);
      if (dt < m) m = dt;
    } else {
      const double dt = CFLSECURITY / sqrt(S + invdtreb_sq);
      if (dt < m) m = dt;
    }
    if (msh < m) m = msh;
    ctl->newdt_raw = m;
  }
  if (hout) {
    // the control block straight into the host's pinned mirror by the first
    // wave: the end of the substep needs no copy of its own
    __syncthreads();
    const unsigned long long *src = (const unsigned long long *)ctl;
    unsigned long long *dst = (unsigned long long *)hout;
    if (threadIdx.x < 64)
      for (int q = threadIdx.x; q < (int)(sizeof(DevCtl) / 8); q += 64) dst[q] = src[q];
  }
}

// ConditionCFL's final reduction (nothing while *guard == 0: the SOR has not
// converged) and, with hflag, the substep's sequence number written into the
// host's flag word after the mirror by the same wave, behind a system-scope
// fence: the host polls the word instead of waiting for an event, whose
// system-scope release at this point held the next kernel back by ~6 us
__device__ __forceinline__ void b_cfl_final(const Geo &g, DevCtl *ctl, const double *__restrict__ pstellar,
                                            int nstellar, double invdtreb_sq, const int *__restrict__ guard,
                                            DevCtl *__restrict__ hout, unsigned *__restrict__ hflag, unsigned seq) {
  if (!guard || *guard != 0) cfl_final_body(g, ctl, pstellar, nstellar, invdtreb_sq, hout);
  if (hflag && threadIdx.x < 64) {
    __threadfence_system();
    if (threadIdx.x == 0) __hip_atomic_store(hflag, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
__global__ void __launch_bounds__(CF_T) k_cfl_final(Geo g, DevCtl *ctl, const double *__restrict__ pstellar,
                                                  int nstellar, double invdtreb_sq, const int *__restrict__ guard,
                                                  DevCtl *__restrict__ hout, unsigned *__restrict__ hflag,
                                                  unsigned seq) {
  b_cfl_final(g, ctl, pstellar, nstellar, invdtreb_sq, guard, hout, hflag, seq);
}


// ---------------------------------------------------------------------------
// FillForcesArrays (Pframeforce.c:43-510).  One thread per cell; per-block
// partial sums of the disk->star and disk->planet accelerations.  PEB adds
// the pebble fluid's accelerations pagr, pagt (vertical extent
// H_peb = H sqrt(alpha/tau)) and, with PebbleGravity, the pebbles' pull on
// the planets; the pebble scale height also decides where the integration
// is local for both fluids (Pframeforce.c:206-214).
struct PebF {
  const double *prho, *tau;
  double *pagr, *pagt;
};

template <bool PEB, int NP>
__device__ __forceinline__ void b_forces(const Geo &g, const double *__restrict__ rho,
                                               const double *__restrict__ cs,
                                               double *__restrict__ agr, double *__restrict__ agt,
                                               const PlanetsK &pk, int npl, double mt,
                                               double *__restrict__ partial, const TqRings &tq, int ntq_rings,
                                               DevCtl *__restrict__ ctl, PebF pf,
                                              const int *__restrict__ guard) {
  __shared__ double red[TB];
  if (guard && *guard == 0) return;
  const int ns = g.ns;
  const int lead = ntq_rings > 0 ? 1 : 0;
  if ((int)blockIdx.x - g.blk0 < lead) {
    // leading block: the torque log's ring means of rho next to the logged
    // planets (UpdateLog, Force.c:41-52), wave w summing rings tq.r[w],
    // tq.r[w + 4], ..., and the sound-speed ring means of the same rings
    // (globcsvec); ring -1 gives 0
    __shared__ double st[TB / 64][RS_CHUNK];
    const int w = threadIdx.x >> 6;
    for (int b = w; b < ntq_rings; b += TB / 64) {
      const int r = tq.r[b];
      double m = 0.0, c = 0.0;
      if (r >= 0) {
        m = wave_ordered_row_sum(rho + (size_t)r * ns, ns, st[w]) / (double)ns;
        c = g.gcs[r];
      }
      if ((threadIdx.x & 63) == 0) {
        ctl->tq_rho[b] = m;
        ctl->tq_cs[b] = c;
      }
    }
    return;
  }
  int i, j, bid;
  if (!rr_cell(g, 0, g.nr, i, j, bid, lead)) return;
  const bool act = j < ns;
  double axs = 0.0, ays = 0.0;
  // NP: size of the per-planet arrays (1 for one planet: fewer registers)
  const int np = npl < NP ? npl : NP;
  double axpl_loc[NP], aypl_loc[NP];
  for (int k = 0; k < NP; k++) { axpl_loc[k] = 0.0; aypl_loc[k] = 0.0; }
  if (act) {
    const size_t l = (size_t)i * ns + j;
    const double omegainv = g.OmegaInv[i];
    const double r2 = g.Rmed[i] * g.Rmed[i];
    double Rsup_Rinf2 = g.Rsup[i] - g.Rinf[i];
    Rsup_Rinf2 *= Rsup_Rinf2;
    // ring average (cs_avr equals globcsvec[i]: the same sequential sum); its
    // vertical integrals were formed with the ring means
    const double H_avr = g.gcs[i] * omegainv * g.sqrt_adiabind_inv;
    double H2, zmax, deltaz;
    const double sigma_avr = g.sig_avr[i], integstar_avr = g.ist_avr[i];
    int IntegrateLocally = 0, Nvert = 10;
    double ax = 0.0, ay = 0.0;
    const double x = g.Rmed[i] * g.cosj[j], y = g.Rmed[i] * g.sinj[j];
    const double mcell = rho[l] * g.Surf[i];
    const double H = cs[l] * omegainv * g.sqrt_adiabind_inv;
    H2 = H * H;
    if (fabs(H - H_avr) / H_avr > 0.05) IntegrateLocally = 1;
    double pmcell = 0.0, Hcorr = 0.0, pH = 0.0;
    if (PEB) {
      pmcell = pf.prho[l] * g.Surf[i];
      Hcorr = sqrt(g.palpha / pf.tau[l]);
      pH = H * Hcorr;
      const double p_H_avr = g.phavr[i];
      if (fabs(pH - p_H_avr) / p_H_avr > 0.05) IntegrateLocally = 1;
    }
    unsigned int withpl = 0;    // bit k: IntegrateWithPlanet[k] (npl <= 8 on this path)
    double s2[NP], hillcut[NP];
    for (int k = 0; k < np; k++) {
      const PlanetK P = pk.p[k];
      double smoothing = 0.0;
      if (P.iring >= 0) smoothing = g.thicksmooth * (g.gcs[P.iring] * g.OmegaInv[P.iring] * g.sqrt_adiabind_inv);
      const double Hpl = smoothing / g.thicksmooth;
      s2[k] = (x - P.x) * (x - P.x) + (y - P.y) * (y - P.y);
      hillcut[k] = 1.0;
      double r2chk_1 = 10.0 * Hpl;
      r2chk_1 *= r2chk_1;
      double r2chk_2 = 5.0 * P.rhill;
      r2chk_2 *= r2chk_2;
      if (mt == 1.0 && (s2[k] < r2chk_1 || s2[k] < r2chk_2)) {
        IntegrateLocally = 1;
        withpl |= 1u << k;
        double d2 = s2[k] + P.z * P.z;
        if (g.exclude_hill) hillcut[k] = 1.0 / (exp(-(sqrt(d2) / P.rhill - g.hillcut) / g.hillcut * 10.0) + 1.0);
      }
      if (s2[k] < P.rmorevert2) Nvert = 30;
    }
    double sigma, integstar, integpl[NP], p_integstar = 0.0, p_integpl[NP];
    for (int k = 0; k < NP; k++) { integpl[k] = 0.0; p_integpl[k] = 0.0; }
    if (IntegrateLocally) {
      zmax = 3.0 * H;
      deltaz = zmax / (double)Nvert;
      sigma = 0.0;
      integstar = 0.0;
      for (int m = 1; m <= Nvert; m++) {
        double znode = ((double)m - 0.5) * deltaz;
        double znode2 = znode * znode;
        double sum = exp(-(0.5 * znode2 / H2));
        sigma += sum;
        double d2 = r2 + znode2;
        double d = sqrt(d2);
        double denom = d * d2;
        integstar += sum / denom;
        if (PEB) {
          d2 = r2 + znode2 * Hcorr * Hcorr;
          d = sqrt(d2);
          denom = d * d2;
          p_integstar += sum / denom;
        }
        for (int k = 0; k < np; k++) {
          if (withpl & (1u << k)) {
            const PlanetK P = pk.p[k];
            double zpl = P.z, taper;
            d2 = s2[k] + (znode - zpl) * (znode - zpl);
            if (d2 < P.rst2) {
              d = sqrt(d2);
              taper = -3.0 * d2 * d2 / P.rst4 + 4.0 * d * d2 / P.rst3;
            } else taper = 1.0;
            d2 += (P.rsm2 > Rsup_Rinf2 ? P.rsm2 : Rsup_Rinf2);
            d = sqrt(d2);
            denom = d * d2;
            integpl[k] += sum / denom * taper;
            if (zpl < -0.05 * H || zpl > 0.05 * H) {
              d2 = s2[k] + (-znode - zpl) * (-znode - zpl);
              if (d2 < P.rst2) {
                d = sqrt(d2);
                taper = -3.0 * d2 * d2 / P.rst4 + 4.0 * d * d2 / P.rst3;
              } else taper = 1.0;
              d2 += (P.rsm2 > Rsup_Rinf2 ? P.rsm2 : Rsup_Rinf2);
              d = sqrt(d2);
              denom = d * d2;
              integpl[k] += sum / denom * taper;
            }
            if (PEB) {
              const double zp = znode * Hcorr;
              d2 = s2[k] + (zp - zpl) * (zp - zpl);
              if (d2 < P.rst2) {
                d = sqrt(d2);
                taper = -3.0 * d2 * d2 / P.rst4 + 4.0 * d * d2 / P.rst3;
              } else taper = 1.0;
              d2 += (P.rsm2 > Rsup_Rinf2 ? P.rsm2 : Rsup_Rinf2);
              d = sqrt(d2);
              denom = d * d2;
              p_integpl[k] += sum / denom * taper;
              if (zpl < -0.05 * pH || zpl > 0.05 * pH) {
                const double zm = -znode * Hcorr;
                d2 = s2[k] + (zm - zpl) * (zm - zpl);
                if (d2 < P.rst2) {
                  d = sqrt(d2);
                  taper = -3.0 * d2 * d2 / P.rst4 + 4.0 * d * d2 / P.rst3;
                } else taper = 1.0;
                d2 += (P.rsm2 > Rsup_Rinf2 ? P.rsm2 : Rsup_Rinf2);
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
      if (PEB) p_integstar = g.pist_avr[i];
    }
    double tmp = integstar / sigma;
    double ar = -g.Rmed[i] * tmp;
    double agr_l = 0.0 + ar;
    axs = mcell * tmp * x;
    ays = mcell * tmp * y;
    double pagr_l = 0.0, pax = 0.0, pay = 0.0;
    if (PEB) {
      tmp = p_integstar / sigma;
      pagr_l = 0.0 + (-g.Rmed[i] * tmp);
    }
    for (int k = 0; k < np; k++) {
      const PlanetK P = pk.p[k];
      double mpl = P.m;
      if (withpl & (1u << k)) {
        if (P.z < -0.05 * H || P.z > 0.05 * H) tmp = (x - P.x) * integpl[k] / (2.0 * sigma);
        else tmp = (x - P.x) * integpl[k] / sigma;
        ax += -mpl * tmp;
        if (g.exclude_hill) tmp *= hillcut[k];
        axpl_loc[k] = mcell * tmp;
        if (P.z < -0.05 * H || P.z > 0.05 * H) tmp = (y - P.y) * integpl[k] / (2.0 * sigma);
        else tmp = (y - P.y) * integpl[k] / sigma;
        ay += -mpl * tmp;
        if (g.exclude_hill) tmp *= hillcut[k];
        aypl_loc[k] = mcell * tmp;
        if (PEB) {
          if (P.z < -0.05 * pH || P.z > 0.05 * pH) tmp = (x - P.x) * p_integpl[k] / (2.0 * sigma);
          else tmp = (x - P.x) * p_integpl[k] / sigma;
          pax += -mpl * tmp;
          if (g.pebble_gravity) {
            if (g.exclude_hill) tmp *= hillcut[k];
            axpl_loc[k] += pmcell * tmp;
          }
          if (P.z < -0.05 * pH || P.z > 0.05 * pH) tmp = (y - P.y) * p_integpl[k] / (2.0 * sigma);
          else tmp = (y - P.y) * p_integpl[k] / sigma;
          pay += -mpl * tmp;
          if (g.pebble_gravity) {
            if (g.exclude_hill) tmp *= hillcut[k];
            aypl_loc[k] += pmcell * tmp;
          }
        }
      } else {
        double smoothing = 0.0;
        if (P.iring >= 0) smoothing = g.thicksmooth * (g.gcs[P.iring] * g.OmegaInv[P.iring] * g.sqrt_adiabind_inv);
        double smoothing2 = smoothing * smoothing;
        mpl *= mt;
        double d2 = s2[k] + P.z * P.z + smoothing2;
        double d = sqrt(d2);
        double denom = d * d2;
        tmp = (x - P.x) / denom;
        ax += -mpl * tmp;
        if (g.exclude_hill) tmp *= hillcut[k];
        axpl_loc[k] = mcell * tmp;
        tmp = (y - P.y) / denom;
        ay += -mpl * tmp;
        if (g.exclude_hill) tmp *= hillcut[k];
        aypl_loc[k] = mcell * tmp;
        if (PEB) {
          d2 += -smoothing2 + smoothing2 * Hcorr * Hcorr;
          d = sqrt(d2);
          denom = d * d2;
          tmp = (x - P.x) / denom;
          pax += -mpl * tmp;
          if (g.pebble_gravity) {
            if (g.exclude_hill) tmp *= hillcut[k];
            axpl_loc[k] += pmcell * tmp;
          }
          tmp = (y - P.y) / denom;
          pay += -mpl * tmp;
          if (g.pebble_gravity) {
            if (g.exclude_hill) tmp *= hillcut[k];
            aypl_loc[k] += pmcell * tmp;
          }
        }
      }
    }
    ar = (x * ax + y * ay) * g.InvRmed[i];
    double at = (x * ay - y * ax) * g.InvRmed[i];
    agr_l += ar;
    agr[l] = agr_l;
    agt[l] = 0.0 + at;
    if (PEB) {
      ar = (x * pax + y * pay) * g.InvRmed[i];
      at = (x * pay - y * pax) * g.InvRmed[i];
      pf.pagr[l] = pagr_l + ar;
      pf.pagt[l] = 0.0 + at;
    }
  }
  // partial sums of this block: [axstar, aystar, axpl[k], aypl[k]], reduced
  // together in one pass, each by block_sum's tree (lanes by xor shuffles,
  // then the waves in order)
  const int nb = g.ncb * g.nr;
  constexpr int NV = 2 + 2 * NP, NW = TB / 64;
  static_assert(NV * NW <= TB, "the reduction rows fit the block's LDS buffer");
  double v[NV];
  v[0] = axs;
  v[1] = ays;
#pragma unroll
  for (int k = 0; k < NP; k++) {
    v[2 + 2 * k] = axpl_loc[k];
    v[3 + 2 * k] = aypl_loc[k];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] += __shfl_xor(v[q], o, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < NV; q++) red[q * NW + (threadIdx.x >> 6)] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < 2 + 2 * np; q++) {
      double r = red[q * NW];
#pragma unroll
      for (int w = 1; w < NW; w++) r += red[q * NW + w];
      partial[(size_t)q * nb + bid] = r;
    }
}
template <bool PEB, int NP>
__global__ void __launch_bounds__(TB) k_forces(Geo g, const double *__restrict__ rho,
                                               const double *__restrict__ cs,
                                               double *__restrict__ agr, double *__restrict__ agt,
                                               const PlanetsK pk, int npl, double mt,
                                               double *__restrict__ partial, TqRings tq, int ntq_rings,
                                               DevCtl *__restrict__ ctl, PebF pf,
                                               const int *__restrict__ guard) {
  b_forces<PEB, NP>(g, rho, cs, agr, agt, pk, npl, mt, partial, tq, ntq_rings, ctl, pf, guard);
}


// reduce the partial sums in block order; one block per quantity
__device__ __forceinline__ void b_forces_reduce(const double *__restrict__ partial, int nb, DevCtl *ctl, BOff bo,
                                                const int *__restrict__ guard) {
  __shared__ double red[RB];
  if (guard && *guard == 0) return;
  const int q = (int)blockIdx.x - bo.blk0;
  double s = 0.0;
  s = strided_sum<RB>(partial + (size_t)q * nb, nb);
  s = block_sum<RB>(s, red);
  if (threadIdx.x == 0) {
    if (q == 0) { ctl->crash_e = 0; ctl->crash_rho = 0; ctl->crash_peb = 0; }   // DetectCrash flags of this substep
    if (q == 0) ctl->axstar = s;
    else if (q == 1) ctl->aystar = s;
    else if ((q & 1) == 0) ctl->axpl[(q - 2) / 2] = s;
    else ctl->aypl[(q - 3) / 2] = s;
  }
}
__global__ void __launch_bounds__(RB) k_forces_reduce(const double *__restrict__ partial, int nb, DevCtl *ctl, BOff bo,
                                                      const int *__restrict__ guard) {
  b_forces_reduce(partial, nb, ctl, bo, guard);
}


// indirect term (Pframeforce.c:488-509)
__device__ __forceinline__ void b_indirect(const Geo &g, double *__restrict__ agr, double *__restrict__ agt,
                                                                                                  const DevCtl *__restrict__ ctl, double indir_px, double indir_py,
                                                 double *__restrict__ pagr, double *__restrict__ pagt,
                                              const int *__restrict__ guard) {
  const int ns = g.ns;
  int i, j, L_;
  if (guard && *guard == 0) return;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  double axindir = -ctl->axstar, ayindir = -ctl->aystar;
  if (g.indirect_term) { axindir += indir_px; ayindir += indir_py; }
  const size_t l = (size_t)i * ns + j;
  const double x = g.Rmed[i] * g.cosj[j], y = g.Rmed[i] * g.sinj[j];
  const double ar = (x * axindir + y * ayindir) * g.InvRmed[i];
  const double at = (x * ayindir - y * axindir) * g.InvRmed[i];
  agr[l] += ar;
  agt[l] += at;
  if (pagr) {
    pagr[l] += ar;
    pagt[l] += at;
  }
}
__global__ void __launch_bounds__(TB) k_indirect(Geo g, double *__restrict__ agr, double *__restrict__ agt,
                                                                                                  const DevCtl *__restrict__ ctl, double indir_px, double indir_py,
                                                 double *__restrict__ pagr, double *__restrict__ pagt,
                                                 const int *__restrict__ guard) {
  b_indirect(g, agr, agt, ctl, indir_px, indir_py, pagr, pagt, guard);
}


// ---------------------------------------------------------------------------
// Pebble fluid (Pebbles.c), the second fluid of the two-fluid runs.

// Start of a substep, one workgroup per ring: PebbleStokesNumbers
// (Pebbles.c:272-293), the ordered ring mean of the Stokes numbers that
// FillForcesArrays uses (Pframeforce.c:151-163) and the per-cell part of
// CriticalCharTime (Pebbles.c:735-770) with its ordered ring mean of the
// pebble vtheta; lanes 0 and 1 of the first wave walk the two ring sums side
// by side.  The ring maximum of the four squared inverse times goes to
// pebmax[i] (ring 0 is not part of CriticalCharTime).
__device__ __forceinline__ void b_peb_cfl(const Geo &g, const double *__restrict__ rho,
                                                const double *__restrict__ psize, double *__restrict__ tau,
                                                const double *__restrict__ pvr, const double *__restrict__ pvt,
                                                const double *__restrict__ vr, const double *__restrict__ vt,
                                                const int *__restrict__ guard) {
  extern __shared__ double sh[];       // 2*ns: tau, pebble vtheta
  __shared__ double red[RB / 64];
  __shared__ double pvtmed;
  if (guard && *guard == 0) return;
  const int i = (int)blockIdx.x - g.blk0, ns = g.ns, tid = threadIdx.x;
  double *s_tau = sh, *s_pvt = sh + ns;
  const size_t base = (size_t)i * ns;
  for (int j = tid; j < ns; j += RB) {
    const size_t l = base + j;
    const double t = 2.0 * g.pbulk * psize[l] / rho[l] * g.sqrt_adiabind_inv;
    tau[l] = t;
    s_tau[j] = t;
    s_pvt[j] = pvt[l];
  }
  __syncthreads();
  if (tid < 2) {
    const double m = ordered_ring_sum(tid == 0 ? s_tau : s_pvt, ns) / (double)ns;
    if (tid == 0) g.tau_avr[i] = m;
    else pvtmed = m;
  }
  __syncthreads();
  double m = -1.0;
  if (i >= 1) {
    const double dxrad = g.Rsup[i] - g.Rinf[i];
    const double dxtheta = g.DxThetaA[i];
    const double pm = pvtmed;
    for (int j = tid; j < ns; j += RB) {
      const size_t l = base + j;
      const double pr = pvr[l], pt = s_pvt[j];
      const double invdt1 = fabs(pr) / dxrad;
      const double vres = g.fast_transport ? pt - pm : pt;
      const double invdt2 = fabs(vres) / dxtheta;
      const double invdt3 = fabs(pr - vr[l]) / dxrad;
      const double invdt4 = fabs(pt - vt[l]) / dxtheta;
      const double S = invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4;
      if (S > m) m = S;
    }
  }
  m = block_max<RB>(m, red);
  if (tid == 0) g.pebmax[i] = m;
}
__global__ void __launch_bounds__(RB) k_peb_cfl(Geo g, const double *__restrict__ rho,
                                                const double *__restrict__ psize, double *__restrict__ tau,
                                                const double *__restrict__ pvr, const double *__restrict__ pvt,
                                                const double *__restrict__ vr, const double *__restrict__ vt,
                                                const int *__restrict__ guard) {
  b_peb_cfl(g, rho, psize, tau, pvr, pvt, vr, vt, guard);
}


// DampPebbles (SideEuler.c:423-454) on the rings of the damping zones; with
// detect, DetectCrashPebbles (SourceEuler.c:274, 48-66) first looks at every
// cell's density before the damping.
__device__ __forceinline__ void b_peb_damp(const Geo &g, double *__restrict__ pdens, double *__restrict__ pvr,
                                                 double *__restrict__ pvt, double dt, int detect, DevCtl *ctl,
                                                 const int *__restrict__ guard) {
  const int ns = g.ns;
  int i, j, L_;
  if (guard && *guard == 0) return;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  if (!detect && !g.pdampring[i]) return;
  const size_t l = (size_t)i * ns + j;
  const double pd = pdens[l];
  if (detect && pd < 0.0) ctl->crash_peb = 1;
  if (!g.pdampring[i]) return;
  const double lambda = g.WaveKiller[i] * dt;
  const double pdens0 = g.peb_dens0[i], pvr0 = g.peb_vr0[i];
  const double pvt0 = g.peb_vt0[i] - g.Rmed[i] * g.omega_frame;
  pdens[l] = (pd + lambda * pdens0) / (1.0 + lambda);
  pvr[l] = (pvr[l] + lambda * pvr0) / (1.0 + lambda);
  pvt[l] = (pvt[l] + lambda * pvt0) / (1.0 + lambda);
}
__global__ void __launch_bounds__(TB) k_peb_damp(Geo g, double *__restrict__ pdens, double *__restrict__ pvr,
                                                 double *__restrict__ pvt, double dt, int detect, DevCtl *ctl,
                                                 const int *__restrict__ guard) {
  b_peb_damp(g, pdens, pvr, pvt, dt, detect, ctl, guard);
}


// SourceTermsPebbles after its DampPebbles (Pebbles.c:549-612): pebble
// accelerations paccr (rings >= 1) and pacct and, with BackReaction, the
// drag on the gas dragr (rings >= 1) and dragt.
__device__ __forceinline__ void b_peb_source(const Geo &g, const double *__restrict__ pdens,
                                                   const double *__restrict__ pvr, const double *__restrict__ pvt,
                                                   const double *__restrict__ tau, const double *__restrict__ vr,
                                                   const double *__restrict__ vt, const double *__restrict__ pagr,
                                                   const double *__restrict__ pagt, double *__restrict__ paccr,
                                                   double *__restrict__ pacct, double *__restrict__ dragr,
                                                   double *__restrict__ dragt) {
  const int ns = g.ns;
  int i, j, L_;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  const size_t l = (size_t)i * ns + j;
  const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
  const size_t ljp = (j == ns - 1) ? (size_t)i * ns : l + 1;
  const double Omega = 1.0 / g.OmegaInv[i];
  if (i >= 1) {
    const size_t lim = l - ns;
    double vt2 = pvt[l] + pvt[ljp] + pvt[lim] + pvt[ljp - ns];
    vt2 = vt2 / 4.0 + g.Rinf[i] * g.omega_frame;
    vt2 = vt2 * vt2;
    paccr[l] = vt2 * g.InvRinf[i] + 0.5 * (pagr[l] + pagr[lim]);
    if (g.back_reaction) {
      const double drag = Omega * (pvr[l] - vr[l]) / tau[l];
      dragr[l] = 0.5 * (pdens[l] + pdens[lim]) * drag;
    }
  }
  pacct[l] = 0.5 * (pagt[l] + pagt[ljm]);
  if (g.back_reaction) {
    const double drag = Omega * (pvt[l] - vt[l]) / tau[l];
    dragt[l] = 0.5 * (pdens[l] + pdens[ljm]) * drag;
  }
}
__global__ void __launch_bounds__(TB) k_peb_source(Geo g, const double *__restrict__ pdens,
                                                   const double *__restrict__ pvr, const double *__restrict__ pvt,
                                                   const double *__restrict__ tau, const double *__restrict__ vr,
                                                   const double *__restrict__ vt, const double *__restrict__ pagr,
                                                   const double *__restrict__ pagt, double *__restrict__ paccr,
                                                   double *__restrict__ pacct, double *__restrict__ dragr,
                                                   double *__restrict__ dragt) {
  b_peb_source(g, pdens, pvr, pvt, tau, vr, vt, pagr, pagt, paccr, pacct, dragr, dragt);
}


// SubStep1Pebbles (Pebbles.c:616-650), the semi-implicit drag step with the
// gas velocities from before SubStep1, then ParticleDiffusion
// (Pebbles.c:654-695), whose corrections depend on the two densities only
// (so one pass over the cells does both).
__device__ __forceinline__ void b_peb_substep1(const Geo &g, double *__restrict__ pvr, double *__restrict__ pvt,
                                                     const double *__restrict__ tau, const double *__restrict__ paccr,
                                                     const double *__restrict__ pacct, const double *__restrict__ accr,
                                                     const double *__restrict__ acct, const double *__restrict__ vr,
                                                     const double *__restrict__ vt, const double *__restrict__ rho,
                                                     const double *__restrict__ pdens, double dt, int diffuse) {
  const int ns = g.ns;
  int i, j, L_;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  const size_t l = (size_t)i * ns + j;
  const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
  const double Omega = 1.0 / g.OmegaInv[i];
  const double t = tau[l];
  const double efac = exp(-dt * Omega / t);
  double nvr = pvr[l];
  if (i > 0) nvr = nvr * efac + accr[l] * dt + (vr[l] + (paccr[l] - accr[l]) * t / Omega) * (1.0 - efac);
  double nvt = pvt[l] * efac + acct[l] * dt + (vt[l] + (pacct[l] - acct[l]) * t / Omega) * (1.0 - efac);
  if (diffuse) {
    if (i >= 1) {
      const size_t lim = l - ns;
      double corr = -fvisc_inf(g, i) / g.schmidt;
      corr *= (rho[l] + rho[lim]) / (pdens[l] + pdens[lim]);
      corr *= (pdens[l] / rho[l] - pdens[lim] / rho[lim]) * g.InvDiffRmed[i];
      nvr += corr;
    }
    const double invdxtheta = g.InvDxTheta[i];
    double corr = -fvisc_med(g, i) / g.schmidt;
    corr *= (rho[l] + rho[ljm]) / (pdens[l] + pdens[ljm]);
    corr *= (pdens[l] / rho[l] - pdens[ljm] / rho[ljm]) * invdxtheta;
    nvt += corr;
  }
  pvr[l] = nvr;
  pvt[l] = nvt;
}
__global__ void __launch_bounds__(TB) k_peb_substep1(Geo g, double *__restrict__ pvr, double *__restrict__ pvt,
                                                     const double *__restrict__ tau, const double *__restrict__ paccr,
                                                     const double *__restrict__ pacct, const double *__restrict__ accr,
                                                     const double *__restrict__ acct, const double *__restrict__ vr,
                                                     const double *__restrict__ vt, const double *__restrict__ rho,
                                                     const double *__restrict__ pdens, double dt, int diffuse) {
  b_peb_substep1(g, pvr, pvt, tau, paccr, pacct, accr, acct, vr, vt, rho, pdens, dt, diffuse);
}


// EvaporatePebbles (Pebbles.c:837-866)
__device__ __forceinline__ void b_peb_evaporate(const Geo &g, double *__restrict__ pdens, const double *__restrict__ T,
                                                      double facc, double evaptemp) {
  const int ns = g.ns;
  int i, j, L_;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  const size_t l = (size_t)i * ns + j;
  if (T[l] > evaptemp) {
    double d = pdens[l] * (1.0 - facc);
    if (d < 1.0e-8) d = 1.0e-8;          // EVAPOREDSIGMA, Pebbles.c:835
    pdens[l] = d;
  }
}
__global__ void __launch_bounds__(TB) k_peb_evaporate(Geo g, double *__restrict__ pdens, const double *__restrict__ T,
                                                      double facc, double evaptemp) {
  b_peb_evaporate(g, pdens, T, facc, evaptemp);
}


// IntegrateColumnMass (Pebbles.c:325-342) over the refinements of Trapzd
// (Pebbles.c:295-320); ok = 0 if it does not converge in TRAPEZMAX steps
#define TRAPEZMAX 16
#define TRAPEZEPS 1.0e-4
__device__ double column_mass(double a, double b, double Hpeb, int &ok) {
  const double H2 = Hpeb * Hpeb;
  double integral = 0.0, iold = 0.0;
  for (int n = 1; n <= TRAPEZMAX; n++) {
    if (n == 1) {
      integral = 0.5 * (b - a) * (exp(-0.5 * a * a / H2) + exp(-0.5 * b * b / H2));
    } else {
      int it = 1;
      for (int j = 1; j < n - 1; j++) it <<= 1;
      const double iter = it;
      const double spacing = (b - a) / iter;
      double x = a + 0.5 * spacing, sum = 0.0;
      for (int j = 1; j <= it; j++) {
        sum += exp(-0.5 * x * x / H2);
        x += spacing;
      }
      integral = 0.5 * (integral + (b - a) * sum / iter);
    }
    if (n > 5)
      if (fabs(integral - iold) < TRAPEZEPS * fabs(iold) || (integral == 0.0 && iold == 0.0)) return integral;
    iold = integral;
  }
  ok = 0;
  return 0.0;
}

// AccreteOntoPlanets (Planet.c:33-163): gas accretion onto the planets with
// an accretion rate, taken in order (each sees the density the previous one
// left).  The window, Roche radius and Kley's fractions come from the host
// with the reference's expressions; every cell of the window is one thread's.
// qtrap/trapzd (Planet.c:176-214) integrate exp(-z^2) to QTRAPEPS = 1e-4.
struct GasAccK {
  int np, pad;
  double x[8], y[8], z[8], rroche[8], facc1[8], facc2[8];
  int imin[8], imax[8], jmin[8], jmax[8];
};
#define GACC_T 256
__device__ double gas_qtrap(double a, double b) {
  double s = 0.0, olds = 0.0;
  for (int n = 1; n <= 16; n++) {
    if (n == 1) {
      s = 0.5 * (b - a) * (exp(-a * a) + exp(-b * b));
    } else {
      int it = 1;
      for (int j = 1; j < n - 1; j++) it <<= 1;
      const double tnm = it, del = (b - a) / tnm;
      double x = a + 0.5 * del, sum = 0.0;
      for (int j = 1; j <= it; j++, x += del) sum += exp(-x * x);
      s = 0.5 * (s + (b - a) * sum / tnm);
    }
    if (n > 5)
      if (fabs(s - olds) < 1.e-4 * fabs(olds) || (s == 0.0 && olds == 0.0)) return s;
    olds = s;
  }
  return s;
}
__global__ void __launch_bounds__(GACC_T) k_gas_accrete(Geo g, GasAccK P, double *__restrict__ dens,
                                                        const double *__restrict__ vrad,
                                                        const double *__restrict__ vtheta,
                                                        const double *__restrict__ cs, DevCtl *ctl) {
  __shared__ double red[GACC_T / 64];
  const int ns = g.ns, tid = threadIdx.x;
  const double frac1 = 0.75, frac2 = 0.45;
  for (int q = 0; q < P.np; q++) {
    const int i0 = P.imin[q], i1 = P.imax[q], j0 = P.jmin[q], nj = P.jmax[q] - P.jmin[q] + 1;
    const int n = (i1 >= i0 && nj > 0) ? (i1 - i0 + 1) * nj : 0;
    const double X = P.x[q], Y = P.y[q], Z = P.z[q], RRoche = P.rroche[q], facc1 = P.facc1[q], facc2 = P.facc2[q];
    double dm = 0.0, dpx = 0.0, dpy = 0.0;
    for (int c = tid; c < n; c += GACC_T) {
      const int i = i0 + c / nj;
      int jf = j0 + c % nj;
      while (jf < 0) jf += ns;
      while (jf >= ns) jf -= ns;
      const size_t l = (size_t)i * ns + jf, lip = l + ns;
      const size_t ljp = (jf == ns - 1) ? (size_t)i * ns : l + 1;
      const double xc = g.Rmed[i] * g.cosj[jf], yc = g.Rmed[i] * g.sinj[jf];
      const double dx = X - xc, dy = Y - yc;
      const double distance = sqrt(dx * dx + dy * dy);
      const double vtcell = 0.5 * (vtheta[l] + vtheta[ljp]) + g.Rmed[i] * g.omega_frame;
      const double vrcell = 0.5 * (vrad[l] + vrad[lip]);
      const double vxcell = (vrcell * xc - vtcell * yc) / g.Rmed[i];
      const double vycell = (vrcell * yc + vtcell * xc) / g.Rmed[i];
      const double H = cs[l] * g.OmegaInv[i] * g.sqrt_adiabind_inv;
      if (distance < frac1 * RRoche) {
        const double dz = sqrt(pow(frac1 * RRoche, 2) - pow(distance, 2));
        const double z1 = Z - dz, z2 = Z + dz;
        double factor = gas_qtrap(z1 / (sqrt(2.0) * H), z2 / (sqrt(2.0) * H));
        factor *= 1.0 / (sqrt(2.0 * DPI) * H);
        const double deltaM = facc1 * factor * dens[l] * g.Surf[i];
        dens[l] *= (1.0 - facc1 * factor);
        dpx += deltaM * vxcell;
        dpy += deltaM * vycell;
        dm += deltaM;
      }
      if (distance < frac2 * RRoche) {
        const double dz = sqrt(pow(frac2 * RRoche, 2) - pow(distance, 2));
        const double z1 = Z - dz, z2 = Z + dz;
        double factor = gas_qtrap(z1 / (sqrt(2.0) * H), z2 / (sqrt(2.0) * H));
        factor *= 1.0 / (sqrt(2.0 * DPI) * H);
        const double deltaM = facc2 * factor * dens[l] * g.Surf[i];
        dens[l] *= (1.0 - facc2);                // the reference removes facc2, not facc2*factor
        dpx += deltaM * vxcell;
        dpy += deltaM * vycell;
        dm += deltaM;
      }
    }
    dm = block_sum<GACC_T>(dm, red);
    dpx = block_sum<GACC_T>(dpx, red);
    dpy = block_sum<GACC_T>(dpy, red);
    if (tid == 0) {
      ctl->gacc[q][0] = dm;
      ctl->gacc[q][1] = dpx;
      ctl->gacc[q][2] = dpy;
    }
    __syncthreads();
  }
}

// planet-side part of AccretePebblesOntoPlanets formed on the host with the
// reference's expressions and summation order (Pebbles.c:374-441: headwind,
// transition mass, the annulus' mean Stokes number and the effective
// accretion radius Reff; cells at the edge of Reff make the accreted mass
// depend on Reff through sqrt(Reff^2 - R^2), so Reff is kept bit-exact)
struct PebAccK {
  int np, pad;
  double x[PEB_MAXPL], y[PEB_MAXPL], z[PEB_MAXPL], m[PEB_MAXPL], om[PEB_MAXPL];
  double reff[PEB_MAXPL], vrel[PEB_MAXPL], mt[PEB_MAXPL];
  int imin[PEB_MAXPL], imax[PEB_MAXPL];
  int jlo[PEB_MAXPL], jn[PEB_MAXPL];   // azimuthal window jlo .. jlo+jn-1 (mod NSEC) holding every cell within Reff
};

// AccretePebblesOntoPlanets (Pebbles.c:400-538), STEPS 3-5: one workgroup
// that takes the planets in order (each sees the pebble density the previous
// one left).  The sums over the cells inside Reff are tree sums.  The mass
// and the x- and y-momentum each planet takes from the disk go to
// ctl->pacc; the host updates the planet as STEP 7 does.
#define ACC_T 1024
__device__ __forceinline__ void b_peb_accrete(const Geo &g, const PebAccK &P, double *__restrict__ pdens,
                                                       const double *__restrict__ pvr, const double *__restrict__ pvt,
                                                       const double *__restrict__ tau, const double *__restrict__ cs,
                                                                                                              double *__restrict__ macc, double dt, DevCtl *ctl) {
  __shared__ double red[ACC_T / 64];
  __shared__ double s_lim;
  const int tid = threadIdx.x, ns = g.ns;
  if (tid == 0) ctl->acc_fail = 0;
  for (int k = 0; k < P.np; k++) {
    const int i0 = P.imin[k], i1 = P.imax[k], jn = P.jn[k], jlo = P.jlo[k];
    const int nann = (i1 >= i0) ? (i1 - i0 + 1) * jn : 0;
    const double X = P.x[k], Y = P.y[k], Z = P.z[k], M = P.m[k];
    // cell q of the window: ring i0 + q / jn, sector jlo + q % jn (mod NSEC)
    auto cell = [&](int q, int &i) -> size_t {
      i = i0 + q / jn;
      int j = jlo + q % jn;
      if (j >= ns) j -= ns;
      return (size_t)i * ns + j;
    };
    // STEP 3: upper limit of the mass each cell below Reff can give
    const double Reff = P.reff[k], Reff2 = Reff * Reff;
    double du = 0.0, hp = 0.0, pdm = 0.0, nc = 0.0;
    for (int q = tid; q < nann; q += ACC_T) {
      int i;
      const size_t l = cell(q, i);
      const int jc = (int)(l - (size_t)i * ns);
      const double xc = g.Rmed[i] * g.cosj[jc], yc = g.Rmed[i] * g.sinj[jc];
      const double R2 = (xc - X) * (xc - X) + (yc - Y) * (yc - Y);
      if (R2 <= Reff2) {
        const double H = cs[l] * g.OmegaInv[i] * g.sqrt_adiabind_inv;
        const double Hpeb = H * sqrt(g.palpha / tau[l]);
        const double zmin = Z - sqrt(Reff2 - R2), zmax = Z + sqrt(Reff2 - R2);
        const double pd = pdens[l];
        const double mcell = g.Surf[i] * pd;
        const double vol = pd / Hpeb * g.sqrt2pi_inv;
        int ok = 1;
        const double cm = column_mass(zmin, zmax, Hpeb, ok);
        if (!ok) ctl->acc_fail = 1;
        double dM = vol * g.Surf[i] * cm;
        if (dM > mcell) dM = 0.99999999 * mcell;
        macc[l] = dM;
        du += dM;
        hp += Hpeb;
        pdm += pd;
        nc += 1.0;
      }
    }
    du = block_sum<ACC_T>(du, red);
    hp = block_sum<ACC_T>(hp, red);
    pdm = block_sum<ACC_T>(pdm, red);
    nc = block_sum<ACC_T>(nc, red);
    // STEP 4: the model accretion rate limits the upper bound
    if (tid == 0) {
      const int ncell = (int)nc;
      if (ncell > 0) {
        hp /= (double)ncell;
        pdm /= (double)ncell;
      }
      const double veloc = (M < P.mt[k]) ? P.vrel[k] : Reff * P.om[k];
      double dM_model;
      if (hp < Reff) dM_model = 2.0 * Reff * veloc * pdm;
      else dM_model = DPI * Reff * Reff * veloc * pdm / hp * g.sqrt2pi_inv;
      dM_model *= dt;
      s_lim = (du > dM_model) ? dM_model / du : 1.0;
    }
    __syncthreads();
    // STEP 5: take the mass from the cells, collect mass and momentum
    const double limiter = s_lim;
    double dm = 0.0, dpx = 0.0, dpy = 0.0;
    for (int q = tid; q < nann; q += ACC_T) {
      int i;
      const size_t l = cell(q, i);
      const double mc = macc[l];
      if (mc > 0.0) {
        const int jc = (int)(l - (size_t)i * ns);
        const double xc = g.Rmed[i] * g.cosj[jc], yc = g.Rmed[i] * g.sinj[jc];
        const double vrc = pvr[l], vtc = pvt[l] + g.Rmed[i] * g.omega_frame;
        const double vxc = (vrc * xc - vtc * yc) * g.InvRmed[i];
        const double vyc = (vrc * yc + vtc * xc) * g.InvRmed[i];
        const double dM = mc * limiter;
        macc[l] = 0.0;
        const double mcell = g.Surf[i] * pdens[l];
        double pn = (mcell - dM) / g.Surf[i];
        if (pn < 0.0) pn = 1.e-20;
        pdens[l] = pn;
        dm += dM;
        dpx += dM * vxc;
        dpy += dM * vyc;
      }
    }
    dm = block_sum<ACC_T>(dm, red);
    dpx = block_sum<ACC_T>(dpx, red);
    dpy = block_sum<ACC_T>(dpy, red);
    if (tid == 0) {
      ctl->pacc[k][0] = dm;
      ctl->pacc[k][1] = dpx;
      ctl->pacc[k][2] = dpy;
    }
    __syncthreads();
  }
}
__global__ void __launch_bounds__(ACC_T) k_peb_accrete(Geo g, PebAccK P, double *__restrict__ pdens,
                                                       const double *__restrict__ pvr, const double *__restrict__ pvt,
                                                       const double *__restrict__ tau, const double *__restrict__ cs,
                                                                                                              double *__restrict__ macc, double dt, DevCtl *ctl) {
  b_peb_accrete(g, P, pdens, pvr, pvt, tau, cs, macc, dt, ctl);
}


// ---------------------------------------------------------------------------
// DampingBoundary (SideEuler.c:368-419) on the rings flagged by the host
// The launch covers only the two bands of damped rings: rings 0 .. nlo-1 and
// hi0 .. NRAD-1 (DRin and DRout bound contiguous bands), one TB-wide column
// block per workgroup.
__device__ __forceinline__ void b_boundary(const Geo &g, double *__restrict__ vrad, double *__restrict__ vtheta,
                                                 double *__restrict__ rho, double *__restrict__ energy, double step,
                                                 const int *__restrict__ guard, int nlo, int hi0) {
  const int ns = g.ns;
  if (guard && *guard == 0) return;
  const int b = (int)blockIdx.x - g.blk0, q = b / g.ncb;
  const int i = q < nlo ? q : hi0 + (q - nlo);
  const int j = (b - q * g.ncb) * TB + (int)threadIdx.x;
  if (i >= g.nr || j >= ns || !g.dampring[i]) return;
  double vrad0 = 0.0, rho0 = 0.0, energy0 = 0.0, vtheta0 = 0.0;
  if (g.damp_init) {
    rho0 = g.SigmaMed[i];
    energy0 = g.EnergyMed[i];
    vtheta0 = g.VthetaMed[i] - g.Rmed[i] * g.omega_frame;
  }
  if (g.damp_flux) {
    vrad0 = -g.gasflux / (2.0 * DPI * g.Rinf[i] * g.SigmaInf[i]);
    rho0 = g.SigmaMed[i];
  }
  const double lambda = g.WaveKiller[i] * step;
  const size_t l = (size_t)i * ns + j;
  vrad[l] = (vrad[l] + lambda * vrad0) / (1.0 + lambda);
  if (g.damp_init) {
    rho[l] = (rho[l] + lambda * rho0) / (1.0 + lambda);
    vtheta[l] = (vtheta[l] + lambda * vtheta0) / (1.0 + lambda);
    energy[l] = (energy[l] + lambda * energy0) / (1.0 + lambda);
  }
  if (g.damp_flux) rho[l] = (rho[l] + lambda * rho0) / (1.0 + lambda);
}
__global__ void __launch_bounds__(TB) k_boundary(Geo g, double *__restrict__ vrad, double *__restrict__ vtheta,
                                                 double *__restrict__ rho, double *__restrict__ energy, double step,
                                                 const int *__restrict__ guard, int nlo, int hi0) {
  b_boundary(g, vrad, vtheta, rho, energy, step, guard, nlo, hi0);
}


// ---------------------------------------------------------------------------
// SubStep1 source terms (SourceEuler.c:319-406) with P = (gamma-1) e on the
// fly (ComputePressureField, SourceEuler.c:740-762) and DetectCrash
// (SourceEuler.c:48-66, 272-273).
__device__ __forceinline__ void b_substep1(const Geo &g, const double *__restrict__ rho,
                                                 const double *__restrict__ energy,
                                                 const double *__restrict__ vrad,
                                                 const double *__restrict__ vtheta,
                                                 const double *__restrict__ agr,
                                                 const double *__restrict__ agt,
                                                 double *__restrict__ vradint, double *__restrict__ vthetaint,
                                                 double dt, DevCtl *ctl, double *__restrict__ accr,
                                                 double *__restrict__ acct, const double *__restrict__ dragr,
                                                 const double *__restrict__ dragt, int fold_ind, double ind_px,
                                                 double ind_py) {
  // accr/acct (pebbles: the gas acceleration, SourceEuler.c:371-405) and
  // dragr/dragt (BackReaction) are null when unused.  fold_ind: agr/agt do
  // not hold the indirect term yet (no k_indirect pass); it is added here to
  // each value read, with the same expression k_indirect adds
  const int ns = g.ns;
  int i, j, L_;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  double axindir = 0.0, ayindir = 0.0;
  if (fold_ind) {
    axindir = -ctl->axstar;
    ayindir = -ctl->aystar;
    if (g.indirect_term) { axindir += ind_px; ayindir += ind_py; }
  }
  auto ind_r = [&](int ii, int jj) {
    const double x = g.Rmed[ii] * g.cosj[jj], y = g.Rmed[ii] * g.sinj[jj];
    return (x * axindir + y * ayindir) * g.InvRmed[ii];
  };
  auto ind_t = [&](int ii, int jj) {
    const double x = g.Rmed[ii] * g.cosj[jj], y = g.Rmed[ii] * g.sinj[jj];
    return (x * ayindir - y * axindir) * g.InvRmed[ii];
  };
  const double gm1 = g.adiabind - 1.0;
  const size_t l = (size_t)i * ns + j;
  const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
  const size_t ljp = (j == ns - 1) ? (size_t)i * ns : l + 1;
  if (energy[l] < 0.0) ctl->crash_e = 1;
  if (rho[l] < 0.0) ctl->crash_rho = 1;
  const double pl = gm1 * energy[l];
  if (i >= 1) {
    const size_t lim = l - ns;
    const double plim = gm1 * energy[lim];
    double gradp = (pl - plim) * 2.0 / (rho[l] + rho[lim]) * g.InvDiffRmed[i];
    double vt2 = vtheta[l] + vtheta[ljp] + vtheta[lim] + vtheta[ljp - ns];
    vt2 = vt2 / 4.0 + g.Rinf[i] * g.omega_frame;
    vt2 = vt2 * vt2;
    double a_l = agr[l], a_lim = agr[lim];
    if (fold_ind) {
      a_l += ind_r(i, j);
      a_lim += ind_r(i - 1, j);
    }
    double tmp = -gradp + 0.5 * (a_l + a_lim) + vt2 * g.InvRinf[i];
    double v = vrad[l] + dt * tmp;
    if (accr) {
      double a = tmp;
      if (dragr) {
        const double t2 = dragr[l] * 2.0 / (rho[l] + rho[lim]);
        v += dt * t2;
        a += t2;
      }
      accr[l] = a;
    }
    vradint[l] = v;
  }
  {
    const double supp_torque = g.supp_torque[i];
    const double invdxtheta = g.InvDxTheta[i];
    const double pljm = gm1 * energy[ljm];
    double gradp = (pl - pljm) * 2.0 / (rho[l] + rho[ljm]) * invdxtheta;
    double t_l = agt[l], t_jm = agt[ljm];
    if (fold_ind) {
      t_l += ind_t(i, j);
      t_jm += ind_t(i, j == 0 ? ns - 1 : j - 1);
    }
    double tmp = -gradp + 0.5 * (t_l + t_jm);
    double v = vtheta[l] + dt * tmp;
    v += dt * supp_torque;
    if (acct) {
      double a = tmp + supp_torque;
      if (dragt) {
        const double t2 = dragt[l] * 2.0 / (rho[l] + rho[ljm]);
        v += dt * t2;
        a += t2;
      }
      acct[l] = a;
    }
    vthetaint[l] = v;
  }
}
__global__ void __launch_bounds__(TB) k_substep1(Geo g, const double *__restrict__ rho,
                                                 const double *__restrict__ energy,
                                                 const double *__restrict__ vrad,
                                                 const double *__restrict__ vtheta,
                                                 const double *__restrict__ agr,
                                                 const double *__restrict__ agt,
                                                 double *__restrict__ vradint, double *__restrict__ vthetaint,
                                                 double dt, DevCtl *ctl, double *__restrict__ accr,
                                                 double *__restrict__ acct, const double *__restrict__ dragr,
                                                 const double *__restrict__ dragt, int fold_ind, double ind_px,
                                                 double ind_py) {
  b_substep1(g, rho, energy, vrad, vtheta, agr, agt, vradint, vthetaint, dt, ctl, accr, acct, dragr, dragt, fold_ind,
             ind_px, ind_py);
}


// UpdateDivVelocAndStressTensor (Viscosity.c:85-162); Drr, Dpp, Drp stay in
// registers.
// divergence and stress tensor of one cell (Viscosity.c:114-160); trp is
// formed only for i >= 1 (has_trp)
struct Stress {
  double dv, trr, tpp, trp;
  bool has_trp;
};

__device__ __forceinline__ Stress stress_cell(const Geo &g, const double *__restrict__ vr,
                                              const double *__restrict__ vt, const double *__restrict__ rho, int i,
                                              int j) {
  const int ns = g.ns;
  const size_t l = (size_t)i * ns + j, lip = l + ns;
  const size_t ljp = (j == ns - 1) ? (size_t)i * ns : l + 1;
  const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
  const double invdphi = g.invdphi, onethird = 1.0 / 3.0;
  const double Drr = (vr[lip] - vr[l]) * g.InvDiffRsup[i];
  const double Dpp = (vt[ljp] - vt[l]) * invdphi * g.InvRmed[i] + 0.5 * (vr[lip] + vr[l]) * g.InvRmed[i];
  double dv = (vr[lip] * g.Rsup[i] - vr[l] * g.Rinf[i]) * g.InvDiffRsup[i] * g.InvRmed[i];
  dv += (vt[ljp] - vt[l]) * invdphi * g.InvRmed[i];
  Stress st;
  st.dv = dv;
  const double visc = fvisc_med(g, i);
  st.trr = 2.0 * rho[l] * visc * (Drr - onethird * dv);
  st.tpp = 2.0 * rho[l] * visc * (Dpp - onethird * dv);
  st.has_trp = i >= 1;
  st.trp = 0.0;
  if (i >= 1) {
    const size_t lim = l - ns, ljmim = ljm - ns;
    const double Drp = 0.5 * (g.Rinf[i] * (vt[l] * g.InvRmed[i] - vt[lim] * g.InvRmed[i - 1]) * g.InvDiffRmed[i] +
                              (vr[l] - vr[ljm]) * invdphi * g.InvRinf[i]);
    const double vinf = fvisc_inf(g, i);
    st.trp = 2.0 * 0.25 * (rho[l] + rho[lim] + rho[ljm] + rho[ljmim]) * vinf * Drp;
  }
  return st;
}

// The first crash's snapshot (SourceEuler.c:275-284): the reference writes
// Rho, Vrad, Vtheta and Temperature as DetectCrash sees them, after the
// boundary conditions and before SubStep1.  SubStep1 (which sets the flags)
// writes only VradInt/VthetaInt, so the next kernel over all cells still
// reads those values and copies them when the flag that decides (the pebble
// test with pebbles, else the energy test) is set.  ctl is null once the run
// has crashed.
struct CrashSnap {
  const DevCtl *ctl;
  int peb;
  const double *src[4];
  double *dst[4];
  // taken by SubStep3 after the fused SubStep1-2 kernel (k_visc_chain): src[1]
  // is vrad before the damping boundary, applied again here where damp is set
  int damp;
  double step;
};

// the damping boundary on vrad alone (k_boundary without DampInit and the
// flux-based damping: the target vrad is 0)
__device__ __forceinline__ double damp_vrad_zero(const Geo &g, int i, double vr, double step) {
  const double vrad0 = 0.0;
  const double lambda = g.WaveKiller[i] * step;
  return (vr + lambda * vrad0) / (1.0 + lambda);
}

__device__ __forceinline__ void b_divstress(const Geo &g, const double *__restrict__ vr,
                                                  const double *__restrict__ vt,
                                                  const double *__restrict__ rho,
                                                  double *__restrict__ divv, double *__restrict__ Trr,
                                                  double *__restrict__ Trp, double *__restrict__ Tpp,
                                                  const CrashSnap &sn) {
  const int ns = g.ns;
  int i, j, L_;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  const size_t l = (size_t)i * ns + j;
  if (sn.ctl && (sn.peb ? sn.ctl->crash_peb : sn.ctl->crash_e)) {
    for (int k = 0; k < 4; k++) sn.dst[k][l] = sn.src[k][l];
  }
  const Stress st = stress_cell(g, vr, vt, rho, i, j);
  if (divv) divv[l] = st.dv;    // SubStep3 forms the divergence again before it is read
  Trr[l] = st.trr;
  Tpp[l] = st.tpp;
  if (st.has_trp) Trp[l] = st.trp;
}
__global__ void __launch_bounds__(TB) k_divstress(Geo g, const double *__restrict__ vr,
                                                  const double *__restrict__ vt,
                                                  const double *__restrict__ rho,
                                                  double *__restrict__ divv, double *__restrict__ Trr,
                                                  double *__restrict__ Trp, double *__restrict__ Tpp,
                                                  CrashSnap sn) {
  b_divstress(g, vr, vt, rho, divv, Trr, Trp, Tpp, sn);
}


// UpdateVelocityWithViscousTerms (Viscosity.c:165-218) + ImposeKeplerianEdges
// (Viscosity.c:221-260, VKepIn/VKepOut from the host)
__device__ __forceinline__ void b_visc_update(const Geo &g, double *__restrict__ vr, double *__restrict__ vt,
                                                    const double *__restrict__ rho,
                                                    const double *__restrict__ Trr,
                                                    const double *__restrict__ Trp,
                                                    const double *__restrict__ Tpp, double DeltaT,
                                                    int kepler, double vkin, double vkout) {
  const int ns = g.ns, nr = g.nr;
  int i, j, L_;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  const size_t l = (size_t)i * ns + j, lip = l + ns;
  const size_t ljp = (j == ns - 1) ? (size_t)i * ns : l + 1;
  const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
  const double invdphi = g.invdphi;
  if (i >= 1 && i < nr - 1)
    vt[l] += DeltaT * g.InvRmed[i] *
             ((g.Rsup[i] * Trp[lip] - g.Rinf[i] * Trp[l]) * g.InvDiffRsup[i] + (Tpp[l] - Tpp[ljm]) * invdphi +
              0.5 * (Trp[l] + Trp[lip])) /
             (0.5 * (rho[l] + rho[ljm]));
  if (i >= 1) {
    const size_t lim = l - ns;
    vr[l] += DeltaT * g.InvRinf[i] *
             ((g.Rmed[i] * Trr[l] - g.Rmed[i - 1] * Trr[lim]) * g.InvDiffRmed[i] + (Trp[ljp] - Trp[l]) * invdphi -
              0.5 * (Tpp[l] + Tpp[lim])) /
             (0.5 * (rho[l] + rho[lim]));
  }
  if (kepler) {
    if (i == 0) vt[l] = vkin - g.Rmed[0] * g.omega_frame;
    if (i == nr - 1) vt[l] = vkout - g.Rmed[nr - 1] * g.omega_frame;
  }
}
__global__ void __launch_bounds__(TB) k_visc_update(Geo g, double *__restrict__ vr, double *__restrict__ vt,
                                                    const double *__restrict__ rho,
                                                    const double *__restrict__ Trr,
                                                    const double *__restrict__ Trp,
                                                    const double *__restrict__ Tpp, double DeltaT,
                                                    int kepler, double vkin, double vkout) {
  b_visc_update(g, vr, vt, rho, Trr, Trp, Tpp, DeltaT, kepler, vkin, vkout);
}


// SubStep2 (SourceEuler.c:417-504): artificial viscosity q computed on the fly
__device__ __forceinline__ double qvnr(double rho, double dv) {
  return (dv < 0.0) ? CVNR * CVNR * rho * dv * dv : 0.0;
}

__device__ __forceinline__ void b_substep2(const Geo &g, const double *__restrict__ rho,
                                                 const double *__restrict__ vrad,
                                                 const double *__restrict__ vtheta,
                                                 const double *__restrict__ energy,
                                                 double *__restrict__ vradnew, double *__restrict__ vthetanew,
                                                 double *__restrict__ energyint, double dt) {
  const int ns = g.ns, nr = g.nr;
  int i, j, L_;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  const size_t l = (size_t)i * ns + j, lip = l + ns;
  const size_t ljp = (j == ns - 1) ? (size_t)i * ns : l + 1;
  const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
  const double dvr = vrad[lip] - vrad[l];
  const double dvt = vtheta[ljp] - vtheta[l];
  const double qr = qvnr(rho[l], dvr);
  const double qt = qvnr(rho[l], dvt);
  if (i >= 1) {
    const size_t lim = l - ns;
    const double qrim = qvnr(rho[lim], vrad[l] - vrad[lim]);
    vradnew[l] = vrad[l] - dt * 2.0 / (rho[l] + rho[lim]) * (qr - qrim) * g.InvDiffRmed[i];
  } else {
    vradnew[l] = 0.0;                       // never written by the reference: stays 0
    vradnew[(size_t)nr * ns + j] = 0.0;     // ghost ring nr likewise
    vthetanew[(size_t)nr * ns + j] = 0.0;
  }
  const double invdxtheta = g.InvDxTheta[i];
  const size_t ljmjp = (size_t)i * ns + j;   // (ljm)+1 == l
  const double qtjm = qvnr(rho[ljm], vtheta[ljmjp] - vtheta[ljm]);
  vthetanew[l] = vtheta[l] - dt * 2.0 / (rho[l] + rho[ljm]) * (qt - qtjm) * invdxtheta;
  energyint[l] = energy[l] - dt * qr * (vrad[lip] - vrad[l]) * g.InvDiffRsup[i] -
                 dt * qt * (vtheta[ljp] - vtheta[l]) * invdxtheta;
}
__global__ void __launch_bounds__(TB) k_substep2(Geo g, const double *__restrict__ rho,
                                                 const double *__restrict__ vrad,
                                                 const double *__restrict__ vtheta,
                                                 const double *__restrict__ energy,
                                                 double *__restrict__ vradnew, double *__restrict__ vthetanew,
                                                 double *__restrict__ energyint, double dt) {
  b_substep2(g, rho, vrad, vtheta, energy, vradnew, vthetanew, energyint, dt);
}


// ---------------------------------------------------------------------------
// SubStep1, ViscousTerms and SubStep2 in one pass, for the gas alone with a
// damping boundary that acts on vrad only (none, or DAMPING without DampInit
// and without the flux-based target): k_boundary, k_substep1, k_divstress,
// k_visc_update, k_substep2 and k_boundary again, fused.  A workgroup owns a
// tile of VC_TI rings x VC_TJ sectors and stages rho, energy, vtheta and the
// accelerations (the indirect term added as SubStep1 adds it) of the tile
// and a 4-cell halo in LDS.  It then forms, each on a region one cell
// smaller than the last: VradInt and VthetaInt (SubStep1, the vrad of the
// damped rings damped first), the stress tensor, the viscous update with
// ImposeKeplerianEdges, and SubStep2 of its own cells, whose vrad is damped
// as the second k_boundary does.  Every value is formed by the expression of
// the separate kernels, so a halo cell holds what its owner computes.  Rings
// 0 and NRAD of VradInt, which SubStep1 does not write, come from the VradInt
// buffer as the separate kernels read them there.  The stress tensor,
// VradInt and VthetaInt stay in LDS.  DetectCrash's flags are set here; the
// crash snapshot is taken by SubStep3 (CrashSnap.damp).
#define VC_TI 16
#define VC_TJ 48
#define VC_T 512                  // threads per workgroup
#define VC_LR (VC_TI + 7)         // rings ib .. ib+VC_LR-1, ib = i0 - 4
#define VC_LW (VC_TJ + 8)         // sectors jb .. jb+VC_LW-1, jb = j0 - 4
#define VC_N (VC_LR * VC_LW)

struct ViscChain {
  const double *rho, *energy, *vrad, *vtheta, *agr, *agt;
  const double *vradint;        // VradInt: rings 0 and NRAD
  double *vradnew, *vthetanew, *energyint;
  DevCtl *ctl;
  double dt;
  int damp;                     // the damping boundary acts on vrad (rings with dampring set)
  int fold_ind;                 // add the indirect term to the accelerations
  double ind_px, ind_py;
  int kepler;                   // ImposeKeplerianEdges
  double vkin, vkout;
};

// ring constants of the tile's rings, held in LDS (index: row of the tile)
enum {
  VR_RINF, VR_RSUP, VR_RMED, VR_IRMED, VR_IRINF, VR_IDRMED, VR_IDRSUP, VR_IDXT, VR_SUPP, VR_VMED, VR_VINF, VR_LAMBDA,
  VR_N
};

// divergence and stress of the cell at LDS index k (ring i, tile row r), as
// stress_cell forms them from the global fields
__device__ __forceinline__ Stress stress_lds(const Geo &g, const double *vr, const double *vt, const double *rho,
                                             const double (*rc)[VC_LR], int i, int r, int k) {
  const int kip = k + VC_LW, kjp = k + 1, kjm = k - 1;
  const double invdphi = g.invdphi, onethird = 1.0 / 3.0;
  const double irmed = rc[VR_IRMED][r], idrsup = rc[VR_IDRSUP][r];
  const double Drr = (vr[kip] - vr[k]) * idrsup;
  const double Dpp = (vt[kjp] - vt[k]) * invdphi * irmed + 0.5 * (vr[kip] + vr[k]) * irmed;
  double dv = (vr[kip] * rc[VR_RSUP][r] - vr[k] * rc[VR_RINF][r]) * idrsup * irmed;
  dv += (vt[kjp] - vt[k]) * invdphi * irmed;
  Stress st;
  st.dv = dv;
  const double visc = rc[VR_VMED][r];
  st.trr = 2.0 * rho[k] * visc * (Drr - onethird * dv);
  st.tpp = 2.0 * rho[k] * visc * (Dpp - onethird * dv);
  st.has_trp = i >= 1;
  st.trp = 0.0;
  if (i >= 1) {
    const int kim = k - VC_LW, kjmim = kjm - VC_LW;
    const double Drp = 0.5 * (rc[VR_RINF][r] * (vt[k] * irmed - vt[kim] * rc[VR_IRMED][r - 1]) * rc[VR_IDRMED][r] +
                              (vr[k] - vr[kjm]) * invdphi * rc[VR_IRINF][r]);
    const double vinf = rc[VR_VINF][r];
    st.trp = 2.0 * 0.25 * (rho[k] + rho[kim] + rho[kjm] + rho[kjmim]) * vinf * Drp;
  }
  return st;
}

__device__ __forceinline__ void b_visc_chain(const Geo &g, const ViscChain &v) {
  // six field arrays over the tile and its halo: rho throughout; the others
  // hold energy, vtheta, agr, agt for SubStep1, then VradInt (in the slot
  // nothing uses before), VthetaInt (energy's), Trp (vtheta's), Trr (agr's)
  // and Tpp (agt's)
  __shared__ double s_rho[VC_N], s_a1[VC_N], s_a2[VC_N], s_a3[VC_N], s_a4[VC_N], s_a5[VC_N];
  __shared__ double s_rc[VR_N][VC_LR];
  __shared__ double s_cos[VC_LW], s_sin[VC_LW];
  const int ns = g.ns, nr = g.nr, tid = threadIdx.x;
  const int ntx = (ns + VC_TJ - 1) / VC_TJ;
  int tx, ty, bid;
  if (!xcd_tile(ntx, (nr + VC_TI - 1) / VC_TI, tx, ty, bid, g.blk0)) return;
  const int i0 = ty * VC_TI, j0 = tx * VC_TJ;
  const int tw = min(VC_TJ, ns - j0);
  const int ib = i0 - 4, jb = j0 - 4;
  const double dt = v.dt, gm1 = g.adiabind - 1.0;
  auto wrapj = [ns](int j) {
    while (j < 0) j += ns;
    while (j >= ns) j -= ns;
    return j;
  };
  // 1. loads: the ring and sector constants first, then every field value of
  // the thread (all in flight together)
  double rcv[VR_N];
  if (tid < VC_LR) {
    const int gi = ib + tid;
#pragma unroll
    for (int q = 0; q < VR_N; q++) rcv[q] = 0.0;
    if (gi >= 0 && gi < nr) {
      rcv[VR_RINF] = g.Rinf[gi];
      rcv[VR_RSUP] = g.Rsup[gi];
      rcv[VR_RMED] = g.Rmed[gi];
      rcv[VR_IRMED] = g.InvRmed[gi];
      rcv[VR_IRINF] = g.InvRinf[gi];
      rcv[VR_IDRMED] = g.InvDiffRmed[gi];
      rcv[VR_IDRSUP] = g.InvDiffRsup[gi];
      rcv[VR_IDXT] = g.InvDxTheta[gi];
      rcv[VR_SUPP] = g.supp_torque[gi];
      rcv[VR_VMED] = fvisc_med(g, gi);
      rcv[VR_VINF] = fvisc_inf(g, gi);
      rcv[VR_LAMBDA] = (v.damp && g.dampring[gi]) ? g.WaveKiller[gi] * dt : -1.0;
    }
  }
  double csj = 0.0, snj = 0.0;
  const int tc = tid - 64;                          // the second wave loads the sector constants
  if (tc >= 0 && tc < VC_LW) {
    const int j = wrapj(jb + tc);
    csj = g.cosj[j];
    snj = g.sinj[j];
  }
  constexpr int N0 = (VC_N + VC_T - 1) / VC_T;
  constexpr int W1 = VC_LW - 2, N1C = W1 * (VC_LR - 1), N1 = (N1C + VC_T - 1) / VC_T;   // SubStep1's region
  constexpr int N4 = (VC_TI * VC_TJ + VC_T - 1) / VC_T;                               // owned cells
  double x0[N0], x1[N0], x2[N0], x3[N0], x4[N0], vr0[N1], e4[N4];
#pragma unroll
  for (int q = 0; q < N0; q++) {
    const int idx = tid + q * VC_T;
    const int r = idx / VC_LW, c = idx - r * VC_LW, gi = ib + r;
    x0[q] = x1[q] = x2[q] = x3[q] = x4[q] = 0.0;
    if (idx < VC_N && gi >= 0 && gi < nr) {
      const size_t l = (size_t)gi * ns + wrapj(jb + c);
      x0[q] = v.rho[l];
      x1[q] = v.energy[l];
      x2[q] = v.vtheta[l];
      x3[q] = v.agr[l];
      x4[q] = v.agt[l];
    }
  }
#pragma unroll
  for (int q = 0; q < N1; q++) {
    const int idx = tid + q * VC_T;
    const int r = 1 + idx / W1, c = 1 + idx % W1, gi = ib + r;
    vr0[q] = 0.0;
    if (idx < N1C && gi >= 0 && gi <= nr) {
      const size_t l = (size_t)gi * ns + wrapj(jb + c);
      vr0[q] = (gi >= 1 && gi < nr) ? v.vrad[l] : v.vradint[l];
    }
  }
#pragma unroll
  for (int q = 0; q < N4; q++) {
    const int idx = tid + q * VC_T;
    const int r = 4 + idx / VC_TJ, c = idx % VC_TJ, gi = ib + r;
    e4[q] = 0.0;
    if (idx < VC_TI * VC_TJ && gi < nr && c < tw) e4[q] = v.energy[(size_t)gi * ns + j0 + c];
  }
  if (tid < VC_LR)
#pragma unroll
    for (int q = 0; q < VR_N; q++) s_rc[q][tid] = rcv[q];
  if (tc >= 0 && tc < VC_LW) {
    s_cos[tc] = csj;
    s_sin[tc] = snj;
  }
  __syncthreads();
  // fields into LDS, the accelerations with the indirect term as SubStep1
  // adds it to each value it reads
  double axindir = 0.0, ayindir = 0.0;
  if (v.fold_ind) {
    axindir = -v.ctl->axstar;
    ayindir = -v.ctl->aystar;
    if (g.indirect_term) { axindir += v.ind_px; ayindir += v.ind_py; }
  }
#pragma unroll
  for (int q = 0; q < N0; q++) {
    const int idx = tid + q * VC_T;
    if (idx >= VC_N) continue;
    const int r = idx / VC_LW, c = idx - r * VC_LW;
    double a = x3[q], t = x4[q];
    if (v.fold_ind) {
      const double rm = s_rc[VR_RMED][r];
      const double x = rm * s_cos[c], y = rm * s_sin[c];
      a += (x * axindir + y * ayindir) * s_rc[VR_IRMED][r];
      t += (x * ayindir - y * axindir) * s_rc[VR_IRMED][r];
    }
    s_rho[idx] = x0[q];
    s_a1[idx] = x1[q];
    s_a2[idx] = x2[q];
    s_a3[idx] = a;
    s_a4[idx] = t;
  }
  __syncthreads();
  // 2. SubStep1 (SourceEuler.c:319-406) on rows 1 .. VC_LR-1, columns 1 ..
  // VC_LW-2, after the damping boundary on vrad; VradInt into s_a5,
  // VthetaInt held until energy's array is free
  {
    const double *s_e = s_a1, *s_vt = s_a2, *s_ar = s_a3, *s_at = s_a4;
    double vti[N1];
#pragma unroll
    for (int q = 0; q < N1; q++) {
      const int idx = tid + q * VC_T;
      const int r = 1 + idx / W1, c = 1 + idx % W1, gi = ib + r;
      vti[q] = 0.0;
      if (idx >= N1C || gi < 0 || gi > nr) continue;
      const int k = r * VC_LW + c;
      if (gi == 0 || gi == nr) {
        s_a5[k] = vr0[q];
      } else {
        const int km = k - VC_LW;
        double vr = vr0[q];
        const double lambda = s_rc[VR_LAMBDA][r];
        if (lambda >= 0.0) {
          const double vrad0 = 0.0;
          vr = (vr + lambda * vrad0) / (1.0 + lambda);
        }
        const double pl = gm1 * s_e[k], plim = gm1 * s_e[km];
        double gradp = (pl - plim) * 2.0 / (s_rho[k] + s_rho[km]) * s_rc[VR_IDRMED][r];
        double vt2 = s_vt[k] + s_vt[k + 1] + s_vt[km] + s_vt[km + 1];
        vt2 = vt2 / 4.0 + s_rc[VR_RINF][r] * g.omega_frame;
        vt2 = vt2 * vt2;
        const double tmp = -gradp + 0.5 * (s_ar[k] + s_ar[km]) + vt2 * s_rc[VR_IRINF][r];
        s_a5[k] = vr + dt * tmp;
      }
      if (gi < nr) {
        const double pl = gm1 * s_e[k], pljm = gm1 * s_e[k - 1];
        const double gradp = (pl - pljm) * 2.0 / (s_rho[k] + s_rho[k - 1]) * s_rc[VR_IDXT][r];
        const double tmp = -gradp + 0.5 * (s_at[k] + s_at[k - 1]);
        double vv = s_vt[k] + dt * tmp;
        vv += dt * s_rc[VR_SUPP][r];
        vti[q] = vv;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N1; q++) {
      const int idx = tid + q * VC_T;
      const int r = 1 + idx / W1, c = 1 + idx % W1, gi = ib + r;
      if (idx < N1C && gi >= 0 && gi < nr) s_a1[r * VC_LW + c] = vti[q];
    }
  }
  __syncthreads();
  double *s_vri = s_a5, *s_vti = s_a1, *s_trp = s_a2, *s_trr = s_a3, *s_tpp = s_a4;
  // 3. stress tensor (Viscosity.c:114-160) on rows 2 .. VC_LR-2, columns 2 .. VC_LW-3
  {
    constexpr int W2 = VC_LW - 4, N2C = W2 * (VC_LR - 3), N2 = (N2C + VC_T - 1) / VC_T;
#pragma unroll
    for (int q = 0; q < N2; q++) {
      const int idx = tid + q * VC_T;
      const int r = 2 + idx / W2, c = 2 + idx % W2, gi = ib + r;
      if (idx >= N2C || gi < 0 || gi >= nr) continue;
      const int k = r * VC_LW + c;
      const Stress st = stress_lds(g, s_vri, s_vti, s_rho, s_rc, gi, r, k);
      s_trr[k] = st.trr;
      s_tpp[k] = st.tpp;
      if (st.has_trp) s_trp[k] = st.trp;
    }
  }
  __syncthreads();
  // 4. UpdateVelocityWithViscousTerms (Viscosity.c:165-218) + ImposeKeplerianEdges
  // on rows 3 .. VC_LR-3, columns 3 .. VC_LW-4, in place
  {
    constexpr int W3 = VC_LW - 6, N3C = W3 * (VC_LR - 5), N3 = (N3C + VC_T - 1) / VC_T;
    const double invdphi = g.invdphi, DeltaT = dt;
#pragma unroll
    for (int q = 0; q < N3; q++) {
      const int idx = tid + q * VC_T;
      const int r = 3 + idx / W3, c = 3 + idx % W3, gi = ib + r;
      if (idx >= N3C || gi < 0 || gi >= nr) continue;
      const int k = r * VC_LW + c, kip = k + VC_LW, kim = k - VC_LW;
      if (gi >= 1 && gi < nr - 1)
        s_vti[k] += DeltaT * s_rc[VR_IRMED][r] *
                    ((s_rc[VR_RSUP][r] * s_trp[kip] - s_rc[VR_RINF][r] * s_trp[k]) * s_rc[VR_IDRSUP][r] +
                     (s_tpp[k] - s_tpp[k - 1]) * invdphi + 0.5 * (s_trp[k] + s_trp[kip])) /
                    (0.5 * (s_rho[k] + s_rho[k - 1]));
      if (gi >= 1)
        s_vri[k] += DeltaT * s_rc[VR_IRINF][r] *
                    ((s_rc[VR_RMED][r] * s_trr[k] - s_rc[VR_RMED][r - 1] * s_trr[kim]) * s_rc[VR_IDRMED][r] +
                     (s_trp[k + 1] - s_trp[k]) * invdphi - 0.5 * (s_tpp[k] + s_tpp[kim])) /
                    (0.5 * (s_rho[k] + s_rho[kim]));
      if (v.kepler) {
        if (gi == 0) s_vti[k] = v.vkin - g.Rmed[0] * g.omega_frame;
        if (gi == nr - 1) s_vti[k] = v.vkout - g.Rmed[nr - 1] * g.omega_frame;
      }
    }
  }
  __syncthreads();
  // 5. SubStep2 (SourceEuler.c:417-504) of the owned cells, DetectCrash's
  // flags, and the damping boundary on the new vrad
  {
    const double *vr = s_vri, *vt = s_vti;
    int ce = 0, cr = 0;
#pragma unroll
    for (int q = 0; q < N4; q++) {
      const int idx = tid + q * VC_T;
      const int r = 4 + idx / VC_TJ, c = 4 + idx % VC_TJ, gi = ib + r;
      if (idx >= VC_TI * VC_TJ || gi >= nr || c - 4 >= tw) continue;
      const int k = r * VC_LW + c, kip = k + VC_LW, kim = k - VC_LW;
      const int j = j0 + c - 4;
      const size_t l = (size_t)gi * ns + j;
      const double el = e4[q];
      if (el < 0.0) ce = 1;
      if (s_rho[k] < 0.0) cr = 1;
      const double dvr = vr[kip] - vr[k];
      const double dvt = vt[k + 1] - vt[k];
      const double qr = qvnr(s_rho[k], dvr);
      const double qt = qvnr(s_rho[k], dvt);
      double vrn;
      if (gi >= 1) {
        const double qrim = qvnr(s_rho[kim], vr[k] - vr[kim]);
        vrn = vr[k] - dt * 2.0 / (s_rho[k] + s_rho[kim]) * (qr - qrim) * s_rc[VR_IDRMED][r];
      } else {
        vrn = 0.0;
        v.vradnew[(size_t)nr * ns + j] = 0.0;
        v.vthetanew[(size_t)nr * ns + j] = 0.0;
      }
      const double invdxtheta = s_rc[VR_IDXT][r];
      const double qtjm = qvnr(s_rho[k - 1], vt[k] - vt[k - 1]);
      v.vthetanew[l] = vt[k] - dt * 2.0 / (s_rho[k] + s_rho[k - 1]) * (qt - qtjm) * invdxtheta;
      v.energyint[l] = el - dt * qr * (vr[kip] - vr[k]) * s_rc[VR_IDRSUP][r] -
                       dt * qt * (vt[k + 1] - vt[k]) * invdxtheta;
      const double lambda = s_rc[VR_LAMBDA][r];
      if (lambda >= 0.0) {
        const double vrad0 = 0.0;
        vrn = (vrn + lambda * vrad0) / (1.0 + lambda);
      }
      v.vradnew[l] = vrn;
    }
    if (ce) v.ctl->crash_e = 1;
    if (cr) v.ctl->crash_rho = 1;
  }
}
__global__ void __launch_bounds__(VC_T) k_visc_chain(Geo g, ViscChain v) { b_visc_chain(g, v); }

__device__ __forceinline__ double flux_limiter(double s) {   // EnergySources.c:684-696
  double tmp;
  if (s <= 2.0) { tmp = 3.0 + sqrt(9.0 + 10.0 * s * s); return 2.0 / tmp; }
  tmp = 10.0 * s + 9.0 + sqrt(81.0 + 180.0 * s);
  return 10.0 / tmp;
}

// SubStep3 up to the matrix, fused (one pass over the grid): the second
// UpdateDivVelocAndStressTensor (SourceEuler.c:302), viscous heating Q+ with
// the ring-0 extrapolation (SourceEuler.c:541-567; ring-0 cells rebuild the
// stress of rings 1 and 2 of their sector), T and cs from EnergyInt
// (EnergySources.c:234-235), MidplaneVolumeDensity + OpacityProfile +
// CalculateQminus, and TemperatureGradient + DiffusionCoefs +
// DiffusionTimestep partials (EnergySources.c:556-653, 833-856) with the
// neighbours' T rebuilt from EnergyInt and rho.  Every value is formed by the
// same expressions as the separate reference steps.
__device__ __forceinline__ double temper_of(const Geo &g, const double *__restrict__ e,
                                            const double *__restrict__ rho, size_t l) {
  const double tmp2 = 1.0 / (rho[l] * 1.0);
  const double tmp1 = e[l] * 1.0 * (g.adiabind - 1.0);
  return tmp1 * tmp2;
}

__device__ __forceinline__ double qplus_of(const Geo &g, int i, double rho, const Stress &st) {
  const double visc = fvisc_med(g, i);
  if (visc != 0.0) {
    double q = 0.5 / visc / rho * (st.trr * st.trr + 2.0 * st.trp * st.trp + st.tpp * st.tpp);
    q += (2.0 / 9.0) * visc * rho * st.dv * st.dv;
    return q;
  }
  return 0.0;
}

template <bool SM>
__device__ __forceinline__ void b_substep3(const Geo &g, const double *__restrict__ vr, const double *__restrict__ vt,
                                                 const double *__restrict__ rho, const double *__restrict__ eint,
                                                 double *__restrict__ divv, double *__restrict__ Trr,
                                                 double *__restrict__ Trp, double *__restrict__ Tpp,
                                                 double *__restrict__ qplus, double *__restrict__ T,
                                                 double *__restrict__ cs, double *__restrict__ opa,
                                                 double *__restrict__ qminus, double *__restrict__ D,
                                                 double *__restrict__ pmin, int fill_opacity, int nsum,
                                                 const CrashSnap &sn) {
  __shared__ double red[TB];
  const int ns = g.ns, nr = g.nr;
  if ((int)blockIdx.x - g.blk0 < nsum) {
    // leading blocks: VMed of ComputeResiduals (TransportEuler.c:86-91) for
    // the coming transport, one ring per wave; vt is final for the transport
    // from here on
    __shared__ double st[TB / 64][RS_CHUNK];
    const int w = threadIdx.x >> 6;
    const int r = ((int)blockIdx.x - g.blk0) * (TB / 64) + w;
    if (r < nr) {
      const double m = wave_ordered_row_sum(vt + (size_t)r * ns, ns, st[w]) / (double)ns;
      if ((threadIdx.x & 63) == 0) g.VMed[r] = m;
    }
    return;
  }
  int i, j, L;
  if (!xcd_cell(g, 0, g.nr, i, j, L, nsum)) return;
  double dtmin = 1.0e38;
  if (j < ns) {
    const size_t l = (size_t)i * ns + j;
    const double rl = rho[l];
    // stress tensor and Q+
    const Stress st = stress_cell(g, vr, vt, rho, i, j);
    divv[l] = st.dv;
    // the tensor itself is not read again before the next substep's
    // ViscousTerms forms it anew: stored only when asked (Trr non-null)
    if (Trr) {
      Trr[l] = st.trr;
      Tpp[l] = st.tpp;
      if (st.has_trp) Trp[l] = st.trp;
    }
    double qp;
    if (i >= 1) {
      qp = qplus_of(g, i, rl, st);
    } else {
      const double v0 = fvisc_med(g, 0);
      const Stress s1 = stress_cell(g, vr, vt, rho, 1, j), s2 = stress_cell(g, vr, vt, rho, 2, j);
      const double q1 = qplus_of(g, 1, rho[l + ns], s1);
      const double q2 = qplus_of(g, 2, rho[l + 2 * ns], s2);
      qp = (v0 != 0.0) ? q1 * exp(log(q1 / q2) * g.lr01 / g.lr12) : 0.0;
    }
    qplus[l] = qp;
    // T, cs
    const double gam = g.adiabind;
    const double tl = temper_of(g, eint, rho, l);
    const double cl = sqrt(gam * (gam - 1.0) * fabs(eint[l]) / rl);
    if (sn.ctl && sn.ctl->crash_e) {
      // the first crash's snapshot after the fused SubStep1-2 kernel: rho,
      // vrad with the damping boundary, vtheta, and T before it is formed here
      double vrs = sn.src[1][l];
      if (sn.damp && g.dampring[i]) vrs = damp_vrad_zero(g, i, vrs, sn.step);
      sn.dst[0][l] = sn.src[0][l];
      sn.dst[1][l] = vrs;
      sn.dst[2][l] = sn.src[2][l];
      sn.dst[3][l] = sn.src[3][l];
    }
    T[l] = tl;
    cs[l] = cl;
    // opacity and Q-
    double o;
    if (g.opacity_law == THORIN_OPA_CONST) {
      o = fill_opacity ? g.parametricopacity * g.opa2cu : opa[l];
    } else {
      const double H = cl * g.OmegaInv[i] * g.sqrt_adiabind_inv;
      const double voldens = rl / H * g.sqrt2pi_inv;
      o = opacity_law<SM>(g, voldens * g.rho2cgs, tl * g.t2si) * g.opa2cu;
    }
    opa[l] = o;
    const double tau = g.opacitydrop * 0.5 * o * rl;
    // EffectiveOpticalDepth (EnergySources.c:701-711)
    const double taueff = 0.375 * tau + (g.stellar_irr ? 0.5 : 0.25 * g.sqrt3) + 0.25 / tau;
    qminus[l] = 2.0 * g.stefan * cube_cr(tl) / taueff;
    // |grad T| at ring ig (clamped to 1..nr-2) from T rebuilt at the four neighbours
    const int ig = (i == 0) ? 1 : ((i == nr - 1) ? nr - 2 : i);
    const size_t lg = (size_t)ig * ns + j;
    const int jp = (j == ns - 1) ? 0 : j + 1;
    const int jm = (j == 0) ? ns - 1 : j - 1;
    const double tg = (ig == i) ? tl : temper_of(g, eint, rho, lg);
    const double t_im = temper_of(g, eint, rho, lg - ns), t_ip = temper_of(g, eint, rho, lg + ns);
    const double t_jm = temper_of(g, eint, rho, (size_t)ig * ns + jm);
    const double t_jp = temper_of(g, eint, rho, (size_t)ig * ns + jp);
    const double invdxtheta = g.InvDxTheta[ig];
    const double gtr_l = (tg - t_im) * g.InvDiffRmed[ig];
    const double gtr_lip = (t_ip - tg) * g.InvDiffRmed[ig + 1];
    const double gtt_l = (tg - t_jm) * invdxtheta;
    const double gtt_ljp = (t_jp - tg) * invdxtheta;
    const double rr = 0.5 * (gtr_lip + gtr_l);
    const double tt = 0.5 * (gtt_ljp + gtt_l);
    const double gradt = sqrt(rr * rr + tt * tt);
    const double H = cl * g.OmegaInv[i] * g.sqrt_adiabind_inv;
    const double voldens = rl / H * g.sqrt2pi_inv;
    const double fac = 1.0 / (voldens * o);
    const double sfl = 4.0 * gradt * fac / tl;
    const double dcoef = flux_limiter(sfl) * 16.0 * g.stefan * cube_cr(tl) * fac;
    D[l] = dcoef;
    double deltar2 = g.Rsup[i] - g.Rinf[i];
    deltar2 = deltar2 * deltar2 / 2.0;
    const double dttmp = deltar2 / dcoef;
    if (dttmp < dtmin) dtmin = dttmp;
  }
  dtmin = block_min<TB>(dtmin, red);
  if (threadIdx.x == 0) pmin[L] = dtmin;
}
template <bool SM>
__global__ void __launch_bounds__(TB) k_substep3(Geo g, const double *__restrict__ vr, const double *__restrict__ vt,
                                                 const double *__restrict__ rho, const double *__restrict__ eint,
                                                 double *__restrict__ divv, double *__restrict__ Trr,
                                                 double *__restrict__ Trp, double *__restrict__ Tpp,
                                                 double *__restrict__ qplus, double *__restrict__ T,
                                                 double *__restrict__ cs, double *__restrict__ opa,
                                                 double *__restrict__ qminus, double *__restrict__ D,
                                                 double *__restrict__ pmin, int fill_opacity, int nsum,
                                                 CrashSnap sn) {
  b_substep3<SM>(g, vr, vt, rho, eint, divv, Trr, Trp, Tpp, qplus, T, cs, opa, qminus, D, pmin, fill_opacity, nsum,
                 sn);
}


// implicit energy matrix (EnergySources.c:256-297) with A and B rebuilt on the fly
__device__ __forceinline__ double coefA(const Geo &g, const double *D, int i, size_t l) {
  return 0.5 * (D[l] + D[l - g.ns]) * g.Rinf[i] * g.InvDiffRmed[i];
}
__device__ __forceinline__ double coefB(const Geo &g, const double *D, const double *cs, int i, int j) {
  const int ns = g.ns;
  const size_t l = (size_t)i * ns + j;
  const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
  const double invdxtheta2 = g.InvDxTheta2[i];
  const double H = cs[l] * g.OmegaInv[i] * g.sqrt_adiabind_inv;
  return 0.5 * (D[l] + D[ljm]) * invdxtheta2 * 2.0 * H;
}

// The matrix kernel also returns the initial SOR residual sum
// (EnergySources.c:435-449): block partials of |T.M - RHS| from the values
// it has just computed.
// Stellar irradiation of cell l of ring i, 1 <= i <= nr-2: CalculateFlaring
// (EnergySources.c:143-165: the sine of the grazing angle from the radial
// derivative of the pressure scale height, 0 for surfaces tilted away; the
// edge rings' values feed no matrix row) and CalculateQirr
// (EnergySources.c:103-135).  pow(Rsup[i], 1.5) is pow_rinf15[i + 1].
__device__ __forceinline__ double qirr_vals(const Geo &g, int i, double csl, double csip, double csim, double opal,
                                            double rhol) {
  const double csin = 0.5 * (csl + csim);
  const double csout = 0.5 * (csl + csip);
  const double Hin = csin * g.pow_rinf15[i] * g.sqrt_adiabind_inv;
  const double Hout = csout * g.pow_rinf15[i + 1] * g.sqrt_adiabind_inv;
  const double dHdr = (Hout - Hin) * g.InvDiffRsup[i];
  const double H = csl * g.OmegaInv[i] * g.sqrt_adiabind_inv;
  double fl = atan(dHdr) - atan((H - 0.4 * g.rstar) * g.InvRmed[i]);
  fl = (fl >= 0.0) ? sin(fl) : 0.0;
  if (!(fl > 0.0)) return 0.0;
  const double r2inv = 1.0 / (g.Rmed[i] * g.Rmed[i]);
  const double tau = g.opacitydrop * 0.5 * opal * rhol;
  const double taueff = 0.375 * tau + 0.5 + 0.25 / tau;
  const double Tirr4 = g.tirr4fac * r2inv * fl;
  return 2.0 * g.stefan * Tirr4 / taueff;
}
__device__ __forceinline__ double qirr_cell(const Geo &g, int i, size_t l, const double *__restrict__ cs,
                                            const double *__restrict__ opa, const double *__restrict__ rho) {
  return qirr_vals(g, i, cs[l], cs[l + g.ns], cs[l - g.ns], opa[l], rho[l]);
}

__device__ __forceinline__ void b_matrix(const Geo &g, const double *__restrict__ rho,
                                               const double *__restrict__ cs, const double *__restrict__ D,
                                               const double *__restrict__ divv,
                                               const double *__restrict__ qminus,
                                               const double *__restrict__ qplus, const double *__restrict__ T,
                                               double *__restrict__ Ml, double *__restrict__ Mlip,
                                               double *__restrict__ Mlim, double *__restrict__ Mljp,
                                               double *__restrict__ Mljm, double *__restrict__ RHS, double dt,
                                               double *__restrict__ presid, const double *__restrict__ opa,
                                               const HeatK &hk) {
  __shared__ double red[TB];
  const int ns = g.ns;
  int i, j, L;                                      // rings 1..nr-2
  if (!xcd_cell(g, 1, g.nr - 2, i, j, L)) return;
  double r0 = 0.0;
  if (j < ns) {
    const size_t l = (size_t)i * ns + j, lip = l + ns, lim = l - ns;
    const int jp = (j == ns - 1) ? 0 : j + 1;
    const size_t ljp = (size_t)i * ns + jp;
    const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
    const double Al = coefA(g, D, i, l), Alip = coefA(g, D, i + 1, lip);
    const double Bl = coefB(g, D, cs, i, j), Bljp = coefB(g, D, cs, i, jp);
    const double H = cs[l] * g.OmegaInv[i] * g.sqrt_adiabind_inv;
    const double bfac = dt / (rho[l] * g.cv);
    const double afac = bfac * g.InvDiffRsup[i] * 2.0 * H / g.Rmed[i];
    const double ml = 1.0 + divv[l] * (g.adiabind - 1.0) * dt + 4.0 * bfac * qminus[l] + afac * (Alip + Al) + bfac * (Bljp + Bl);
    const double mlip = -afac * Alip, mlim = -afac * Al, mljp = -bfac * Bljp, mljm = -bfac * Bl;
    const double tl = T[l];
    double rhs = tl + bfac * (qplus[l] + 3.0 * qminus[l] * tl);
    if (g.stellar_irr) rhs += bfac * qirr_cell(g, i, l, cs, opa, rho);   // EnergySources.c:290
    for (int k = 0; k < hk.n; k++)                                          // EnergySources.c:291-294
      if ((int)l == hk.idx[k]) rhs += bfac * hk.q[k];
    Ml[l] = ml;
    Mlip[l] = mlip;
    Mlim[l] = mlim;
    Mljp[l] = mljp;
    Mljm[l] = mljm;
    RHS[l] = rhs;
    r0 = fabs(tl * ml + T[lip] * mlip + T[lim] * mlim + T[ljp] * mljp + T[ljm] * mljm - rhs);
  }
  r0 = block_sum<TB>(r0, red);
  if (threadIdx.x == 0) presid[L] = r0;
}
__global__ void __launch_bounds__(TB) k_matrix(Geo g, const double *__restrict__ rho,
                                               const double *__restrict__ cs, const double *__restrict__ D,
                                               const double *__restrict__ divv,
                                               const double *__restrict__ qminus,
                                               const double *__restrict__ qplus, const double *__restrict__ T,
                                               double *__restrict__ Ml, double *__restrict__ Mlip,
                                               double *__restrict__ Mlim, double *__restrict__ Mljp,
                                               double *__restrict__ Mljm, double *__restrict__ RHS, double dt,
                                               double *__restrict__ presid, const double *__restrict__ opa,
                                               HeatK hk) {
  b_matrix(g, rho, cs, D, divv, qminus, qplus, T, Ml, Mlip, Mlim, Mljp, Mljm, RHS, dt, presid, opa, hk);
}


// ---------------------------------------------------------------------------
// SOR (EnergySources.c:406-511).  Red-black in place: within one pass every
// updated cell reads only cells of the other colour, so the parallel sweep
// equals the reference's sequential sweep value for value.  Pass 0 updates
// cells with (i+j) odd (jchess1st = 0 at i = 1), pass 1 cells with (i+j) even.
__device__ __forceinline__ double sor_resid(const double *T, const double *Ml, const double *Mlip,
                                            const double *Mlim, const double *Mljp, const double *Mljm,
                                            const double *RHS, size_t l, size_t lip, size_t lim, size_t ljp,
                                            size_t ljm) {
  return T[l] * Ml[l] + T[lip] * Mlip[l] + T[lim] * Mlim[l] + T[ljp] * Mljp[l] + T[ljm] * Mljm[l] - RHS[l];
}




// NIT red-black SOR iterations (EnergySources.c:455-493) per launch.  A
// workgroup owns a tile of SR_TI rings x SR_TJ sectors and stages T with a
// halo of 2*NIT cells in LDS.  Half-sweep s (red for even s, black for odd s)
// is applied to the cells of its colour within 2*NIT-1-s cells of the tile:
// a halo cell is updated with the same inputs and arithmetic as its owner
// (the cells it reads were updated one half-sweep earlier one cell further
// out), so after the last half-sweep the owned cells hold exactly the
// iterates of a sequential sweep.  Iterate n lives in X[n % 4] (X = T0 .. T3;
// the initial T in T0): a launch from iteration it reads X[it % 4] and writes
// X[(it+k+1) % 4] after its k-th iteration (k < NIT <= 3, so no launch writes
// the buffer other workgroups stage their halos from), and a solve that stops
// inside the launch has its iterate (ctl->sor_cur names the buffer).  The
// stop-test partials of each iteration are summed over the owned cells in a
// fixed order (through LDS), the same for every NIT, so the iteration count
// does not depend on how many iterations a launch runs.
//
// A square tile keeps the halo small (a 32x32 tile with a 6-cell halo
// updates 1.35 cells per owned cell over three iterations, an 8x128 one
// 1.70).  Each thread holds the matrix rows of its cells of either colour
// in registers.  A colour's cells are numbered owned cells first, then the
// rings of cells at Chebyshev distance 1, 2, ... from the tile, so that the
// cells a half-sweep updates (distance <= its halo) are a prefix of the
// numbering: the waves past that prefix skip the half-sweep as a whole.
#define SR_TI 32
#define SR_TJ 32
// the FP32 one-sweep tiles (k_mp_sweep) keep their own shape
#define SOR_TI 8
#define SOR_TJ 128
// threads per workgroup of k_sor_iter: 512 for three iterations per launch
// (two cells of each colour per thread), 256 otherwise
#define SOR_NT(nit) ((nit) == 3 ? 2 * TB : TB)
#define SOR_LW (SOR_TJ + 4)
// cells per thread of the one-sweep tiles (FP32 sweeps): red with a one-cell halo, black owned
#define SOR_NRED (((SOR_TI + 2) * ((SOR_TJ + 3) / 2) + TB - 1) / TB)
#define SOR_NBLK ((SOR_TI * ((SOR_TJ + 1) / 2) + TB - 1) / TB)

// one cell of a half-sweep: its matrix row is loaded before the tile's T is
// staged, so that all global loads of the workgroup are in flight together
struct SorCell {
  double ml, mip, mim, mjp, mjm, rhs;
  int gl;   // offset of the cell in the field (written for owned cells)
  // packed: bits 0-11 the LDS index of the cell, 12-14 its distance from
  // the owned tile (0: owned), 16-25 its index among the owned cells (owned
  // only); -1: no cell
  int pk;
  __device__ __forceinline__ int k() const { return pk & 0xFFF; }
  __device__ __forceinline__ int d() const { return (pk >> 12) & 7; }
  __device__ __forceinline__ int o() const { return (pk >> 16) & 0x3FF; }
};
static_assert((SR_TI + 12) * (SR_TJ + 12) <= 0xFFF + 1 && SR_TI * SR_TJ <= 0x3FF + 1,
              "SorCell packs the LDS index in 12 bits and the owned index in 10");
static_assert(SR_TI % 2 == 0 && SR_TJ % 2 == 0, "the slot counts of a colour assume even tile sides");

// The energy system the SOR solves (EnergySources.c:256-297): rows ml..mjm
// and right-hand side rhs.  With form, every launch of a solve forms the
// rows of its staged cells from the system's inputs with k_matrix's
// expressions (rho, cs, D, div v, Q-, Q+, T, the opacity for the stellar
// irradiation, the accretion heating sources; D and cs staged in LDS with
// T): the first one its right-hand side too, from the initial T, which it
// stores for the owned cells (the later launches load it; their T is an
// iterate).  The rows themselves are never stored.
struct SorSys {
  double *ml, *mip, *mim, *mjp, *mjm, *rhs;
  const double *rho, *cs, *D, *divv, *qminus, *qplus, *opa;
  double dt;
  HeatK heat;
  int form;
};

__device__ __forceinline__ void sor_load(SorCell &x, const SorSys &S, size_t l) {
  x.ml = S.ml[l];
  x.mip = S.mip[l];
  x.mim = S.mim[l];
  x.mjp = S.mjp[l];
  x.mjm = S.mjm[l];
  x.rhs = S.rhs[l];
}

// k_matrix's row of cell l (ring i) from D, cs and T staged in LDS (index k,
// rows of LW) and the cell's own inputs: A at the radial interfaces i and
// i+1, B at the azimuthal interfaces j and j+1 (the scale height of the
// cell on the interface's upper side), the same expressions in the same
// order as coefA, coefB, b_matrix and qirr_cell
struct SorIn {                  // a cell's own inputs of its row
  double rho, qminus, divv, qplus, opa;
};
__device__ __forceinline__ SorIn sor_in(const Geo &g, const SorSys &S, size_t l) {
  SorIn v;
  v.rho = S.rho[l];
  v.qminus = S.qminus[l];
  v.divv = S.divv[l];
  v.qplus = S.qplus[l];
  v.opa = g.stellar_irr ? S.opa[l] : 0.0;
  return v;
}
template <int LW, bool RHS>
__device__ __forceinline__ void sor_form_lds(SorCell &x, const Geo &g, const SorSys &S, const SorIn &in,
                                             const double *sd, const double *sc, const double *t, int i, size_t l,
                                             int k) {
  const double Dl = sd[k];
  const double Al = 0.5 * (Dl + sd[k - LW]) * g.Rinf[i] * g.InvDiffRmed[i];
  const double Alip = 0.5 * (sd[k + LW] + Dl) * g.Rinf[i + 1] * g.InvDiffRmed[i + 1];
  const double invdxtheta2 = g.InvDxTheta2[i];
  const double csl = sc[k];
  const double H = csl * g.OmegaInv[i] * g.sqrt_adiabind_inv;
  const double Bl = 0.5 * (Dl + sd[k - 1]) * invdxtheta2 * 2.0 * H;
  const double Hjp = sc[k + 1] * g.OmegaInv[i] * g.sqrt_adiabind_inv;
  const double Bljp = 0.5 * (sd[k + 1] + Dl) * invdxtheta2 * 2.0 * Hjp;
  const double rl = in.rho, qm = in.qminus;
  const double bfac = S.dt / (rl * g.cv);
  const double afac = bfac * g.InvDiffRsup[i] * 2.0 * H / g.Rmed[i];
  x.ml = 1.0 + in.divv * (g.adiabind - 1.0) * S.dt + 4.0 * bfac * qm + afac * (Alip + Al) + bfac * (Bljp + Bl);
  x.mip = -afac * Alip;
  x.mim = -afac * Al;
  x.mjp = -bfac * Bljp;
  x.mjm = -bfac * Bl;
  if (!RHS) return;
  const double tl = t[k];
  double rhs = tl + bfac * (in.qplus + 3.0 * qm * tl);
  if (g.stellar_irr) rhs += bfac * qirr_vals(g, i, csl, sc[k + LW], sc[k - LW], in.opa, rl);   // EnergySources.c:290
  for (int q = 0; q < S.heat.n; q++)                                                             // EnergySources.c:291-294
    if ((int)l == S.heat.idx[q]) rhs += bfac * S.heat.q[q];
  x.rhs = rhs;
}

// one cell of a half-sweep; an owned cell's new value is also its iterate in
// Tout (no later half-sweep of the iteration changes it) and its stop-test
// terms go to sacc / schg
template <int LW>
__device__ __forceinline__ void sor_update(const SorCell &x, double *t, double omega, double *sacc, int *schg,
                                           double *__restrict__ Tout) {
  const int k = x.k();
  const double resid = t[k] * x.ml + t[k + LW] * x.mip + t[k - LW] * x.mim + t[k + 1] * x.mjp +
                       t[k - 1] * x.mjm - x.rhs;
  const double dtemp = omega * resid / x.ml;
  const double tn = t[k] - dtemp;
  t[k] = tn;
  if (x.d() == 0) {
    sacc[x.o()] = fabs(resid);
    schg[x.o()] = (dtemp > tn * SORPRECISION) ? 1 : 0;
    Tout[x.gl] = tn;
  }
}

// cells of one colour of a full tile: SR_TI*SR_TJ/2 owned ones, and in the
// ring at distance d two rows of SR_TJ+2d cells and two columns of
// SR_TI+2d-2 cells, half of each of that colour
__host__ __device__ constexpr int sr_upto(int h) {
  int n = SR_TI * SR_TJ / 2;
  for (int d = 1; d <= h; d++) n += SR_TJ + SR_TI + 4 * d - 2;
  return n;
}

template <int NIT, int NT = TB>
struct SorShape {
  static constexpr int H = 2 * NIT;                        // staged halo
  static constexpr int LW = SR_TJ + 2 * H;                 // LDS row
  static constexpr int LR = SR_TI + 2 * H;                 // LDS rows
  static constexpr int HR = 2 * NIT - 1, HB = 2 * NIT - 2; // halo of the red / black cells
  static constexpr int NRED = (sr_upto(HR) + NT - 1) / NT;
  static constexpr int NBLK = (sr_upto(HB) + NT - 1) / NT;
};

// Cell s of colour c (c = 1: i+j odd, red) of the tile at rings i0.., sectors
// j0.. (P0 = (i0+j0) & 1) with ti x tw owned cells: row rho and column kappa
// relative to the tile's first cell, distance d from the tile.  The slots
// follow a full tile; a narrower or shorter last tile leaves some empty
// (false).  Rows of a ring run over columns -d .. tw+d-1 (top, then bottom),
// its columns over rows -d+1 .. ti+d-2 (left, then right); in each, the
// first cell of colour c is at offset 0 or 1 and the others follow every
// second cell.
template <int H>
__device__ __forceinline__ bool sr_slot(int s, int c, int P0, int ti, int tw, int &rho, int &kappa, int &d) {
  constexpr int OW = SR_TJ / 2;
  if (s < SR_TI * OW) {
    const int r = s / OW, m = s - r * OW;
    rho = r;
    d = 0;
    kappa = ((P0 + r + c) & 1) + 2 * m;
    return r < ti && kappa < tw;
  }
  s -= SR_TI * OW;
#pragma unroll
  for (int dd = 1; dd <= H; dd++) {
    const int nrow = SR_TJ / 2 + dd, ncol = SR_TI / 2 + dd - 1;
    if (s < 2 * nrow + 2 * ncol) {
      d = dd;
      if (s < nrow) {
        rho = -dd;
        kappa = -dd + ((P0 + c) & 1) + 2 * s;
        return kappa < tw + dd;
      }
      s -= nrow;
      if (s < nrow) {
        rho = ti + dd - 1;
        kappa = -dd + ((P0 + ti - 1 + c) & 1) + 2 * s;
        return kappa < tw + dd;
      }
      s -= nrow;
      if (s < ncol) {
        kappa = -dd;
        rho = -dd + 1 + ((P0 + 1 + c) & 1) + 2 * s;
        return rho < ti + dd - 1;
      }
      s -= ncol;
      kappa = tw + dd - 1;
      rho = -dd + 1 + ((P0 + tw + c) & 1) + 2 * s;
      return rho < ti + dd - 1;
    }
    s -= 2 * nrow + 2 * ncol;
  }
  return false;
}

// (mresp non-null, a solve's first launch: the tiles' partials of the
// initial residual are read with the others and added into normres0 the
// same way)
template <int NIT, int NT>
__device__ __forceinline__ void sor_stop_test(const double *__restrict__ psum, const int *__restrict__ pcnt, int nb,
                                              DevCtl *ctl, int maxiter, int it, double eps, double *red,
                                              const double *__restrict__ mresp) {
  // the partial sums of all NIT iterations are read together (eight loads
  // of each array in flight per thread); each thread adds its partials in
  // block order as a strided sum does
  double s[NIT], c[NIT];
  int ci[NIT];
  double r0 = 0.0;
#pragma unroll
  for (int k = 0; k < NIT; k++) { s[k] = 0.0; ci[k] = 0; }
  int b = threadIdx.x;
  for (; b + 7 * NT < nb; b += 8 * NT) {
    double v[NIT][8], v0[8];
    int w[NIT][8];
#pragma unroll
    for (int u = 0; u < 8; u++) v0[u] = mresp ? mresp[b + u * NT] : 0.0;
#pragma unroll
    for (int k = 0; k < NIT; k++)
#pragma unroll
      for (int u = 0; u < 8; u++) {
        v[k][u] = psum[(size_t)k * nb + b + u * NT];
        w[k][u] = pcnt[(size_t)k * nb + b + u * NT];
      }
#pragma unroll
    for (int u = 0; u < 8; u++) r0 += v0[u];
#pragma unroll
    for (int k = 0; k < NIT; k++)
#pragma unroll
      for (int u = 0; u < 8; u++) {
        s[k] += v[k][u];
        ci[k] += w[k][u];
      }
  }
  for (; b < nb; b += NT) {
    if (mresp) r0 += mresp[b];
#pragma unroll
    for (int k = 0; k < NIT; k++) {
      s[k] += psum[(size_t)k * nb + b];
      ci[k] += pcnt[(size_t)k * nb + b];
    }
  }
  // every sum in one pass, each in block_sum's order (red holds NT doubles)
  static_assert((2 * NIT + 1) * (NT / 64) <= NT, "red holds the wave partials of every stop-test sum");
  double v[2 * NIT + 1];
#pragma unroll
  for (int k = 0; k < NIT; k++) {
    v[2 * k] = s[k];
    v[2 * k + 1] = (double)ci[k];
  }
  v[2 * NIT] = r0;
  block_sums<NT, 2 * NIT + 1>(v, red);
#pragma unroll
  for (int k = 0; k < NIT; k++) {
    s[k] = v[2 * k];
    c[k] = v[2 * k + 1];
  }
  if (mresp && threadIdx.x == 0) ctl->sor_normres0 = v[2 * NIT];
  if (threadIdx.x == 0) {
    for (int k = 0; k < NIT; k++) {
      const int n = it + k + 1;
      ctl->sor_iter = n;
      ctl->sor_cur = n % 4;
      ctl->sor_normres = s[k];
      if (s[k] < eps * ctl->sor_normres0 || s[k] < SORMIN || c[k] == 0.0) {
        ctl->sor_done = 1;
        ctl->sor_niter = n;
        break;
      } else if (n >= maxiter) {
        ctl->sor_done = 1;
        ctl->sor_niter = maxiter + 1;
        break;
      }
    }
  }
}

// residual T.M - RHS of a cell from the T in LDS (EnergySources.c:435-449)
template <int LW>
__device__ __forceinline__ double sor_resid_lds(const SorCell &x, const double *t) {
  const int k = x.k();
  return t[k] * x.ml + t[k + LW] * x.mip + t[k - LW] * x.mim + t[k + 1] * x.mjp + t[k - 1] * x.mjm - x.rhs;
}

// FM: 0 the rows are loaded; 1 the first launch of a solve that forms the
// rows and right-hand sides (storing the owned cells' right-hand sides); 2 a
// later launch that forms the rows and loads the right-hand sides.  Each its
// own instance, so that the launches that load keep their registers.
template <int NIT, int NT = TB, int FM = 0>
__device__ __forceinline__ void b_sor_iter(const Geo &g, double *__restrict__ T0, double *__restrict__ T1,
                                           double *__restrict__ T2, double *__restrict__ T3, const SorSys &Sy,
                                           double omega, double *__restrict__ psum, int *__restrict__ pcnt,
                                           DevCtl *__restrict__ ctl, int it, double *__restrict__ mresp, int nbs) {
  static_assert(NIT >= 1 && NIT <= 3, "four iterate buffers hold at most three iterations per launch");
  static_assert(NT % TB == 0, "the partial sums keep the pattern of TB threads");
  typedef SorShape<NIT, NT> S;
  constexpr int STG = S::LR * S::LW;                 // staged places
  constexpr int NOWN = SR_TI * SR_TJ;
  // the stop-test arrays (|residual| and changed flags of the owned cells
  // per iteration) share their LDS with the staged D and cs of a forming
  // launch, which are read before the first of them is written
  constexpr int RAW = (NIT * NOWN * 3 / 2 > 2 * STG) ? NIT * NOWN * 3 / 2 : 2 * STG;
  __shared__ double t[STG];
  __shared__ double sraw[RAW];
  __shared__ double red[NT];
  double(*sacc)[NOWN] = reinterpret_cast<double(*)[NOWN]>(sraw);
  int(*schg)[NOWN] = reinterpret_cast<int(*)[NOWN]>(sraw + NIT * NOWN);
  double *sd = sraw, *sc = sraw + STG;
  static_assert(2 * NIT * (NT / 64) <= NT, "red holds the wave partials of the stop-test sums");
  const int ns = g.ns, nr = g.nr;
  const int ntx = (ns + SR_TJ - 1) / SR_TJ;
  int tx, ty, bid;
  if (!xcd_tile(ntx, (nr - 2 + SR_TI - 1) / SR_TI, tx, ty, bid, g.blk0)) return;
  const int i0 = 1 + ty * SR_TI;
  const int i1 = min(i0 + SR_TI, nr - 1);           // owned rings [i0, i1) within [1, nr-2]
  const int j0 = tx * SR_TJ;
  const int tw = min(SR_TJ, ns - j0);                // owned sectors [j0, j0+tw)
  const int ti = i1 - i0;
  const int tid = threadIdx.x;
  if (it == 0) {
    // first iteration of a solve: reset the control words (k_sor_check forms
    // normres0 from this launch's tile partials), and give T1 .. T3 the rings
    // no sweep writes (0, NRAD-1 and the ghost ring)
    if (bid == 0 && tid == 0) {
      ctl->sor_done = 0;
      ctl->sor_niter = 0;
      ctl->sor_iter = 0;
      ctl->sor_cur = 0;
    }
    if (ty == 0)
      for (int j = j0 + tid; j < j0 + tw; j += NT) {
        T1[j] = T0[j];
        T2[j] = T0[j];
        T3[j] = T0[j];
      }
    if (i1 == nr - 1)
      for (int j = j0 + tid; j < j0 + tw; j += NT)
        for (int r = nr - 1; r <= nr; r++) {
          T1[(size_t)r * ns + j] = T0[(size_t)r * ns + j];
          T2[(size_t)r * ns + j] = T0[(size_t)r * ns + j];
          T3[(size_t)r * ns + j] = T0[(size_t)r * ns + j];
        }
  } else if (ctl->sor_done) {
    return;
  }
  double *X[4] = {T0, T1, T2, T3};
  const double *__restrict__ Tin = X[it % 4];
  constexpr bool form = FM != 0;
  SorCell rc[S::NRED], bc[S::NBLK];
  const int P0 = (i0 + j0) & 1;
  // the cells of the thread's slots; their rows are loaded here (issued
  // before the staging loads), or formed after the staging
  auto setup = [&](SorCell *cs, int ncell, auto hc, int c) {
    constexpr int h = decltype(hc)::value;
#pragma unroll
    for (int q = 0; q < ncell; q++) {
      cs[q].pk = -1;
      int rho, kappa, dd;
      if (!sr_slot<h>(tid + q * NT, c, P0, ti, tw, rho, kappa, dd)) continue;
      const int i = i0 + rho;
      if (i < 1 || i > nr - 2) continue;
      int j = j0 + kappa;
      while (j < 0) j += ns;
      while (j >= ns) j -= ns;
      cs[q].pk = ((rho + S::H) * S::LW + (kappa + S::H)) | (dd << 12) | ((dd == 0 ? rho * tw + kappa : 0) << 16);
      cs[q].gl = i * ns + j;
      if (FM == 0) sor_load(cs[q], Sy, (size_t)i * ns + j);
      if (FM == 2) cs[q].rhs = Sy.rhs[(size_t)i * ns + j];
    }
  };
  setup(rc, S::NRED, std::integral_constant<int, S::HR>(), 1);
  setup(bc, S::NBLK, std::integral_constant<int, S::HB>(), 0);
  // stage rings i0-H .. i1+H-1 and sectors j0-H .. j0+tw+H-1 (periodic in
  // phi) over LDS rows of the full width: every load of the thread is issued
  // before the first LDS store (one memory latency, not one per load).
  // Places outside the tile's rings or the grid are never read; they get 0.
  constexpr int NSTG = (STG + NT - 1) / NT;
  const int sw = tw + 2 * S::H;
  // (forming: T, D and cs in one batch, then the inputs of every slot cell
  // before the first row is formed)
  constexpr int NSRC = form ? 3 : 1;
  {
    const double *src[3] = {Tin, Sy.D, Sy.cs};
    double *dst[3] = {t, sd, sc};
    double sv[NSRC][NSTG];
#pragma unroll
    for (int q = 0; q < NSTG; q++) {
      const int idx = tid + q * NT;
      const int r = idx / S::LW, c = idx - r * S::LW;
      const int i = i0 - S::H + r;
#pragma unroll
      for (int a = 0; a < NSRC; a++) sv[a][q] = 0.0;
      if (r < ti + 2 * S::H && c < sw && i >= 0 && i <= nr - 1) {
        int j = j0 - S::H + c;
        while (j < 0) j += ns;
        while (j >= ns) j -= ns;
#pragma unroll
        for (int a = 0; a < NSRC; a++) sv[a][q] = src[a][(size_t)i * ns + j];
      }
    }
#pragma unroll
    for (int q = 0; q < NSTG; q++) {
      const int idx = tid + q * NT;
      if (idx < STG)
#pragma unroll
        for (int a = 0; a < NSRC; a++) dst[a][idx] = sv[a][q];
    }
  }
  if constexpr (form) {
    SorIn ri[S::NRED], bi[S::NBLK];
#pragma unroll
    for (int q = 0; q < S::NRED; q++)
      if (rc[q].pk >= 0) ri[q] = sor_in(g, Sy, (size_t)rc[q].gl);
#pragma unroll
    for (int q = 0; q < S::NBLK; q++)
      if (bc[q].pk >= 0) bi[q] = sor_in(g, Sy, (size_t)bc[q].gl);
    __syncthreads();
    // the rows of the staged cells (and, first launch, their right-hand
    // sides: the owned cells' are stored)
    auto form_rows = [&](SorCell *cs, const SorIn *in, int ncell) {
#pragma unroll
      for (int q = 0; q < ncell; q++) {
        if (cs[q].pk < 0) continue;
        const size_t l = (size_t)cs[q].gl;
        const int k = cs[q].k();
        const int i = i0 + k / S::LW - S::H;        // the cell's ring from its LDS row
        sor_form_lds<S::LW, FM == 1>(cs[q], g, Sy, in[q], sd, sc, t, i, l, k);
        if (FM == 1 && cs[q].d() == 0) Sy.rhs[l] = cs[q].rhs;
      }
    };
    form_rows(rc, ri, S::NRED);
    form_rows(bc, bi, S::NBLK);
  }
  __syncthreads();              // (forming: every read of sd, sc before sacc is written)
  const int nown = ti * tw;
  if (it == 0) {
    // the solve's initial residual sum|T.M - RHS| (EnergySources.c:435-451)
    // over the owned cells from the staged T, added as the stop-test sums
    // are; k_sor_check adds the tiles' partials in tile order
#pragma unroll
    for (int q = 0; q < S::NRED; q++)
      if (rc[q].pk >= 0 && rc[q].d() == 0) sacc[0][rc[q].o()] = fabs(sor_resid_lds<S::LW>(rc[q], t));
#pragma unroll
    for (int q = 0; q < S::NBLK; q++)
      if (bc[q].pk >= 0 && bc[q].d() == 0) sacc[0][bc[q].o()] = fabs(sor_resid_lds<S::LW>(bc[q], t));
    __syncthreads();
    double r[1] = {0.0};
    for (int idx = tid; tid < TB && idx < nown; idx += TB) r[0] += sacc[0][idx];
    block_sums<NT, 1>(r, red);      // (its closing barrier: every read of sacc done before the sweeps write it)
    if (tid == 0) mresp[bid] = r[0];
  }
#pragma unroll
  for (int k = 0; k < NIT; k++) {
    const int hr = 2 * NIT - 1 - 2 * k, hb = hr - 1;
    // iterate it+k+1 of the owned cells goes to Tout as they are updated
    double *__restrict__ Tout = X[(it + k + 1) % 4];
#pragma unroll
    for (int q = 0; q < S::NRED; q++)
      if (rc[q].pk >= 0 && rc[q].d() <= hr) sor_update<S::LW>(rc[q], t, omega, sacc[k], schg[k], Tout);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < S::NBLK; q++)
      if (bc[q].pk >= 0 && bc[q].d() <= hb) sor_update<S::LW>(bc[q], t, omega, sacc[k], schg[k], Tout);
    __syncthreads();
  }
  // stop-test partials of the NIT iterations: each of the first TB threads
  // adds the owned cells tid, tid+TB, ... in order, then the block sums
  {
    double v[2 * NIT];
#pragma unroll
    for (int k = 0; k < NIT; k++) {
      double acc = 0.0;
      int chg = 0;
      for (int idx = tid; tid < TB && idx < nown; idx += TB) {
        acc += sacc[k][idx];
        chg += schg[k][idx];
      }
      v[2 * k] = acc;
      v[2 * k + 1] = (double)chg;
    }
    block_sums<NT, 2 * NIT>(v, red);
    if (tid == 0)
#pragma unroll
      for (int k = 0; k < NIT; k++) {
        psum[(size_t)k * nbs + bid] = v[2 * k];
        pcnt[(size_t)k * nbs + bid] = (int)v[2 * k + 1];
      }
  }
}

template <int NIT, int NT = TB, int FM = 0>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 2 * TB ? 4 : 1)))
k_sor_iter(Geo g, double *__restrict__ T0, double *__restrict__ T1, double *__restrict__ T2, double *__restrict__ T3,
           SorSys sy, double omega, double *__restrict__ psum, int *__restrict__ pcnt, DevCtl *__restrict__ ctl, int it,
           double *__restrict__ mresp, int nbs) {
  b_sor_iter<NIT, NT, FM>(g, T0, T1, T2, T3, sy, omega, psum, pcnt, ctl, it, mresp, nbs);
}

// stop test of iterations it+1 .. it+NIT (EnergySources.c:487-493), partial
// sums added in block order; the solve ends at the first iteration that
// meets it
template <int NIT>
__device__ __forceinline__ void b_sor_check(const double *__restrict__ psum, const int *__restrict__ pcnt, int nb,
                                            DevCtl *ctl, int maxiter, int it, double eps,
                                            const double *__restrict__ mresp) {
  __shared__ double red[RB];
  // at iteration 0 the launch has just cleared sor_done, and the initial
  // residual normres0 (EnergySources.c:435-451) is formed from its tile
  // partials in tile order
  if (it != 0 && ctl->sor_done) return;
  sor_stop_test<NIT, RB>(psum, pcnt, nb, ctl, maxiter, it, eps, red, it == 0 ? mresp : nullptr);
}

template <int NIT>
__global__ void __launch_bounds__(RB) k_sor_check(const double *__restrict__ psum, const int *__restrict__ pcnt,
                                                  int nb, DevCtl *ctl, int maxiter, int it, double eps,
                                                  const double *__restrict__ mresp) {
  b_sor_check<NIT>(psum, pcnt, nb, ctl, maxiter, it, eps, mresp);
}



// ---------------------------------------------------------------------------
// Radial transport: ComputeLRMomenta (TransportEuler.c:318-347) + OneWindRad
// (49-65) with ComputeStarRad (232-273), VanLeerRadial (384-415) and
// DivisePolarGrid (SideEuler.c:82-99) for the six conserved quantities.
// A workgroup handles a tile of RT_J sectors x RI rings (8, 16 or 32) for
// all quantities: one wave per quantity.  First the whole workgroup stages
// rho (rings i0-2 .. i0+RI+1) and vr (i0-3 .. i0+RI+2) of the tile in LDS and
// forms from them the density's slopes drho (rings i0-1 .. i0+RI) and star
// values RhoStar (interfaces i0 .. i0+RI), which every quantity's flux
// uses; then each wave walks its columns outward: at ring i it forms
// W(i+2) = Q/(RhoInt+1e-20), the slope dq(i+1), the star value qs at
// interface i+1 and the flux F(i+1) = dt*dtheta*Radii[i+1]*qs*RhoStar*vr,
// and writes Q(i) + (F(i) - F(i+1))*InvSurf[i].  The reference evaluates that
// product for the lower interface of ring i with Rinf[i] and for the upper
// one with Rsup[i] = Radii[i+1]: both are F.  Momenta are rebuilt from rho,
// vr, vtheta exactly as ComputeLRMomenta forms them.  Rings outside the grid
// read as 0 (as the reference's dq = 0 and qs = 0 there make them); a cell's
// result does not depend on RI.  Tiles are numbered so that XCD x (= block %
// 8) gets a contiguous band of them.
#define RC_NQ 6
#define RT_J 64                  // sectors per tile: one wave
#define RT_T (RT_J * RC_NQ)      // threads per workgroup: one wave per quantity

struct RadQ {
  const double *rho, *vr, *vt, *e;
  // the energy from the SOR's last iterate (EnergyNew = rho*T*CV with the
  // edge rings' T taken from rings 1 and NRAD-2, EnergySources.c:319-334):
  // T buffer ctl->sor_cur of tsrc, instead of e (tsrc[0] null: read e)
  const double *tsrc[4];
  const DevCtl *ctl;
  double cv;
  double *out[RC_NQ];          // RadMomP, RadMomM, ThetaMomP, ThetaMomM, Energy, Rho
  const int *guard;             // skip the launch while *guard == 0
  double of;
  int ntile_j, ntiles;          // tiles of RT_J sectors x rows rings
  int nq;                       // 6; 5 for the pebble fluid (no energy: quantity 4 skipped)
  int rows;                     // rings per tile (8, 16 or 32)
};

// ring geometry of one radial tile and the shared density values, in LDS.
// Geometry entry r is ring i0-4+r (0 outside 0..NRAD); rho entry a is ring
// i0-2+a, vr entry a ring i0-3+a, drho entry a ring i0-1+a, rs entry a
// interface i0+a (values outside the grid are 0).
template <int RI>
struct RadGeo {
  static constexpr int G = RI + 6;
  double rmed[G], idr[G], rinf[G], isurf[G];
  double rho[RI + 4][RT_J], vr[RI + 6][RT_J], drho[RI + 2][RT_J], rs[RI + 1][RT_J];
};

// slope of ComputeStarRad at ring k (dq = 0 on the first and last ring)
__device__ __forceinline__ double rad_dq_s(int k, int nr, double idr_k, double idr_k1, double wm, double w0,
                                           double wp) {
  if (k <= 0 || k >= nr - 1) return 0.0;
  const double dqm = (w0 - wm) * idr_k;
  const double dqp = (wp - w0) * idr_k1;
  return (dqp * dqm > 0.0) ? 2.0 * dqp * dqm / (dqp + dqm) : 0.0;
}

// star value of ComputeStarRad at interface k (0 on ring 0 and the ghost ring)
__device__ __forceinline__ double rad_star_s(int k, int nr, double rm_km, double rm_k, double rm_kp, double wm,
                                             double w0, double dqm, double dq0, double vrk, double dt) {
  if (k < 1 || k > nr - 1) return 0.0;
  if (vrk > 0.0) return wm + (rm_k - rm_km - vrk * dt) * 0.5 * dqm;
  return w0 - (rm_kp - rm_k + vrk * dt) * 0.5 * dq0;
}

// one column (tile column c, sector j) of quantity W (0 RadMomP, 1 RadMomM,
// 2 ThetaMomP, 3 ThetaMomM, 4 Energy, 5 Rho) over rings i0 .. i1-1
// (i1 - i0 <= RI).  The walk has the fixed length RI + 4 and is unrolled:
// the loads of later rings go out ahead of the arithmetic and the rolling
// values stay in named registers; rings past i1 (a partial last tile) are
// walked on clamped data and store nothing.
template <int W, int RI>
__device__ __forceinline__ void rad_column(const Geo &g, const RadQ &q, const RadGeo<RI> &sg, double dt, int c,
                                           int j, int i0, int i1) {
  const int ns = g.ns, nr = g.nr;
  const int jp = (j == ns - 1) ? 0 : j + 1;
  const int gb = i0 - 4;                        // ring of geometry entry 0
  const double dtheta = 2.0 * DPI / (double)ns;
  const double *__restrict__ pvt = q.vt;
  const double *__restrict__ pe = q.e;
  const double *__restrict__ pt = nullptr;
  if (W == 4 && q.tsrc[0]) {
    const int cur = q.ctl->sor_cur;   // (a select: no indexing of the argument block)
    pt = cur == 0 ? q.tsrc[0] : (cur == 1 ? q.tsrc[1] : (cur == 2 ? q.tsrc[2] : q.tsrc[3]));
  }
  double *__restrict__ out = q.out[W];
  // rolling state for output ring i: W at i, i+1; Q at i, i+1; dq at i;
  // flux at i
  double w0 = 0.0, w1 = 0.0, Q0 = 0.0, Q1 = 0.0, dq0 = 0.0, F0 = 0.0;
#pragma unroll
  for (int t = 0; t < RI + 4; t++) {
    const int i = gb + t;
    const int k = i + 2;                       // ring entering the window
    const int kc = min(max(k, 0), nr - 1);
    const double rho2 = sg.rho[t][c];          // 0 outside the grid
    double x_l = 0.0;
    if (W == 2) x_l = pvt[kc * ns + jp];
    if (W == 3) x_l = pvt[kc * ns + j];
    if (W == 4) {
      if (pt) {
        const int kt = kc == 0 ? 1 : (kc == nr - 1 ? nr - 2 : kc);
        x_l = pt[kt * ns + j];
      } else {
        x_l = pe[kc * ns + j];
      }
    }
    const int r = t + 1;                       // geometry entry of ring i+1
    double Q2 = 0.0;
    if (k >= 0 && k < nr) {
      if (W == 0) Q2 = rho2 * sg.vr[t + 2][c];
      if (W == 1) Q2 = rho2 * sg.vr[t + 1][c];
      if (W == 2 || W == 3) Q2 = rho2 * (x_l + sg.rmed[r + 1] * q.of) * sg.rmed[r + 1];
      if (W == 4) Q2 = pt ? rho2 * x_l * q.cv : x_l;
      if (W == 5) Q2 = rho2;
    }
    const double w2 = Q2 / (rho2 + 1e-20);
    // dq at ring i+1 is used from t = 2 on, the flux at interface i+1 from t = 3
    double dq1 = 0.0, F1 = 0.0;
    if (t >= 2) dq1 = rad_dq_s(i + 1, nr, sg.idr[r], sg.idr[r + 1], w0, w1, w2);
    if (t >= 3) {
      const double vr1 = sg.vr[t][c];
      const double qs = rad_star_s(i + 1, nr, sg.rmed[r - 1], sg.rmed[r], sg.rmed[r + 1], w0, w1, dq0, dq1, vr1, dt);
      F1 = (i + 1 >= 1 && i + 1 <= nr - 1) ? dt * dtheta * sg.rinf[r] * qs * sg.rs[t - 3][c] * vr1 : 0.0;
    }
    if (t >= 4 && i < i1) {
      double varq = F0;
      varq -= F1;
      out[i * ns + j] = Q0 + varq * sg.isurf[r - 1];
    }
    w0 = w1; w1 = w2;
    Q0 = Q1; Q1 = Q2;
    dq0 = dq1; F0 = F1;
  }
}

template <int RI>
__device__ __forceinline__ void b_transport_rad(const Geo &g, const RadQ &q, double dt) {
  __shared__ RadGeo<RI> sg;
  if (q.guard && *q.guard == 0) return;
  const int b = (int)blockIdx.x - g.blk0, slot = b & 7;
  const int per = (q.ntiles + 7) >> 3;
  const int tile = slot * per + (b >> 3);
  if (tile >= q.ntiles) return;                // whole workgroup
  const int tj = tile % q.ntile_j, ti = tile / q.ntile_j;
  const int i0 = ti * RI, i1 = min(i0 + RI, g.nr);
  const int nr = g.nr, ns = g.ns, j0 = tj * RT_J;
  for (int t = threadIdx.x; t < RadGeo<RI>::G; t += RT_T) {
    const int k = i0 - 4 + t;
    const bool ok = (k >= 0 && k <= nr);
    sg.rmed[t] = ok ? g.Rmed[k] : 0.0;
    sg.idr[t] = ok ? g.InvDiffRmed[k] : 0.0;
    sg.rinf[t] = ok ? g.Rinf[k] : 0.0;
    sg.isurf[t] = ok ? g.InvSurf[k] : 0.0;
  }
  // rho of rings 0..NRAD-1 and vr of rings 0..NRAD (the ghost ring); 0
  // elsewhere and on sectors past NSEC.  Every thread issues all its loads
  // (at clamped, in-bounds addresses) before the first LDS write.
  {
    constexpr int NE = ((RI + 6) * RT_J + RT_T - 1) / RT_T;
    double rv[NE], vv[NE];
#pragma unroll
    for (int u = 0; u < NE; u++) {
      const int e = threadIdx.x + u * RT_T;
      const int a = e / RT_J, j = j0 + e % RT_J;
      const int jc = min(j, ns - 1);
      const int kr = min(max(i0 - 2 + a, 0), nr - 1), kv = min(max(i0 - 3 + a, 0), nr);
      rv[u] = q.rho[kr * ns + jc];
      vv[u] = q.vr[kv * ns + jc];
    }
#pragma unroll
    for (int u = 0; u < NE; u++) {
      const int e = threadIdx.x + u * RT_T;
      const int a = e / RT_J, c = e % RT_J, j = j0 + c;
      const int kr = i0 - 2 + a, kv = i0 - 3 + a;
      if (a < RI + 4) sg.rho[a][c] = (j < ns && kr >= 0 && kr < nr) ? rv[u] : 0.0;
      if (a < RI + 6) sg.vr[a][c] = (j < ns && kv >= 0 && kv <= nr) ? vv[u] : 0.0;
    }
  }
  __syncthreads();
  // the density's slopes (ComputeStarRad on rho) ...
  for (int e = threadIdx.x; e < (RI + 2) * RT_J; e += RT_T) {
    const int a = e / RT_J, c = e % RT_J, m = i0 - 1 + a, r = a + 3;   // geometry entry of ring m
    sg.drho[a][c] = rad_dq_s(m, nr, sg.idr[r], sg.idr[r + 1], sg.rho[a][c], sg.rho[a + 1][c], sg.rho[a + 2][c]);
  }
  __syncthreads();
  // ... and its star values RhoStar at interfaces i0 .. i0+RI
  for (int e = threadIdx.x; e < (RI + 1) * RT_J; e += RT_T) {
    const int a = e / RT_J, c = e % RT_J, m = i0 + a, r = a + 4;
    sg.rs[a][c] = rad_star_s(m, nr, sg.rmed[r - 1], sg.rmed[r], sg.rmed[r + 1], sg.rho[a + 1][c], sg.rho[a + 2][c],
                             sg.drho[a][c], sg.drho[a + 1][c], sg.vr[a + 3][c], dt);
  }
  __syncthreads();
  const int c = threadIdx.x % RT_J, j = j0 + c;
  const int which = threadIdx.x / RT_J;
  if (q.nq == RC_NQ - 1 && which == 4) return;   // the pebble fluid has no energy
  if (j >= ns) return;
  switch (which) {
    case 0: rad_column<0, RI>(g, q, sg, dt, c, j, i0, i1); break;
    case 1: rad_column<1, RI>(g, q, sg, dt, c, j, i0, i1); break;
    case 2: rad_column<2, RI>(g, q, sg, dt, c, j, i0, i1); break;
    case 3: rad_column<3, RI>(g, q, sg, dt, c, j, i0, i1); break;
    case 4: rad_column<4, RI>(g, q, sg, dt, c, j, i0, i1); break;
    default: rad_column<5, RI>(g, q, sg, dt, c, j, i0, i1); break;
  }
}
template <int RI>
__global__ void __launch_bounds__(RT_T) k_transport_rad(Geo g, RadQ q, double dt) {
  b_transport_rad<RI>(g, q, dt);
}


// Ordered ring means out[r] = (sum over j in order of f[r][j]) / NSEC
// (mpi_make1Dprofile order, mpiTasks.c:26-32) for `rows` rings per
// workgroup: the workgroup stages its rings whole in LDS (one coalesced pass,
// all loads in flight together), then lane r of the first wave walks ring r.
// The chain of NSEC dependent additions per ring is the cost; a workgroup's
// rings share it.
__device__ __forceinline__ void b_ring_means(const double *__restrict__ f, int nr, int ns, int rows,
                                             double *__restrict__ out, const int *__restrict__ guard, BOff bo) {
  extern __shared__ double sh[];       // rows * ns
  if (guard && *guard == 0) return;
  const int r0 = ((int)blockIdx.x - bo.blk0) * rows;
  const int nrow = min(rows, nr - r0);
  const int n = nrow * ns;
  const double *__restrict__ src = f + (size_t)r0 * ns;
  for (int k = threadIdx.x; k < n; k += 256) sh[k] = src[k];
  __syncthreads();
  if (threadIdx.x < nrow) {
    const double m = ordered_ring_sum(sh + (size_t)threadIdx.x * ns, ns) / (double)ns;
    out[r0 + threadIdx.x] = m;
  }
}
__global__ void __launch_bounds__(256) k_ring_means(const double *__restrict__ f, int nr, int ns, int rows,
                                                     double *__restrict__ out, const int *__restrict__ guard,
                                                     BOff bo) {
  b_ring_means(f, nr, ns, rows, out, guard, bo);
}


// Azimuthal transport, ring-local: ComputeResiduals (TransportEuler.c:70-111),
// QuantitiesAdvection with the residual and then the uniform velocity
// (159-175; ComputeStarTheta 275-316, VanLeerTheta 417-445) and AdvectSHIFT
// (113-136).  One workgroup per ring; a thread keeps its NPT cells
// j = tid + k*TT_T of every quantity, of rho before and after the first pass,
// of both RhoStar arrays and of the residual velocity in registers; two LDS
// arrays exchange work values and fluxes with the neighbouring cells.  NPT is
// the smallest count with NSEC <= TT_T*NPT, so that no register holds a cell
// outside the ring.
#define TT_T 512
#define TT_NPT 6                // largest instance: NSEC <= 3072

// dq of ComputeStarTheta at x from the work values in w
__device__ __forceinline__ double th_dq(const double *w, int x, int ns, double invdxtheta) {
  const int xm = (x == 0) ? ns - 1 : x - 1, xp = (x == ns - 1) ? 0 : x + 1;
  const double w0 = w[x];
  const double dqm = (w0 - w[xm]);
  const double dqp = (w[xp] - w0);
  return ((dqp * dqm > 0.0) && (dqp + dqm != 0.0)) ? dqp * dqm / (dqp + dqm) * invdxtheta : 0.0;
}

// star value at j (ComputeStarTheta's second loop); only the upwind cell's
// slope is formed
__device__ __forceinline__ double th_star(const double *w, int j, int ns, double ksi, double dxtheta, double invdxtheta) {
  const int jm = (j == 0) ? ns - 1 : j - 1;
  const bool up = ksi > 0.0;
  const int x = up ? jm : j;
  const double d = th_dq(w, x, ns, invdxtheta);
  const double wx = w[x];
  return up ? wx + (dxtheta - ksi) * d : wx - (dxtheta + ksi) * d;
}

struct ThGeom {
  int ns;
  double dt, dxtheta, invdxtheta, dxrad, invsurf;
};

// one VanLeerTheta (q held in registers): W = q/(rint+1e-20), star values,
// fluxes dxrad*qs*RhoStar*v, update q += (F[j] - F[j+1]) * invsurf
template <int NPT>
__device__ __forceinline__ void th_pass(const ThGeom &t, double *qv, const double *rint, const double *rs,
                                        const double *vel, double vconst, bool uniform, double *s_w, double *s_f) {
  const int tid = threadIdx.x, ns = t.ns;
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int j = tid + k * TT_T;
    if (j < ns) s_w[j] = qv[k] / (rint[k] + 1e-20);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int j = tid + k * TT_T;
    if (j < ns) {
      const double v = uniform ? vconst : vel[k];
      const double qs = th_star(s_w, j, ns, v * t.dt, t.dxtheta, t.invdxtheta);
      s_f[j] = t.dxrad * qs * rs[k] * v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int j = tid + k * TT_T;
    if (j < ns) {
      const int jp = (j == ns - 1) ? 0 : j + 1;
      double varq = s_f[j];
      varq -= s_f[jp];
      qv[k] = qv[k] + varq * t.invsurf;
    }
  }
}

template <int NPT>
__device__ __forceinline__ void b_transport_theta(const Geo &g, const double *__restrict__ vt,
                                                          const double *__restrict__ xrp, const double *__restrict__ xrm,
                                                          const double *__restrict__ xtp, const double *__restrict__ xtm,
                                                          const double *__restrict__ xe, const double *__restrict__ xrho,
                                                          double *__restrict__ orp, double *__restrict__ orm,
                                                          double *__restrict__ otp, double *__restrict__ otm,
                                                          double *__restrict__ oe, double *__restrict__ orho,
                                                          double dt, const int *__restrict__ guard, int nq) {
  // nq quantities after rho: 5 (momenta and energy), 4 for the pebble fluid
  extern __shared__ double sh[];       // 2*ns
  if (guard && *guard == 0) return;
  const int i = (int)blockIdx.x - g.blk0, ns = g.ns, tid = threadIdx.x;
  double *s_w = sh, *s_f = sh + ns;
  const size_t base = (size_t)i * ns;
  __shared__ double vmed, vtemp;
  __shared__ int nshift;
  double vtr[NPT], r0[NPT], r1[NPT], rs1[NPT], rs2[NPT], qv[NPT];
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int j = tid + k * TT_T;
    vtr[k] = 0.0;
    r0[k] = 0.0;
    if (j < ns) {
      vtr[k] = vt[base + j];
      r0[k] = xrho[base + j];
      s_w[j] = vtr[k];
      s_f[j] = r0[k];
    }
  }
  __syncthreads();
  if (tid == 0) {
    const double vm = g.VMed[i];         // ordered ring mean of vtheta (k_ring_means)
    const double Ntilde = vm * g.InvRmed[i] * dt * (double)ns / 2.0 / DPI;
    const double Nround = floor(Ntilde + 0.5);
    vmed = vm;
    nshift = (int)(long)Nround;
    vtemp = g.fast_transport ? (Ntilde - Nround) * g.Rmed[i] * (1.0 / dt) * (2.0 * DPI / (double)ns) : 0.0;
    g.Nshift[i] = nshift;
  }
  __syncthreads();
  const bool fast = g.fast_transport != 0;
  ThGeom t;
  t.ns = ns;
  t.dt = dt;
  t.dxtheta = 2.0 * DPI / (double)ns * g.Rmed[i];
  t.invdxtheta = 1.0 / t.dxtheta;
  t.dxrad = (g.Rsup[i] - g.Rinf[i]) * dt;
  t.invsurf = 1.0 / g.Surf[i];
  const double vt2 = vtemp;
  // residual velocity and RhoStar of the first pass (ComputeStarTheta(Rho): base rho itself, in s_f)
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int j = tid + k * TT_T;
    if (fast) vtr[k] = vtr[k] - vmed;
    rs1[k] = (j < ns) ? th_star(s_f, j, ns, vtr[k] * dt, t.dxtheta, t.invdxtheta) : 0.0;
    qv[k] = r0[k];
  }
  th_pass<NPT>(t, qv, r0, rs1, vtr, 0.0, false, s_w, s_f);             // VanLeerTheta(Rho), residual velocity
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int j = tid + k * TT_T;
    r1[k] = qv[k];
    if (j < ns) s_w[j] = qv[k];
  }
  __syncthreads();
  if (fast) {
#pragma unroll
    for (int k = 0; k < NPT; k++) {
      const int j = tid + k * TT_T;
      rs2[k] = (j < ns) ? th_star(s_w, j, ns, vt2 * dt, t.dxtheta, t.invdxtheta) : 0.0;
    }
    __syncthreads();
    th_pass<NPT>(t, qv, r1, rs2, nullptr, vt2, true, s_w, s_f);       // VanLeerTheta(Rho), uniform velocity
  }
  const int sh_n = fast ? nshift : 0;
  // AdvectSHIFT: out[j] = q[j - Nshift], written as out[j + Nshift] = q[j]
  // (a rotation of the ring: each cell goes to one place, the stores of a
  // wave stay contiguous but for the wrap)
  int sh0 = sh_n % ns;
  if (sh0 < 0) sh0 += ns;
  auto store_shifted = [&](double *out) {
#pragma unroll
    for (int k = 0; k < NPT; k++) {
      const int j = tid + k * TT_T;
      if (j < ns) {
        int jo = j + sh0;
        if (jo >= ns) jo -= ns;
        out[base + jo] = qv[k];
      }
    }
  };
  store_shifted(orho);
  const double *ins[5] = {xrp, xrm, xtp, xtm, xe};
  double *outs[5] = {orp, orm, otp, otm, oe};
  double nxt[NPT];             // the next quantity, loaded one quantity ahead
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int j = tid + k * TT_T;
    nxt[k] = (j < ns) ? xrp[base + j] : 0.0;
  }
  for (int which = 0; which < nq; which++) {
#pragma unroll
    for (int k = 0; k < NPT; k++) qv[k] = nxt[k];
    if (which < nq - 1) {
      const double *src = ins[which + 1];
#pragma unroll
      for (int k = 0; k < NPT; k++) {
        const int j = tid + k * TT_T;
        nxt[k] = (j < ns) ? src[base + j] : 0.0;
      }
    }
    // (the passes' LDS arrays are each written only after the barrier that
    // follows every read of them in the pass before)
    th_pass<NPT>(t, qv, r0, rs1, vtr, 0.0, false, s_w, s_f);
    if (fast) th_pass<NPT>(t, qv, r1, rs2, nullptr, vt2, true, s_w, s_f);
    store_shifted(outs[which]);
  }
}
template <int NPT>
__global__ void __launch_bounds__(TT_T) k_transport_theta(Geo g, const double *__restrict__ vt,
                                                          const double *__restrict__ xrp, const double *__restrict__ xrm,
                                                          const double *__restrict__ xtp, const double *__restrict__ xtm,
                                                          const double *__restrict__ xe, const double *__restrict__ xrho,
                                                          double *__restrict__ orp, double *__restrict__ orm,
                                                          double *__restrict__ otp, double *__restrict__ otm,
                                                          double *__restrict__ oe, double *__restrict__ orho,
                                                          double dt, const int *__restrict__ guard, int nq) {
  b_transport_theta<NPT>(g, vt, xrp, xrm, xtp, xtm, xe, xrho, orp, orm, otp, otm, oe, orho, dt, guard, nq);
}


// ComputeVelocities (TransportEuler.c:350-381)
__device__ __forceinline__ void b_velocities(const Geo &g, const double *__restrict__ rho,
                                                   const double *__restrict__ rp, const double *__restrict__ rm,
                                                   const double *__restrict__ tp, const double *__restrict__ tm,
                                                   double *__restrict__ vr, double *__restrict__ vt,
                                                   const int *__restrict__ guard) {
  const int ns = g.ns;
  int i, j, L_;
  if (guard && *guard == 0) return;
  if (!xcd_cell(g, 0, g.nr, i, j, L_) || j >= ns) return;
  const size_t l = (size_t)i * ns + j;
  const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
  if (i == 0) vr[l] = 0.0;
  else {
    const size_t lim = l - ns;
    vr[l] = (rp[lim] + rm[l]) / (rho[l] + rho[lim] + 1e-20);
  }
  vt[l] = (tp[ljm] + tm[l]) / (rho[l] + rho[ljm] + 1e-15) / g.Rmed[i] - g.Rmed[i] * g.omega_frame;
}
__global__ void __launch_bounds__(TB) k_velocities(Geo g, const double *__restrict__ rho,
                                                   const double *__restrict__ rp, const double *__restrict__ rm,
                                                   const double *__restrict__ tp, const double *__restrict__ tm,
                                                   double *__restrict__ vr, double *__restrict__ vt,
                                                   const int *__restrict__ guard) {
  b_velocities(g, rho, rp, rm, tp, tm, vr, vt, guard);
}


// GasTotalMass partial sums: per ring, sequential in j (SideEuler.c:35-39)
__global__ void __launch_bounds__(RB) k_ring_mass(Geo g, const double *__restrict__ rho, double *__restrict__ out) {
  extern __shared__ double sh[];
  const int i = blockIdx.x, ns = g.ns;
  for (int j = threadIdx.x; j < ns; j += RB) sh[j] = g.Surf[i] * rho[(size_t)i * ns + j];
  __syncthreads();
  if (threadIdx.x == 0) out[i] = ordered_ring_sum(sh, ns);
}

__global__ void k_pressure(Geo g, const double *__restrict__ e, double *__restrict__ p, size_t n) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = (g.adiabind - 1.0) * e[k];
}

// ===========================================================================
// host side

static inline void swapb(thorin_dev *d, int a, int b) {
  double *t = d->b[a];
  d->b[a] = d->b[b];
  d->b[b] = t;
}

static int need_dev(thorin_ctx *c) {
  if (!c || !c->dev) {
    thorin_set_error("this context has no device (created with thorin_ctx_create_host or failed)");
    return -1;
  }
  return 0;
}

// dynamic LDS of k_eos_cfl: the vtheta and cs rings, and T when the Flock
// alpha needs its ring mean
static inline size_t eos_lds(const thorin_dev *d, int do_eos) {
  return sizeof(double) * (size_t)(d->ns + EC_PAD) * ((d->g.alpha_flock && do_eos) ? 3 : 2);
}

// cells per thread of the end-of-substep kernels: the smallest instance that
// holds NSEC (an ensemble runs its disks with one instance)
static inline int eos_ec(const thorin_dev *d) {
  int ec = d->ns <= 3 * RB ? 3 : (d->ns <= 6 * RB ? 6 : 12);
  if (ec < d->ec_min) ec = d->ec_min;
  return ec;
}
// SubStep3 with the SM03 law compiled in only where the disk uses it
#define SUB3_LAUNCH(grid, block, shmem, ...)                                              \
  do {                                                                                    \
    if (d->g.opacity_law == THORIN_OPA_SM03) LAUNCH(k_substep3<true>, grid, block, shmem, __VA_ARGS__); \
    else LAUNCH(k_substep3<false>, grid, block, shmem, __VA_ARGS__);                      \
  } while (0)
#define EOS_LAUNCH(KERN, grid, block, shmem, ...)                          \
  do {                                                                    \
    const int ec_ = eos_ec(d);                                            \
    if (ec_ == 3) LAUNCH(KERN<3>, grid, block, shmem, __VA_ARGS__);       \
    else if (ec_ == 6) LAUNCH(KERN<6>, grid, block, shmem, __VA_ARGS__);  \
    else LAUNCH(KERN<12>, grid, block, shmem, __VA_ARGS__);               \
  } while (0)

// 1D grid of the XCD-banded cell blocks over `rings` rings (xcd_cell), and
// the number of logical blocks (= block partials a reduction kernel writes)
static inline int cell_count(const thorin_dev *d, int rings) { return d->g.ncb * rings; }
static inline dim3 cell_grid(const thorin_dev *d, int rings) {
  return dim3((unsigned)((cell_count(d, rings) + 7) / 8 * 8));
}

// k_boundary: the two bands of damped rings (at least one block)
static inline dim3 bnd_grid(const thorin_dev *d) {
  const int rings = d->damp_nlo + (d->nr - d->damp_hi0);
  const int nb = rings * ((d->ns + TB - 1) / TB);
  return dim3((unsigned)(nb > 0 ? nb : 1));
}

extern "C" int thorin_dev_create(thorin_ctx *c) {
  thorin_dev *d = (thorin_dev *)calloc(1, sizeof(thorin_dev));
  if (!d) { thorin_set_error("out of host memory"); return -1; }
  c->dev = d;
  d->nr = c->nr;
  d->ns = c->ns;
  d->nf = c->nfield;
  if (c->ns % 2 != 0) {
    thorin_set_error("NSEC must be even: the red-black SOR sweep of the energy equation alternates colours "
                     "around each ring");
    return -1;
  }
  if (c->ns > TT_T * TT_NPT) {
    thorin_set_error("NSEC too large for the azimuthal transport (max 3072)");
    return -1;
  }
  if ((c->ns + TB - 1) / TB * TB / 2 < c->ns / 2) return -1;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    thorin_set_error("no HIP device available: the FARGO_THORIN GPU path needs an MI355X (gfx950)");
    return -1;
  }
  {
    int dev = 0;
    hipDeviceProp_t prop;
    HCHECK(hipGetDevice(&dev));
    HCHECK(hipGetDeviceProperties(&prop, dev));
    const size_t need = (size_t)2 * c->ns * sizeof(double);
    if (need > (size_t)prop.sharedMemPerBlock) {
      char m[200];
      snprintf(m, sizeof m, "NSEC %d needs %zu B of LDS per workgroup for the azimuthal transport; device allows %zu",
               c->ns, need, (size_t)prop.sharedMemPerBlock);
      thorin_set_error(m);
      return -1;
    }
    const void *tt[TT_NPT] = {(const void *)k_transport_theta<1>, (const void *)k_transport_theta<2>,
                              (const void *)k_transport_theta<3>, (const void *)k_transport_theta<4>,
                              (const void *)k_transport_theta<5>, (const void *)k_transport_theta<6>};
    for (int k = 0; k < TT_NPT; k++)
      HCHECK(hipFuncSetAttribute(tt[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    // k_ring_means: one ring per workgroup (the most workgroups staging and
    // walking at once), two for the narrow rings of small grids
    d->rs_rows = c->ns < 512 ? 2 : 1;
    // launch shapes (the results do not depend on them): what the autotuner
    // picks at 1024x1536 and for the 16-disk ensemble
    d->sor_nit = 3;
    d->rad_rows = 8;
    d->visc_chain = 1;
    d->sor_assembly = 1;
    const size_t rs_lds = sizeof(double) * (size_t)d->rs_rows * c->ns;
    if (rs_lds > (size_t)prop.sharedMemPerBlock) {
      thorin_set_error("NSEC too large for the ring-mean kernel's LDS");
      return -1;
    }
    HCHECK(hipFuncSetAttribute((const void *)k_ring_means, hipFuncAttributeMaxDynamicSharedMemorySize, (int)rs_lds));
    HCHECK(hipFuncSetAttribute((const void *)k_peb_cfl, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    // k_eos_cfl: up to three rows of NSEC + EC_PAD
    const size_t eos_need = sizeof(double) * 3 * (size_t)(c->ns + EC_PAD);
    if (eos_need > (size_t)prop.sharedMemPerBlock) {
      thorin_set_error("NSEC too large for the end-of-substep kernel's LDS");
      return -1;
    }
    const void *ek[6] = {(const void *)k_eos_cfl<3>,     (const void *)k_eos_cfl<6>,
                         (const void *)k_eos_cfl<12>,    (const void *)k_eos_cfl_vel<3>,
                         (const void *)k_eos_cfl_vel<6>, (const void *)k_eos_cfl_vel<12>};
    for (int k = 0; k < 6; k++)
      HCHECK(hipFuncSetAttribute(ek[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)eos_need));
  }
  HCHECK(hipStreamCreateWithFlags(&d->s, hipStreamNonBlocking));
  d->prof_pend = new std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>>();
  d->prof_pool = new std::vector<hipEvent_t>();
  const size_t fb = (d->nf + 64) * sizeof(double);
  for (int k = 0; k < NBUF; k++) {
    HCHECK(hipMalloc(&d->b[k], fb));
    HCHECK(hipMemsetAsync(d->b[k], 0, fb, d->s));
  }
  // 1D arrays: 43 slots of (nr+3) doubles
  const int n1 = c->nr + 3;
  HCHECK(hipMalloc(&d->d1base, sizeof(double) * (size_t)n1 * 43));
  HCHECK(hipMemsetAsync(d->d1base, 0, sizeof(double) * (size_t)n1 * 43, d->s));
  double *p1 = d->d1base;
  const double *src[21] = {c->Rinf, c->Rsup, c->Rmed, c->Surf, c->InvRmed, c->InvSurf, c->InvDiffRmed,
                           c->InvDiffRsup, c->InvRinf, c->OmegaInv, c->SigmaMed, c->SigmaInf, c->EnergyMed,
                           c->VthetaMed, c->WaveKiller, c->pow_rmed15, c->pow_rinf15, c->cav_med, c->cav_inf,
                           c->cfl_pow2, c->supp_torque};
  for (int k = 0; k < 21; k++) HCHECK(hipMemcpyAsync(p1 + (size_t)k * n1, src[k], sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  Geo &g = d->g;
  memset(&g, 0, sizeof(g));
  g.nr = c->nr;
  g.ns = c->ns;
  g.ncb = (c->ns + TB - 1) / TB;
  g.Rinf = p1; g.Rsup = p1 + n1; g.Rmed = p1 + 2 * n1; g.Surf = p1 + 3 * n1; g.InvRmed = p1 + 4 * n1;
  g.InvSurf = p1 + 5 * n1; g.InvDiffRmed = p1 + 6 * n1; g.InvDiffRsup = p1 + 7 * n1; g.InvRinf = p1 + 8 * n1;
  g.OmegaInv = p1 + 9 * n1; g.SigmaMed = p1 + 10 * n1; g.SigmaInf = p1 + 11 * n1; g.EnergyMed = p1 + 12 * n1;
  g.VthetaMed = p1 + 13 * n1; g.WaveKiller = p1 + 14 * n1; g.pow_rmed15 = p1 + 15 * n1;
  g.pow_rinf15 = p1 + 16 * n1; g.cav_med = p1 + 17 * n1; g.cav_inf = p1 + 18 * n1; g.cfl_pow2 = p1 + 19 * n1;
  g.supp_torque = p1 + 20 * n1;
  g.gcs = p1 + 21 * n1; g.gtemp = p1 + 22 * n1; g.Vmoy = p1 + 23 * n1; g.VMed = p1 + 24 * n1;
  g.ringmax = p1 + 25 * n1;
  d->ringtmp = p1 + 26 * n1;
  g.sig_avr = p1 + 27 * n1;
  g.ist_avr = p1 + 28 * n1;
  g.peb_dens0 = p1 + 29 * n1;
  g.peb_vr0 = p1 + 30 * n1;
  g.peb_vt0 = p1 + 31 * n1;
  g.tau_avr = p1 + 32 * n1;
  g.phavr = p1 + 33 * n1;
  g.pist_avr = p1 + 34 * n1;
  g.pebmax = p1 + 35 * n1;
  g.pvmed = p1 + 36 * n1;
  {
    std::vector<double> az((size_t)3 * n1, 0.0);
    const int ns = c->ns;
    for (int i = 0; i < n1; i++) {
      const double dxtheta = 2.0 * DPI / (double)ns * c->Rmed[i];
      az[i] = 1.0 / dxtheta;
      az[(size_t)n1 + i] = 1.0 / dxtheta / dxtheta;
      az[(size_t)2 * n1 + i] = c->Rmed[i] * 2.0 * DPI / (double)ns;
    }
    HCHECK(hipMemcpyAsync(p1 + (size_t)37 * n1, az.data(), sizeof(double) * 3 * n1, hipMemcpyHostToDevice, d->s));
    HCHECK(hipStreamSynchronize(d->s));        // az is freed at the end of this block
    g.InvDxTheta = p1 + 37 * n1;
    g.InvDxTheta2 = p1 + 38 * n1;
    g.DxThetaA = p1 + 39 * n1;
    g.invdphi = 1.0 / (2.0 * M_PI / (double)ns);
  }
  int *ip;
  HCHECK(hipMalloc(&ip, sizeof(int) * (size_t)n1 * 3));
  g.Nshift = ip;
  g.dampring = ip + n1;
  g.pdampring = ip + 2 * n1;
  {
    // rings of DampingBoundary (Damping YES) and of DampPebbles (always)
    std::vector<int> dr(2 * n1, 0);
    const thorin_params *p = &c->p;
    double DRin = p->rmin * p->dampingrminfrac, DRout = p->rmax * p->dampingrmaxfrac;
    for (int i = 0; i < c->nr; i++) {
      const int z = (c->Rmed[i] < DRin || c->Rmed[i] > DRout) ? 1 : 0;
      dr[i] = p->damping_on ? z : 0;
      dr[n1 + i] = z;
    }
    // the bands k_boundary is launched over: rings below the first undamped
    // ring and from the last undamped ring on (all rings when none is undamped)
    int lo = 0, hi = c->nr;
    while (lo < c->nr && dr[lo]) lo++;
    if (lo < c->nr) {
      while (hi > lo && dr[hi - 1]) hi--;
      for (int i = lo; i < hi; i++)
        if (dr[i]) { lo = 0; hi = 0; break; }   // damped rings inside: cover the whole grid
    }
    d->damp_nlo = lo;
    d->damp_hi0 = hi;
    HCHECK(hipMemcpyAsync(ip + n1, dr.data(), sizeof(int) * 2 * n1, hipMemcpyHostToDevice, d->s));
    HCHECK(hipStreamSynchronize(d->s));
  }
  const thorin_params *p = &c->p;
  g.adiabind = p->adiabind;
  g.sqrt_adiabind_inv = c->k.sqrt_adiabind_inv;
  g.cv = c->k.cv;
  g.viscosity = p->viscosity;
  g.alphavisc = p->alphaviscosity;
  g.alphamri = p->alphamri;
  g.alphadead = p->alphadead;
  g.tmri = p->tmri;
  g.twidth = p->twidth;
  g.t2si = c->k.t2si;
  g.opa2cu = c->k.opa2cu;
  g.rho2cgs = c->k.rho2cgs;
  g.stefan = c->k.stefan;
  g.sqrt2pi_inv = c->k.sqrt2pi_inv;
  g.sqrt3 = c->k.sqrt3;
  g.opacitydrop = p->opacitydrop;
  g.parametricopacity = p->parametricopacity;
  g.powerlawopacity = p->powerlawopacity;
  g.gasflux = p->gasflux;
  g.omega_frame = c->st.omega_frame;
  g.lr01 = c->k.lr01;
  g.lr12 = c->k.lr12;
  g.thicksmooth = p->thicknesssmoothing;
  g.visc_alpha = p->viscosity_alpha;
  g.alpha_flock = p->alpha_flock_on;
  g.fast_transport = p->fast_transport;
  g.damp_init = p->damp_init;
  g.damp_flux = p->damp_flux;
  g.opacity_law = p->opacity_law;
  g.stellar_irr = p->stellar_irradiation;
  {
    // CalculateFlaring's R_star and CalculateQirr's factor (EnergySources.c:111-117, 152)
    g.rstar = p->stellarradius * 6.957e8 / TH_AU_SI;
    double Teff4 = p->effectivetemperature / TH_T2SI;
    Teff4 = pow(Teff4, 4.0);
    double Rstar2 = 6.957e8 / TH_AU_SI;
    Rstar2 = pow(p->stellarradius * Rstar2, 2.0);
    g.tirr4fac = (1.0 - p->discalbedo) * Teff4 * Rstar2;
  }
  g.indirect_term = p->indirect_term;
  g.exclude_hill = p->exclude_hill;
  g.hillcut = p->hillcut;
  g.pebbles = p->pebbles;
  g.back_reaction = p->back_reaction;
  g.pebble_gravity = p->pebble_gravity;
  g.palpha = p->pebblealpha;
  g.pbulk = p->pebblebulkdens / TH_RHO2CGS;
  g.schmidt = p->schmidtnumber;
  // LP85 thresholds exactly as the reference's expressions (Opacity.c:205-220)
  g.lp_t1 = pow(2.0e-4 / 2.0e16, -1.0 / 9.0);
  g.lp_t2 = pow(2.0e16 / 5.0e-3, 1.0 / 8.0);
  g.lp_t3a = pow(5.0e-3 / 2.0e34, -1.0 / 10.0);
  g.lp_t4 = pow(2.0e34 / 2.0e-8, 1.0 / 12.0);
  g.lp_t5a = pow(2.0e-8 / 1.0e-36, 1.0 / 7.0);
  g.lp_t6a = pow(1.0e-36 / 1.5e20, -1.0 / 12.5);
  HCHECK(hipMalloc(&d->ctl, sizeof(DevCtl)));
  HCHECK(hipMemsetAsync(d->ctl, 0, sizeof(DevCtl), d->s));
  HCHECK(hipHostMalloc(&d->hctl, sizeof(DevCtl), hipHostMallocDefault));
  HCHECK(hipHostMalloc((void **)&d->hflag, 64, hipHostMallocDefault));
  *(volatile unsigned *)d->hflag = 0;
  d->flag_seq = 0;
  memset(d->hctl, 0, sizeof(DevCtl));
  d->nblk_cells = ((c->ns + TB - 1) / TB) * c->nr;
  HCHECK(hipMalloc(&d->red, sizeof(double) * (size_t)d->nblk_cells * (2 + 2 * 8)));
  const int half = c->ns / 2;
  d->nblk_sor = ((half + TB - 1) / TB) * (c->nr - 2);
  // the SOR's stop-test partials: one per tile and iteration of a launch (up to 3)
  {
    const size_t nbs = (size_t)((c->ns + SR_TJ - 1) / SR_TJ) * (size_t)((c->nr - 2 + SR_TI - 1) / SR_TI);
    const size_t np = 3 * (nbs > (size_t)d->nblk_cells ? nbs : (size_t)d->nblk_cells);
    HCHECK(hipMalloc(&d->sorp, sizeof(double) * np));
    HCHECK(hipMalloc(&d->sorc, sizeof(int) * np));
  }
  HCHECK(hipMalloc(&d->pstellar, sizeof(double) * (size_t)d->nblk_cells));
  {
    const int nbs = ((c->ns + SR_TJ - 1) / SR_TJ) * ((c->nr - 2 + SR_TI - 1) / SR_TI);
    HCHECK(hipMalloc(&d->mresp, sizeof(double) * (size_t)std::max(d->nblk_cells, nbs)));
  }
  HCHECK(hipHostMalloc(&d->h_part, sizeof(double) * (size_t)d->nblk_cells, hipHostMallocDefault));
  HCHECK(hipEventCreate(&d->ev_t0));
  HCHECK(hipEventCreate(&d->ev_t1));
  HCHECK(hipGetDevice(&d->device));
  // cos / sin of the sector angles
  HCHECK(hipMalloc(&d->trig, sizeof(double) * 2 * (size_t)d->ns));
  HCHECK(hipMemcpyAsync(d->trig, c->h_trig, sizeof(double) * 2 * (size_t)d->ns, hipMemcpyHostToDevice, d->s));
  d->g.cosj = d->trig;
  d->g.sinj = d->trig + d->ns;
  if (c->p.opacity_law == THORIN_OPA_SM03) {
    const int rows = c->sm_rows, cols = c->sm_cols;
    const size_t n = (size_t)rows + cols + (size_t)rows * cols;
    HCHECK(hipMalloc(&d->sm, sizeof(double) * n));
    HCHECK(hipMemcpyAsync(d->sm, c->sm_pgrid, sizeof(double) * rows, hipMemcpyHostToDevice, d->s));
    HCHECK(hipMemcpyAsync(d->sm + rows, c->sm_tgrid, sizeof(double) * cols, hipMemcpyHostToDevice, d->s));
    HCHECK(hipMemcpyAsync(d->sm + rows + cols, c->sm_ross, sizeof(double) * rows * cols, hipMemcpyHostToDevice,
                          d->s));
    d->g.sm_pg = d->sm;
    d->g.sm_tg = d->sm + rows;
    d->g.sm_ross = d->sm + rows + cols;
    d->g.sm_rows = rows;
    d->g.sm_cols = cols;
    d->g.sm_fac = ((double)rows) / log10(c->sm_pgrid[rows - 1] / c->sm_pgrid[0]);
  }
  HCHECK(hipStreamSynchronize(d->s));
  return 0;
}

extern "C" void thorin_dev_destroy(thorin_ctx *c) {
  thorin_dev *d = c->dev;
  if (!d) return;
  if (d->s) (void)hipStreamSynchronize(d->s);
  for (int k = 0; k < NBUF; k++) if (d->b[k]) (void)hipFree(d->b[k]);
  if (d->d1base) (void)hipFree(d->d1base);
  if (d->g.Nshift) (void)hipFree(d->g.Nshift);
  if (d->ctl) (void)hipFree(d->ctl);
  if (d->hctl) (void)hipHostFree(d->hctl);
  if (d->hflag) (void)hipHostFree(d->hflag);
  if (d->red) (void)hipFree(d->red);
  if (d->sorp) (void)hipFree(d->sorp);
  if (d->trig) (void)hipFree(d->trig);
  if (d->sm) (void)hipFree(d->sm);
  if (d->sorc) (void)hipFree(d->sorc);
  if (d->pstellar) (void)hipFree(d->pstellar);
  if (d->mresp) (void)hipFree(d->mresp);
  if (d->acc_stage) (void)hipHostFree(d->acc_stage);
  {
    float *f[8] = {d->m32.ml, d->m32.mip, d->m32.mim, d->m32.mjp, d->m32.mjm, d->m32.r, d->m32.e0, d->m32.e1};
    for (int k = 0; k < 8; k++)
      if (f[k]) (void)hipFree(f[k]);
  }
  if (d->prof_pend) {
    for (auto &r : *d->prof_pend) { (void)hipEventDestroy(r.second.first); (void)hipEventDestroy(r.second.second); }
    delete d->prof_pend;
  }
  if (d->prof_pool) {
    for (auto e : *d->prof_pool) (void)hipEventDestroy(e);
    delete d->prof_pool;
  }
  if (d->ev_t0) (void)hipEventDestroy(d->ev_t0);
  if (d->ev_t1) (void)hipEventDestroy(d->ev_t1);
  if (d->h_part) (void)hipHostFree(d->h_part);
  if (d->s) (void)hipStreamDestroy(d->s);
  free(d);
  c->dev = NULL;
}

static int upload_1d_profiles(thorin_ctx *c) {
  thorin_dev *d = c->dev;
  const int n1 = c->nr + 3;
  double *p1 = d->d1base;
  HCHECK(hipMemcpyAsync(p1 + 10 * n1, c->SigmaMed, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(p1 + 11 * n1, c->SigmaInf, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(p1 + 12 * n1, c->EnergyMed, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(p1 + 13 * n1, c->VthetaMed, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->g.gcs, c->globcsvec, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->g.gtemp, c->globtempervec, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  return 0;
}

extern "C" int thorin_dev_upload_init(thorin_ctx *c) {
  thorin_dev *d = c->dev;
  if (upload_1d_profiles(c)) return -1;
  const size_t nb = sizeof(double) * d->nf;
  HCHECK(hipMemcpyAsync(d->b[B_DENS], c->h_dens, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_VRAD], c->h_vrad, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_VTHETA], c->h_vtheta, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_ENERGY], c->h_energy, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_LABEL], c->h_label, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_CS], c->h_cs, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_TEMPER], c->h_temper, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_PRESSURE], c->h_press, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipStreamSynchronize(d->s));
  d->cfl_ready = 0;
  d->spec.valid = 0;
  d->opa_filled = 0;
  return 0;
}

// pebble fields and damping targets of thorin_init_pebbles
extern "C" int thorin_dev_upload_pebbles(thorin_ctx *c) {
  thorin_dev *d = c->dev;
  const size_t nb = sizeof(double) * (size_t)c->nr * c->ns;
  const int n1 = c->nr + 3;
  HCHECK(hipMemcpyAsync(d->b[B_PDENS], c->h_pdens, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_PVR], c->h_pvr, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_PVT], c->h_pvt, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_PSIZE], c->h_psize, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[B_TAU], c->h_tau, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync(d->b[THORIN_F_ETA], c->h_eta, nb, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync((double *)d->g.peb_dens0, c->PebDensInit, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync((double *)d->g.peb_vr0, c->PebVradInit, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  HCHECK(hipMemcpyAsync((double *)d->g.peb_vt0, c->PebVthetaInit, sizeof(double) * n1, hipMemcpyHostToDevice, d->s));
  HCHECK(hipStreamSynchronize(d->s));
  d->cfl_ready = 0;
  d->spec.valid = 0;
  return 0;
}

extern "C" int thorin_upload_field(thorin_ctx *c, int field, const double *host, size_t count) {
  if (!c || !c->dev || field < 0 || field >= THORIN_NFIELDS || !host) { thorin_set_error("bad upload args"); return -1; }
  thorin_dev *d = c->dev;
  if (count > d->nf) count = d->nf;
  HCHECK(hipMemcpyAsync(d->b[field], host, sizeof(double) * count, hipMemcpyHostToDevice, d->s));
  HCHECK(hipStreamSynchronize(d->s));
  d->cfl_ready = 0;
  d->spec.valid = 0;
  return 0;
}

extern "C" int thorin_download_field(thorin_ctx *c, int field, double *host, size_t count) {
  if (c && c->dev && host && field >= THORIN_X_ML && field <= THORIN_X_CRASH_TEMPER) {
    static const int xb[] = {B_ML, B_MLIP, B_MLIM, B_MLJP, B_MLJM, B_RHS, B_DIFF,
                             B_CRASH_DENS, B_CRASH_VRAD, B_CRASH_VTHETA, B_CRASH_TEMPER};
    thorin_dev *d = c->dev;
    if (count > d->nf) count = d->nf;
    HCHECK(hipMemcpyAsync(host, d->b[xb[field - THORIN_X_ML]], sizeof(double) * count, hipMemcpyDeviceToHost,
                          d->s));
    HCHECK(hipStreamSynchronize(d->s));
    return 0;
  }
  if (!c || !c->dev || field < 0 || field >= THORIN_NFIELDS || !host) { thorin_set_error("bad download args"); return -1; }
  thorin_dev *d = c->dev;
  if (count > d->nf) count = d->nf;
  if (field == THORIN_F_PRESSURE) {
    const size_t n = d->nf;
    hipLaunchKernelGGL(k_pressure, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->s, d->g, d->b[B_ENERGY],
                       d->b[B_PRESSURE], n);
    HCHECK(hipGetLastError());
  }
  HCHECK(hipMemcpyAsync(host, d->b[field], sizeof(double) * count, hipMemcpyDeviceToHost, d->s));
  HCHECK(hipStreamSynchronize(d->s));
  return 0;
}

extern "C" int thorin_sync(thorin_ctx *c) {
  if (need_dev(c)) return -1;
  HCHECK(hipStreamSynchronize(c->dev->s));
  return 0;
}

// The device math library as the kernels call it (same translation unit,
// same flags), one value per thread.
__global__ void k_math_eval(int fn, const double *__restrict__ x, const double *__restrict__ y,
                            double *__restrict__ out, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double a = x[k];
  double r;
  switch (fn) {
    case THORIN_MATH_POW: r = pow(a, y[k]); break;
    case THORIN_MATH_EXP: r = exp(a); break;
    case THORIN_MATH_LOG: r = log(a); break;
    case THORIN_MATH_LOG10: r = log10(a); break;
    case THORIN_MATH_TANH: r = tanh(a); break;
    case THORIN_MATH_SIN: r = sin(a); break;
    case THORIN_MATH_ATAN2: r = atan2(a, y[k]); break;
    case THORIN_MATH_ASIN: r = asin(a); break;
    case THORIN_MATH_ATAN: r = atan(a); break;
    default: r = sqrt(a); break;
  }
  out[k] = r;
}

// the context's opacity law (cgs in, cgs out), as the energy kernels call it
__global__ void k_opacity_eval(Geo g, const double *__restrict__ rho, const double *__restrict__ T,
                               double *__restrict__ out, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = opacity_law(g, rho[k], T[k]);
}

extern "C" int thorin_opacity_eval(thorin_ctx *c, const double *rho_cgs, const double *T_K, double *kappa_cgs,
                                   size_t n) {
  if (need_dev(c)) return -1;
  if (!rho_cgs || !T_K || !kappa_cgs || n > (size_t)1 << 26) {
    thorin_set_error("thorin_opacity_eval: needs rho, T, out and at most 2^26 values");
    return -1;
  }
  if (n == 0) return 0;
  thorin_dev *d = c->dev;
  const size_t nb = n * sizeof(double);
  double *dr = nullptr, *dt = nullptr, *dk = nullptr;
  int rc = -1;
  hipError_t e = hipMalloc((void **)&dr, nb);
  if (e == hipSuccess) e = hipMalloc((void **)&dt, nb);
  if (e == hipSuccess) e = hipMalloc((void **)&dk, nb);
  if (e == hipSuccess) e = hipMemcpyAsync(dr, rho_cgs, nb, hipMemcpyHostToDevice, d->s);
  if (e == hipSuccess) e = hipMemcpyAsync(dt, T_K, nb, hipMemcpyHostToDevice, d->s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_opacity_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->s, d->g, dr, dt, dk,
                       (int)n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(kappa_cgs, dk, nb, hipMemcpyDeviceToHost, d->s);
  if (e == hipSuccess) e = hipStreamSynchronize(d->s);
  if (e == hipSuccess) {
    rc = 0;
  } else {
    char m[200];
    snprintf(m, sizeof m, "thorin_opacity_eval: HIP error %s", hipGetErrorString(e));
    thorin_set_error(m);
  }
  if (dr) (void)hipFree(dr);
  if (dt) (void)hipFree(dt);
  if (dk) (void)hipFree(dk);
  return rc;
}

extern "C" int thorin_math_eval(int fn, const double *x, const double *y, double *out, size_t n) {
  if (fn < 0 || fn >= THORIN_MATH_N) {
    thorin_set_error("thorin_math_eval: unknown function");
    return -1;
  }
  const int two = (fn == THORIN_MATH_POW || fn == THORIN_MATH_ATAN2);
  if (!x || !out || (two && !y) || n > (size_t)1 << 26) {
    thorin_set_error("thorin_math_eval: needs x, out (and y for pow, atan2) and at most 2^26 values");
    return -1;
  }
  if (n == 0) return 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    thorin_set_error("no HIP device available: the FARGO_THORIN GPU path needs an MI355X (gfx950)");
    return -1;
  }
  const size_t nb = n * sizeof(double);
  double *dx = nullptr, *dy = nullptr, *dout = nullptr;
  int rc = -1;
  hipError_t e = hipMalloc((void **)&dx, nb);
  if (e == hipSuccess && two) e = hipMalloc((void **)&dy, nb);
  if (e == hipSuccess) e = hipMalloc((void **)&dout, nb);
  if (e == hipSuccess) e = hipMemcpy(dx, x, nb, hipMemcpyHostToDevice);
  if (e == hipSuccess && two) e = hipMemcpy(dy, y, nb, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_math_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, fn, dx, dy, dout, (int)n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dout, nb, hipMemcpyDeviceToHost);
  if (e == hipSuccess) {
    rc = 0;
  } else {
    char m[200];
    snprintf(m, sizeof m, "thorin_math_eval: HIP error %s", hipGetErrorString(e));
    thorin_set_error(m);
  }
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (dout) (void)hipFree(dout);
  return rc;
}


__global__ void k_int_to_double(const int *__restrict__ a, double *__restrict__ b, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = (double)a[k];
}

// 1D device profiles: 0 globcsvec, 1 globtempervec, 2 Vmoy (ConditionCFL),
// 3 VMed and 4 Nshift (ComputeResiduals) of the last substep
extern "C" int thorin_ring_profile(thorin_ctx *c, int which, double *out) {
  if (need_dev(c)) return -1;
  thorin_dev *d = c->dev;
  const double *src = NULL;
  switch (which) {
    case 0: src = d->g.gcs; break;
    case 1: src = d->g.gtemp; break;
    case 2: src = d->g.Vmoy; break;
    case 3: src = d->g.VMed; break;
    case 4:
      hipLaunchKernelGGL(k_int_to_double, dim3((d->nr + 255) / 256), dim3(256), 0, d->s, d->g.Nshift, d->ringtmp, d->nr);
      HCHECK(hipGetLastError());
      src = d->ringtmp;
      break;
    default: thorin_set_error("unknown ring profile"); return -1;
  }
  HCHECK(hipMemcpyAsync(out, src, sizeof(double) * d->nr, hipMemcpyDeviceToHost, d->s));
  HCHECK(hipStreamSynchronize(d->s));
  return 0;
}

extern "C" int thorin_gas_total_mass(thorin_ctx *c, double *out) {
  if (need_dev(c)) return -1;
  thorin_dev *d = c->dev;
  std::vector<double> rings(d->nr);
  hipLaunchKernelGGL(k_ring_mass, dim3(d->nr), dim3(RB), sizeof(double) * d->ns, d->s, d->g, d->b[B_DENS], d->ringtmp);
  HCHECK(hipGetLastError());
  HCHECK(hipMemcpyAsync(rings.data(), d->ringtmp, sizeof(double) * d->nr, hipMemcpyDeviceToHost, d->s));
  HCHECK(hipStreamSynchronize(d->s));
  double t = 0.0;   // the reference keeps one running sum over all cells; rings are summed in order here
  for (int i = 0; i < d->nr; i++) t += rings[i];
  *out = t;
  return 0;
}

static hipEvent_t prof_event(thorin_dev *d) {
  if (!d->prof_pool->empty()) {
    hipEvent_t e = d->prof_pool->back();
    d->prof_pool->pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Adds the timed launches whose end event has completed to the totals.
// drain: wait for all of them first (when the totals are read); otherwise
// the stream is not synchronised (the launches enqueued ahead of the host,
// such as the next substep's forces, keep running) and the pending ones,
// in stream order, stay for a later flush.
static int prof_flush(thorin_dev *d, int drain = 0) {
  if (!d->prof_pend || d->prof_pend->empty()) return 0;
  if (drain) HCHECK(hipStreamSynchronize(d->s));
  size_t k = 0;
  for (; k < d->prof_pend->size(); k++) {
    auto &r = (*d->prof_pend)[k];
    if (!drain) {
      const hipError_t q = hipEventQuery(r.second.second);
      if (q == hipErrorNotReady) break;
      HCHECK(q);
    }
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, r.second.first, r.second.second));
    d->prof_ms[r.first] += ms;
    d->prof_cnt[r.first] += 1;
    d->prof_pool->push_back(r.second.first);
    d->prof_pool->push_back(r.second.second);
  }
  d->prof_pend->erase(d->prof_pend->begin(), d->prof_pend->begin() + (long)k);
  return 0;
}

// launch on the context's stream; with the timer on, bracket the kernels of
// the selected ids with HIP events on that same stream
// ---------------------------------------------------------------------------
// Ensemble execution (thorin_ensemble_*).
//
// Each disk of an ensemble runs its own AlgoGas host logic on a host thread
// of its own, but issues nothing to the GPU: kernel launches, copies and
// stream syncs go through dev_launch / dev_copy / dev_sync, which in ensemble
// mode append to the disk's queue (Rec) and, for a sync, block the thread.
// When every disk thread waits or has finished, the ensemble merges the
// queues into one stream: the head operations that launch the same kernel
// become ONE launch over the disks' grids laid end to end (k_batch: a
// workgroup finds its disk from the block offsets and runs the kernel body
// with that disk's arguments; the disk's Geo / BOff tell the body its first
// workgroup), copies run as recorded.  Every disk's own order of operations
// is kept, and each workgroup runs exactly the code and arguments of the
// plain launch, so every disk computes what its own AlgoGas computes, bit for
// bit.

template <typename... P>
std::tuple<P...> params_of(void (*)(P...));

template <int NT, auto Body, typename... P>
__global__ void __launch_bounds__(NT) k_batch(const std::tuple<P...> *__restrict__ tab,
                                              const int *__restrict__ off, int n) {
  // the disk is the last one whose first block is at or before this block
  // (off ascending, off[0] = 0): a binary search on scalar loads
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  std::apply(Body, tab[lo]);
}

template <int NT, auto Body, typename T>
struct BatchKernel;
template <int NT, auto Body, typename... P>
struct BatchKernel<NT, Body, std::tuple<P...>> {
  static constexpr auto fn = &k_batch<NT, Body, P...>;
};

// kernels whose body can run in a batched launch
template <auto K>
struct Batch {
  static constexpr bool ok = false;
};
#define BATCH_BODY(KERN, BODY, NT)             \
  template <>                                  \
  struct Batch<&KERN> {                        \
    static constexpr bool ok = true;           \
    static constexpr auto body = &BODY;        \
    static constexpr int nt = NT;              \
  };
BATCH_BODY(k_eos_cfl<3>, b_eos_cfl<3>, RB)
BATCH_BODY(k_eos_cfl<6>, b_eos_cfl<6>, RB)
BATCH_BODY(k_eos_cfl<12>, b_eos_cfl<12>, RB)
BATCH_BODY(k_eos_cfl_vel<3>, b_eos_cfl_vel<3>, RB)
BATCH_BODY(k_eos_cfl_vel<6>, b_eos_cfl_vel<6>, RB)
BATCH_BODY(k_eos_cfl_vel<12>, b_eos_cfl_vel<12>, RB)
BATCH_BODY(k_cfl_final, b_cfl_final, CF_T)
BATCH_BODY((k_forces<false, 1>), (b_forces<false, 1>), TB)
BATCH_BODY((k_forces<false, 8>), (b_forces<false, 8>), TB)
BATCH_BODY((k_forces<true, 1>), (b_forces<true, 1>), TB)
BATCH_BODY((k_forces<true, 8>), (b_forces<true, 8>), TB)
BATCH_BODY(k_forces_reduce, b_forces_reduce, RB)
BATCH_BODY(k_indirect, b_indirect, TB)
BATCH_BODY(k_peb_cfl, b_peb_cfl, RB)
BATCH_BODY(k_peb_damp, b_peb_damp, TB)
BATCH_BODY(k_peb_source, b_peb_source, TB)
BATCH_BODY(k_peb_substep1, b_peb_substep1, TB)
BATCH_BODY(k_peb_evaporate, b_peb_evaporate, TB)
BATCH_BODY(k_peb_accrete, b_peb_accrete, ACC_T)
BATCH_BODY(k_boundary, b_boundary, TB)
BATCH_BODY(k_substep1, b_substep1, TB)
BATCH_BODY(k_divstress, b_divstress, TB)
BATCH_BODY(k_visc_update, b_visc_update, TB)
BATCH_BODY(k_substep2, b_substep2, TB)
BATCH_BODY(k_visc_chain, b_visc_chain, VC_T)
BATCH_BODY(k_substep3<false>, b_substep3<false>, TB)
BATCH_BODY(k_substep3<true>, b_substep3<true>, TB)
BATCH_BODY(k_matrix, b_matrix, TB)
BATCH_BODY((k_sor_iter<1, SOR_NT(1), 0>), (b_sor_iter<1, SOR_NT(1), 0>), SOR_NT(1))
BATCH_BODY((k_sor_iter<2, SOR_NT(2), 0>), (b_sor_iter<2, SOR_NT(2), 0>), SOR_NT(2))
BATCH_BODY((k_sor_iter<3, SOR_NT(3), 0>), (b_sor_iter<3, SOR_NT(3), 0>), SOR_NT(3))
BATCH_BODY((k_sor_iter<1, SOR_NT(1), 1>), (b_sor_iter<1, SOR_NT(1), 1>), SOR_NT(1))
BATCH_BODY((k_sor_iter<2, SOR_NT(2), 1>), (b_sor_iter<2, SOR_NT(2), 1>), SOR_NT(2))
BATCH_BODY((k_sor_iter<3, SOR_NT(3), 1>), (b_sor_iter<3, SOR_NT(3), 1>), SOR_NT(3))
BATCH_BODY((k_sor_iter<1, SOR_NT(1), 2>), (b_sor_iter<1, SOR_NT(1), 2>), SOR_NT(1))
BATCH_BODY((k_sor_iter<2, SOR_NT(2), 2>), (b_sor_iter<2, SOR_NT(2), 2>), SOR_NT(2))
BATCH_BODY((k_sor_iter<3, SOR_NT(3), 2>), (b_sor_iter<3, SOR_NT(3), 2>), SOR_NT(3))
BATCH_BODY(k_sor_check<1>, b_sor_check<1>, RB)
BATCH_BODY(k_sor_check<2>, b_sor_check<2>, RB)
BATCH_BODY(k_sor_check<3>, b_sor_check<3>, RB)
BATCH_BODY(k_transport_rad<8>, b_transport_rad<8>, RT_T)
BATCH_BODY(k_transport_rad<16>, b_transport_rad<16>, RT_T)
BATCH_BODY(k_transport_rad<32>, b_transport_rad<32>, RT_T)
BATCH_BODY(k_ring_means, b_ring_means, 256)
BATCH_BODY(k_transport_theta<1>, b_transport_theta<1>, TT_T)
BATCH_BODY(k_transport_theta<2>, b_transport_theta<2>, TT_T)
BATCH_BODY(k_transport_theta<3>, b_transport_theta<3>, TT_T)
BATCH_BODY(k_transport_theta<4>, b_transport_theta<4>, TT_T)
BATCH_BODY(k_transport_theta<5>, b_transport_theta<5>, TT_T)
BATCH_BODY(k_transport_theta<6>, b_transport_theta<6>, TT_T)
BATCH_BODY(k_velocities, b_velocities, TB)
#undef BATCH_BODY

static inline void set_blk0(Geo &g, int o) { g.blk0 = o; }
static inline void set_blk0(BOff &b, int o) { b.blk0 = o; }
template <typename T>
static inline void set_blk0(T &, int) {}

// staging of the batched launches' argument tables: pinned host memory and
// its device copy, one host-to-device copy per flush
struct Stage {
  char *h = nullptr;
  char *dv = nullptr;
  size_t cap = 0, used = 0;
  size_t take(size_t n) {
    const size_t o = used;
    used += (n + 255) & ~(size_t)255;
    return o;
  }
};

struct RecOp;
typedef std::function<int(hipStream_t, const char *)> IssueFn;
// plan: appends the tables of a group of operations of one kernel to the
// host image of the stage and returns what issues them (given the device
// address of the stage)
// The words appended to sig describe everything the issued work depends on
// besides the table contents (kernel, grid, offsets, copy addresses): two
// flushes with equal signatures issue the same graph.
typedef IssueFn (*PlanFn)(RecOp **ops, int n, std::vector<char> &tables, std::vector<unsigned long long> &sig);

struct RecOp {
  int type;                     // 0 kernel, 1 copy
  const void *key;              // kernel identity: the batching group
  PlanFn plan;
  std::vector<char> args;       // the kernel's argument tuple
  dim3 grid, block;
  size_t shmem;
  void *dst;
  const void *src;
  size_t bytes;
  hipMemcpyKind kind;
};

// The end-of-substep marker of a recorded queue: what comes before it is what
// the disk's host side waits for; what follows (the next substep's forces)
// may still run while the host works.  Merged like a kernel (one group when
// every disk has reached it), issued as an event the flush then waits for.
static const int ens_marker_key = 0;

struct EnsGraph {               // a captured flush
  std::vector<unsigned long long> sig;
  hipGraph_t g;
  hipGraphExec_t x;
  unsigned long long used;
};

struct thorin_ensemble;
struct Rec {                    // one disk's queue
  std::vector<RecOp> q;
  thorin_ensemble *e;
  int idx;
};

struct thorin_ensemble {
  int n;
  thorin_ctx **c;
  hipStream_t s;
  std::mutex m;
  std::condition_variable cv_coord, cv_disk;
  int nwait, nfinished;
  unsigned long gen;
  int err;
  std::vector<Rec> recs;
  Stage st;
  long long launches, batched_ops;   // statistics: launches issued, kernel operations they carried
  int use_graphs;                    // replay flushes as HIP graphs
  std::vector<struct EnsGraph> graphs;
  std::vector<std::vector<unsigned long long>> seen;   // signatures met once
  unsigned long long tick;
  long long captures, replays;
  // the disks' host threads, one per disk for the ensemble's life: woken
  // for each DT (job), they run thorin_algogas on their disk
  hipEvent_t mark;               // the end-of-substep markers' event
  int marked;                    // markers in the flush being planned (one per disk at most)
  std::vector<std::thread> workers;
  std::condition_variable cv_job;
  unsigned long job;             // DTs started
  bool quit;
  thorin_planets *const *j_sys;
  thorin_nbody_fn j_nbody;
  void *const *j_user;
  thorin_step_log *j_logs;
  std::vector<int> j_rc;
  std::string j_err;             // the first failing disk's error text
};

// the table of a group: n argument tuples, then n+1 block offsets; the host
// image is built in `tables`, the device address is the stage's base + offset
template <auto K>
static IssueFn plan_batch(RecOp **ops, int n, std::vector<char> &tables, std::vector<unsigned long long> &sig) {
  using Tup = decltype(params_of(K));
  const size_t o = tables.size();
  const size_t tb = (sizeof(Tup) * (size_t)n + 15) & ~(size_t)15;
  const size_t need = (tb + sizeof(int) * (size_t)(n + 1) + 255) & ~(size_t)255;
  tables.resize(o + need);
  Tup *t = reinterpret_cast<Tup *>(tables.data() + o);
  int *off = reinterpret_cast<int *>(tables.data() + o + tb);
  int b = 0;
  size_t shm = 0;
  for (int k = 0; k < n; k++) {
    new (&t[k]) Tup(*reinterpret_cast<const Tup *>(ops[k]->args.data()));
    std::apply([&](auto &...a) { (set_blk0(a, b), ...); }, t[k]);
    off[k] = b;
    b += (int)ops[k]->grid.x;
    shm = ops[k]->shmem > shm ? ops[k]->shmem : shm;
  }
  off[n] = b;
  const dim3 block = ops[0]->block;
  const int nb = b;
  constexpr auto fn = BatchKernel<Batch<K>::nt, Batch<K>::body, Tup>::fn;
  if (shm > 65536) {
    static size_t set = 0;
    if (shm > set &&
        hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) == hipSuccess)
      set = shm;
  }
  const unsigned long long w[] = {(unsigned long long)(uintptr_t)K, (unsigned long long)n, (unsigned long long)nb,
                                  (unsigned long long)block.x, (unsigned long long)shm, (unsigned long long)o};
  sig.insert(sig.end(), w, w + 6);
  return [=](hipStream_t s, const char *dev) -> int {
    hipLaunchKernelGGL(fn, dim3((unsigned)nb), block, shm, s, reinterpret_cast<const Tup *>(dev + o),
                       reinterpret_cast<const int *>(dev + o + tb), n);
    HCHECK(hipGetLastError());
    return 0;
  };
}

// a kernel without a batched body: each disk's launch as recorded
template <auto K>
static IssueFn plan_plain(RecOp **ops, int n, std::vector<char> &, std::vector<unsigned long long> &sig) {
  using Tup = decltype(params_of(K));
  std::vector<Tup> args;
  std::vector<std::pair<dim3, std::pair<dim3, size_t>>> cfg;
  for (int k = 0; k < n; k++) {
    args.push_back(*reinterpret_cast<const Tup *>(ops[k]->args.data()));
    cfg.push_back({ops[k]->grid, {ops[k]->block, ops[k]->shmem}});
  }
  // the arguments are baked into the launches: such a flush is never replayed
  sig.push_back(~0ull);
  return [=](hipStream_t s, const char *) -> int {
    for (size_t k = 0; k < args.size(); k++)
      std::apply([&](auto... a) { hipLaunchKernelGGL(K, cfg[k].first, cfg[k].second.first, cfg[k].second.second, s, a...); },
                 args[k]);
    HCHECK(hipGetLastError());
    return 0;
  };
}

static int rec_wait(Rec *r);

template <auto K, typename... A>
static int dev_launch(thorin_dev *d, const char *name, dim3 grid, dim3 block, size_t shmem, A &&...a) {
  using Tup = decltype(params_of(K));
  if (d->rec) {
    RecOp op;
    op.type = 0;
    op.key = (const void *)K;
    op.grid = grid;
    op.block = block;
    op.shmem = shmem;
    op.args.resize(sizeof(Tup));
    new (op.args.data()) Tup(std::forward<A>(a)...);
    if constexpr (Batch<K>::ok) op.plan = &plan_batch<K>;
    else op.plan = &plan_plain<K>;
    d->rec->q.push_back(std::move(op));
    return 0;
  }
  int kid = d->prof_on ? kid_of(name) : -1;
  hipEvent_t ea = nullptr, eb = nullptr;
  if (kid >= 0 && ((d->prof_mask >> kid) & 1ull)) {
    ea = prof_event(d);
    eb = prof_event(d);
    if (ea && eb) (void)hipEventRecord(ea, d->s);
  }
  hipLaunchKernelGGL(K, grid, block, shmem, d->s, std::forward<A>(a)...);
  if (ea && eb) {
    (void)hipEventRecord(eb, d->s);
    d->prof_pend->push_back(std::make_pair(kid, std::make_pair(ea, eb)));
  }
  return 0;
}

static int dev_copy(thorin_dev *d, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  if (d->rec) {
    RecOp op;
    op.type = 1;
    op.key = nullptr;
    op.plan = nullptr;
    op.dst = dst;
    op.src = src;
    op.bytes = bytes;
    op.kind = kind;
    d->rec->q.push_back(std::move(op));
    return 0;
  }
  HCHECK(hipMemcpyAsync(dst, src, bytes, kind, d->s));
  return 0;
}

static int dev_sync(thorin_dev *d) {
  if (d->rec) return rec_wait(d->rec);
  HCHECK(hipStreamSynchronize(d->s));
  return 0;
}

// the end of the substep's own work.  One disk: k_cfl_final writes the
// substep's sequence number into the pinned flag word, and dev_wait_mark
// polls it (no event: recording one there gives k_cfl_final a system-scope
// release that holds the next kernel back).  In an ensemble: a marker op,
// recorded as an event by the merged flush.
static int dev_mark(thorin_dev *d) {
  if (d->rec) {
    RecOp op;
    op.type = 0;
    op.key = (const void *)&ens_marker_key;
    op.plan = nullptr;            // issued by the flush itself
    op.grid = dim3(1);
    op.block = dim3(1);
    op.shmem = 0;
    d->rec->q.push_back(std::move(op));
  }
  return 0;
}

// wait for the marked end of the substep (dev_mark).  Should the stream
// finish without the flag word showing the substep's number, that is an
// error (checked after 2 ms, then every ~1000 polls).
static int dev_wait_mark(thorin_dev *d) {
  if (d->rec) return rec_wait(d->rec);
  const unsigned seq = d->flag_seq;
  volatile unsigned *f = d->hflag;
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned long k = 1; *f != seq; k++) {
    __builtin_ia32_pause();
    if ((k & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
      const hipError_t e = hipStreamQuery(d->s);
      if (e == hipSuccess && *f != seq) {
        thorin_set_error("the end-of-substep kernel finished without writing the host flag");
        return -1;
      }
      if (e != hipSuccess && e != hipErrorNotReady) HCHECK(e);
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return 0;
}

#define LAUNCH(kern, grid, block, shmem, ...)                                       \
  do {                                                                              \
    if (dev_launch<&kern>(d, #kern, grid, block, shmem, __VA_ARGS__)) return -1;    \
  } while (0)
#define DCOPY(dst, src, bytes, kind)                                 \
  do {                                                               \
    if (dev_copy(d, dst, src, bytes, kind)) return -1;               \
  } while (0)
#define DSYNC()                  \
  do {                           \
    if (dev_sync(d)) return -1;  \
  } while (0)

extern "C" int thorin_ring_means(thorin_ctx *c, int field, double *out) {
  if (need_dev(c)) return -1;
  thorin_dev *d = c->dev;
  if (field < 0 || field >= THORIN_NFIELDS) return -1;
  LAUNCH(k_ring_means, dim3((unsigned)((d->nr + d->rs_rows - 1) / d->rs_rows)), dim3(256),
         sizeof(double) * (size_t)d->rs_rows * d->ns, (const double *)d->b[field], d->nr, d->ns, d->rs_rows,
         d->ringtmp, (const int *)nullptr, BOff{0});
  DCOPY(out, d->ringtmp, sizeof(double) * d->nr, hipMemcpyDeviceToHost);
  DSYNC();
  return 0;
}

extern "C" int thorin_prof_enable(thorin_ctx *c, int on, unsigned long long mask) {
  if (need_dev(c)) return -1;
  thorin_dev *d = c->dev;
  if (prof_flush(d, 1)) return -1;
  d->prof_on = on;
  d->prof_mask = mask;
  for (int k = 0; k < 64; k++) { d->prof_ms[k] = 0.0; d->prof_cnt[k] = 0; }
  return 0;
}

extern "C" int thorin_prof_read(thorin_ctx *c, int kid, double *total_ms, long long *count) {
  if (need_dev(c)) return -1;
  if (kid < 0 || kid >= 64) return -1;
  if (prof_flush(c->dev, 1)) return -1;
  *total_ms = c->dev->prof_ms[kid];
  *count = c->dev->prof_cnt[kid];
  return 0;
}

extern "C" const char *thorin_prof_kernel_name(int kid) {
  int n = 0;
  while (KNAMES[n]) n++;
  return (kid >= 0 && kid < n) ? KNAMES[kid] : NULL;
}

// --------------------------------------------------------------------------
// SOR driver

// enqueue iterations it0 .. it0+count-1, each launch of iterations followed
// by its stop test: launches of sor_nit iterations while that many remain,
// then one launch of the rest
static int sor_launch(thorin_ctx *c, double omega, int it0, int &count, double eps = SOREPS,
                      int maxiter = SORMAXITERS) {
  thorin_dev *d = c->dev;
  const Geo &g = d->g;
  const int nr = d->nr, ns = d->ns;
  // the energy system: its rows, or (sys_form) their inputs for the first
  // launch of a solve; every solve's first launch forms the initial residual
  SorSys sy;
  sy.ml = d->b[B_ML];
  sy.mip = d->b[B_MLIP];
  sy.mim = d->b[B_MLIM];
  sy.mjp = d->b[B_MLJP];
  sy.mjm = d->b[B_MLJM];
  sy.rhs = d->b[B_RHS];
  sy.rho = d->b[B_DENS];
  sy.cs = d->b[B_CS];
  sy.D = d->b[B_DIFF];
  sy.divv = d->b[B_DIVV];
  sy.qminus = d->b[B_QMINUS];
  sy.qplus = d->b[B_QPLUS];
  sy.opa = d->b[B_OPACITY];
  sy.dt = d->sys_dt;
  sy.heat = d->sys_heat;
  sy.form = d->sys_form;
  const int nbs = ((ns + SR_TJ - 1) / SR_TJ) * ((nr - 2 + SR_TI - 1) / SR_TI);
  const dim3 gs((unsigned)((nbs + 7) / 8 * 8));
  // the iterations queued beyond the expected count are a margin: a launch
  // after the solve's end returns at once, so the rest goes in one launch
  // (an empty launch and its stop test cost about 9 us at 1024x1536)
  const int nmax = (d->sor_nit >= 1 && d->sor_nit <= 3) ? d->sor_nit : 1;
#define SOR_LAUNCH(N, F)                                                                                          \
  do {                                                                                                            \
    LAUNCH((k_sor_iter<N, SOR_NT(N), F>), gs, dim3(SOR_NT(N)), 0, g, d->b[B_TEMPER], d->b[B_TEMPER2],      \
           d->b[B_TEMPER3], d->b[B_TEMPER4], sy, omega, d->sorp, d->sorc, d->ctl, it, d->mresp, nbs);             \
    LAUNCH(k_sor_check<N>, dim3(1), dim3(RB), 0, d->sorp, d->sorc, nbs, d->ctl, maxiter, it, eps,                  \
           (const double *)d->mresp);                                                                             \
  } while (0)
  // (a stop test made by the last workgroup of each single-iteration launch
  // needs a device-scope fence per workgroup: a single iteration then took
  // three times as long as with the separate check)
  for (int it = it0; it < it0 + count;) {
    const int nit = std::min(nmax, it0 + count - it);
    if (!sy.form) {
      if (nit == 3) SOR_LAUNCH(3, 0);
      else if (nit == 2) SOR_LAUNCH(2, 0);
      else SOR_LAUNCH(1, 0);
    } else if (it == 0) {           // the solve's first launch: rows and right-hand sides
      if (nit == 3) SOR_LAUNCH(3, 1);
      else if (nit == 2) SOR_LAUNCH(2, 1);
      else SOR_LAUNCH(1, 1);
    } else {
      if (nit == 3) SOR_LAUNCH(3, 2);
      else if (nit == 2) SOR_LAUNCH(2, 2);
      else SOR_LAUNCH(1, 2);
    }
    it += nit;
  }
#undef SOR_LAUNCH
  HCHECK(hipGetLastError());
  return 0;
}

// Run the solve to its end with host checks between batches, the first
// `launched` iterations being already enqueued.  The last iterate stays in
// B_TEMPER .. B_TEMPER4 as ctl->sor_cur says (the radial transport forms
// EnergyNew from it).
static int sor_finish(thorin_ctx *c, double omega, int launched, int *niter_out, double eps = SOREPS,
                      int maxiter = SORMAXITERS, int first_batch = 0) {
  thorin_dev *d = c->dev;
  int batch = first_batch > 0 ? first_batch : (c->st.sor_niter_last > 0 ? c->st.sor_niter_last + 1 : 8);
  if (batch > 256) batch = 256;
  if (launched == 0) {
    if (sor_launch(c, omega, 0, batch, eps, maxiter)) return -1;
    launched = batch;
  }
  for (;;) {
    DCOPY(d->hctl, d->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost);
    DSYNC();
    if (prof_flush(d)) return -1;
    if (d->hctl->sor_done) {
      *niter_out = d->hctl->sor_niter;
      return 0;
    }
    if (launched > maxiter + 2) { *niter_out = maxiter + 1; return 0; }
    batch = batch < 64 ? 2 * batch : 128;
    if (sor_launch(c, omega, launched, batch, eps, maxiter)) return -1;
    launched += batch;
  }
}

static int sor_solve(thorin_ctx *c, double omega, int *niter_out) { return sor_finish(c, omega, 0, niter_out); }

// IterateRelaxationParameter (EnergySources.c:340-403)
static int iterate_relaxation(thorin_ctx *c) {
  thorin_dev *d = c->dev;
  double omegamin = 1.0, omega = 1.0, omegamax = OMEGAUPPER;
  int nsplit = 9, Niter, count = 0;
  const size_t nb = sizeof(double) * d->nf;
  DCOPY(d->b[B_TBACK], d->b[B_TEMPER], nb, hipMemcpyDeviceToDevice);
  c->st.sor_domega = (omegamax - omegamin) / nsplit;
  c->st.sor_niter_best = SORMAXITERS;
  c->st.sor_omega_best = 1.0;
  int saved_last = c->st.sor_niter_last;
  for (;;) {
    for (int n = 0; n <= nsplit; n++) {
      c->st.sor_niter_last = 0;
      if (sor_solve(c, omega, &Niter)) return -1;
      if (Niter == c->st.sor_niter_best) count++;
      if (Niter < c->st.sor_niter_best) { c->st.sor_niter_best = Niter; c->st.sor_omega_best = omega; }
      omega += c->st.sor_domega;
      DCOPY(d->b[B_TEMPER], d->b[B_TBACK], nb, hipMemcpyDeviceToDevice);
    }
    if (c->st.sor_niter_best == SORMAXITERS) {
      c->st.sor_niter_last = saved_last;
      thorin_set_error("SOR did not converge when estimating the initial over-relaxation parameter");
      return -1;
    }
    if (count == nsplit + 1) {
      c->st.sor_domega = omegamax - omegamin;
      return 0;
    }
    count = 0;
    omegamin = c->st.sor_omega_best - c->st.sor_domega;
    omegamax = c->st.sor_omega_best + c->st.sor_domega;
    if (omegamin < 1.0) omegamin = 1.0;
    if (omegamax > OMEGAUPPER) omegamax = OMEGAUPPER;
    c->st.sor_domega = (omegamax - omegamin) / nsplit;
    omega = omegamin;
  }
}

// MinStepForRebound (ReboundInterface.c:360-385)
static double minstep_rebound(const thorin_planets *sys) {
  double dt = 1.e30;
  for (int i = 0; i < sys->nb; i++)
    for (int j = i + 1; j < sys->nb; j++) {
      double r2 = (sys->x[i] - sys->x[j]) * (sys->x[i] - sys->x[j]) + (sys->y[i] - sys->y[j]) * (sys->y[i] - sys->y[j]) +
                  (sys->z[i] - sys->z[j]) * (sys->z[i] - sys->z[j]);
      double rv = (sys->x[i] - sys->x[j]) * (sys->vx[i] - sys->vx[j]) + (sys->y[i] - sys->y[j]) * (sys->vy[i] - sys->vy[j]) +
                  (sys->z[i] - sys->z[j]) * (sys->vz[i] - sys->vz[j]);
      double dtnew = fabs(r2 / rv);
      if (dtnew < dt) dt = dtnew;
    }
  dt *= 0.25;
  dt *= dt;
  return 1.0 / dt;
}



// AlgoGas (SourceEuler.c:210-317): one DT
// AccretePebblesOntoPlanets STEPS 0-2 on the host (Pebbles.c:374-441): the
// planet's scalars, vt1D next to it (ordered ring means of the gas vtheta,
// mpi_make1Dprofile order), the annulus of its Hill sphere, the annulus'
// mean Stokes number in the reference's sequential order and Reff.  The two
// vtheta rings and the annulus' Stokes numbers come from the device (the
// values the substep starts with).
static int peb_accrete_host(thorin_ctx *c, const thorin_planets *sys, PebAccK *P) {
  thorin_dev *d = c->dev;
  const int nr = d->nr, ns = d->ns;
  struct Pl { double Rp, Rhill, frac; int iv, imin, imax; } pl[PEB_MAXPL];
  int lo = nr, hi = -1;
  P->np = sys->nb;
  for (int k = 0; k < sys->nb; k++) {
    const double X = sys->x[k], Y = sys->y[k], Z = sys->z[k], M = sys->mass[k];
    double Rp = X * X + Y * Y;
    const double Rp3D = sqrt(Rp + Z * Z);
    Rp = sqrt(Rp);
    pl[k].Rp = Rp;
    pl[k].Rhill = Rp3D * pow(1.0 / 3.0 * M, 1.0 / 3.0);
    P->x[k] = X; P->y[k] = Y; P->z[k] = Z; P->m[k] = M;
    P->om[k] = pow(Rp, -1.5);
    double ifrac;   // GetGlobalIFrac (LowTasks.c:21-31)
    if (Rp < c->GlobalRmed[0]) ifrac = 0.0;
    else if (Rp > c->GlobalRmed[nr - 1]) ifrac = (double)nr - 1.0;
    else {
      int i = 0;
      while (c->GlobalRmed[i] <= Rp) i++;
      ifrac = (double)i + (Rp - c->GlobalRmed[i - 1]) / (c->GlobalRmed[i] - c->GlobalRmed[i - 1]) - 1.0;
    }
    pl[k].frac = ifrac - floor(ifrac);
    pl[k].iv = (int)ifrac;
    int imax = nr - 1, imin = 0;
    while (imax >= imin && c->Rinf[imax] > (Rp + pl[k].Rhill)) imax--;
    while (imin <= imax && c->Rsup[imin] < (Rp - pl[k].Rhill)) imin++;
    pl[k].imin = imin;
    pl[k].imax = imax;
    P->imin[k] = imin;
    P->imax[k] = imax;
    if (imin <= imax) { lo = imin < lo ? imin : lo; hi = imax > hi ? imax : hi; }
  }
  // staging: vtheta rings iv, iv+1 of each planet, then the Stokes rings lo..hi
  const size_t nv = (size_t)2 * sys->nb * ns, nt = hi >= lo ? (size_t)(hi - lo + 1) * ns : 0;
  if (nv + nt > d->acc_stage_n) {
    if (d->acc_stage) HCHECK(hipHostFree(d->acc_stage));
    d->acc_stage = nullptr;
    HCHECK(hipHostMalloc(&d->acc_stage, sizeof(double) * (nv + nt), hipHostMallocDefault));
    d->acc_stage_n = nv + nt;
  }
  double *sv = d->acc_stage, *st = d->acc_stage + nv;
  for (int k = 0; k < sys->nb; k++)
    for (int q = 0; q < 2; q++) {
      const int r = pl[k].iv + q;
      if (r < nr)
        DCOPY(sv + (size_t)(2 * k + q) * ns, d->b[B_VTHETA] + (size_t)r * ns, sizeof(double) * ns,
                              hipMemcpyDeviceToHost);
    }
  if (nt)
    DCOPY(st, d->b[B_TAU] + (size_t)lo * ns, sizeof(double) * nt, hipMemcpyDeviceToHost);
  DSYNC();
  for (int k = 0; k < sys->nb; k++) {
    double vt1[2];
    for (int q = 0; q < 2; q++) {
      vt1[q] = 0.0;   // vt1D[nr] stays 0 (mpi_make1Dprofile fills rings 0..nr-1)
      if (pl[k].iv + q < nr) {
        const double *row = sv + (size_t)(2 * k + q) * ns;
        double s = 0.0;
        for (int j = 0; j < ns; j++) s += row[j];
        vt1[q] = s / (double)ns;
      }
    }
    const double X = P->x[k], Y = P->y[k], M = P->m[k], Rp = pl[k].Rp;
    const double VX = sys->vx[k], VY = sys->vy[k], VZ = sys->vz[k];
    const double flow = vt1[0] * (1.0 - pl[k].frac) + vt1[1] * pl[k].frac + Rp * c->st.omega_frame;
    const double flowx = -flow * Y / Rp, flowy = flow * X / Rp;
    const double vrel = sqrt((VX - flowx) * (VX - flowx) + (VY - flowy) * (VY - flowy) + VZ * VZ);
    const double Mt = sqrt(1.0 / 3.0) * pow(vrel, 3.0) / P->om[k];
    int ncell = 0;
    double tau_mean = 0.0;
    for (int i = pl[k].imin; i <= pl[k].imax; i++) {
      const double *row = st + (size_t)(i - lo) * ns;
      for (int j = 0; j < ns; j++) { tau_mean += row[j]; ncell++; }
    }
    tau_mean /= (double)ncell;
    double Reff, fac;
    if (M < Mt) {
      const double Rbondi = M / vrel / vrel;
      const double tbondi = Rbondi / vrel;
      fac = sqrt(tau_mean * pow(Rp, 1.5) / tbondi);
      Reff = Rbondi * fac;
    } else {
      fac = pow(tau_mean / 0.1, 1.0 / 3.0);
      Reff = pl[k].Rhill * fac;
    }
    if (Reff > pl[k].Rhill) Reff = pl[k].Rhill;
    P->reff[k] = Reff;
    // Sectors that can hold a cell centre within Reff of the planet: a cell
    // at radius r and angle dphi from it lies at least r |sin dphi| away (and
    // at least r away for |dphi| >= pi/2), so with Reff < r_min every such
    // cell has |dphi| <= asin(Reff / r_min); two sectors of margin.
    P->jlo[k] = 0;
    P->jn[k] = ns;
    if (pl[k].imin <= pl[k].imax && Reff < c->Rmed[pl[k].imin]) {
      const double dphi = 2.0 * TH_PI / (double)ns;
      const int w = (int)(asin(Reff / c->Rmed[pl[k].imin]) / dphi) + 2;
      if (2 * w + 1 < ns) {
        double ph = atan2(Y, X);
        if (ph < 0.0) ph += 2.0 * TH_PI;
        const int jc = (int)floor(ph / dphi + 0.5) % ns;
        P->jlo[k] = ((jc - w) % ns + ns) % ns;
        P->jn[k] = 2 * w + 1;
      }
    }
    P->vrel[k] = vrel;
    P->mt[k] = Mt;
  }
  return 0;
}

// Transport (TransportEuler.c:34-47) of one fluid: radial van Leer of the
// momenta, [energy,] density into the X buffers, azimuthal pass with the
// FARGO shift back into the fields and the momenta, ComputeVelocities.  be < 0:
// no energy (TransportPebbles).  g.VMed is the fluid's vtheta ring mean.
// with_vel = 0: the velocities are left to the end-of-substep kernel (k_eos_cfl_vel)
static int transport(thorin_dev *d, const Geo &g, const RadQ &q, int brho, int bvr, int bvt, int be, double dt,
                     const int *guard, int with_vel = 1) {
  const int nr = d->nr, ns = d->ns;
  const dim3 gall = cell_grid(d, nr), tb(TB);
  const dim3 grad((unsigned)((q.ntiles + 7) / 8 * 8)), trad(RT_T);
  if (q.rows == 8) LAUNCH(k_transport_rad<8>, grad, trad, 0, g, q, dt);
  else if (q.rows == 32) LAUNCH(k_transport_rad<32>, grad, trad, 0, g, q, dt);
  else LAUNCH(k_transport_rad<16>, grad, trad, 0, g, q, dt);
  double *xe = be >= 0 ? d->b[B_XE] : nullptr, *oe = be >= 0 ? d->b[be] : nullptr;
  const int nq = be >= 0 ? 5 : 4;
#define TT_LAUNCH(NPT)                                                                                        \
  LAUNCH(k_transport_theta<NPT>, dim3(nr), dim3(TT_T), sizeof(double) * 2 * ns, g, d->b[bvt], d->b[B_XRP],     \
         d->b[B_XRM], d->b[B_XTP], d->b[B_XTM], xe, d->b[B_XRHO], d->b[B_RP], d->b[B_RM], d->b[B_TP], d->b[B_TM], \
         oe, d->b[brho], dt, guard, nq)
  int npt = (ns + TT_T - 1) / TT_T;
  if (npt < d->tt_npt_min) npt = d->tt_npt_min;   // an ensemble runs its disks with one instance
  switch (npt) {
    case 1: TT_LAUNCH(1); break;
    case 2: TT_LAUNCH(2); break;
    case 3: TT_LAUNCH(3); break;
    case 4: TT_LAUNCH(4); break;
    case 5: TT_LAUNCH(5); break;
    default: TT_LAUNCH(6); break;
  }
#undef TT_LAUNCH
  if (with_vel)
    LAUNCH(k_velocities, gall, tb, 0, g, d->b[brho], d->b[B_RP], d->b[B_RM], d->b[B_TP], d->b[B_TM], d->b[bvr],
           d->b[bvt], guard);
  return 0;
}


// FillForcesArrays (Pframeforce.c:43-510) for the planets sys at MassTaper mt:
// planet constants, the torque log's rings (want_tq: UpdateLog, Force.c:27-70,
// at the first substep of a DT) and DampingTW04's rings, then k_forces,
// k_forces_reduce, k_indirect.  guard: as the post-SOR kernels.
static bool force_set_matches(const ForceSet &f, const thorin_planets *sys, double mt, int want_tq, double t) {
  if (f.nb != sys->nb || f.want_tq != want_tq || f.mt != mt || f.t != t || sys->nb > 8) return false;
  for (int k = 0; k < sys->nb; k++)
    if (f.x[k] != sys->x[k] || f.y[k] != sys->y[k] || f.z[k] != sys->z[k] || f.m[k] != sys->mass[k] ||
        f.feel[k] != sys->feeldisk[k])
      return false;
  return true;
}

static int launch_forces(thorin_ctx *c, const thorin_planets *sys, double mt, int want_tq, const int *guard,
                         ForceSet *fs) {
  thorin_dev *d = c->dev;
  const thorin_params *p = &c->p;
  const int nr = d->nr;
  const int peb = p->pebbles;
  const dim3 tb(TB), rb(RB);
  const dim3 gall = cell_grid(d, nr);
  memset(fs, 0, sizeof *fs);
  fs->nb = sys->nb;
  fs->want_tq = want_tq;
  fs->mt = mt;
  fs->t = c->st.physical_time;
  for (int k = 0; k < sys->nb && k < 8; k++) {
    fs->x[k] = sys->x[k];
    fs->y[k] = sys->y[k];
    fs->z[k] = sys->z[k];
    fs->m[k] = sys->mass[k];
    fs->feel[k] = sys->feeldisk[k];
  }
  double indir_px = 0.0, indir_py = 0.0;
  PlanetsK pks;
  memset(&pks, 0, sizeof pks);
  {
    double axs = 0.0, ays = 0.0;
    for (int k = 0; k < sys->nb; k++) {
      PlanetK &P = pks.p[k];
      const double xpl = sys->x[k], ypl = sys->y[k], zpl = sys->z[k];
      double mpl = sys->mass[k];
      const double dpl = sqrt(xpl * xpl + ypl * ypl + zpl * zpl);
      P.x = xpl; P.y = ypl; P.z = zpl; P.m = mpl;
      P.rhill = dpl * pow((1.0 / 3.0 * mpl), 1.0 / 3.0);
      const double rst = 0.5 * P.rhill;
      P.rst2 = rst * rst;
      P.rst3 = P.rst2 * rst;
      P.rst4 = P.rst3 * rst;
      const double rsm = 0.05 * P.rhill;
      P.rsm2 = rsm * rsm;
      P.rmorevert2 = 2.0 * P.rhill;
      P.rmorevert2 *= P.rmorevert2;
      // ring of ThicknessSmoothing (Force.c:85-90)
      const double r = sqrt(xpl * xpl + ypl * ypl);
      P.iring = -1;
      if (r >= c->Rinf[0] && r <= c->Rsup[nr - 1]) {
        int i = 0;
        while (c->Rsup[i] < r) i++;
        P.iring = i;
      }
      if (p->indirect_term) {
        const double invdpl3 = 1.0 / (dpl * dpl * dpl);
        mpl *= mt;
        axs += mpl * invdpl3 * xpl;
        ays += mpl * invdpl3 * ypl;
      }
    }
    indir_px = -axs;
    indir_py = -ays;
  }
  // UpdateLog (Force.c:27-70, called from AdvanceSystemFromDisk) at the first
  // substep of the DT: the planet side is taken now, the two ring means per
  // planet are formed by the force kernel's leading block and the rows are
  // completed after the end-of-substep sync
  int ntq = 0;
  TqRings rr;
  if (want_tq) {
    ntq = sys->nb < 8 ? sys->nb : 8;
    for (int k = 0; k < ntq; k++) {
      const double x = sys->x[k], y = sys->y[k];
      const double r = sqrt(x * x + y * y);
      double ifrac;   // GetGlobalIFrac (LowTasks.c:21-31)
      if (r < c->GlobalRmed[0]) ifrac = 0.0;
      else if (r > c->GlobalRmed[nr - 1]) ifrac = (double)nr - 1.0;
      else {
        int i = 0;
        while (c->GlobalRmed[i] <= r) i++;
        ifrac = (double)i + (r - c->GlobalRmed[i - 1]) / (c->GlobalRmed[i] - c->GlobalRmed[i - 1]) - 1.0;
      }
      const int ii = (int)ifrac;
      fs->tq_t[k] = c->st.physical_time;
      fs->tq_x[k] = x;
      fs->tq_y[k] = y;
      fs->tq_m[k] = sys->mass[k];
      fs->tq_frac[k] = ifrac - floor(ifrac);
      rr.r[2 * k] = ii;
      rr.r[2 * k + 1] = ii + 1 < nr ? ii + 1 : -1;
    }
  }
  // DampingTW04 (Pframeforce.c:571-603) needs the ring means of rho and cs
  // at the ring of each planet that feels the disk: formed by the same
  // leading block, after the torque log's rings
  int nrings = 2 * ntq, dslot[8];
  for (int k = 0; k < 8; k++) dslot[k] = -1;
  for (int k = 0; k < sys->nb && k < 8; k++) {
    if (!sys->feeldisk[k] || !(p->verticaldamping > 0.0)) continue;
    const double x = sys->x[k], y = sys->y[k];
    const double r = sqrt(x * x + y * y);
    if (r >= c->Rinf[0] && r <= c->Rsup[nr - 1]) {
      int i = 0;
      while (c->Rsup[i] < r) i++;
      dslot[k] = nrings;
      rr.r[nrings++] = i;
    }
  }
  PebF pf;
  pf.prho = d->b[B_PDENS];
  pf.tau = d->b[B_TAU];
  pf.pagr = d->b[B_PAGR];
  pf.pagt = d->b[B_PAGT];
  const dim3 gf(gall.x + (nrings ? 1u : 0u));
#define FORCES_LAUNCH(K)                                                                                  \
  LAUNCH(K, gf, tb, 0, d->g, d->b[B_DENS], d->b[B_CS], d->b[B_AGR], d->b[B_AGT], pks, sys->nb, mt, d->red, rr, \
         nrings, d->ctl, pf, guard)
  // one planet (or none): the per-planet arrays of the kernel shrink to one
  // entry, which halves its registers
  if (peb && sys->nb <= 1) FORCES_LAUNCH((k_forces<true, 1>));
  else if (peb) FORCES_LAUNCH((k_forces<true, 8>));
  else if (sys->nb <= 1) FORCES_LAUNCH((k_forces<false, 1>));
  else FORCES_LAUNCH((k_forces<false, 8>));
#undef FORCES_LAUNCH
  LAUNCH(k_forces_reduce, dim3(2 + 2 * sys->nb), rb, 0, d->red, d->nblk_cells, d->ctl, BOff{0}, guard);
  // the gas-only path adds the indirect term in k_substep1 (the only reader
  // of agr/agt); the pebble fluid's accelerations need it stored
  if (peb)
    LAUNCH(k_indirect, gall, tb, 0, d->g, d->b[B_AGR], d->b[B_AGT], d->ctl, indir_px, indir_py, d->b[B_PAGR],
           d->b[B_PAGT], guard);
  HCHECK(hipGetLastError());
  fs->ntq = ntq;
  for (int k = 0; k < 8; k++) fs->dslot[k] = dslot[k];
  fs->indir_px = indir_px;
  fs->indir_py = indir_py;
  fs->valid = 1;
  return 0;
}

// Heating1 (Planet.c:222-246): the accretion luminosity of planet k spread
// over the cell that holds it, tapered in over 4 HEATINGDELAY DTs
static void heating1(const thorin_ctx *c, HeatK *h, double Mplanet, double dMplanet, double Xplanet, double Yplanet,
                     double dt, int k) {
  const thorin_params *p = &c->p;
  const int nr = c->nr, ns = c->ns;
  const double Rperp = sqrt(Xplanet * Xplanet + Yplanet * Yplanet);
  if (Rperp < c->Rinf[0] || Rperp > c->Rsup[nr - 1]) return;
  const double tdelay = 4.0 * p->dt * (double)p->heatingdelay;
  double taper = c->st.physical_time / tdelay;
  if (taper > 1.0) taper = 1.0;
  const double plrho = p->planetarydensity / TH_RHO2CGS;
  const double plradius = pow(3.0 * Mplanet / (4.0 * TH_PI * plrho), 1.0 / 3.0);
  const double L = Mplanet * dMplanet / dt / plradius;
  double ang = atan2(Yplanet, Xplanet);
  if (ang < 0.0) ang += 2.0 * TH_PI;
  int i = 0;
  while (c->Rsup[i] < Rperp) i++;
  int j = (int)floor((double)ns / 2.0 / TH_PI * ang + 0.5);
  if (j == ns) j = 0;
  h->idx[k] = j + i * ns;
  const double dE = L * taper / c->Surf[i];
  h->q[k] += dE;
}

extern "C" int thorin_algogas(thorin_ctx *c, thorin_planets *sys, thorin_nbody_fn nbody, void *nbody_user,
                              thorin_step_log *log) {
  if (need_dev(c)) return -1;
  if (!sys) { thorin_set_error("thorin_algogas: null planetary system"); return -1; }
  thorin_dev *d = c->dev;
  const thorin_params *p = &c->p;
  const int nr = d->nr, ns = d->ns;
  const double DT = p->dt;
  double dtemp = 0.0, dt = DT;
  int rc = 0;
  if (sys->nb > 8) { thorin_set_error("this build supports at most 8 planets in the force kernel"); return -1; }
  const int peb = p->pebbles;
  if (peb && !c->peb_ready) {
    thorin_set_error("PebbleAccretion YES: the pebble disk must be set up with thorin_init_pebbles first");
    return -1;
  }
  if (log) {   // the per-substep arrays are written as substeps are added (nsub says how many)
    log->nsub = 0;
    log->crashed_any = 0;
    log->lost_mass = 0.0;
    log->crash_dumped = 0;
    log->crash_time = 0.0;
    log->tq_rows = 0;
  }
  d->g.omega_frame = c->st.omega_frame;
  const dim3 tb(TB), rb(RB);
  const dim3 gall = cell_grid(d, nr);
  int gacc = 0;                          // a planet accretes gas (AccreteOntoPlanets)
  for (int k = 0; k < sys->nb; k++) gacc |= sys->acc[k] > 1e-10;
  while (dtemp < 0.999999999 * DT) {
    double mt = c->st.physical_time / (p->masstaper * 2.0 * M_PI);
    c->st.mass_taper = (mt > 1.0 ? 1.0 : pow(sin(mt * M_PI / 2.0), 2.0));
    mt = c->st.mass_taper;
    // ---- ConditionCFL: the per-cell minima of this substep were formed at
    // the end of the previous one (k_eos_cfl) and read back with its sync;
    // after an (re)initialisation or a change of the planets they are
    // formed here.
    const double invdtreb_sq = minstep_rebound(sys);
    if (!d->cfl_ready || invdtreb_sq != d->cfl_invdtreb) {
      if (peb)   // PebbleStokesNumbers + CriticalCharTime (SourceEuler.c:236, 241)
        LAUNCH(k_peb_cfl, dim3(nr), rb, sizeof(double) * 2 * ns, d->g, d->b[B_DENS], d->b[B_PSIZE], d->b[B_TAU],
               d->b[B_PVR], d->b[B_PVT], d->b[B_VRAD], d->b[B_VTHETA], (const int *)nullptr);
      EOS_LAUNCH(k_eos_cfl, dim3((nr + 7) / 8 * 8), rb, eos_lds(d, 0), d->g, d->b[B_DENS], d->b[B_ENERGY],
             d->b[B_TEMPER], d->b[B_CS], d->b[B_VRAD], d->b[B_VTHETA], 0, (const int *)nullptr);
      LAUNCH(k_cfl_final, dim3(1), dim3(CF_T), 0, d->g, d->ctl, (const double *)nullptr, 0, invdtreb_sq,
             (const int *)nullptr, (DevCtl *)nullptr, (unsigned *)nullptr, 0u);
      HCHECK(hipGetLastError());
      DCOPY(d->hctl, d->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost);
      DSYNC();
      d->cfl_ready = 1;
      d->cfl_invdtreb = invdtreb_sq;
    }
    double newdt = d->hctl->newdt_raw, deltaT = DT - dtemp;
    if (newdt > c->st.dt_stellar) {
      double a = 1.0 * c->st.dt_stellar, b = p->minimumstep;
      double mx = a > b ? a : b;
      newdt = newdt < mx ? newdt : mx;
    }
    newdt = newdt < p->maximumstep ? newdt : p->maximumstep;
    if (c->st.cfl_count == 0) {
      if (p->firststep > 0.0) newdt = p->firststep;
      c->st.cfl_count = 1;
    }
    if (deltaT < newdt) newdt = deltaT;
    if (!(newdt > 0.0) || !(deltaT / newdt <= (double)THORIN_MAX_CFL_N)) {
      thorin_set_errorf("ConditionCFL gave dt = %g for a remaining step of %g: the step count is not finite or "
                       "exceeds %d; the substep is refused", newdt, deltaT, THORIN_MAX_CFL_N);
      return -1;
    }
    const int N = (int)(ceil(deltaT / newdt));
    const double t_sub = c->st.physical_time;
    dt = (DT - dtemp) / (double)N;
    // ---- InitHeating (SourceEuler.c:246, Planet.c:251-258)
    HeatK heat;
    memset(&heat, 0, sizeof heat);
    for (int k = 0; k < 8; k++) heat.idx[k] = -1;
    // ---- AccreteOntoPlanets (SourceEuler.c:247, Planet.c:33-163): the
    // planet side with the reference's expressions, the cells on the device
    if (gacc) {
      GasAccK P;
      memset(&P, 0, sizeof P);
      int idx[8];
      for (int k = 0; k < sys->nb; k++) {
        if (!(sys->acc[k] > 1e-10)) continue;
        const int q = P.np++;
        idx[q] = k;
        const double facc = fabs(dt) * (sys->acc[k]);
        double facc1 = 1.0 / 3.0 * facc, facc2 = 2.0 / 3.0 * facc;
        facc1 = facc1 < 1.0 ? facc1 : 1.0;
        facc2 = facc2 < 1.0 ? facc2 : 1.0;
        const double X = sys->x[k], Y = sys->y[k], Z = sys->z[k], M = sys->mass[k];
        const double Rperp = sqrt(X * X + Y * Y);
        const double Rplanet = sqrt(X * X + Y * Y + Z * Z);
        const double RRoche = pow((1.0 / 3.0 * M * c->st.mass_taper), (1.0 / 3.0)) * Rplanet;
        int i_min = 0, i_max = nr - 1;
        while (i_min < nr && c->Rsup[i_min] < Rperp - RRoche) i_min++;
        while ((c->Rinf[i_max] > Rperp + RRoche) && (i_max > 0)) i_max--;
        const double angle = atan2(Y, X);
        P.jmin[q] = (int)((double)ns / 2.0 / TH_PI * (angle - 2.0 * RRoche / Rperp));
        P.jmax[q] = (int)((double)ns / 2.0 / TH_PI * (angle + 2.0 * RRoche / Rperp));
        if (P.jmax[q] - P.jmin[q] + 1 > ns) {
          thorin_set_error("gas accretion: the planet's Roche window spans more than a ring");
          return -1;
        }
        P.imin[q] = i_min;
        P.imax[q] = i_max < nr - 1 ? i_max : nr - 1;
        P.x[q] = X; P.y[q] = Y; P.z[q] = Z;
        P.rroche[q] = RRoche;
        P.facc1[q] = facc1;
        P.facc2[q] = facc2;
      }
      LAUNCH(k_gas_accrete, dim3(1), dim3(GACC_T), 0, d->g, P, d->b[B_DENS], (const double *)d->b[B_VRAD],
             (const double *)d->b[B_VTHETA], (const double *)d->b[B_CS], d->ctl);
      HCHECK(hipGetLastError());
      DCOPY(d->hctl, d->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost);
      DSYNC();
      for (int q = 0; q < P.np; q++) {
        const int k = idx[q];
        double M = sys->mass[k];
        if (p->gas_accret_heating && k < 8)
          heating1(c, &heat, M, d->hctl->gacc[q][0], sys->x[k], sys->y[k], dt, k);
        double PX = M * sys->vx[k], PY = M * sys->vy[k];
        PX += d->hctl->gacc[q][1];
        PY += d->hctl->gacc[q][2];
        M += d->hctl->gacc[q][0];
        if (sys->feeldisk[k]) {
          sys->vx[k] = PX / M;
          sys->vy[k] = PY / M;
        }
        sys->mass[k] = M;
      }
    }
    // ---- AccretePebblesOntoPlanets, EvaporatePebbles (SourceEuler.c:249-250)
    if (peb && p->accret_heating) d->heat_max = sys->nb;   // Pebbles.c:379
    if (peb && sys->nb > 0) {
      PebAccK P;
      memset(&P, 0, sizeof P);
      if (peb_accrete_host(c, sys, &P)) return -1;
      LAUNCH(k_peb_accrete, dim3(1), dim3(ACC_T), 0, d->g, P, d->b[B_PDENS], d->b[B_PVR], d->b[B_PVT], d->b[B_TAU],
             d->b[B_CS], d->b[B_MACC], dt, d->ctl);
      HCHECK(hipGetLastError());
      DCOPY(d->hctl, d->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost);
      DSYNC();
      if (d->hctl->acc_fail) {
        thorin_set_error("pebble accretion: integration of the column mass did not converge");
        return -1;
      }
      // STEP 7 (Pebbles.c:529-537)
      for (int k = 0; k < sys->nb; k++) {
        double M = sys->mass[k];
        double PX = sys->vx[k] * M, PY = sys->vy[k] * M;
        const double dM = d->hctl->pacc[k][0];
        // OPTIONAL STEP 6 (Pebbles.c:525-528)
        if (p->accret_heating && k < 8) heating1(c, &heat, M, dM, sys->x[k], sys->y[k], dt, k);
        PX += d->hctl->pacc[k][1];
        PY += d->hctl->pacc[k][2];
        M += dM;
        sys->mass[k] = M;
        sys->solid[k] = sys->solid[k] + dM;
        sys->vx[k] = PX / M;
        sys->vy[k] = PY / M;
      }
    }
    if (peb && p->evaporationrate > 0.0) {
      double facc = p->evaporationrate * dt;
      if (facc > 1.0) facc = 1.0;
      LAUNCH(k_peb_evaporate, gall, tb, 0, d->g, d->b[B_PDENS], d->b[B_TEMPER], facc,
             p->evaporationtemperature / c->k.t2si);
    }
    dtemp += dt;
    // ---- FillForcesArrays (the planets' constants go to the kernel by value)
    const int want_tq = p->write_torque && log && log->nsub == 0 && sys->nb > 0;
    ForceSet fset;
    if (!peb && !gacc && d->spec.valid && force_set_matches(d->spec, sys, mt, want_tq, c->st.physical_time)) {
      fset = d->spec;           // enqueued at the end of the previous substep
    } else {
      if (launch_forces(c, sys, mt, want_tq, nullptr, &fset)) return -1;
    }
    d->spec.valid = 0;
    const ForceSet *fs = &fset;
    const int ntq = fs->ntq;
    // ---- AdvanceSystemFromDisk (host; needs the reductions only when a planet feels the disk)
    int anyfeel = 0;
    for (int k = 0; k < sys->nb; k++) anyfeel |= sys->feeldisk[k];
    if (anyfeel) {
      DCOPY(d->hctl, d->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost);
      DSYNC();
      for (int k = 0; k < sys->nb; k++) { sys->ax[k] = d->hctl->axpl[k]; sys->ay[k] = d->hctl->aypl[k]; }
    }
    if (anyfeel) {
      const double indx = -d->hctl->axstar, indy = -d->hctl->aystar;
      for (int k = 0; k < sys->nb; k++) {
        if (!sys->feeldisk[k]) continue;
        // DampingTW04 (Pframeforce.c:571-603), with the ring means of the
        // planet's ring from the force kernel's leading block
        double damp = 0.0, x = sys->x[k], y = sys->y[k], z = sys->z[k], m = sys->mass[k], vz = sys->vz[k];
        double r = sqrt(x * x + y * y);
        if (k < 8 && fs->dslot[k] >= 0) {
          const double rmean = d->hctl->tq_rho[fs->dslot[k]], csmean = d->hctl->tq_cs[fs->dslot[k]];
          double cs4inv = pow(csmean, -4.0), Om = pow(r, -1.5);
          damp = p->verticaldamping * m * rmean * Om * cs4inv * (-2.176 * vz - 0.871 * z * Om);
        }
        sys->vx[k] += dt * sys->ax[k];
        sys->vy[k] += dt * sys->ay[k];
        sys->vz[k] += dt * (sys->az[k] + damp);
        sys->vx[k] += dt * indx;
        sys->vy[k] += dt * indy;
      }
    }
    // ---- host N-body (AdvanceSystemRebound, RotatePsys, SynchronizeFargoRebound,
    // SourceEuler.c:260-269) comes next in the reference.  No kernel from
    // here to the end of the substep reads the planets (the forces were
    // formed above; in the fixed frame nothing else is), so the substep's
    // kernels are enqueued first and the host steps the planets while they
    // run; k_cfl_final and the next substep's forces, which use the new
    // planets, are enqueued after it.
    const double t_nbody = c->st.physical_time;
    int nbody_done = 0;
    auto step_planets = [&]() -> int {
      if (nbody_done || !nbody) return 0;
      nbody_done = 1;
      int e = nbody(nbody_user, sys, t_nbody, dt, c->st.omega_frame * dt);
      if (e) {
        char m[600];
        snprintf(m, sizeof m, "N-body step failed (%d): %s", e, thorin_last_error());
        thorin_set_error(m);
        return -1;
      }
      return 0;
    };
    // ---- ApplyBoundaryCondition, DetectCrash, ComputePressureField, SubStep1,
    // ViscousTerms, SubStep2, ActualiseGas, ApplyBoundaryCondition.  The gas
    // alone with a damping boundary on vrad only runs them as one kernel
    // (k_visc_chain), which leaves the crash snapshot to SubStep3.
    const int kepler = (!p->damping_on) || (p->damping_on && p->damp_vrad);
    double vkin = 0.0, vkout = 0.0;
    if (kepler) {
      // ImposeKeplerianEdges (Viscosity.c:236-243), host libm
      vkin = sqrt(1.0 * 1.0 / c->Rmed[0]) *
             sqrt(1.0 - pow(thorin_aspectratio_host(c, c->Rmed[0]), 2.0) * pow(c->Rmed[0], 2.0 * p->flaringindex) *
                            (1. + p->sigmaslope - 2.0 * p->flaringindex));
      vkout = sqrt(1.0 * 1.0 / c->Rmed[nr - 1]) *
              sqrt(1.0 - pow(thorin_aspectratio_host(c, c->Rmed[nr - 1]), 2.0) *
                             pow(c->Rmed[nr - 1], 2.0 * p->flaringindex) * (1. + p->sigmaslope - 2.0 * p->flaringindex));
    }
    const bool fuse_chain = !peb && !(p->damping_on && (p->damp_init || p->damp_flux)) && d->visc_chain;
    CrashSnap snap;
    memset(&snap, 0, sizeof snap);
    CrashSnap snap3;              // taken by SubStep3 after the fused kernel
    memset(&snap3, 0, sizeof snap3);
    if (fuse_chain) {
      ViscChain vc;
      vc.rho = d->b[B_DENS];
      vc.energy = d->b[B_ENERGY];
      vc.vrad = d->b[B_VRAD];
      vc.vtheta = d->b[B_VTHETA];
      vc.agr = d->b[B_AGR];
      vc.agt = d->b[B_AGT];
      vc.vradint = d->b[B_VRADINT];
      vc.vradnew = d->b[B_VRADNEW];
      vc.vthetanew = d->b[B_VTHETANEW];
      vc.energyint = d->b[B_EINT];
      vc.ctl = d->ctl;
      vc.dt = dt;
      vc.damp = p->damping_on;
      vc.fold_ind = 1;
      vc.ind_px = fs->indir_px;
      vc.ind_py = fs->indir_py;
      vc.kepler = kepler;
      vc.vkin = vkin;
      vc.vkout = vkout;
      const int ntile = ((ns + VC_TJ - 1) / VC_TJ) * ((nr + VC_TI - 1) / VC_TI);
      LAUNCH(k_visc_chain, dim3((unsigned)((ntile + 7) / 8 * 8)), dim3(VC_T), 0, d->g, vc);
      if (c->st.already_crashed == 0) {
        // after the swaps below the substep's old vrad and vtheta are in B_VRADNEW and B_VTHETANEW
        snap3.ctl = d->ctl;
        snap3.damp = p->damping_on;
        snap3.step = dt;
        const int sb[4] = {B_DENS, B_VRAD, B_VTHETA, B_TEMPER}, db[4] = {B_CRASH_DENS, B_CRASH_VRAD, B_CRASH_VTHETA,
                                                                         B_CRASH_TEMPER};
        for (int k = 0; k < 4; k++) { snap3.src[k] = d->b[sb[k]]; snap3.dst[k] = d->b[db[k]]; }
      }
      swapb(d, B_VRAD, B_VRADNEW);
      swapb(d, B_VTHETA, B_VTHETANEW);
    } else {
      if (p->damping_on)
        LAUNCH(k_boundary, bnd_grid(d), tb, 0, d->g, d->b[B_VRAD], d->b[B_VTHETA], d->b[B_DENS], d->b[B_ENERGY], dt,
               (const int *)nullptr, d->damp_nlo, d->damp_hi0);
      if (peb) {
        // DetectCrashPebbles, then SourceTermsPebbles (SourceEuler.c:274, 291)
        LAUNCH(k_peb_damp, gall, tb, 0, d->g, d->b[B_PDENS], d->b[B_PVR], d->b[B_PVT], dt, 1, d->ctl,
               (const int *)nullptr);
        LAUNCH(k_peb_source, gall, tb, 0, d->g, d->b[B_PDENS], d->b[B_PVR], d->b[B_PVT], d->b[B_TAU], d->b[B_VRAD],
               d->b[B_VTHETA], d->b[B_PAGR], d->b[B_PAGT], d->b[B_PACCR], d->b[B_PACCT], d->b[B_DRAGR], d->b[B_DRAGT]);
      }
      const int br = peb && p->back_reaction;
      LAUNCH(k_substep1, gall, tb, 0, d->g, d->b[B_DENS], d->b[B_ENERGY], d->b[B_VRAD], d->b[B_VTHETA], d->b[B_AGR],
             d->b[B_AGT], d->b[B_VRADINT], d->b[B_VTHETAINT], dt, d->ctl, peb ? d->b[B_GACCR] : (double *)nullptr,
             peb ? d->b[B_GACCT] : (double *)nullptr, br ? d->b[B_DRAGR] : (const double *)nullptr,
             br ? d->b[B_DRAGT] : (const double *)nullptr, peb ? 0 : 1, fs->indir_px, fs->indir_py);
      if (peb)   // SubStep1Pebbles + ParticleDiffusion (SourceEuler.c:293-296)
        LAUNCH(k_peb_substep1, gall, tb, 0, d->g, d->b[B_PVR], d->b[B_PVT], d->b[B_TAU], d->b[B_PACCR], d->b[B_PACCT],
               d->b[B_GACCR], d->b[B_GACCT], d->b[B_VRAD], d->b[B_VTHETA], d->b[B_DENS], d->b[B_PDENS], dt,
               p->diffusive_particles);
      if (c->st.already_crashed == 0) {
        snap.ctl = d->ctl;
        snap.peb = peb;
        const int sb[4] = {B_DENS, B_VRAD, B_VTHETA, B_TEMPER}, db[4] = {B_CRASH_DENS, B_CRASH_VRAD, B_CRASH_VTHETA,
                                                                         B_CRASH_TEMPER};
        for (int k = 0; k < 4; k++) { snap.src[k] = d->b[sb[k]]; snap.dst[k] = d->b[db[k]]; }
      }
      LAUNCH(k_divstress, gall, tb, 0, d->g, d->b[B_VRADINT], d->b[B_VTHETAINT], d->b[B_DENS], (double *)nullptr,
             d->b[B_TAURR], d->b[B_TAURP], d->b[B_TAUPP], snap);
      LAUNCH(k_visc_update, gall, tb, 0, d->g, d->b[B_VRADINT], d->b[B_VTHETAINT], d->b[B_DENS], d->b[B_TAURR],
             d->b[B_TAURP], d->b[B_TAUPP], dt, kepler, vkin, vkout);
      // ---- SubStep2, ActualiseGas, ApplyBoundaryCondition
      LAUNCH(k_substep2, gall, tb, 0, d->g, d->b[B_DENS], d->b[B_VRADINT], d->b[B_VTHETAINT], d->b[B_ENERGY],
             d->b[B_VRADNEW], d->b[B_VTHETANEW], d->b[B_EINT], dt);
      swapb(d, B_VRAD, B_VRADNEW);
      swapb(d, B_VTHETA, B_VTHETANEW);
      if (p->damping_on)
        LAUNCH(k_boundary, bnd_grid(d), tb, 0, d->g, d->b[B_VRAD], d->b[B_VTHETA], d->b[B_DENS], d->b[B_ENERGY], dt,
               (const int *)nullptr, d->damp_nlo, d->damp_hi0);
    }
    HCHECK(hipGetLastError());
    // ---- SubStep3: Q+, implicit radiative diffusion
    int niter = 0;
    double omega_used = 0.0;
    {
      const int nsum = (nr + TB / 64 - 1) / (TB / 64);   // ring-mean blocks, one ring per wave
      const int nsum8 = (nsum + 7) / 8 * 8;
      SUB3_LAUNCH(dim3(gall.x + (unsigned)nsum8), tb, 0, d->g, d->b[B_VRAD], d->b[B_VTHETA], d->b[B_DENS],
                  d->b[B_EINT], d->b[B_DIVV], (double *)nullptr, (double *)nullptr, (double *)nullptr, d->b[B_QPLUS],
                  d->b[B_TEMPER], d->b[B_CS], d->b[B_OPACITY], d->b[B_QMINUS], d->b[B_DIFF], d->pstellar,
                  d->opa_filled ? 0 : 1, nsum8, snap3);
    }
    d->opa_filled = 1;
    {
      const dim3 gm = cell_grid(d, nr - 2);
      heat.n = d->heat_max < 8 ? d->heat_max : 8;       // heatsrc_max as the matrix reads it
      // the energy system (EnergySources.c:256-297): the SOR launches form
      // its rows from SubStep3's fields (no matrix pass), or (sor_assembly 0)
      // the matrix pass assembles it
      d->sys_dt = dt;
      d->sys_heat = heat;
      if (d->sor_assembly) {
        d->sys_form = 1;
      } else {
        LAUNCH(k_matrix, gm, tb, 0, d->g, d->b[B_DENS], d->b[B_CS], d->b[B_DIFF], d->b[B_DIVV], d->b[B_QMINUS],
               d->b[B_QPLUS], d->b[B_TEMPER], d->b[B_ML], d->b[B_MLIP], d->b[B_MLIM], d->b[B_MLJP], d->b[B_MLJM],
               d->b[B_RHS], dt, d->mresp, (const double *)d->b[B_OPACITY], heat);
        d->sys_form = 0;
      }
    }
    HCHECK(hipGetLastError());
    if (c->st.sor_first) {
      c->st.sor_first = 0;
      if (iterate_relaxation(c)) return -1;
      c->st.sor_omega = c->st.sor_omega_best;
      c->st.sor_niter_last = c->st.sor_niter_best;
    }
    c->st.sor_omega += c->st.sor_domega;
    if (c->st.sor_omega > OMEGAUPPER) { c->st.sor_omega = OMEGAUPPER; c->st.sor_domega = -c->st.sor_domega; }
    if (c->st.sor_omega < 1.0) { c->st.sor_omega = 1.0; c->st.sor_domega = -c->st.sor_domega; }
    omega_used = c->st.sor_omega;
    // The iterations are enqueued without a host round trip: one more than
    // the longest of the last four solves took (at least 2), each ending in
    // the device-side stop test (iterations after convergence do nothing).  Everything after the solve
    // is enqueued behind them and does nothing unless the solve converged
    // (guard = ctl->sor_done); in the rare case it needed more iterations the
    // solve is finished with host checks and the rest of the substep is
    // enqueued again from the same buffers.
    int sor_k = c->st.sor_niter_last;
    for (int k = 0; k < 4; k++) sor_k = d->niter_hist[k] > sor_k ? d->niter_hist[k] : sor_k;
    sor_k += 1;
    if (sor_k < 2) sor_k = 2;
    if (sor_k < d->sor_queue_min) sor_k = d->sor_queue_min;   // an ensemble queues alike for all disks
    if (sor_k > SORMAXITERS + 2) sor_k = SORMAXITERS + 2;
    if (sor_launch(c, c->st.sor_omega, 0, sor_k)) return -1;
    c->st.physical_time += dt;
    if (p->lambdadoubling != 0.0) {
      thorin_cavity_factors(c);
      const int n1 = nr + 3;
      DCOPY((double *)d->g.cav_med, c->cav_med, sizeof(double) * n1, hipMemcpyHostToDevice);
      DCOPY((double *)d->g.cav_inf, c->cav_inf, sizeof(double) * n1, hipMemcpyHostToDevice);
    }
    const int *guard = &d->ctl->sor_done;
    // a gas-only substep with no damping or ZEROVRAD damping forms the
    // velocities and their damping in its end-of-substep kernel
    const bool fuse_vel = !peb && (!p->damping_on || (p->damp_vrad && !p->damp_init && !p->damp_flux));
    double *bsave[NBUF];
    memcpy(bsave, d->b, sizeof bsave);
    auto after_sor = [&]() -> int {
      // ---- EnergyNew = rho T CV from the SOR's last iterate, formed where the
      // radial transport reads the energy; Transport
      {
        RadQ q = {};
        q.rho = d->b[B_DENS];
        q.vr = d->b[B_VRAD];
        q.vt = d->b[B_VTHETA];
        q.e = nullptr;
        q.tsrc[0] = d->b[B_TEMPER];
        q.tsrc[1] = d->b[B_TEMPER2];
        q.tsrc[2] = d->b[B_TEMPER3];
        q.tsrc[3] = d->b[B_TEMPER4];
        q.ctl = d->ctl;
        q.cv = d->g.cv;
        const int outs[RC_NQ] = {B_XRP, B_XRM, B_XTP, B_XTM, B_XE, B_XRHO};
        for (int k = 0; k < RC_NQ; k++) q.out[k] = d->b[outs[k]];
        q.guard = guard;
        q.of = c->st.omega_frame;
        q.ntile_j = (ns + RT_J - 1) / RT_J;
        q.rows = d->rad_rows;
        q.ntiles = q.ntile_j * ((nr + q.rows - 1) / q.rows);
        q.nq = RC_NQ;
        if (transport(d, d->g, q, B_DENS, B_VRAD, B_VTHETA, B_ENERGY, dt, guard, !fuse_vel)) return -1;
      }
      if (p->damping_on && !fuse_vel)
        LAUNCH(k_boundary, bnd_grid(d), tb, 0, d->g, d->b[B_VRAD], d->b[B_VTHETA], d->b[B_DENS], d->b[B_ENERGY], dt,
               guard, d->damp_nlo, d->damp_hi0);
      if (peb) {
        // ---- EvolvePebbleDisk (SourceEuler.c:312, Pebbles.c:699-708): damping,
        // TransportPebbles (TransportEuler.c:449-505: the gas scheme without
        // energy; ComputeResiduals with the pebble vtheta's ring mean, whose
        // Nshift overwrites the gas one as in the reference), damping; then
        // the next substep's Stokes numbers and pebble CFL limit
        LAUNCH(k_peb_damp, gall, tb, 0, d->g, d->b[B_PDENS], d->b[B_PVR], d->b[B_PVT], dt, 0, d->ctl, guard);
        LAUNCH(k_ring_means, dim3((unsigned)((nr + d->rs_rows - 1) / d->rs_rows)), dim3(256),
               sizeof(double) * (size_t)d->rs_rows * ns, (const double *)d->b[B_PVT], nr, ns, d->rs_rows,
               d->g.pvmed, guard, BOff{0});
        Geo gp = d->g;
        gp.VMed = d->g.pvmed;
        RadQ q = {};
        q.rho = d->b[B_PDENS];
        q.vr = d->b[B_PVR];
        q.vt = d->b[B_PVT];
        q.e = nullptr;
        const int outs[RC_NQ] = {B_XRP, B_XRM, B_XTP, B_XTM, B_XE, B_XRHO};
        for (int k = 0; k < RC_NQ; k++) q.out[k] = d->b[outs[k]];
        q.guard = guard;
        q.of = c->st.omega_frame;
        q.ntile_j = (ns + RT_J - 1) / RT_J;
        q.rows = d->rad_rows;
        q.ntiles = q.ntile_j * ((nr + q.rows - 1) / q.rows);
        q.nq = RC_NQ - 1;
        if (transport(d, gp, q, B_PDENS, B_PVR, B_PVT, -1, dt, guard)) return -1;
        LAUNCH(k_peb_damp, gall, tb, 0, d->g, d->b[B_PDENS], d->b[B_PVR], d->b[B_PVT], dt, 0, d->ctl, guard);
        LAUNCH(k_peb_cfl, dim3(nr), rb, sizeof(double) * 2 * ns, d->g, d->b[B_DENS], d->b[B_PSIZE], d->b[B_TAU],
               d->b[B_PVR], d->b[B_PVT], d->b[B_VRAD], d->b[B_VTHETA], guard);
      }
      // ---- end of the substep: T, cs and their ring means, and the CFL
      // minima of the next substep (its planets are those of now)
      if (fuse_vel) {
        const VelSrc vs = {d->b[B_RP], d->b[B_RM], d->b[B_TP], d->b[B_TM], dt, p->damping_on};
        EOS_LAUNCH(k_eos_cfl_vel, dim3((nr + 7) / 8 * 8), rb, eos_lds(d, 1), d->g, d->b[B_DENS], d->b[B_ENERGY], d->b[B_TEMPER],
               d->b[B_CS], d->b[B_VRAD], d->b[B_VTHETA], guard, vs);
      } else {
        EOS_LAUNCH(k_eos_cfl, dim3((nr + 7) / 8 * 8), rb, eos_lds(d, 1), d->g, d->b[B_DENS], d->b[B_ENERGY],
               d->b[B_TEMPER], d->b[B_CS], d->b[B_VRAD], d->b[B_VTHETA], 1, guard);
      }
      // the planets' step, overlapped with the kernels enqueued so far
      if (step_planets()) return -1;
      d->cfl_invdtreb = minstep_rebound(sys);
      // one sync per substep: SOR result, crash flags, dt_stellar, the next
      // CFL minimum, the planet accelerations, written to the pinned mirror
      // by k_cfl_final.  It does nothing while the SOR has not converged:
      // the mirror then reads sor_done = 0 and the solve is finished from
      // the host.
      d->hctl->sor_done = 0;
      // outside an ensemble the host polls the flag word (dev_wait_mark)
      unsigned *hflag = d->rec ? nullptr : d->hflag;
      const unsigned seq = d->rec ? 0u : ++d->flag_seq;
      LAUNCH(k_cfl_final, dim3(1), dim3(CF_T), 0, d->g, d->ctl, (const double *)d->pstellar, d->nblk_cells,
             d->cfl_invdtreb, guard, d->hctl, hflag, seq);
      HCHECK(hipGetLastError());
      d->spec.valid = 0;
      if (peb || gacc) {   // (the gas accretion changes the density before the forces)
        DSYNC();
        return 0;
      }
      // The next substep's FillForcesArrays is enqueued behind the sync point
      // and runs while the host reads the mirror and steps the planets: its
      // inputs (the gas density and sound speed of now, the planets as they
      // are now, MassTaper at the new time) are those the next substep would
      // use unless the host changes the planets, which the next substep
      // detects by comparing the planet set before it uses these forces.
      if (dev_mark(d)) return -1;
      {
        double mtn = c->st.physical_time / (p->masstaper * 2.0 * M_PI);
        mtn = (mtn > 1.0 ? 1.0 : pow(sin(mtn * M_PI / 2.0), 2.0));
        const int tqn = !(dtemp < 0.999999999 * DT) && p->write_torque && sys->nb > 0;
        if (launch_forces(c, sys, mtn, tqn, guard, &d->spec)) return -1;
      }
      if (dev_wait_mark(d)) return -1;
      if (!d->hctl->sor_done) d->spec.valid = 0;   // the guarded launches did nothing
      return 0;
    };
    if (after_sor()) return -1;
    if (d->hctl->sor_done) {
      niter = d->hctl->sor_niter;
    } else {
      if (sor_finish(c, c->st.sor_omega, sor_k, &niter)) return -1;
      memcpy(d->b, bsave, sizeof bsave);
      if (after_sor()) return -1;
    }
    if (niter > SORMAXITERS) rc = 2;
    if (niter > c->st.sor_niter_last) c->st.sor_domega = -c->st.sor_domega;
    c->st.sor_niter_last = niter;
    d->niter_hist[d->nsolve++ & 3] = niter;
    if (log && log->nsub < THORIN_MAXSUB) {
      log->cfl_n[log->nsub] = N;
      log->dt[log->nsub] = dt;
      log->sor_niter[log->nsub] = niter;
      log->sor_omega[log->nsub] = omega_used;
      log->crashed[log->nsub] = 0;
      log->nsub++;
    }
    d->cfl_ready = 1;
    c->st.dt_stellar = d->hctl->dt_stellar;
    for (int k = 0; k < sys->nb; k++) { sys->ax[k] = d->hctl->axpl[k]; sys->ay[k] = d->hctl->aypl[k]; }
    if (ntq) {
      log->tq_rows = ntq;
      for (int k = 0; k < ntq; k++) {
        const double x = fs->tq_x[k], y = fs->tq_y[k], m = fs->tq_m[k], frac = fs->tq_frac[k];
        const double ax = d->hctl->axpl[k], ay = d->hctl->aypl[k];
        const double r2 = x * x + y * y, r = sqrt(r2);
        const double rhopl = d->hctl->tq_rho[2 * k] * (1.0 - frac) + d->hctl->tq_rho[2 * k + 1] * frac;
        const double cspl = d->hctl->tq_cs[2 * k] * (1.0 - frac) + d->hctl->tq_cs[2 * k + 1] * frac;
        const double omega = pow(r, -1.5);
        const double h = cspl / (omega * r) * c->k.sqrt_adiabind_inv;
        double norm = m * omega * r2 / h;
        norm *= norm;
        norm *= rhopl;
        const double tq = x * ay - y * ax;
        const double row[7] = {fs->tq_t[k], tq, tq * m * p->adiabind / norm, x, y, ax, ay};
        for (int q = 0; q < 7; q++) log->tq[k][q] = row[q];
      }
    }
    if (prof_flush(d)) return -1;
    // DetectCrash (SourceEuler.c:272-274): the last test of the substep decides
    if (peb ? d->hctl->crash_peb : d->hctl->crash_e) {
      if (c->st.already_crashed == 0 && log) {   // k_divstress took the snapshot
        log->crash_dumped = 1;
        log->crash_time = t_sub;
      }
      c->st.already_crashed++;
      if (log) { log->crashed[log->nsub - 1] = 1; log->crashed_any = 1; }
      if (!rc) rc = 1;
    }
  }
  return rc;
}

// ===========================================================================
// Ensemble of disks advanced together (see the recorder above).

static int rec_wait(Rec *r) {
  thorin_ensemble *e = r->e;
  std::unique_lock<std::mutex> lk(e->m);
  const unsigned long g = e->gen;
  e->nwait++;
  e->cv_coord.notify_one();
  e->cv_disk.wait(lk, [&] { return e->gen != g; });
  return e->err ? -1 : 0;
}

// Merge the recorded queues into one stream.  At every step the copies at the
// heads run first; then, among the kernels at the heads, one that no other
// disk still has further down its queue is launched for every disk whose head
// it is (the largest such group); if there is none, the largest group.  With
// the disks' sequences alike this launches each kernel once for all of them,
// and a disk whose solve needs more SOR iterations runs those alone before
// the rest of the substep is batched again.
static int ens_flush(thorin_ensemble *e) {
  const int n = e->n;
  e->marked = 0;
  // disks with work in this flush: the flush may end at the markers' event
  // only when each of them ends its part with a marker (else a disk waits
  // for work, such as a copy of its control block, queued after the event)
  int active = 0;
  for (int k = 0; k < n; k++) active += e->recs[k].q.empty() ? 0 : 1;
  std::vector<size_t> pos((size_t)n, 0);
  std::vector<std::unordered_map<const void *, int>> left((size_t)n);
  for (int k = 0; k < n; k++)
    for (auto &op : e->recs[k].q)
      if (op.type == 0) left[k][op.key]++;
  std::vector<char> tables;
  std::vector<IssueFn> issue;
  std::vector<unsigned long long> sig;
  std::vector<RecOp *> grp;
  long long nops = 0;
  for (;;) {
    for (int k = 0; k < n; k++) {
      auto &q = e->recs[k].q;
      while (pos[k] < q.size() && q[pos[k]].type == 1) {
        const RecOp &op = q[pos[k]++];
        void *dst = op.dst;
        const void *src = op.src;
        const size_t bytes = op.bytes;
        const hipMemcpyKind kind = op.kind;
        const unsigned long long w[] = {1ull, (unsigned long long)(uintptr_t)dst, (unsigned long long)(uintptr_t)src,
                                        (unsigned long long)bytes, (unsigned long long)kind};
        sig.insert(sig.end(), w, w + 5);
        issue.push_back([=](hipStream_t s, const char *) -> int {
          HCHECK(hipMemcpyAsync(dst, src, bytes, kind, s));
          return 0;
        });
      }
    }
    const void *best = nullptr;
    int best_n = 0, best_ready = 0;
    for (int k = 0; k < n; k++) {
      auto &q = e->recs[k].q;
      if (pos[k] >= q.size()) continue;
      const void *key = q[pos[k]].key;
      int cnt = 0, ready = 1;
      for (int m = 0; m < n; m++) {
        auto &qm = e->recs[m].q;
        if (pos[m] < qm.size() && qm[pos[m]].key == key) cnt++;
        else {
          auto it = left[m].find(key);
          if (it != left[m].end() && it->second > 0) ready = 0;
        }
      }
      if ((ready > best_ready) || (ready == best_ready && cnt > best_n)) {
        best = key;
        best_n = cnt;
        best_ready = ready;
      }
    }
    if (!best) break;
    grp.clear();
    for (int k = 0; k < n; k++) {
      auto &q = e->recs[k].q;
      if (pos[k] < q.size() && q[pos[k]].key == best) {
        grp.push_back(&q[pos[k]++]);
        left[k][best]--;
      }
    }
    if (best == (const void *)&ens_marker_key) {
      // an event record (the last one recorded follows every disk's work
      // before its marker); such a flush is not captured as a graph
      e->marked += (int)grp.size();
      sig.push_back(~0ull);
      issue.push_back([e](hipStream_t s, const char *) -> int {
        HCHECK(hipEventRecord(e->mark, s));
        return 0;
      });
    } else {
      issue.push_back(grp[0]->plan(grp.data(), (int)grp.size(), tables, sig));
      e->launches++;
      nops += (long long)grp.size();
    }
  }
  e->batched_ops += nops;
  if (tables.size() > e->st.cap) {
    if (e->st.h) HCHECK(hipHostFree(e->st.h));
    if (e->st.dv) HCHECK(hipFree(e->st.dv));
    e->st.h = nullptr;
    e->st.dv = nullptr;
    size_t cap = 1 << 20;
    while (cap < tables.size()) cap *= 2;
    HCHECK(hipHostMalloc((void **)&e->st.h, cap, hipHostMallocDefault));
    HCHECK(hipMalloc((void **)&e->st.dv, cap));
    e->st.cap = cap;
  }
  if (!tables.empty()) memcpy(e->st.h, tables.data(), tables.size());
  auto enqueue = [&]() -> int {
    if (!tables.empty()) HCHECK(hipMemcpyAsync(e->st.dv, e->st.h, tables.size(), hipMemcpyHostToDevice, e->s));
    for (auto &f : issue)
      if (f(e->s, e->st.dv)) return -1;
    return 0;
  };
  const bool replayable = e->use_graphs && !issue.empty() &&
                          std::find(sig.begin(), sig.end(), ~0ull) == sig.end();
  if (!replayable) {
    if (enqueue()) return -1;
  } else {
    // the flush as a graph: the tables' upload, the merged launches and the
    // copies; replayed while the signature repeats (the usual substep)
    sig.push_back((unsigned long long)tables.size());
    sig.push_back((unsigned long long)(uintptr_t)e->st.dv);
    sig.push_back((unsigned long long)(uintptr_t)e->st.h);
    EnsGraph *G = nullptr;
    for (auto &x : e->graphs)
      if (x.sig == sig) G = &x;
    // a flush is captured the second time its signature comes: one-off
    // flushes (the first DT's over-relaxation scan) run as they are
    bool seen = G != nullptr;
    for (size_t k = 0; !seen && k < e->seen.size(); k++) seen = e->seen[k] == sig;
    if (!seen) {
      if (e->seen.size() >= 16) e->seen.erase(e->seen.begin());
      e->seen.push_back(sig);
      if (enqueue()) return -1;
    } else if (G) {
      e->replays++;
    } else {
      if (e->graphs.size() >= 8) {    // evict the least recently used
        size_t v = 0;
        for (size_t k = 1; k < e->graphs.size(); k++)
          if (e->graphs[k].used < e->graphs[v].used) v = k;
        (void)hipGraphExecDestroy(e->graphs[v].x);
        (void)hipGraphDestroy(e->graphs[v].g);
        e->graphs.erase(e->graphs.begin() + (long)v);
      }
      EnsGraph x;
      x.sig = sig;
      HCHECK(hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
      const int r = enqueue();
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(e->s, &g);
      if (r) {
        if (g) (void)hipGraphDestroy(g);
        return -1;
      }
      HCHECK(ce);
      x.g = g;
      const hipError_t ie = hipGraphInstantiate(&x.x, g, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        HCHECK(ie);
      }
      e->graphs.push_back(std::move(x));
      G = &e->graphs.back();
      e->captures++;
    }
    if (G) {
      G->used = ++e->tick;
      HCHECK(hipGraphLaunch(G->x, e->s));
    }
  }
  // every disk's part ending in its end-of-substep marker: the disks wait
  // for the last marker (the next substeps' forces queued behind it keep
  // running); else for everything issued
  if (e->marked > 0 && e->marked == active) HCHECK(hipEventSynchronize(e->mark));
  else HCHECK(hipStreamSynchronize(e->s));
  for (int k = 0; k < n; k++) e->recs[k].q.clear();
  return 0;
}


// disk k's host thread: one thorin_algogas per job; its errors (kept per
// thread) are copied out under the lock
static void ens_worker(thorin_ensemble *e, int k) {
  unsigned long seen = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(e->m);
      e->cv_job.wait(lk, [&] { return e->quit || e->job != seen; });
      if (e->quit) return;
      seen = e->job;
    }
    const int r = thorin_algogas(e->c[k], e->j_sys[k], e->j_nbody, e->j_user ? e->j_user[k] : nullptr,
                                 e->j_logs ? &e->j_logs[k] : nullptr);
    std::lock_guard<std::mutex> lk(e->m);
    e->j_rc[k] = r;
    if (r < 0 && e->j_err.empty()) e->j_err = "disk " + std::to_string(k) + ": " + thorin_last_error();
    e->nfinished++;
    e->cv_coord.notify_one();
  }
}

extern "C" int thorin_ensemble_create(thorin_ctx *const *disks, int n, thorin_ensemble **out) {
  if (!out) { thorin_set_error("thorin_ensemble_create: null output"); return -1; }
  *out = nullptr;
  if (!disks || n < 1 || n > THORIN_MAXDISKS) {
    thorin_set_error("thorin_ensemble_create: between 1 and THORIN_MAXDISKS contexts are needed");
    return -1;
  }
  for (int k = 0; k < n; k++) {
    if (need_dev(disks[k])) return -1;
    for (int m = 0; m < k; m++)
      if (disks[m] == disks[k]) { thorin_set_error("thorin_ensemble_create: a context appears twice"); return -1; }
    if (disks[k]->dev->rec) { thorin_set_error("thorin_ensemble_create: a context belongs to a running ensemble"); return -1; }
  }
  thorin_ensemble *e = new (std::nothrow) thorin_ensemble();
  if (!e) { thorin_set_error("out of host memory"); return -1; }
  e->n = n;
  e->c = new thorin_ctx *[n];
  for (int k = 0; k < n; k++) e->c[k] = disks[k];
  e->recs.resize((size_t)n);
  for (int k = 0; k < n; k++) { e->recs[k].e = e; e->recs[k].idx = k; }
  e->use_graphs = 1;
  // one instance of the azimuthal transport for all disks
  int npt = 1;
  for (int k = 0; k < n; k++) {
    const int v = (disks[k]->dev->ns + TT_T - 1) / TT_T;
    npt = v > npt ? v : npt;
  }
  for (int k = 0; k < n; k++) disks[k]->dev->tt_npt_min = npt;
  // and of the end-of-substep kernels
  int ec = 3;
  for (int k = 0; k < n; k++) ec = std::max(ec, eos_ec(disks[k]->dev));
  for (int k = 0; k < n; k++) disks[k]->dev->ec_min = ec;
  if (hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking) != hipSuccess) {
    thorin_set_error("thorin_ensemble_create: stream creation failed");
    delete[] e->c;
    delete e;
    return -1;
  }
  if (hipEventCreateWithFlags(&e->mark, hipEventDisableTiming) != hipSuccess) {
    thorin_set_error("thorin_ensemble_create: event creation failed");
    (void)hipStreamDestroy(e->s);
    delete[] e->c;
    delete e;
    return -1;
  }
  // work queued on the disks' own streams comes first
  for (int k = 0; k < n; k++)
    if (hipStreamSynchronize(disks[k]->dev->s) != hipSuccess) {
      thorin_set_error("thorin_ensemble_create: a disk's stream failed");
      (void)hipEventDestroy(e->mark);
      (void)hipStreamDestroy(e->s);
      delete[] e->c;
      delete e;
      return -1;
    }
  e->job = 0;
  e->quit = false;
  e->j_rc.assign((size_t)n, 0);
  try {
    for (int k = 0; k < n; k++) e->workers.emplace_back(ens_worker, e, k);
  } catch (...) {
    {
      std::lock_guard<std::mutex> lk(e->m);
      e->quit = true;
    }
    e->cv_job.notify_all();
    for (auto &t : e->workers) t.join();
    (void)hipEventDestroy(e->mark);
    (void)hipStreamDestroy(e->s);
    delete[] e->c;
    delete e;
    thorin_set_error("thorin_ensemble_create: the disks' threads could not be started");
    return -1;
  }
  *out = e;
  return 0;
}

extern "C" void thorin_ensemble_destroy(thorin_ensemble *e) {
  if (!e) return;
  {
    std::lock_guard<std::mutex> lk(e->m);
    e->quit = true;
  }
  e->cv_job.notify_all();
  for (auto &t : e->workers) t.join();
  (void)hipStreamSynchronize(e->s);
  for (int k = 0; k < e->n; k++)
    if (e->c[k] && e->c[k]->dev) {
      e->c[k]->dev->tt_npt_min = 0;
      e->c[k]->dev->ec_min = 0;
    }
  for (auto &x : e->graphs) {
    (void)hipGraphExecDestroy(x.x);
    (void)hipGraphDestroy(x.g);
  }
  if (e->st.h) (void)hipHostFree(e->st.h);
  if (e->st.dv) (void)hipFree(e->st.dv);
  (void)hipEventDestroy(e->mark);
  (void)hipStreamDestroy(e->s);
  delete[] e->c;
  delete e;
}

extern "C" int thorin_ensemble_algogas(thorin_ensemble *e, thorin_planets *const *sys, thorin_nbody_fn nbody,
                                       void *const *nbody_user, thorin_step_log *logs, int *rc_out) {
  if (!e || !sys) { thorin_set_error("thorin_ensemble_algogas: null argument"); return -1; }
  const int n = e->n;
  for (int k = 0; k < n; k++)
    if (!sys[k]) { thorin_set_error("thorin_ensemble_algogas: null planetary system"); return -1; }
  e->nwait = 0;
  e->nfinished = 0;
  e->err = 0;
  // every disk queues as many SOR iterations as the disk that expects the
  // most (those after its convergence do nothing): the merged substeps
  // then repeat from DT to DT and replay as graphs
  int qmin = 0;
  for (int k = 0; k < n; k++) {
    const thorin_dev *d = e->c[k]->dev;
    int q = e->c[k]->st.sor_niter_last;
    for (int m = 0; m < 4; m++) q = d->niter_hist[m] > q ? d->niter_hist[m] : q;
    q += 1;
    qmin = q > qmin ? q : qmin;
  }
  if (qmin > 256) qmin = 256;
  for (int k = 0; k < n; k++) {
    e->recs[k].q.clear();
    e->c[k]->dev->rec = &e->recs[k];
    e->c[k]->dev->sor_queue_min = qmin;
  }
  {
    std::lock_guard<std::mutex> lk(e->m);
    e->j_sys = sys;
    e->j_nbody = nbody;
    e->j_user = nbody_user;
    e->j_logs = logs;
    e->j_err.clear();
    for (int k = 0; k < n; k++) e->j_rc[k] = 0;
    e->job++;
  }
  e->cv_job.notify_all();
  int status = 0;
  for (;;) {
    std::unique_lock<std::mutex> lk(e->m);
    e->cv_coord.wait(lk, [&] { return e->nwait + e->nfinished == n; });
    bool pending = false;
    for (int k = 0; k < n; k++) pending |= !e->recs[k].q.empty();
    if (e->nfinished == n && !pending) break;
    // every disk thread waits or has finished: the queues are the coordinator's
    lk.unlock();
    const int f = ens_flush(e);
    lk.lock();
    if (f) {
      e->err = 1;
      status = -1;
      for (int k = 0; k < n; k++) e->recs[k].q.clear();
    }
    e->nwait = 0;
    e->gen++;
    e->cv_disk.notify_all();
  }
  for (int k = 0; k < n; k++) {
    e->c[k]->dev->rec = nullptr;
    e->c[k]->dev->sor_queue_min = 0;
  }
  int worst = 0;
  for (int k = 0; k < n; k++) {
    const int r = e->j_rc[k];
    if (rc_out) rc_out[k] = r;
    if (r < 0) status = -1;
    else if (r > worst) worst = r;
  }
  if (!e->j_err.empty()) thorin_set_error(e->j_err.c_str());
  return status ? -1 : worst;
}

extern "C" int thorin_ensemble_stats(const thorin_ensemble *e, long long *launches, long long *kernel_ops,
                                     long long *graph_captures, long long *graph_replays) {
  if (!e) return -1;
  if (launches) *launches = e->launches;
  if (kernel_ops) *kernel_ops = e->batched_ops;
  if (graph_captures) *graph_captures = e->captures;
  if (graph_replays) *graph_replays = e->replays;
  return 0;
}

extern "C" int thorin_ensemble_set_graphs(thorin_ensemble *e, int on) {
  if (!e) return -1;
  e->use_graphs = on ? 1 : 0;
  return 0;
}

// ===========================================================================
// Autotuner: the launch shapes that leave every result unchanged -- SOR
// iterations per launch (1 or 2, the stop-test sums do not depend on it) and
// rings per tile of the radial transport (8, 16, 32) -- timed on this disk's
// grid with HIP events, the fastest kept.

// A diagonally dominant five-point system for timing the SOR launches whose
// solution (-5000) lies far below the temperature: Gauss-Seidel sweeps
// (omega = 1) approach it monotonically from above, so every cell keeps
// changing and the stop test never ends the timed iterations early.
__global__ void __launch_bounds__(TB) k_tune_system(double *__restrict__ ml, double *__restrict__ mip,
                                                    double *__restrict__ mim, double *__restrict__ mjp,
                                                    double *__restrict__ mjm, double *__restrict__ rhs,
                                                    const double *__restrict__ T, size_t n) {
  const size_t l = (size_t)blockIdx.x * TB + threadIdx.x;
  if (l >= n) return;
  ml[l] = 1.0;
  mip[l] = -0.2;
  mim[l] = -0.2;
  mjp[l] = -0.2;
  mjm[l] = -0.2;
  rhs[l] = -1000.0;
  (void)T;
}

static int tune_time(thorin_dev *d, const std::function<int()> &work, int reps, double *ms_out) {
  hipEvent_t a, b;
  HCHECK(hipEventCreate(&a));
  HCHECK(hipEventCreate(&b));
  double best = 1e30;
  for (int r = 0; r < reps; r++) {
    HCHECK(hipEventRecord(a, d->s));
    if (work()) return -1;
    HCHECK(hipEventRecord(b, d->s));
    HCHECK(hipEventSynchronize(b));
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, a, b));
    best = ms < best ? ms : best;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms_out = best;
  return 0;
}

// The SOR iterations a solve is expected to take: the most of the last
// solves (as the queue of a substep counts them), 3 before any solve.
static int sor_expected(const thorin_ctx *c) {
  int e = c->st.sor_niter_last;
  for (int k = 0; k < 4; k++) e = std::max(e, c->dev->niter_hist[k]);
  return e > 0 ? e : 3;
}

// The launch shape that finishes a solve of `expect` iterations first: the
// queue runs expect / n launches of n iterations and one launch of the rest
// (launches queued after the stop return at once).
static int sor_pick(const double *us_per_iter, int expect) {
  int best = 1;
  double tbest = 0.0;
  for (int n = 1; n <= 3; n++) {
    const int rest = expect % n;
    const double t = us_per_iter[n - 1] * (double)((expect / n) * n) + (rest ? us_per_iter[rest - 1] * rest : 0.0);
    if (n == 1 || t < tbest) { best = n; tbest = t; }
  }
  return best;
}

extern "C" int thorin_autotune(thorin_ctx *c, thorin_tuning *out) {
  if (need_dev(c)) return -1;
  thorin_dev *d = c->dev;
  d->spec.valid = 0;                     // forces enqueued for a next substep are redone
  if (d->rec) { thorin_set_error("thorin_autotune: the context is inside an ensemble step"); return -1; }
  const int nr = d->nr, ns = d->ns;
  const size_t nb = sizeof(double) * d->nf;
  thorin_tuning t;
  memset(&t, 0, sizeof t);
  HCHECK(hipStreamSynchronize(d->s));
  // the SOR launches overwrite the temperature buffers and the matrix: the
  // temperature is kept in B_TBACK and restored, the matrix is assembled
  // again by the next substep (or thorin_rad_system)
  HCHECK(hipMemcpyAsync(d->b[B_TBACK], d->b[B_TEMPER], nb, hipMemcpyDeviceToDevice, d->s));
  hipLaunchKernelGGL(k_tune_system, dim3((unsigned)((d->nf + TB - 1) / TB)), dim3(TB), 0, d->s, d->b[B_ML],
                     d->b[B_MLIP], d->b[B_MLIM], d->b[B_MLJP], d->b[B_MLJM], d->b[B_RHS],
                     (const double *)d->b[B_TEMPER], d->nf);
  HCHECK(hipGetLastError());
  d->sys_form = 0;
  d->rad_ready = 0;
  const int saved_nit = d->sor_nit, saved_rows = d->rad_rows;
  const int iters = 6;          // whole launches of 1, 2 and 3 iterations
  for (int v = 0; v < 3; v++) {
    d->sor_nit = v + 1;
    double ms = 0.0;
    auto work = [&]() -> int {
      HCHECK(hipMemcpyAsync(d->b[B_TEMPER], d->b[B_TBACK], nb, hipMemcpyDeviceToDevice, d->s));
      int cnt = iters;
      return sor_launch(c, 1.0, 0, cnt, 0.0, 1 << 30);
    };
    if (work()) return -1;      // warm-up
    if (tune_time(d, work, 3, &ms)) return -1;
    // the copy of T (one pass) is timed too; it is the same for all
    t.sor_us_per_iter[v] = 1e3 * ms / iters;
  }
  RadQ q = {};
  q.rho = d->b[B_DENS];
  q.vr = d->b[B_VRAD];
  q.vt = d->b[B_VTHETA];
  q.e = d->b[B_ENERGY];
  const int outs[RC_NQ] = {B_XRP, B_XRM, B_XTP, B_XTM, B_XE, B_XRHO};
  for (int k = 0; k < RC_NQ; k++) q.out[k] = d->b[outs[k]];
  q.guard = nullptr;
  q.of = c->st.omega_frame;
  q.ntile_j = (ns + RT_J - 1) / RT_J;
  q.nq = RC_NQ;
  const int rows[3] = {8, 16, 32};
  for (int v = 0; v < 3; v++) {
    q.rows = rows[v];
    q.ntiles = q.ntile_j * ((nr + q.rows - 1) / q.rows);
    const dim3 grad((unsigned)((q.ntiles + 7) / 8 * 8)), trad(RT_T);
    const double dt = c->p.dt;
    auto work = [&]() -> int {
      if (q.rows == 8) LAUNCH(k_transport_rad<8>, grad, trad, 0, d->g, q, dt);
      else if (q.rows == 32) LAUNCH(k_transport_rad<32>, grad, trad, 0, d->g, q, dt);
      else LAUNCH(k_transport_rad<16>, grad, trad, 0, d->g, q, dt);
      HCHECK(hipGetLastError());
      return 0;
    };
    double ms = 0.0;
    if (work()) return -1;
    if (tune_time(d, work, 3, &ms)) return -1;
    t.rad_us[v] = 1e3 * ms;
  }
  HCHECK(hipMemcpyAsync(d->b[B_TEMPER], d->b[B_TBACK], nb, hipMemcpyDeviceToDevice, d->s));
  HCHECK(hipStreamSynchronize(d->s));
  (void)saved_nit;
  (void)saved_rows;
  t.sor_iters_per_launch = sor_pick(t.sor_us_per_iter, sor_expected(c));
  int bi = 0;
  for (int v = 1; v < 3; v++)
    if (t.rad_us[v] < t.rad_us[bi]) bi = v;
  t.rad_rows = rows[bi];
  t.fused_chain = d->visc_chain;
  t.sor_assembly = d->sor_assembly;
  d->sor_nit = t.sor_iters_per_launch;
  d->rad_rows = t.rad_rows;
  if (out) *out = t;
  return 0;
}

extern "C" int thorin_set_tuning(thorin_ctx *c, const thorin_tuning *t) {
  if (need_dev(c)) return -1;
  if (!t || t->sor_iters_per_launch < 1 || t->sor_iters_per_launch > 3 ||
      (t->rad_rows != 8 && t->rad_rows != 16 && t->rad_rows != 32) || (t->fused_chain != 0 && t->fused_chain != 1) ||
      (t->sor_assembly != 0 && t->sor_assembly != 1)) {
    thorin_set_error("thorin_set_tuning: sor_iters_per_launch must be 1, 2 or 3, rad_rows 8, 16 or 32, "
                     "fused_chain and sor_assembly 0 or 1");
    return -1;
  }
  c->dev->sor_nit = t->sor_iters_per_launch;
  c->dev->rad_rows = t->rad_rows;
  c->dev->visc_chain = t->fused_chain;
  c->dev->sor_assembly = t->sor_assembly;
  return 0;
}

extern "C" int thorin_get_tuning(thorin_ctx *c, thorin_tuning *t) {
  if (need_dev(c) || !t) return -1;
  memset(t, 0, sizeof *t);
  t->sor_iters_per_launch = c->dev->sor_nit;
  t->rad_rows = c->dev->rad_rows;
  t->fused_chain = c->dev->visc_chain;
  t->sor_assembly = c->dev->sor_assembly;
  return 0;
}

// one choice for all disks (a merged launch runs one kernel): the one with
// the least time summed over the disks
extern "C" int thorin_ensemble_autotune(thorin_ensemble *e, thorin_tuning *out) {
  if (!e) return -1;
  thorin_tuning sum;
  memset(&sum, 0, sizeof sum);
  for (int k = 0; k < e->n; k++) {
    thorin_tuning t;
    if (thorin_autotune(e->c[k], &t)) return -1;
    for (int v = 0; v < 3; v++) sum.sor_us_per_iter[v] += t.sor_us_per_iter[v];
    for (int v = 0; v < 3; v++) sum.rad_us[v] += t.rad_us[v];
  }
  const int rows[3] = {8, 16, 32};
  int expect = 0;
  for (int k = 0; k < e->n; k++) expect = std::max(expect, sor_expected(e->c[k]));
  sum.sor_iters_per_launch = sor_pick(sum.sor_us_per_iter, expect);
  int bi = 0;
  for (int v = 1; v < 3; v++)
    if (sum.rad_us[v] < sum.rad_us[bi]) bi = v;
  sum.rad_rows = rows[bi];
  sum.fused_chain = e->n > 0 ? e->c[0]->dev->visc_chain : 1;
  sum.sor_assembly = e->n > 0 ? e->c[0]->dev->sor_assembly : 1;
  for (int k = 0; k < e->n; k++)
    if (thorin_set_tuning(e->c[k], &sum)) return -1;
  if (out) *out = sum;
  return 0;
}

// ===========================================================================
// The energy equation's implicit solve on its own (EnergySources.c:200-335),
// for solver studies: the plain FP64 red-black SOR of the reference, and FP32
// red-black SOR correction sweeps inside FP64 iterative refinement.
//
// Refinement round: r = A.T - RHS in FP64 (rings 1..NRAD-2, T of rings 0 and
// NRAD-1 fixed), then A.e = -r solved approximately by `inner` FP32 red-black
// sweeps from e = 0 (e = 0 on the fixed rings), then T += e in FP64.  The
// rounds stop when sum|r| <= tol * sum|r0|.

// FP64 residual r = A.T - RHS of rings 1..nr-2 as FP32, with block partials of
// sum|r| (FP64)
__global__ void __launch_bounds__(TB) k_mp_resid(Geo g, const double *__restrict__ T,
                                                 const double *__restrict__ Ml, const double *__restrict__ Mlip,
                                                 const double *__restrict__ Mlim, const double *__restrict__ Mljp,
                                                 const double *__restrict__ Mljm, const double *__restrict__ RHS,
                                                 float *__restrict__ r32, double *__restrict__ partial) {
  __shared__ double red[TB / 64];
  const int ns = g.ns;
  int i, j, L;
  if (!xcd_cell(g, 1, g.nr - 2, i, j, L)) return;
  double a = 0.0;
  if (j < ns) {
    const size_t l = (size_t)i * ns + j;
    const size_t ljp = (j == ns - 1) ? (size_t)i * ns : l + 1;
    const size_t ljm = (j == 0) ? (size_t)i * ns + ns - 1 : l - 1;
    const double r = sor_resid(T, Ml, Mlip, Mlim, Mljp, Mljm, RHS, l, l + ns, l - ns, ljp, ljm);
    r32[l] = (float)r;
    a = fabs(r);
  }
  a = block_sum<TB>(a, red);
  if (threadIdx.x == 0) partial[L] = a;
}

// FP32 copies of the five matrix diagonals (rings 1..nr-2)
__global__ void __launch_bounds__(TB) k_mp_convert(Geo g, const double *__restrict__ Ml,
                                                   const double *__restrict__ Mlip, const double *__restrict__ Mlim,
                                                   const double *__restrict__ Mljp, const double *__restrict__ Mljm,
                                                   Mat32 m) {
  const int ns = g.ns;
  int i, j, L;
  if (!xcd_cell(g, 1, g.nr - 2, i, j, L) || j >= ns) return;
  const size_t l = (size_t)i * ns + j;
  m.ml[l] = (float)Ml[l];
  m.mip[l] = (float)Mlip[l];
  m.mim[l] = (float)Mlim[l];
  m.mjp[l] = (float)Mljp[l];
  m.mjm[l] = (float)Mljm[l];
}

// T += e on rings 1..nr-2
__global__ void __launch_bounds__(TB) k_mp_update(Geo g, double *__restrict__ T, const float *__restrict__ e) {
  const int ns = g.ns;
  int i, j, L;
  if (!xcd_cell(g, 1, g.nr - 2, i, j, L) || j >= ns) return;
  const size_t l = (size_t)i * ns + j;
  T[l] += (double)e[l];
}

// one FP32 red-black sweep of A.e = -r, tiled like k_sor_iter (red
// half-sweep also on the one-cell halo); ein -> eout
struct SorCell32 {
  float ml, mip, mim, mjp, mjm, r;
  int k;
};

__global__ void __launch_bounds__(TB) k_mp_sweep(Geo g, const float *__restrict__ ein, float *__restrict__ eout,
                                                 Mat32 m, float omega) {
  __shared__ float t[(SOR_TI + 4) * SOR_LW];
  const int ns = g.ns, nr = g.nr;
  const int ntx = (ns + SOR_TJ - 1) / SOR_TJ;
  int tx, ty, bid;
  if (!xcd_tile(ntx, (nr - 2 + SOR_TI - 1) / SOR_TI, tx, ty, bid)) return;
  const int i0 = 1 + ty * SOR_TI, i1 = min(i0 + SOR_TI, nr - 1);
  const int j0 = tx * SOR_TJ, tw = min(SOR_TJ, ns - j0), ti = i1 - i0, tid = threadIdx.x;
  SorCell32 rc[SOR_NRED], bc[SOR_NBLK];
  const int rw = (tw + 3) / 2, bw = (tw + 1) / 2;
  auto load = [&](SorCell32 &x, size_t l, int k) {
    x.ml = m.ml[l]; x.mip = m.mip[l]; x.mim = m.mim[l]; x.mjp = m.mjp[l]; x.mjm = m.mjm[l]; x.r = m.r[l];
    x.k = k;
  };
#pragma unroll
  for (int q = 0; q < SOR_NRED; q++) {
    rc[q].k = -1;
    const int sl = tid + q * TB;
    if (sl < (ti + 2) * rw) {
      const int r = 1 + sl / rw, mm = sl - (r - 1) * rw;
      const int i = i0 - 2 + r;
      const int c = (((i + j0 - 1) & 1) ? 1 : 2) + 2 * mm;
      if (c <= tw + 2 && i >= 1 && i <= nr - 2) {
        int j = j0 - 2 + c;
        if (j < 0) j += ns;
        if (j >= ns) j -= ns;
        load(rc[q], (size_t)i * ns + j, r * SOR_LW + c);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < SOR_NBLK; q++) {
    bc[q].k = -1;
    const int sl = tid + q * TB;
    if (sl < ti * bw) {
      const int r = 2 + sl / bw, mm = sl - (r - 2) * bw;
      const int i = i0 - 2 + r;
      const int c = (((i + j0) & 1) ? 3 : 2) + 2 * mm;
      if (c <= tw + 1) load(bc[q], (size_t)i * ns + (j0 - 2 + c), r * SOR_LW + c);
    }
  }
  for (int idx = tid; idx < (ti + 4) * (tw + 4); idx += TB) {
    const int r = idx / (tw + 4), c = idx - r * (tw + 4);
    const int i = i0 - 2 + r;
    float v = 0.0f;                       // rings 0 and nr-1: e = 0
    if (i >= 1 && i <= nr - 2) {
      int j = j0 - 2 + c;
      if (j < 0) j += ns;
      if (j >= ns) j -= ns;
      v = ein[(size_t)i * ns + j];
    }
    t[r * SOR_LW + c] = v;
  }
  __syncthreads();
  auto upd = [&](const SorCell32 &x) {
    const int k = x.k;
    const float res = t[k] * x.ml + t[k + SOR_LW] * x.mip + t[k - SOR_LW] * x.mim + t[k + 1] * x.mjp +
                      t[k - 1] * x.mjm + x.r;
    t[k] = t[k] - omega * res / x.ml;
  };
#pragma unroll
  for (int q = 0; q < SOR_NRED; q++)
    if (rc[q].k >= 0) upd(rc[q]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SOR_NBLK; q++)
    if (bc[q].k >= 0) upd(bc[q]);
  __syncthreads();
  for (int idx = tid; idx < ti * tw; idx += TB) {
    const int r = 2 + idx / tw, c = 2 + idx % tw;
    eout[(size_t)(i0 - 2 + r) * ns + (j0 - 2 + c)] = t[r * SOR_LW + c];
  }
}

// Several FP32 red-black sweeps per launch (temporal blocking).  A
// workgroup owns an interior tile of MB_TI rings x MB_TJ sectors and stages
// e of the tile plus a halo of MB_H = 2*MB_K cells in LDS; each half-sweep
// leaves the outermost not-yet-updated layer stale, so after up to MB_K
// sweeps (2*MB_K half-sweeps) the interior is exactly what MB_K sweeps of
// k_mp_sweep give.  Thread t of MB_T keeps the coefficients of its 16 cells
// in registers: rows (t >> 6) + (MB_T/64)s, s = 0..7, columns 2(t & 63) and
// +1 (one cell of each colour per row).  Colour order as k_mp_sweep: (i + j)
// odd first.  512 threads on 64 x 128 tiles (a 48 x 112 interior) held to
// four waves per SIMD (two workgroups per CU): 1.2x the rate of 256 threads
// on 32 x 128 tiles (16 x 112 interiors, three workgroups per CU).
#define MB_K 4
#define MB_H (2 * MB_K)
#define MB_T 512
#define MB_ER (16 * MB_T / 128)
#define MB_EW 128
#define MB_LW (MB_EW + 4)
#define MB_TI (MB_ER - 2 * MB_H)
#define MB_TJ (MB_EW - 2 * MB_H)
__global__ void __launch_bounds__(MB_T) __attribute__((amdgpu_waves_per_eu(4))) k_mp_sweeps(Geo g, const float *__restrict__ ein, float *__restrict__ eout,
                                                  Mat32 m, float omega, int nsw) {
  __shared__ float t[MB_ER * MB_LW];
  const int ns = g.ns, nr = g.nr;
  const int ntx = (ns + MB_TJ - 1) / MB_TJ, nty = (nr - 2 + MB_TI - 1) / MB_TI;
  int tx, ty, bid;
  if (!xcd_tile(ntx, nty, tx, ty, bid)) return;
  const int i0 = 1 + ty * MB_TI, j0 = tx * MB_TJ;
  const int ti = min(MB_TI, nr - 1 - i0), tw = min(MB_TJ, ns - j0);
  const int tid = threadIdx.x, cp = 2 * (tid & 63), rb = tid >> 6;
  for (int idx = tid; idx < MB_ER * MB_EW; idx += MB_T) {
    const int r = idx / MB_EW, c = idx % MB_EW;
    const int i = i0 - MB_H + r;
    const int j = ((j0 - MB_H + c) % ns + ns) % ns;
    float v = 0.0f;                       // rings 0 and nr-1 (and outside): e = 0
    if (i >= 1 && i <= nr - 2) v = ein[(size_t)i * ns + j];
    t[r * MB_LW + c] = v;
  }
  float ml[16], mip[16], mim[16], mjp[16], mjm[16], rr[16];
  unsigned act = 0u, odd = 0u;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int r = rb + (MB_T / 64) * (q >> 1), c = cp + (q & 1);
    const int i = i0 - MB_H + r;
    const int j = ((j0 - MB_H + c) % ns + ns) % ns;
    ml[q] = 1.0f; mip[q] = mim[q] = mjp[q] = mjm[q] = rr[q] = 0.0f;
    if (i >= 1 && i <= nr - 2 && r >= 1 && r <= MB_ER - 2 && c >= 1 && c <= MB_EW - 2) {
      const size_t l = (size_t)i * ns + j;
      ml[q] = m.ml[l]; mip[q] = m.mip[l]; mim[q] = m.mim[l]; mjp[q] = m.mjp[l]; mjm[q] = m.mjm[l]; rr[q] = m.r[l];
      act |= 1u << q;
      if ((i + j) & 1) odd |= 1u << q;
    }
  }
  __syncthreads();
  for (int sw = 0; sw < nsw; sw++) {
#pragma unroll
    for (int colour = 0; colour < 2; colour++) {
      const unsigned sel = act & (colour == 0 ? odd : ~odd);
#pragma unroll
      for (int q = 0; q < 16; q++) {
        if (sel & (1u << q)) {
          const int k = (rb + (MB_T / 64) * (q >> 1)) * MB_LW + cp + (q & 1);
          const float res = t[k] * ml[q] + t[k + MB_LW] * mip[q] + t[k - MB_LW] * mim[q] + t[k + 1] * mjp[q] +
                            t[k - 1] * mjm[q] + rr[q];
          t[k] = t[k] - omega * res / ml[q];
        }
      }
      __syncthreads();
    }
  }
  for (int idx = tid; idx < ti * tw; idx += MB_T) {
    const int r = idx / tw, c = idx - r * tw;
    eout[(size_t)(i0 + r) * ns + (j0 + c)] = t[(MB_H + r) * MB_LW + MB_H + c];
  }
}

// sum|A.T - RHS| over rings 1..nr-2 (block partials summed in order on the host)
static int mp_residual(thorin_ctx *c, float *r32, double *out) {
  thorin_dev *d = c->dev;
  const int nr = d->nr;
  const dim3 gc = cell_grid(d, nr - 2);
  const int nb = cell_count(d, nr - 2);
  hipLaunchKernelGGL(k_mp_resid, gc, dim3(TB), 0, d->s, d->g, (const double *)d->b[B_TEMPER],
                     (const double *)d->b[B_ML], (const double *)d->b[B_MLIP], (const double *)d->b[B_MLIM],
                     (const double *)d->b[B_MLJP], (const double *)d->b[B_MLJM], (const double *)d->b[B_RHS], r32,
                     d->mresp);
  HCHECK(hipGetLastError());
  HCHECK(hipMemcpyAsync(d->h_part, d->mresp, sizeof(double) * nb, hipMemcpyDeviceToHost, d->s));
  HCHECK(hipStreamSynchronize(d->s));
  double t = 0.0;
  for (int k = 0; k < nb; k++) t += d->h_part[k];
  *out = t;
  return 0;
}

extern "C" int thorin_rad_system(thorin_ctx *c, double dt, double *dt_used) {
  if (need_dev(c)) return -1;
  thorin_dev *d = c->dev;
  d->spec.valid = 0;                     // forces enqueued for a next substep are redone
  const int nr = d->nr, ns = d->ns;
  const size_t nb = sizeof(double) * d->nf;
  const dim3 gall = cell_grid(d, nr), tb(TB), rb(RB);
  if (dt <= 0.0) {
    // the hydro CFL step of the current fields (ConditionCFL without planets)
    Geo gg = d->g;
    gg.pebbles = 0;
    EOS_LAUNCH(k_eos_cfl, dim3((nr + 7) / 8 * 8), rb, eos_lds(d, 0), gg, d->b[B_DENS], d->b[B_ENERGY], d->b[B_TEMPER],
           d->b[B_CS], d->b[B_VRAD], d->b[B_VTHETA], 0, (const int *)nullptr);
    LAUNCH(k_cfl_final, dim3(1), dim3(CF_T), 0, gg, d->ctl, (const double *)nullptr, 0, 0.0, (const int *)nullptr,
           (DevCtl *)nullptr, (unsigned *)nullptr, 0u);
    HCHECK(hipMemcpyAsync(d->hctl, d->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, d->s));
    HCHECK(hipStreamSynchronize(d->s));
    dt = d->hctl->newdt_raw;
    (void)ns;
  }
  if (dt_used) *dt_used = dt;
  HCHECK(hipMemcpyAsync(d->b[B_EINT], d->b[B_ENERGY], nb, hipMemcpyDeviceToDevice, d->s));
  SUB3_LAUNCH(gall, tb, 0, d->g, d->b[B_VRAD], d->b[B_VTHETA], d->b[B_DENS], d->b[B_EINT], d->b[B_DIVV],
              d->b[B_TAURR], d->b[B_TAURP], d->b[B_TAUPP], d->b[B_QPLUS], d->b[B_TEMPER], d->b[B_CS],
              d->b[B_OPACITY], d->b[B_QMINUS], d->b[B_DIFF], d->pstellar, d->opa_filled ? 0 : 1, 0, CrashSnap{});
  d->opa_filled = 1;
  const dim3 gm = cell_grid(d, nr - 2);
  LAUNCH(k_matrix, gm, tb, 0, d->g, d->b[B_DENS], d->b[B_CS], d->b[B_DIFF], d->b[B_DIVV], d->b[B_QMINUS],
         d->b[B_QPLUS], d->b[B_TEMPER], d->b[B_ML], d->b[B_MLIP], d->b[B_MLIM], d->b[B_MLJP], d->b[B_MLJM],
         d->b[B_RHS], dt, d->mresp, (const double *)d->b[B_OPACITY], HeatK{});
  d->sys_form = 0;
  HCHECK(hipMemcpyAsync(d->b[B_TBACK], d->b[B_TEMPER], nb, hipMemcpyDeviceToDevice, d->s));
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(d->s));
  d->rad_ready = 1;
  return 0;
}

extern "C" int thorin_dev_system(thorin_ctx *c, thorin_sysptrs *p) {
  if (need_dev(c)) return -1;
  thorin_dev *d = c->dev;
  d->spec.valid = 0;                     // forces enqueued for a next substep are redone
  if (!d->rad_ready) { thorin_set_error("no energy system assembled (thorin_rad_system)"); return -1; }
  p->nr = d->nr;
  p->ns = d->ns;
  p->ml = d->b[B_ML]; p->mlip = d->b[B_MLIP]; p->mlim = d->b[B_MLIM]; p->mljp = d->b[B_MLJP];
  p->mljm = d->b[B_MLJM]; p->rhs = d->b[B_RHS];
  p->T = d->b[B_TEMPER];
  p->T0 = d->b[B_TBACK];
  p->stream = (void *)d->s;
  return 0;
}

// zero fill of FP32 planes (a kernel, so that no memset node is recorded
// where a caller captures the stream)
__global__ void __launch_bounds__(256) k_zero_f32(float *__restrict__ p, size_t n) {
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) p[k] = 0.0f;
}

static void zero_f32(thorin_dev *d, float *p, size_t n) {
  size_t nb = (n + 255) / 256;
  const unsigned g = (unsigned)(nb < 4096 ? (nb > 0 ? nb : 1) : 4096);
  hipLaunchKernelGGL(k_zero_f32, dim3(g), dim3(256), 0, d->s, p, n);
}

// The solve of thorin_rad_solve and thorin_rad_solve_noalloc on the FP32
// planes m (matrix diagonals, residual, correction double buffer).
static int rad_solve_on(thorin_ctx *c, const thorin_solve_opts *o, thorin_solve_result *res, Mat32 m) {
  thorin_dev *d = c->dev;
  memset(res, 0, sizeof *res);
  const int nr = d->nr, ns = d->ns;
  const size_t nb = sizeof(double) * d->nf;
  const int maxit = o->max_iters > 0 ? o->max_iters : 100000;
  HCHECK(hipMemcpyAsync(d->b[B_TEMPER], d->b[B_TBACK], nb, hipMemcpyDeviceToDevice, d->s));
  double r0 = 0.0;
  if (mp_residual(c, m.r, &r0)) return -1;
  res->resid0 = r0;
  if (o->warm && o->method != 1) {
    thorin_set_error("thorin_rad_solve: the learned warm start is used with method 1");
    return -1;
  }
  HCHECK(hipEventRecord(d->ev_t0, d->s));
  double rcur = r0;
  if (o->warm) {
    // the network's guess replaces T0 only if it lowers the FP64 residual;
    // the target stays tol * r0 of T0
    if (thorin_cnn_predict_dev(o->warm, c, d->b[B_TEMPER2])) return -1;
    HCHECK(hipMemcpyAsync(d->b[B_TEMPER], d->b[B_TEMPER2], nb, hipMemcpyDeviceToDevice, d->s));
    double rg = 0.0;
    if (mp_residual(c, m.r, &rg)) return -1;
    res->resid_warm = rg;
    if (rg < r0) {
      res->warm_used = 1;
      rcur = rg;
    } else {
      HCHECK(hipMemcpyAsync(d->b[B_TEMPER], d->b[B_TBACK], nb, hipMemcpyDeviceToDevice, d->s));
      if (mp_residual(c, m.r, &rcur)) return -1;
    }
  }
  if (o->method == 0) {
    // the reference's SOR with stop test sum|r| (in-sweep) < tol * normres0
    int niter = 0;
    if (sor_finish(c, o->omega, 0, &niter, o->tol, maxit, 16)) return -1;
    res->sweeps = niter > maxit ? maxit : niter;
    // the last iterate into B_TEMPER (it is in T0 .. T3 by sor_cur)
    if (d->hctl->sor_cur)
      HCHECK(hipMemcpyAsync(d->b[B_TEMPER], d->b[d->hctl->sor_cur == 1 ? B_TEMPER2 : (d->hctl->sor_cur == 2 ? B_TEMPER3 : B_TEMPER4)], nb,
                            hipMemcpyDeviceToDevice, d->s));
    res->rounds = 1;
  } else if (o->method == 1) {
    const dim3 gm = cell_grid(d, nr - 2), tb(TB);
    const int nbs = ((ns + SOR_TJ - 1) / SOR_TJ) * ((nr - 2 + SOR_TI - 1) / SOR_TI);
    const dim3 gs((unsigned)((nbs + 7) / 8 * 8));
    const int inner = o->inner_sweeps > 0 ? o->inner_sweeps : 8;
    // sweeps per launch: 1 = k_mp_sweep; 2..MB_K = k_mp_sweeps (LDS-blocked)
    const int blk = o->sweep_block > 0 ? (o->sweep_block < MB_K ? o->sweep_block : MB_K) : MB_K;
    const int ntb = ((ns + MB_TJ - 1) / MB_TJ) * ((nr - 2 + MB_TI - 1) / MB_TI);
    const dim3 gb((unsigned)((ntb + 7) / 8 * 8));
    hipLaunchKernelGGL(k_mp_convert, gm, tb, 0, d->s, d->g, (const double *)d->b[B_ML],
                       (const double *)d->b[B_MLIP], (const double *)d->b[B_MLIM], (const double *)d->b[B_MLJP],
                       (const double *)d->b[B_MLJM], m);
    while (rcur > o->tol * r0 && res->sweeps < maxit) {
      // r32 holds A.T - RHS of the current T; solve A.e = -r from e = 0
      zero_f32(d, m.e0, d->nf);
      zero_f32(d, m.e1, d->nf);
      float *ein = m.e0, *eout = m.e1;
      for (int k = 0; k < inner;) {
        const int nsw = blk > 1 ? (inner - k < blk ? inner - k : blk) : 1;
        if (nsw > 1)
          hipLaunchKernelGGL(k_mp_sweeps, gb, dim3(MB_T), 0, d->s, d->g, (const float *)ein, eout, m, (float)o->omega,
                             nsw);
        else
          hipLaunchKernelGGL(k_mp_sweep, gs, tb, 0, d->s, d->g, (const float *)ein, eout, m, (float)o->omega);
        float *t = ein; ein = eout; eout = t;
        k += nsw;
      }
      hipLaunchKernelGGL(k_mp_update, gm, tb, 0, d->s, d->g, d->b[B_TEMPER], (const float *)ein);
      HCHECK(hipGetLastError());
      res->sweeps += inner;
      res->rounds++;
      if (mp_residual(c, m.r, &rcur)) return -1;
    }
  } else {
    thorin_set_error("thorin_rad_solve: unknown method");
    return -1;
  }
  HCHECK(hipEventRecord(d->ev_t1, d->s));
  HCHECK(hipEventSynchronize(d->ev_t1));
  float ms = 0.f;
  HCHECK(hipEventElapsedTime(&ms, d->ev_t0, d->ev_t1));
  res->ms = ms;
  double rfin = 0.0;
  if (mp_residual(c, m.r, &rfin)) return -1;
  res->resid = rfin;
  res->converged = rfin <= o->tol * r0;
  return 0;
}

static int rad_solve_check(thorin_ctx *c, const thorin_solve_opts *o, thorin_solve_result *res, const char *who) {
  if (need_dev(c)) return -1;
  if (!c->dev->rad_ready) {
    thorin_set_errorf("%s: no system assembled (thorin_rad_system)", who);
    return -1;
  }
  if (!o || !res || o->tol <= 0.0 || o->omega <= 0.0 || o->omega >= 2.0) {
    thorin_set_errorf("%s: need tol > 0 and 0 < omega < 2", who);
    return -1;
  }
  return 0;
}

extern "C" int thorin_rad_solve(thorin_ctx *c, const thorin_solve_opts *o, thorin_solve_result *res) {
  if (rad_solve_check(c, o, res, "thorin_rad_solve")) return -1;
  thorin_dev *d = c->dev;
  d->spec.valid = 0;                     // forces enqueued for a next substep are redone
  if (!d->m32.ml) {
    float **f[8] = {&d->m32.ml, &d->m32.mip, &d->m32.mim, &d->m32.mjp, &d->m32.mjm, &d->m32.r, &d->m32.e0,
                    &d->m32.e1};
    for (int k = 0; k < 8; k++) {
      HCHECK(hipMalloc(f[k], sizeof(float) * d->nf));
      zero_f32(d, *f[k], d->nf);
    }
  }
  return rad_solve_on(c, o, res, d->m32);
}

// ---------------------------------------------------------------------------
// The energy solve without allocation (thorin_rad_solve_noalloc).  Its FP32
// planes (matrix diagonals, residual, correction double buffer) are slices of
// one file-scope device array; each device holds its own copy, whose address
// is looked up once per device with that device current.
#define NOALLOC_PLANE_ALIGN 64   // floats: 256-byte aligned planes
#define NOALLOC_PLANE ((size_t)THORIN_NOALLOC_MAX_CELLS + NOALLOC_PLANE_ALIGN)
static __device__ __align__(256) float scratch[8 * NOALLOC_PLANE];
#define NOALLOC_MAXDEV 64
static std::mutex g_scratch_mu;
static float *g_scratch_dev[NOALLOC_MAXDEV];   // device address of scratch, per device

static int scratch_for_device(int dev, float **out) {
  if (dev < 0 || dev >= NOALLOC_MAXDEV) {
    thorin_set_errorf("thorin_rad_solve_noalloc: device %d beyond the %d the library keeps", dev, NOALLOC_MAXDEV);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = 0;
    HCHECK(hipGetDevice(&cur));
    if (cur != dev) HCHECK(hipSetDevice(dev));
    float *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      thorin_set_errorf("thorin_rad_solve_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return -1;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

extern "C" int thorin_rad_solve_noalloc(thorin_ctx *c, const thorin_solve_opts *o, thorin_solve_result *res) {
  if (rad_solve_check(c, o, res, "thorin_rad_solve_noalloc")) return -1;
  thorin_dev *d = c->dev;
  if (o->warm) {
    thorin_set_error("thorin_rad_solve_noalloc: no warm start (the network allocates its activations)");
    return -1;
  }
  if (d->nf > (size_t)THORIN_NOALLOC_MAX_CELLS) {
    thorin_set_errorf("thorin_rad_solve_noalloc: (NRAD+1)*NSEC = %zu exceeds THORIN_NOALLOC_MAX_CELLS = %d", d->nf,
                      THORIN_NOALLOC_MAX_CELLS);
    return -1;
  }
  if (d->prof_on) {
    thorin_set_error("thorin_rad_solve_noalloc: kernel timing (thorin_prof_enable) creates events; switch it off");
    return -1;
  }
  float *base = nullptr;
  if (scratch_for_device(d->device, &base)) return -1;
  d->spec.valid = 0;
  const size_t stride = (d->nf + NOALLOC_PLANE_ALIGN - 1) / NOALLOC_PLANE_ALIGN * NOALLOC_PLANE_ALIGN;
  Mat32 m;
  float **f[8] = {&m.ml, &m.mip, &m.mim, &m.mjp, &m.mjm, &m.r, &m.e0, &m.e1};
  for (int k = 0; k < 8; k++) *f[k] = base + (size_t)k * stride;
  // the array keeps the last call's contents (of any grid): the planes start
  // at zero, as the allocated ones of thorin_rad_solve do
  zero_f32(d, base, 8 * stride);
  return rad_solve_on(c, o, res, m);
}
