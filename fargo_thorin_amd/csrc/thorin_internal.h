/*
 * thorin_internal.h -- context shared by the host C code (gcc -std=c99, no
 * FMA contraction, like the reference's build) and the HIP device code.
 */
#ifndef THORIN_INTERNAL_H
#define THORIN_INTERNAL_H
#include "../../include/thorin_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* fondam.h:11-39; the unit macros are evaluated on the host (gcc folds the
 * pow() calls exactly as in the reference build) and passed to kernels. */
#define TH_PI 3.14159265358979323844
#define TH_MSOL_SI 1.98855e30
#define TH_GM_SI 1.32712440018e20
#define TH_AU_SI 149597870700.0
#define TH_R_STANDARD 8.3144598
#define TH_MMW 2.4
#define TH_R_SI (TH_R_STANDARD / (TH_MMW * 0.001))
#define TH_OPA2CU (TH_MSOL_SI * 1000.0 / pow(TH_AU_SI * 100.0, 2.0))
#define TH_T2SI (TH_GM_SI / TH_R_SI / TH_AU_SI)
#define TH_TIME2SI (sqrt(pow(TH_AU_SI, 3.0) / TH_GM_SI))
#define TH_RHO2CGS (TH_MSOL_SI * 1000.0 / pow(TH_AU_SI * 100.0, 3.0))
#define TH_STEFANBOLTZMANN \
  (5.670367e-8 * (1.0 / TH_MSOL_SI) * pow(1.0 / TH_TIME2SI, -3.0) * pow(1.0 / TH_T2SI, -4.0))
#define TH_SQRT2PI_INV (1.0 / sqrt(2.0 * TH_PI))

/* constants of the hot path that kernels need, computed once on the host */
typedef struct thorin_consts {
  double opa2cu, t2si, rho2cgs, stefan, sqrt2pi_inv, sqrt3;
  double sqrt_adiabind_inv, cv;
  double lr01, lr12;       /* log(r0/r1), log(r1/r2) of SourceEuler.c:563 */
} thorin_consts;

struct thorin_dev;   /* device mirror, defined in thorin_dev.hip */

struct thorin_ctx {
  thorin_params p;
  int nr, ns;
  size_t nfield;             /* doubles per device field: (nr+1)*ns */
  /* host 1D geometry and profiles (nr+2 entries) */
  double *Radii, *Rinf, *Rsup, *Rmed, *Surf, *InvRmed, *InvSurf, *InvDiffRmed, *InvDiffRsup;
  double *InvRinf, *OmegaInv, *Rmed2, *GlobalRmed, *SigmaMed, *SigmaInf, *EnergyMed;
  double *VthetaMed, *WaveKiller, *globcsvec, *globtempervec;
  /* per-ring factors evaluated with the host libm so that kernels only do
   * IEEE +,-,*,/ and sqrt on them */
  double *pow_rmed15, *pow_rinf15;   /* pow(Rmed,1.5), pow(Rinf,1.5) (Viscosity.c:37) */
  double *cav_med, *cav_inf;         /* cavity multiplier of FViscosity (Viscosity.c:44-52) */
  double *cfl_pow2;                  /* pow(min2(dxrad,dxtheta),2) (SourceEuler.c:648) */
  double *supp_torque;               /* IMPOSEDDISKDRIFT*.5*pow(Rmed,-2.5+SIGMASLOPE) */
  thorin_consts k;
  thorin_state st;
  /* host staging copies used during initialisation (nfield doubles each) */
  double *h_dens, *h_vrad, *h_vtheta, *h_energy, *h_label, *h_cs, *h_press, *h_temper;
  double *h_abs, *h_ord;
  double *h_trig;               /* cos(2 pi j/NSEC), then sin (InitComputeAccel's angles) */
  int sm_rows, sm_cols;          /* OpacityTable SM03: the gas table's pressure and temperature grids */
  double *sm_pgrid, *sm_tgrid, *sm_ross;   /* and its Rosseland values (InitOpacities, Opacity2.c:252-300) */
  /* pebble fluid (allocated by thorin_init_pebbles): host copies of the
   * initial pebble fields and the damping targets PebDensInit, PebVradInit,
   * PebVthetaInit (global.h, Pebbles.c:190-193, 206-232) */
  double *h_pdens, *h_pvr, *h_pvt, *h_psize, *h_tau, *h_eta;
  double *PebDensInit, *PebVradInit, *PebVthetaInit;
  int peb_ready;
  int t0_mode;                  /* the last initial state (thorin_init_state mode 0 or 1) + 1; 0: none */
  struct thorin_dev *dev;
};

void thorin_set_error(const char *msg);
void thorin_set_errorf(const char *fmt, ...);
double thorin_fviscosity_host(const thorin_ctx *c, double rad);
double thorin_aspectratio_host(const thorin_ctx *c, double rad);
void thorin_cavity_factors(thorin_ctx *c);

/* device side (thorin_dev.hip) */
int thorin_dev_create(thorin_ctx *c);
void thorin_dev_destroy(thorin_ctx *c);
int thorin_dev_upload_init(thorin_ctx *c);
int thorin_dev_upload_pebbles(thorin_ctx *c);

/* device pointers of the energy system that thorin_rad_system assembled
 * (rings 1..nr-2 of nr x ns doubles), the current T and the T0 it started
 * from, and the context's HIP stream */
typedef struct thorin_sysptrs {
  int nr, ns;
  const double *ml, *mlip, *mlim, *mljp, *mljm, *rhs;
  const double *T, *T0;
  void *stream;
} thorin_sysptrs;
int thorin_dev_system(thorin_ctx *c, thorin_sysptrs *p);
int thorin_cnn_predict_dev(thorin_cnn *n, thorin_ctx *c, double *Tg);

#ifdef __cplusplus
}
#endif
#endif
