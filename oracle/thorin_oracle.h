/*
 * thorin_oracle.h -- CPU oracle for the FARGO_THORIN hydro step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a sequential C99 restatement of the
 * reference's AlgoGas (src_main/SourceEuler.c:210-317) and everything it
 * calls for the in-scope options.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product (fargo_thorin_amd/libthorin_amd.so) never links or calls it.
 *
 * Parity status of each function is stated in oracle/thorin_oracle.c next to
 * the function.
 */
#ifndef THORIN_ORACLE_H
#define THORIN_ORACLE_H
#include "../include/thorin_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

int  oracle_create(const thorin_params *p, oracle_ctx **out);
void oracle_destroy(oracle_ctx *o);
/* InitializeFromFile route of Initialization (Init.c:86-108,131-133):
 * arrays are nrad*nsec, temper is converted to energy. */
int  oracle_init_from_arrays(oracle_ctx *o, const double *dens, const double *vrad,
                             const double *vtheta, const double *temper);
/* Pebble disk (Init.c:138-144): mode 0 = coagulation-drift equilibrium
 * from the current gas state (EquilPebbleDisk); mode 1 = restart from the
 * given nrad*nsec arrays (RestartPebbleDisk + BckpFieldsForBC). */
int  oracle_init_pebbles(oracle_ctx *o, int mode, const double *pdens, const double *pvrad,
                         const double *pvtheta, const double *psize);
int  oracle_algogas(oracle_ctx *o, thorin_planets *sys, thorin_nbody_fn nbody, void *user,
                    thorin_step_log *log);
/* field ids of thorin_amd.h; count = (nrad+1)*nsec doubles */
int  oracle_get_field(oracle_ctx *o, int field, double *out, size_t count);
int  oracle_set_field(oracle_ctx *o, int field, const double *in, size_t count);
int  oracle_get_state(oracle_ctx *o, thorin_state *s);
int  oracle_set_state(oracle_ctx *o, const thorin_state *s);
/* 1D arrays: 0 Radii, 1 Rinf, 2 Rsup, 3 Rmed, 4 Surf, 5 InvDiffRmed, 6 SigmaMed,
 * 7 SigmaInf, 8 EnergyMed, 9 WaveKiller, 10 globcsvec, 11 globtempervec,
 * 12 VthetaMed, 13 Nshift (as double), 14 VMed, 15 PebDensInit, 16 PebVradInit,
 * 17 PebVthetaInit; count = nrad+1 */
int  oracle_get_1d(oracle_ctx *o, int which, double *out, size_t count);
double oracle_gas_total_mass(oracle_ctx *o);
/* Individual stages for unit tests of the GPU kernels. */
int  oracle_condition_cfl(oracle_ctx *o, double deltaT, double invdtreb_sq, double *newdt);

#ifdef __cplusplus
}
#endif
#endif
