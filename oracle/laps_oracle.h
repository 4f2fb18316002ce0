/*
 * laps_oracle.h -- CPU restatement (plain C, FP64) of the reference LAPS
 * compressible Hall-MHD time step. TEST INFRASTRUCTURE ONLY: the reference used
 * by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
 * The product library (laps_amd/liblaps_amd.so) never links or calls it.
 *
 * Every function follows a reference subroutine of
 * chenshihelio/LAPS src_compressible/ (cited per function in laps_oracle.c).
 * Array layouts are the reference's Fortran orders: real (nx,ny,nz,nvar),
 * spectral (nx/2+1,ny,nz,nvar) interleaved complex, x fastest.
 */
#ifndef LAPS_ORACLE_H
#define LAPS_ORACLE_H

#include <stdint.h>
#include "../include/laps.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

oracle* oracle_create(const laps_params* p);
void oracle_destroy(oracle* o);
void oracle_set_threads(int nthreads);

/* state management (same meaning as the laps_* entry points) */
int oracle_set_state(oracle* o, const double* uu, const double* uu_prim);
int oracle_set_spectral(oracle* o, const double* uhat);
int oracle_set_time(oracle* o, double time, double dt_prev);
int oracle_set_fixed_dt(oracle* o, double dt);
int oracle_step(oracle* o, int nsteps);
int oracle_get_scalars(oracle* o, double* time, double* dt, double* radius,
                       int64_t* istep, int32_t* status);
int oracle_get_state(oracle* o, double* uu, double* uu_prim);
int oracle_get_spectral(oracle* o, double* uhat);
int oracle_cfl(oracle* o, double* dtmin_raw, double* dt_next);
int oracle_moments(oracle* o, double ave[8], double var[8], double rho_u2[3]);
int oracle_max_divb(oracle* o, double* out);

/* pieces exposed for the pinning tests */
int oracle_rhs(oracle* o, double* fnl_out);               /* calc_flux..calc_rhs on the current state */
int oracle_flux(oracle* o, double* flux_out);             /* calc_flux real fluxes (nx,ny,nz,19) */
int oracle_dealias_apply(oracle* o, double* uhat_inout);  /* dealias on an (nx/2+1,ny,nz,8) array */
int oracle_filters(oracle* o, double* fx, double* fy, double* fz);
int oracle_ksquare(oracle* o, double* ksq_out);
void oracle_fft1d(int n, int sign, const double* in, double* out);
int oracle_fft_r2c(oracle* o, int nfields, const double* real_in, double* spec_out);
int oracle_fft_c2r(oracle* o, int nfields, const double* spec_in, double* real_out);
void oracle_cons_from_prim(const laps_params* p, int64_t npts, double* uu, double* uu_prim);

#ifdef __cplusplus
}
#endif
#endif
