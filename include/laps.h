/*
 * laps.h -- C ABI of the MI355X-native LAPS time step.
 *
 * The reference (chenshihelio/LAPS, src_compressible/) has no plugin or
 * operator API: its hot path is a set of argument-less Fortran module
 * subroutines acting on module-global arrays, called from `program mhd`.
 * Each entry point below replaces one of those subroutines (or a fixed
 * sequence of them); the citation names the reference file:line it replaces.
 * The Fortran host (host/mhd_host.f90) binds these through iso_c_binding;
 * Python binds them through ctypes (laps_amd/runtime.py).
 *
 * Conventions
 *  - Scalars by value, arrays as caller-owned contiguous host buffers in the
 *    reference's Fortran order: real fields (nx,ny,nz,nvar), x fastest;
 *    spectral fields (nx/2+1,ny,nz,nvar) of interleaved complex doubles.
 *  - The context owns every device allocation. One host thread per context,
 *    one HIP stream per context.
 *  - Every call returns LAPS_OK (0) or a negative LAPS_E* code; the message
 *    is available from laps_last_error(). Nothing aborts across the ABI.
 *  - laps_step() is asynchronous; every laps_get_* / reduction call
 *    synchronises the context's stream before returning.
 */
#ifndef LAPS_H
#define LAPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LAPS_OK = 0,
    LAPS_EINVAL = -1,       /* bad argument or unsupported parameter      */
    LAPS_ENOMEM = -2,       /* device allocation failed                   */
    LAPS_EDEVICE = -3,      /* HIP runtime error                          */
    LAPS_ESTATE = -4,       /* call out of order (e.g. step before state) */
};

enum { LAPS_FP64 = 0, LAPS_MIXED = 1 };

/* status word returned by laps_get_scalars */
enum {
    LAPS_RUNNING = 0,
    LAPS_DT_STOP = 1,       /* dt < 1e-8: the step was not taken (mhd.f90:216-242) */
};

/*
 * Run parameters: the reference's namelist variables that the time step
 * reads (mhd.f90:30-41), with the same meaning and units.
 */
typedef struct laps_params {
    int32_t nx, ny, nz;             /* &grid                                   */
    int32_t precision;              /* LAPS_FP64 or LAPS_MIXED                 */
    double Lx, Ly, Lz;              /* &grid                                   */
    double adiabatic_index;         /* &phys                                   */
    double cfl;                     /* &numerical                              */
    int32_t dealias_option;         /* &numerical: 1 = 2/3 sphere, 2 = filter  */
    int32_t if_resis, if_resis_exp; /* &phys / &numerical                      */
    int32_t if_visc, if_visc_exp;
    int32_t if_conserve_background; /* &numerical                              */
    double afx, afy, afz;           /* &numerical (compact filter alpha)       */
    double resistivity, viscosity;  /* &phys                                   */
    int32_t if_hall, if_aeb;        /* &Hall, &AEB                             */
    int32_t if_corotating, device;  /* &AEB; HIP device ordinal                */
    double ion_inertial_length;     /* &Hall                                   */
    double radius0, Ur0;            /* &AEB                                    */
    double corotating_angle;        /* &AEB                                    */
} laps_params;

typedef struct laps_ctx laps_ctx;

/* Module defaults of the reference: mhdinit.f90:5-27, dealiasing.f90:8-10,
 * AEBmod.f90:10-12, mhd.f90:21. */
void laps_default_params(laps_params* p);

/* Allocates device state for one periodic box. Replaces parallel_start,
 * fftw_initialize, grid_initialize, arrays_initialize, AEB_initialize and
 * dealias_initialize (mhd.f90:58-99; parallel.f90:28-212; fftw.f90:21-39;
 * mhdinit.f90:58-181; AEBmod.f90:16-44; dealiasing.f90:13-67).
 * Supported lengths per axis: 16, 32, 64, 96, 128, 256, 512. */
int laps_create(const laps_params* p, laps_ctx** out);

/* Frees every device allocation of the context (fftw_finalize, fftw.f90:225). */
void laps_destroy(laps_ctx* ctx);

/* Message of the last failing call on this thread. */
const char* laps_last_error(void);

/* Uploads the conserved state uu (nx,ny,nz,8) and primitives uu_prim
 * (nx,ny,nz,4) as produced by initial_calc_conserve_variable
 * (mhdinit.f90:1038-1056) and forms the un-dealiased spectral state
 * uu_fourier = FFT(uu)/N (transform_uu_real_to_fourier, fftw.f90:42-71;
 * order of mhd.f90:121-122). The first substage of the next step uses these
 * real arrays directly, as the reference does. Resets time, istep, the
 * previous-dt memory used by the CFL hysteresis (dt = 0, mhd.f90:135) and
 * the stop status. */
int laps_set_state(laps_ctx* ctx, const double* uu, const double* uu_prim);

/* Sets the simulation time (restart path: t_restart, mhd.f90:101-108) and
 * the previous dt used by the hysteresis; the radius follows evolve_radius
 * (AEBmod.f90:56-73). */
int laps_set_time(laps_ctx* ctx, double time, double dt_prev);

/* dt > 0: every following step uses this fixed dt and skips the CFL
 * reduction; dt <= 0: CFL-controlled dt (vardt, mhd.f90:328-429). */
int laps_set_fixed_dt(laps_ctx* ctx, double dt);

/* CFL reduction of vardt (mhd.f90:328-429) on the current real-space state,
 * without changing the context: dtmin_raw is the minimum before the cfl
 * factor, dt_next the dt the next step would use after the 2% hysteresis. */
int laps_cfl(laps_ctx* ctx, double* dtmin_raw, double* dt_next);

/* Advances nsteps RK3 steps on the device. One step is the main-loop body
 * of mhd.f90:216-285: the dt<1e-8 stop test, evolve (three substages of
 * calc_flux, transform_flux_real_to_fourier, calc_rhs, rkt, dealias,
 * transform_uu_fourier_to_real, update_uu_prim_from_uu; mhd.f90:298-326),
 * time += dt, istep += 1, evolve_radius(time), and the vardt of the next
 * step (performed at the start of that step on the same state). */
int laps_step(laps_ctx* ctx, int nsteps);

/* Synchronises and reads the step scalars: time, the dt of the last step
 * taken (before any step: 0), radius, istep and the status word. */
int laps_get_scalars(laps_ctx* ctx, double* time, double* dt, double* radius,
                     int64_t* istep, int32_t* status);

/* Real-space state uu (nx,ny,nz,8) and uu_prim (nx,ny,nz,4), as left by
 * transform_uu_fourier_to_real + update_uu_prim_from_uu (fftw.f90:73-103,
 * mhdrhs.f90:282-294). Either pointer may be NULL. */
int laps_get_state(laps_ctx* ctx, double* uu, double* uu_prim);

/* Output staging (the write side of output_uu, mhdoutput.f90:72-131; the restart files of
 * restart.f90:17-63 have the same payload). laps_stage_output forms the real-space state of
 * the current step on the device and queues its copy into a pinned host buffer of the context
 * on a stream of its own, so that laps_step calls made after it run while the copy proceeds.
 * The payload is the (nx,ny,nz,8) FP64 array output_uu writes: rho, rho u (or u when
 * primitive != 0, output_primitive), B, e (or p). laps_staged_output waits for the copy and
 * returns a pointer to that payload (valid until the next laps_stage_output or
 * laps_destroy) with the step's time, dt and istep. */
int laps_stage_output(laps_ctx* ctx, int primitive);
int laps_staged_output(laps_ctx* ctx, const double** payload, double* time, double* dt, int64_t* istep);

/* Spectral state uu_fourier (nx/2+1,ny,nz,8), interleaved complex. */
int laps_get_spectral(laps_ctx* ctx, double* uhat);

/* calc_rms (mhdrms.f90:53-126): means of (rho,ux,uy,uz,Bx,By,Bz,p), their
 * variances <f^2>-<f>^2 and <rho (u_i - <u_i>)^2>, reduced in FP64. */
int laps_moments(laps_ctx* ctx, double ave[8], double var[8], double rho_u2[3]);

/* Budgets; no counterpart in the reference, whose diagnostics are calc_rms
 * and calc_max_divB (mhdrms.f90:53-126, mhd.f90:522-570). Volume averages in
 * FP64: out[0] <e>, out[1] kinetic <(rho u).u>/2, out[2] magnetic <|B|^2>/2,
 * out[3] internal <p>/(gamma-1), out[4] cross helicity <u.B>, and by Parseval
 * over the spectrum without k = 0: out[5] magnetic helicity <A.B> with the
 * Coulomb-gauge potential A_hat = i k x B_hat / |k|^2, out[6] current
 * helicity <J.B>, J_hat = i k x B_hat (k with the expanding-box scaling). */
#define LAPS_NBUDGET 7
int laps_budgets(laps_ctx* ctx, double out[LAPS_NBUDGET]);

/* calc_max_divB (mhd.f90:522-570): max over modes of |k . B_hat|. */
int laps_max_divb(laps_ctx* ctx, double* out);

/* 3D transforms with the reference's conventions (fftw.f90:42-103):
 * forward = x r2c, y c2c, z c2c, each divided by n along its axis, sign -1;
 * inverse = unnormalised, sign +1. nfields fields of (nx,ny,nz) real /
 * (nx/2+1,ny,nz) complex. Used by the parity tests. */
int laps_fft_r2c(laps_ctx* ctx, int nfields, const double* real_in, double* spec_out);
int laps_fft_c2r(laps_ctx* ctx, int nfields, const double* spec_in, double* real_out);

/* Blocks until all work queued on the context's stream has finished. */
int laps_sync(laps_ctx* ctx);

/* Instrumentation; no counterpart in the reference, whose only timing is
 * system_clock around the whole loop (mhd.f90:56, :125-129, :431-457).
 * laps_timing_enable(ctx, 1) resets the counters and brackets every kernel
 * of laps_step with HIP events on the context's stream; laps_timing_read
 * returns, for kernel class `index` (0, 1, ... until *name is NULL), its
 * total device time, the algorithmic bytes its launches had to move (reads
 * plus writes of their buffers) and the launch count. */
int laps_timing_enable(laps_ctx* ctx, int enable);
int laps_timing_read(laps_ctx* ctx, int index, const char** name, double* total_ms,
                     double* total_bytes, int64_t* launches);

/* The context's HIP stream (a hipStream_t), for callers that order their
 * own work after laps_step. No counterpart in the reference. */
int laps_get_stream(laps_ctx* ctx, void** stream);

/* Ensembles of independent boxes. No counterpart in the reference, which runs
 * one box per MPI job (mhd.f90:56-167); callers: parameter scans and
 * ensemble statistics, many small boxes on one GPU.
 *
 * laps_ens_create makes nbox contexts from params[0..nbox-1] (any mix of
 * grids, physics switches and precisions; one device). laps_ens_box returns
 * box i, a context owned by the ensemble: every per-context call above
 * (laps_set_state, laps_get_*, diagnostics, laps_set_fixed_dt) applies to it,
 * and each box keeps its own CFL dt, time and stop status. laps_ens_step
 * advances every box by nsteps RK3 steps: a first step after any box's
 * laps_set_state runs per box; then the steps of all boxes are captured as
 * hipGraphs of 8 steps and of 1 step (each launch group a branch, one launch
 * per kernel for all boxes of the group) and replayed, 8 steps per launch
 * while at least 8 remain. Like laps_step it is asynchronous; per-box reads
 * synchronise with it. laps_ens_destroy destroys the boxes too. */
typedef struct laps_ens laps_ens;
int laps_ens_create(int nbox, const laps_params* params, laps_ens** ens);
void laps_ens_destroy(laps_ens* ens);
int laps_ens_size(const laps_ens* ens);
/* Number of launch groups: boxes with identical parameters are stepped together, one launch
 * per kernel for the whole group (their step buffers sit at a fixed stride in pools of the
 * ensemble); a box whose parameters match no other is a group of one. */
int laps_ens_groups(const laps_ens* ens);
laps_ctx* laps_ens_box(laps_ens* ens, int i);
int laps_ens_step(laps_ens* ens, int nsteps);
int laps_ens_sync(laps_ens* ens);

/* Learned subgrid closure. No counterpart in the reference, whose fluxes are
 * the resolved ones only (mhdrhs.f90:60-119); callers: coarse-grid runs
 * that stand in for finer ones. Features per point (28): rho, u, B and
 * their central-difference gradients d/dx, d/dy, d/dz (rho, ux, uy, uz,
 * Bx, By, Bz), each (f - mean)/std. Network 28 -> 64 (tanh) -> 64 (tanh) ->
 * 18 in FP32: subgrid Reynolds stress (xx, xy, xz, yy, yz, zz), Maxwell
 * stress (same order), EMF (x, y, z) and energy flux (x, y, z), each times
 * ystd. Parameter vector: W1[28][64], b1[64], W2[64][64], b2[64], W3[64][18],
 * b3[18] (the network, LAPS_MLP_NPARAM floats), then xmean[28], xstd[28],
 * ystd[18]. laps_closure_set (NULL: off) makes every following step evaluate
 * the network on the state at the start of the step and add -div(R - M + q I)
 * to the momentum, -curl(EMF) to the induction and -div(energy flux) to the
 * energy right-hand side of its three substages, with q the subgrid part of
 * p + B^2/2. AEB is not supported. laps_closure_features returns the raw
 * (unnormalised) features (nx*ny*nz, 28); laps_closure_eval the normalised
 * outputs (nx*ny*nz, 18) and, if rhs != NULL, the RHS term (nx/2+1, ny, nz, 7:
 * momentum, B, energy) of the current state.
 * laps_mlp_*: the network on caller samples (rows of 28 inputs, 18 targets)
 * on device `device`: forward; gradient of the mean squared error with its
 * loss; `steps` Adam steps (lr, betas 0.9/0.999) over consecutive batches of
 * the (caller-shuffled) samples, with the loss of each batch. */
#define LAPS_MLP_NPARAM 7186
#define LAPS_CLOSURE_NPARAM (LAPS_MLP_NPARAM + 74)
int laps_closure_set(laps_ctx* ctx, const float* params, int nparams);
int laps_closure_features(laps_ctx* ctx, float* features);
int laps_closure_eval(laps_ctx* ctx, float* outputs, double* rhs);
/* Training data of the closure, no counterpart in the reference: the box filter (mean over r^3
 * fine cells, r dividing the grid) of the current state and of its flux products, formed on
 * the device. out (nx/r, ny/r, nz/r, 26): rho, rho u (3), B (3), e; rho u_i u_j and B_i B_j
 * (xx, xy, xz, yy, yz, zz each); E = -u x B + (d_i/rho) J x B (3, J = curl B, d_i = 0 without
 * Hall); the energy flux (e + p + B^2/2) u - (u.B) B (3) (the fluxes of mhdrhs.f90:60-119). */
int laps_closure_filter(laps_ctx* ctx, int r, double* out);
/* The same training data under the sharp spectral cut-off of a coarser grid (no counterpart in
 * the reference): the modes kx < ncx/2, |ky| < ncy/2, |kz| < ncz/2 of the fine context `ctx`,
 * taken to the points of the coarse context `coarse` (same device, no finer grid) by its own
 * inverse transform. out (ncx, ncy, ncz, 52): fields 0-25 the cut of the 26 fields above;
 * fields 26-51 the cut of the same 26 fields formed, on the fine grid and so without aliasing,
 * from the cut state (u = cut(rho u) / cut(rho), p from the cut e) and its current. Subgrid
 * targets are the differences of fields 8-25 and 34-51. */
int laps_closure_filter_spectral(laps_ctx* ctx, laps_ctx* coarse, double* out);
int laps_mlp_forward(int device, int64_t n, const float* X, const float* params, float* Y);
int laps_mlp_grad(int device, int64_t n, const float* X, const float* T, const float* params, float* grad,
                  float* loss);
int laps_mlp_train(int device, int64_t n, const float* X, const float* T, float* params, int steps, int batch,
                   float lr, float* loss_hist);

/* Gradient of the network without allocation. No counterpart in the reference; callers that
 * cannot allocate device memory: code under HIP graph capture, tight training loops, harnesses
 * that forbid hipMalloc. Same result as laps_mlp_grad, but X (n, 28), T (n, 18), params,
 * grad (LAPS_MLP_NPARAM) and loss (1 float, may be NULL) are DEVICE pointers of `device`, and
 * the call is asynchronous on `stream` (a hipStream_t; NULL: the null stream). The activations
 * and deltas live in a fixed device array of the library (LAPS_MLP_NOALLOC_MAX samples, 9.6 MB
 * on each device it runs on, held until the process ends): n above the limit is LAPS_EINVAL.
 * Calls on one device share that array, so calls that overlap on one device race. */
#define LAPS_MLP_NOALLOC_MAX 8192
int laps_mlp_grad_noalloc(int device, int64_t n, const float* X, const float* T, const float* params,
                          float* grad, float* loss, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* LAPS_H */
