// internal.hpp -- host-side declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "fft_dev.hpp"

namespace laps {

// Strided batched C2C along one axis. Lines run along the axis with element
// stride `ls`; `ncol` adjacent lines (stride 1) form a column tile; two batch
// dimensions select the tile. Stored positions may be pruned (cnt < N).
struct Ctl;

struct StridedArgs {
    int ncol;
    int in_cnt, out_cnt;
    long long in_ls, out_ls;
    int nb1;
    long long in_b1, out_b1;
    int nb2;
    long long in_b2, out_b2;
    // optional grid shift (all three or none): the input at position p of
    // column col in batch slab b1 is multiplied by ph[p] * phc[col] * phb[b1];
    // with two chains only chain 1 is shifted
    const cpx<double>* ph;
    const cpx<double>* phc;
    const cpx<double>* phb;
    // nch = 2: a second chain of the same shape at in + in_ch, out + out_ch in
    // the same launch (0 or 1 counts as one chain); pair: chain 0 and chain 1 of a
    // slab run side by side on one XCD (for inputs the two chains share)
    int nch;
    long long in_ch, out_ch;
    bool pair;
    // dual (nch = 2, both chains read the same input, chain 1 shifted): one work-group runs
    // both chains of its slab, loading the input once
    bool dual;
    // optional rows of the dealiasing mask (mask 1: 2/3 sphere, 3: phase-shift ellipsoid) for
    // a kept-box state that is zero outside the kept set: a row (kx = col, my) with no kept
    // mode even at mz = 0 is all zero. row_role 1 (z inverse, slab b1 = stored ky row of
    // cy_cnt): such rows are neither loaded nor stored; 2 (y inverse): they are taken as
    // zero without loading them. The test is integer arithmetic on the grid lengths.
    int mask, row_role, cy_cnt;
    long long mX, mY, mZ;
    // batched boxes (an ensemble group of identical contexts whose buffers sit at a fixed
    // stride): the launch covers nbox boxes (grid y), box b at in + b * in_bx, out + b * out_bx
    int nbox;
    long long in_bx, out_bx;
    // curl (z inverse of a kept-box state with Hall): input fields 8-10 are not read; they are
    // J = i k' x B formed from fields 4-6 as they are loaded (k' of the kept-box tables kw*,
    // with the expanding-box / corotating scaling of the box's control block, one per grid row)
    int curl;
    const double* kwx;
    const double* kwy;
    const double* kwz;
    const Ctl* kctl;
    double k_r0, k_cos, k_sin;
    int k_corot;
};

// true when no mode of row (mx, my) is kept: its mz = 0 mode is outside the mask (the
// sphere test is strict, so rows at exactly r = 1/3 count as kept)
__host__ __device__ __forceinline__ bool mask_row_dead(int mask, long long X, long long Y, long long Z,
                                                       long long mx, long long my) {
    if (mask == 3) return 9 * (mx * mx * Y * Y * Z * Z + my * my * X * X * Z * Z) >= 2 * X * X * Y * Y * Z * Z;
    if (mask == 1) return 9 * (mx * mx * Y * Y + my * my * X * X) > X * X * Y * Y;
    return false;
}

// Rows of an x transform: `nrows` rows (even), complex rows hold `kx` stored
// coefficients with stride `crow`, real rows hold N values with stride `rrow`.
struct XArgs {
    long long nrows;
    int kx;
    long long crow, rrow;
};

// Device control block of one box (scalars of mhd.f90 / AEBmod.f90).
// The CFL minimum of a step is folded by 64-bit atomicMin on the bit patterns
// of non-negative doubles, spread over kCflSlots addresses (work-group index
// modulo kCflSlots) so that the waves of a pass do not queue on one address;
// k_ctl_finalize folds the slots.
constexpr int kCflSlots = 256;
struct Ctl {
    double time, dt, dt_prev, radius, tau, r0r;
    double cc[3], dd[3], ts[3];
    unsigned long long dtmin_bits;
    unsigned long long dtmin_slot[kCflSlots];
    long long istep;
    int status;
    int ksq_updated;
};

// Kept-mode box of the spectral state: kx in [0,KX), ky/kz stored compactly.
// Rows are stored with stride KXP >= KX (KXP a multiple of 16 for the kept
// box, so every 128-byte segment of a row is one cache line); the entries
// kx in [KX, KXP) are padding and stay zero.
struct Box {
    int KX, CY, CZ, KXP;
    __host__ __device__ long long modes() const { return (long long)KXP * CY * CZ; }
};

// Per-box wave-number and dealias tables (device pointers).
struct BoxTables {
    const double* wx;      // [KX]
    const double* wy;      // [CY]
    const double* wz;      // [CZ]
    const int* kzlim;      // [CY*KX], option 1: keep |mz| <= kzlim (-1: none)
    const int* ylim;       // [KX]: the rows (kx, my) holding a kept mode are |my| <= ylim
    const double* fx;      // [KX] option 2 filters
    const double* fy;      // [CY]
    const double* fz;      // [CZ]
    const int* mz;         // [CZ] signed integer index of stored z row
};

struct PhysParams {
    double gamma, di, nu, eta, radius0, Ur, cos_c, sin_c, cfl, fixed_dt;
    double dx, dy, dz;
    int hall, aeb, corot_k, corot_ksq, visc, visc_exp, resis, resis_exp, conserve_bg, dealias;
};

template <typename T>
struct Twiddles {
    const cpx<T>* x;
    const cpx<T>* y;
    const cpx<T>* z;
};

// --- FFT launchers (fft_kernels.hip) ---
// variant bit 0: 0 two stages where possible (Plan), 1 radices <= 8 (PlanLow); bit 1 (inverse
// passes only): the other tile width (half the columns, or twice as many where the default has
// fewer than 16)
template <typename TI, typename T>
hipError_t launch_strided(int n, int dir, const cpx<TI>* in, cpx<T>* out, const StridedArgs& a,
                          const cpx<T>* tw, T scale, hipStream_t s, int variant = 0);
template <typename T>
hipError_t launch_x_c2r(int n, const cpx<T>* in, T* out, const XArgs& a, const cpx<T>* tw, hipStream_t s);
template <typename T>
hipError_t launch_x_r2c(int n, const T* in, cpx<T>* out, const XArgs& a, const cpx<T>* tw, hipStream_t s);

// --- fused x pass (xfused.hip) ---
struct XFArgs {
    long long nrows;          // nz*ny rows per field
    long long npts;           // nx*ny*nz
    const void* st;           // state spectral rows (T complex): kx_s stored of row stride rs_s
    int kx_s, rs_s;
    long long fs_s;           // field stride of st (elements)
    const void* jb;           // Hall current spectral rows
    int kx_j, rs_j;
    long long fs_j;
    const double* S;          // first substage: uploaded uu(8) + uu_prim(4), FP64 real
    void* out;                // flux spectral rows (T complex)
    int kx_o, rs_o;
    long long fs_o;
    // nch = 2: a second chain of rows at st + st_ch, jb + jb_ch, out + out_ch (elements) in the
    // same launch, without the CFL reduction (0 or 1 counts as one chain)
    int nch;
    long long st_ch, jb_ch, out_ch;
    // batched boxes (grid y): box b at st + b * st_bx, jb + b * jb_bx, out + b * out_bx and
    // control block ctl + b
    int nbox;
    long long st_bx, jb_bx, out_bx;
};
template <typename T>
hipError_t launch_xfused(int nx, bool first, bool cfl, const XFArgs& a, const cpx<T>* tw, const PhysParams& pp,
                         Ctl* ctl, hipStream_t s);

// --- forward y / z passes with the RHS and RK3 folded in (fwd_fused.hip) ---
struct YGroup {
    int kind;     // 0 divergence (3 in, 2 out), 1 induction (3 in, 5 out), 2 copy (1 in, 1 out),
                  // 3 symmetric momentum flux tensor (6 in, 6 out)
    int in[6];
    int out[6];
};
struct YGArgs {
    int ncol, cy;
    long long rs;
    long long in_fs;
    long long out_fs;
    int ngroups;
    YGroup g[8];
    // nch = 2: a second chain at in + in_ch, out + out_ch in the same launch
    int nch;
    long long in_ch, out_ch;
    // optional [ncol]: output rows (kx, my) with |my| > ylim[kx] hold no kept mode and
    // are not stored (the z pass does not read them)
    const int* ylim;
    // batched boxes (grid y): box b at in + b * in_bx, out + b * out_bx, control block ctl + b
    int nbox;
    long long in_bx, out_bx;
};
struct ZTerm {
    int field;
    int mult;     // 0: 1, -1: -i kz', +1: +i kz'
};
struct ZVar {
    int nterm;
    ZTerm t[3];
};
struct ZArgs {
    int KX, CY, CZ;
    long long rs;
    long long in_fs;
    long long ufs;
    ZVar v[8];
    const void* Ds;           // optional second input (T complex, layout of D) from the shifted grid:
    const cpx<double>* ph;    // fnl = (Z(D) + conj(ph[p] phc[kx] phb[cy]) Z(Ds)) / 2 at output mode
    const cpx<double>* phc;   // (kx, cy, p)
    const cpx<double>* phb;
    const cpx<double>* clos;  // optional learned-closure RHS [7][modes] added to momentum, B and e (v = 1..7)
    cpx<float>* Uf;           // optional FP32 copy of the updated state (mixed precision)
    // with a mask, write zero to the modes outside the kept set (U, Uf): needed once after a
    // state upload, which leaves un-dealiased values there; later passes find them zero
    int clear_outside;
    // batched boxes (grid y): box b at D/Ds + b * d_bx, U + b * u_bx, Fp + b * fp_bx,
    // Uf + b * uf_bx, control block ctl + b
    int nbox;
    long long d_bx, u_bx, fp_bx, uf_bx;
};
template <typename T>
hipError_t launch_yfwd_group(int ny, const cpx<T>* in, cpx<T>* out, const YGArgs& a, const cpx<T>* tw,
                             const double* wxt, const double* wyf, const PhysParams& pp, const Ctl* ctl, int nz,
                             hipStream_t s);
template <typename T>
hipError_t launch_zfwd_rhs(int nz, const cpx<T>* D, cpx<double>* U, cpx<T>* Fp, const ZArgs& a,
                           const cpx<T>* tw, const BoxTables& tb, const double* wzf, const PhysParams& pp,
                           const Ctl* ctl, int irk, T scale, hipStream_t s);

// --- physics launchers (physics_kernels.hip) ---
// The step's control kernels take nbox batched boxes (ensemble groups): control block ctl + b.
// J = i k' x B of one box (the first substage after a state upload, the closure's filters;
// later substages form J in the z inverse)
hipError_t launch_curl(const cpx<double>* B, cpx<double>* J, long long fstride, Box b, BoxTables t,
                       const PhysParams& pp, const Ctl* ctl, hipStream_t s);
hipError_t launch_ctl_begin(Ctl* ctl, hipStream_t s, int nbox = 1);
hipError_t launch_ctl_finalize(Ctl* ctl, const PhysParams& pp, hipStream_t s, int nbox = 1);
hipError_t launch_ctl_end(Ctl* ctl, const PhysParams& pp, hipStream_t s, int nbox = 1);
template <typename T>
hipError_t launch_prim(const T* R, double* P, long long npts, double gamma, hipStream_t s);
hipError_t launch_cfl(const double* U8, const double* P4, long long npts, const PhysParams& pp, Ctl* ctl,
                      hipStream_t s);
hipError_t launch_moments(const double* U8, const double* P4, long long npts, double* partial, int nblk,
                          hipStream_t s);
hipError_t launch_moments2(const double* U8, const double* P4, long long npts, const double* ave, double* partial,
                           int nblk, hipStream_t s);
hipError_t launch_budget_real(const double* U8, const double* P4, long long npts, double gamma, double* partial,
                              int nblk, hipStream_t s);
hipError_t launch_budget_spec(const cpx<double>* U, long long fs, Box b, BoxTables t, const PhysParams& pp,
                              const Ctl* ctl, int nx, double* partial, int nblk, hipStream_t s);
hipError_t launch_divb(const cpx<double>* U, long long fstride, Box b, BoxTables t, const PhysParams& pp,
                       const Ctl* ctl, unsigned long long* out, hipStream_t s);
hipError_t launch_gather_box(const cpx<double>* full, cpx<double>* box, int nf, long long full_stride,
                             long long box_stride, int nxh, int ny, int nz, Box b, hipStream_t s);
template <typename TI, typename TO>
hipError_t launch_convert(const TI* in, TO* out, long long n, hipStream_t s);

// --- learned closure (closure.hip) ---
constexpr int kClIn = 28, kClH = 64, kClOut = 18;
constexpr int kClNParam = kClIn * kClH + kClH + kClH * kClH + kClH + kClH * kClOut + kClOut;
constexpr int kClNNorm = 2 * kClIn + kClOut;  // xmean, xstd, ystd
constexpr int kClFilt = 26;  // box-filtered fields of laps_closure_filter
// field q (0..25, the numbering of launch_cl_filter) at every point of S/J, unfiltered
hipError_t launch_cl_point(const double* S, const double* J, long long N, int q, double di, double* out,
                           hipStream_t s);
// sharp spectral cut-off between full spectra: dir > 0 fine F -> coarse C, dir < 0 C -> F
hipError_t launch_spec_cut(cpx<double>* F, int nyf, int nzf, int kxpf, cpx<double>* C, int nxc, int nyc, int nzc,
                           int kxpc, int dir, hipStream_t s);
hipError_t launch_cl_filter(const double* S, const double* J, int nx, int ny, int nz, int r, double di, double* out,
                            hipStream_t s);
hipError_t launch_cl_features(const double* S, int nx, int ny, int nz, double dx, double dy, double dz,
                              const float* norm, float* X, hipStream_t s);
hipError_t launch_mlp_fwd(const float* X, long long n, const float* P, float* Y, float* H1, float* H2,
                          hipStream_t s);
hipError_t launch_mlp_grad(const float* X, const float* T, long long n, const float* P, const float* Y,
                           const float* H1, const float* H2, float* scratch, float* G, float* loss, hipStream_t s);
hipError_t launch_loss_mean(const float* sum, long long n, float* out, hipStream_t s);
hipError_t launch_adam(float* P, const float* G, float* m, float* v, int n, float lr, float b1t, float b2t,
                       hipStream_t s);
hipError_t launch_cl_terms(const float* Y, long long N, const float* ystd, double gamma, double* C, hipStream_t s);
hipError_t launch_cl_rhs(const cpx<double>* Ch, long long M, int nxh, int ny, const double* wx, const double* wy,
                         const double* wz, cpx<double>* R, hipStream_t s);

}  // namespace laps
