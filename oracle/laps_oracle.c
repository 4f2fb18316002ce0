/*
 * laps_oracle.c -- CPU restatement (plain C, FP64, OpenMP over pencils) of the
 * reference LAPS compressible Hall-MHD RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY. This file gives the reference results for the HIP path:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Nothing in laps_amd/ links, loads or calls it.
 *
 * Citations are to chenshihelio/LAPS, src_compressible/<file>:<line>.
 * The reference builds with -r8 (every REAL/COMPLEX is 8-byte); arithmetic
 * below keeps the reference's evaluation order where it is written out.
 *
 * Third-party boundary: the reference's 1D transforms are FFTW3 plans
 * (fftw.f90:27-33, version unpinned, not vendored, absent from this image).
 * They are restated here from the DFT definition with FFTW's conventions:
 * forward sign -1, unnormalised; c2r reads only the real parts of the DC and
 * Nyquist coefficients. Pinned by the tests against numpy.fft and a
 * brute-force DFT.
 */
#include "laps_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cx;

static inline cx cx_add(cx a, cx b) { cx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cx cx_sub(cx a, cx b) { cx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cx cx_mul(cx a, cx b) {
    cx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cx cx_scale(cx a, double s) { cx r = {a.re * s, a.im * s}; return r; }
static inline cx cx_divr(cx a, double d) { cx r = {a.re / d, a.im / d}; return r; }

/* ------------------------------------------------------------------------ */
/* 1D complex DFT: iterative radix-2 for 2^k, one radix-3 split for 3*2^k.   */
/* ------------------------------------------------------------------------ */

#define PI_L 3.141592653589793238462643383279502884L

/* exp(sign * 2 pi i t / n) with the angle reduced to the first octant in
 * long double, so quarter-turn values are exact. */
static cx unit_root(long t, long n, int sign) {
    t %= n;
    if (t < 0) t += n;
    long q8 = 8 * t;
    long oct = q8 / n;               /* theta = (pi/4) * (oct + rem/n) */
    long rem = q8 - oct * n;
    long double a = (oct % 2 == 0) ? PI_L / 4.0L * (long double)rem / (long double)n
                                   : PI_L / 4.0L * (long double)(n - rem) / (long double)n;
    long double c = cosl(a), s = sinl(a), cr, sr;
    switch (oct) {
        case 0: cr = c;  sr = s;  break;   /* a          */
        case 1: cr = s;  sr = c;  break;   /* pi/2 - a   */
        case 2: cr = -s; sr = c;  break;   /* pi/2 + a   */
        case 3: cr = -c; sr = s;  break;   /* pi - a     */
        case 4: cr = -c; sr = -s; break;   /* pi + a     */
        case 5: cr = -s; sr = -c; break;   /* 3pi/2 - a  */
        case 6: cr = s;  sr = -c; break;   /* 3pi/2 + a  */
        default: cr = c; sr = -s; break;   /* 2pi - a    */
    }
    cx r = {(double)cr, (double)(sign * sr)};
    return r;
}

/* cached forward roots exp(-2 pi i t / n), t < n */
#define MAX_TW 16
static int tw_n[MAX_TW];
static cx* tw_tab[MAX_TW];

static const cx* roots(int n) {
    const cx* found = NULL;
#pragma omp critical(oracle_roots)
    {
        for (int i = 0; i < MAX_TW && tw_tab[i]; ++i)
            if (tw_n[i] == n) { found = tw_tab[i]; break; }
        if (!found) {
            int slot = 0;
            while (slot < MAX_TW && tw_tab[slot]) ++slot;
            if (slot == MAX_TW) slot = 0;   /* recycle (never happens for <= 16 sizes) */
            cx* t = (cx*)malloc(sizeof(cx) * (size_t)n);
            for (int k = 0; k < n; ++k) t[k] = unit_root(k, n, -1);
            tw_n[slot] = n;
            tw_tab[slot] = t;
            found = t;
        }
    }
    return found;
}

static inline cx root_pow(const cx* w, long t, int n, int sign) {
    cx r = w[t % n];
    if (sign > 0) r.im = -r.im;
    return r;
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* in-place radix-2 DIT on x[0..n); n power of two; w = roots(n) */
static void fft_pow2(cx* x, int n, int sign, const cx* w) {
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cx t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                cx tw = root_pow(w, (long)k * step, n, sign);
                cx u = x[i + k];
                cx v = cx_mul(x[i + k + half], tw);
                x[i + k] = cx_add(u, v);
                x[i + k + half] = cx_sub(u, v);
            }
        }
    }
}

/* out = DFT_sign(in), unnormalised; n = 2^k or 3*2^k; work >= n entries */
static void fft_any(const cx* in, cx* out, int n, int sign, cx* work) {
    const cx* w = roots(n);
    if (is_pow2(n)) {
        memcpy(out, in, sizeof(cx) * (size_t)n);
        fft_pow2(out, n, sign, w);
        return;
    }
    int m = n / 3;
    const cx* wm = roots(m);
    /* x_r[j] = in[3j + r] ; X[k] = sum_r W_n^{rk} X_r[k mod m] */
    for (int r = 0; r < 3; ++r) {
        for (int j = 0; j < m; ++j) work[r * m + j] = in[3 * j + r];
        fft_pow2(work + r * m, m, sign, wm);
    }
    for (int k = 0; k < n; ++k) {
        cx acc = work[k % m];
        acc = cx_add(acc, cx_mul(root_pow(w, k, n, sign), work[m + k % m]));
        acc = cx_add(acc, cx_mul(root_pow(w, 2L * k, n, sign), work[2 * m + k % m]));
        out[k] = acc;
    }
}

void oracle_fft1d(int n, int sign, const double* in, double* out) {
    cx* work = (cx*)malloc(sizeof(cx) * (size_t)n);
    fft_any((const cx*)in, (cx*)out, n, sign, work);
    free(work);
}

/* ------------------------------------------------------------------------ */
/* Oracle state                                                             */
/* ------------------------------------------------------------------------ */

struct oracle {
    laps_params p;
    int nx, ny, nz, nxh;
    size_t N, M;
    double pi, dx, dy, dz;
    double *wx, *wy, *wz;        /* wave_number_x/y/z (mhdinit.f90:83-111) */
    double* ksq;                 /* k_square (nxh,ny,nz)                   */
    double *filtx, *filty, *filtz;
    double *uu, *prim, *flux, *J;
    cx *uhat, *fhat, *jhat, *fnl, *fnl_rk;
    /* phase-shift dealiasing (option 3): e^{i pi m/n} per axis and the shifted-grid buffers */
    cx *phx, *phy, *phz;
    double *uu_s, *prim_s;
    cx *uhat_s, *fhat_s;
    double time, dt, dt_last, radius, Ur, tau, cos_c, sin_c, fixed_dt;
    double cc1[3], dd1[3], tstep[3];
    int64_t istep;
    int32_t status;
};

static int g_threads = 0;
void oracle_set_threads(int n) { g_threads = n; }

#define IDX3(o, ix, iy, iz) ((size_t)(ix) + (size_t)(o)->nx * ((size_t)(iy) + (size_t)(o)->ny * (size_t)(iz)))
#define SDX3(o, ix, iy, iz) ((size_t)(ix) + (size_t)(o)->nxh * ((size_t)(iy) + (size_t)(o)->ny * (size_t)(iz)))

/* wave numbers: mhdinit.f90:79-111 ; 2*pi*(i-1)/L for i <= n/2+1, else 2*pi*(i-1-n)/L */
static void wave_numbers(double* w, int n, double L, double pi) {
    for (int i = 1; i <= n; ++i) {
        int m = (i <= n / 2 + 1) ? (i - 1) : (i - 1 - n);
        w[i - 1] = 2 * pi * (double)m / L;
    }
}

/* update_ksquare (AEBmod.f90:87-125) */
static void update_ksquare(oracle* o) {
    const double r0r = o->p.radius0 / o->radius;
    for (int iz = 0; iz < o->nz; ++iz)
        for (int iy = 0; iy < o->ny; ++iy)
            for (int ix = 0; ix < o->nxh; ++ix) {
                double kx = o->wx[ix], ky = o->wy[iy], kz = o->wz[iz];
                double v;
                if (o->p.if_corotating) {
                    double c = o->cos_c, s = o->sin_c;
                    v = kx * kx * (c * c + (s * o->p.radius0 / o->radius) * (s * o->p.radius0 / o->radius))
                        + ky * ky * (s * s + (c * o->p.radius0 / o->radius) * (c * o->p.radius0 / o->radius))
                        + kx * ky * 2 * c * s * (1 - r0r * r0r)
                        + (kz * o->p.radius0 / o->radius) * (kz * o->p.radius0 / o->radius);
                } else {
                    double a = ky * o->p.radius0 / o->radius, b = kz * o->p.radius0 / o->radius;
                    v = kx * kx + a * a + b * b;
                }
                o->ksq[SDX3(o, ix, iy, iz)] = v;
            }
}

/* AEB_calc + evolve_radius (AEBmod.f90:46-73) */
static void evolve_radius(oracle* o, double t) {
    o->radius = o->p.radius0 + o->Ur * t;
    o->tau = o->radius / o->Ur;
    update_ksquare(o);
}

/* dealias_initialize option 2 filter (dealiasing.f90:31-65) */
static double compact_filter(double w, double L, int n, double af) {
    double aj = (5.0 + 6.0 * af) / 8.0;
    double bj = (1.0 + 2.0 * af) / 2.0;
    double cj = -(1.0 - 2 * af) / 8.0;
    double wl = w * L / n;
    return (aj + bj * cos(wl) + cj * cos(2 * wl)) / (1 + 2 * af * cos(wl));
}

oracle* oracle_create(const laps_params* p) {
    oracle* o = (oracle*)calloc(1, sizeof(oracle));
    if (!o) return NULL;
    o->p = *p;
    o->nx = p->nx; o->ny = p->ny; o->nz = p->nz; o->nxh = p->nx / 2 + 1;
    o->N = (size_t)p->nx * p->ny * p->nz;
    o->M = (size_t)o->nxh * p->ny * p->nz;
    o->pi = 3.141592653589793;   /* mhdinit.f90:7 */
    o->dx = p->Lx / p->nx; o->dy = p->Ly / p->ny; o->dz = p->Lz / p->nz;
    o->wx = (double*)malloc(sizeof(double) * p->nx);
    o->wy = (double*)malloc(sizeof(double) * p->ny);
    o->wz = (double*)malloc(sizeof(double) * p->nz);
    wave_numbers(o->wx, p->nx, p->Lx, o->pi);
    wave_numbers(o->wy, p->ny, p->Ly, o->pi);
    wave_numbers(o->wz, p->nz, p->Lz, o->pi);
    o->ksq = (double*)malloc(sizeof(double) * o->M);
    o->uu = (double*)calloc(8 * o->N, sizeof(double));
    o->prim = (double*)calloc(4 * o->N, sizeof(double));
    o->flux = (double*)calloc(19 * o->N, sizeof(double));
    o->J = (double*)calloc(3 * o->N, sizeof(double));
    o->uhat = (cx*)calloc(8 * o->M, sizeof(cx));
    o->fhat = (cx*)calloc(19 * o->M, sizeof(cx));
    o->jhat = (cx*)calloc(3 * o->M, sizeof(cx));
    o->fnl = (cx*)calloc(8 * o->M, sizeof(cx));
    o->fnl_rk = (cx*)calloc(8 * o->M, sizeof(cx));
    if (!o->uu || !o->prim || !o->flux || !o->J || !o->uhat || !o->fhat || !o->jhat || !o->fnl ||
        !o->fnl_rk || !o->ksq) {
        oracle_destroy(o);
        return NULL;
    }
    /* grid_initialize k_square (mhdinit.f90:114-122) */
    for (int iz = 0; iz < o->nz; ++iz)
        for (int iy = 0; iy < o->ny; ++iy)
            for (int ix = 0; ix < o->nxh; ++ix)
                o->ksq[SDX3(o, ix, iy, iz)] = o->wx[ix] * o->wx[ix] + o->wy[iy] * o->wy[iy] + o->wz[iz] * o->wz[iz];
    /* mhd.f90:88-90 and AEB_initialize (AEBmod.f90:16-44) */
    double Ur0 = p->if_aeb ? p->Ur0 : 0.0;
    o->Ur = Ur0;
    o->radius = p->radius0;
    o->tau = o->radius / o->Ur;
    double ang = p->if_corotating ? p->corotating_angle : 0.0;
    o->cos_c = cos(ang);
    o->sin_c = sin(ang);
    /* dealias_initialize (dealiasing.f90:13-67) */
    o->filtx = (double*)malloc(sizeof(double) * o->nxh);
    o->filty = (double*)malloc(sizeof(double) * o->ny);
    o->filtz = (double*)malloc(sizeof(double) * o->nz);
    for (int i = 0; i < o->nxh; ++i) o->filtx[i] = compact_filter(o->wx[i], p->Lx, p->nx, p->afx);
    for (int i = 0; i < o->ny; ++i) o->filty[i] = compact_filter(o->wy[i], p->Ly, p->ny, p->afy);
    for (int i = 0; i < o->nz; ++i) o->filtz[i] = compact_filter(o->wz[i], p->Lz, p->nz, p->afz);
    o->fixed_dt = 0.0;
    if (p->dealias_option == 3) {
        /* half a cell along every axis: e^{i k delta} = e^{i pi m/n} per axis, signed m
         * with the +Nyquist convention of the wave numbers */
        o->phx = (cx*)malloc(sizeof(cx) * o->nxh);
        o->phy = (cx*)malloc(sizeof(cx) * o->ny);
        o->phz = (cx*)malloc(sizeof(cx) * o->nz);
        o->uu_s = (double*)calloc(8 * o->N, sizeof(double));
        o->prim_s = (double*)calloc(4 * o->N, sizeof(double));
        o->uhat_s = (cx*)calloc(8 * o->M, sizeof(cx));
        o->fhat_s = (cx*)calloc(19 * o->M, sizeof(cx));
        if (!o->phx || !o->phy || !o->phz || !o->uu_s || !o->prim_s || !o->uhat_s || !o->fhat_s) {
            oracle_destroy(o);
            return NULL;
        }
        for (int i = 0; i < o->nxh; ++i) { o->phx[i].re = cos(o->pi * i / o->nx); o->phx[i].im = sin(o->pi * i / o->nx); }
        for (int i = 0; i < o->ny; ++i) {
            int m = (i <= o->ny / 2) ? i : i - o->ny;
            o->phy[i].re = cos(o->pi * m / o->ny); o->phy[i].im = sin(o->pi * m / o->ny);
        }
        for (int i = 0; i < o->nz; ++i) {
            int m = (i <= o->nz / 2) ? i : i - o->nz;
            o->phz[i].re = cos(o->pi * m / o->nz); o->phz[i].im = sin(o->pi * m / o->nz);
        }
    }
    return o;
}

void oracle_destroy(oracle* o) {
    if (!o) return;
    free(o->wx); free(o->wy); free(o->wz); free(o->ksq);
    free(o->filtx); free(o->filty); free(o->filtz);
    free(o->uu); free(o->prim); free(o->flux); free(o->J);
    free(o->uhat); free(o->fhat); free(o->jhat); free(o->fnl); free(o->fnl_rk);
    free(o->phx); free(o->phy); free(o->phz); free(o->uu_s); free(o->prim_s); free(o->uhat_s); free(o->fhat_s);
    free(o);
}

/* ------------------------------------------------------------------------ */
/* 3D transforms (fftw.f90:42-222)                                          */
/* ------------------------------------------------------------------------ */

static int nthreads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* transform_uu_real_to_fourier for one field: x r2c /nx, y c2c /ny, z c2c /nz */
static void fwd3d(const oracle* o, const double* f, cx* out) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nxh = o->nxh;
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
#pragma omp parallel num_threads(nthreads())
    {
        cx* a = (cx*)malloc(sizeof(cx) * nmax);
        cx* b = (cx*)malloc(sizeof(cx) * nmax);
        cx* w = (cx*)malloc(sizeof(cx) * nmax);
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy) {
                for (int ix = 0; ix < nx; ++ix) { a[ix].re = f[IDX3(o, ix, iy, iz)]; a[ix].im = 0.0; }
                fft_any(a, b, nx, -1, w);
                for (int ix = 0; ix < nxh; ++ix) out[SDX3(o, ix, iy, iz)] = cx_divr(b[ix], (double)nx);
            }
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; ++iz)
            for (int ix = 0; ix < nxh; ++ix) {
                for (int iy = 0; iy < ny; ++iy) a[iy] = out[SDX3(o, ix, iy, iz)];
                fft_any(a, b, ny, -1, w);
                for (int iy = 0; iy < ny; ++iy) out[SDX3(o, ix, iy, iz)] = cx_divr(b[iy], (double)ny);
            }
#pragma omp for collapse(2) schedule(static)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nxh; ++ix) {
                for (int iz = 0; iz < nz; ++iz) a[iz] = out[SDX3(o, ix, iy, iz)];
                fft_any(a, b, nz, -1, w);
                for (int iz = 0; iz < nz; ++iz) out[SDX3(o, ix, iy, iz)] = cx_divr(b[iz], (double)nz);
            }
        free(a); free(b); free(w);
    }
}

/* transform_uu_fourier_to_real for one field: z c2c, y c2c, x c2r, unnormalised.
 * c2r uses the Hermitian extension and only Re of DC and Nyquist (FFTW c2r). */
static void inv3d(const oracle* o, const cx* in, double* f, cx* scratch) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nxh = o->nxh;
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    cx* t = scratch; /* (nxh,ny,nz) */
    memcpy(t, in, sizeof(cx) * o->M);
#pragma omp parallel num_threads(nthreads())
    {
        cx* a = (cx*)malloc(sizeof(cx) * nmax);
        cx* b = (cx*)malloc(sizeof(cx) * nmax);
        cx* w = (cx*)malloc(sizeof(cx) * nmax);
#pragma omp for collapse(2) schedule(static)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nxh; ++ix) {
                for (int iz = 0; iz < nz; ++iz) a[iz] = t[SDX3(o, ix, iy, iz)];
                fft_any(a, b, nz, +1, w);
                for (int iz = 0; iz < nz; ++iz) t[SDX3(o, ix, iy, iz)] = b[iz];
            }
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; ++iz)
            for (int ix = 0; ix < nxh; ++ix) {
                for (int iy = 0; iy < ny; ++iy) a[iy] = t[SDX3(o, ix, iy, iz)];
                fft_any(a, b, ny, +1, w);
                for (int iy = 0; iy < ny; ++iy) t[SDX3(o, ix, iy, iz)] = b[iy];
            }
#pragma omp for collapse(2) schedule(static)
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy) {
                for (int k = 0; k < nxh; ++k) a[k] = t[SDX3(o, k, iy, iz)];
                a[0].im = 0.0;
                if (nx % 2 == 0) a[nx / 2].im = 0.0;
                for (int k = nxh; k < nx; ++k) { a[k].re = a[nx - k].re; a[k].im = -a[nx - k].im; }
                fft_any(a, b, nx, +1, w);
                for (int ix = 0; ix < nx; ++ix) f[IDX3(o, ix, iy, iz)] = b[ix].re;
            }
        free(a); free(b); free(w);
    }
}

int oracle_fft_r2c(oracle* o, int nf, const double* real_in, double* spec_out) {
    for (int f = 0; f < nf; ++f) fwd3d(o, real_in + (size_t)f * o->N, (cx*)spec_out + (size_t)f * o->M);
    return 0;
}

int oracle_fft_c2r(oracle* o, int nf, const double* spec_in, double* real_out) {
    cx* s = (cx*)malloc(sizeof(cx) * o->M);
    for (int f = 0; f < nf; ++f) inv3d(o, (const cx*)spec_in + (size_t)f * o->M, real_out + (size_t)f * o->N, s);
    free(s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Physics                                                                  */
/* ------------------------------------------------------------------------ */

/* spectral derivative factors i*k with AEB / corotating scaling
 * (mhdrhs.f90:191-204, :313-326; mhd.f90:541-554) */
static void kvec(const oracle* o, int ix, int iy, int iz, cx* kx, cx* ky, cx* kz) {
    cx z;
    z.re = 0.0; z.im = o->wz[iz];
    *kz = cx_divr(cx_scale(z, o->p.radius0), o->radius);
    z.re = 0.0; z.im = o->wy[iy];
    *ky = cx_divr(cx_scale(z, o->p.radius0), o->radius);
    kx->re = 0.0; kx->im = o->wx[ix];
    if (o->p.if_aeb && o->p.if_corotating) {
        cx ikx = {0.0, o->wx[ix]}, iky = {0.0, o->wy[iy]};
        *kx = cx_add(cx_scale(ikx, o->cos_c), cx_scale(iky, o->sin_c));
        cx m = cx_add(cx_scale(cx_scale(ikx, -1.0), o->sin_c), cx_scale(iky, o->cos_c));
        *ky = cx_divr(cx_scale(m, o->p.radius0), o->radius);
    }
}

/* calc_current_density_real (mhdrhs.f90:296-362) */
static void calc_current_density_real(oracle* o) {
    const cx* bx = o->uhat + 4 * o->M;
    const cx* by = o->uhat + 5 * o->M;
    const cx* bz = o->uhat + 6 * o->M;
#pragma omp parallel for collapse(2) num_threads(nthreads())
    for (int iz = 0; iz < o->nz; ++iz)
        for (int iy = 0; iy < o->ny; ++iy)
            for (int ix = 0; ix < o->nxh; ++ix) {
                cx kx, ky, kz;
                kvec(o, ix, iy, iz, &kx, &ky, &kz);
                size_t s = SDX3(o, ix, iy, iz);
                o->jhat[s] = cx_sub(cx_mul(ky, bz[s]), cx_mul(kz, by[s]));
                o->jhat[o->M + s] = cx_sub(cx_mul(kz, bx[s]), cx_mul(kx, bz[s]));
                o->jhat[2 * o->M + s] = cx_sub(cx_mul(kx, by[s]), cx_mul(ky, bx[s]));
            }
    oracle_fft_c2r(o, 3, (const double*)o->jhat, o->J);
}

/* calc_flux (mhdrhs.f90:21-124) ; flux(:,19) holds expand_term with AEB */
static void calc_flux(oracle* o) {
    if (o->p.if_hall) calc_current_density_real(o);
    const size_t N = o->N;
    const double g = o->p.adiabatic_index, di = o->p.ion_inertial_length, tau = o->tau;
    double* F = o->flux;
    const double* uu = o->uu;
    const double* pr = o->prim;
    const double* J = o->J;
    const int hall = o->p.if_hall, aeb = o->p.if_aeb;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (size_t i = 0; i < N; ++i) {
        double p = pr[3 * N + i];
        double Bx = uu[4 * N + i], By = uu[5 * N + i], Bz = uu[6 * N + i];
        double ux = pr[i], uy = pr[N + i], uz = pr[2 * N + i];
        double rho = uu[i];
        double ptot = p + 0.5 * (Bx * Bx + By * By + Bz * Bz);
        double udotb = ux * Bx + uy * By + uz * Bz;
        double mx = uu[N + i], my = uu[2 * N + i], mz = uu[3 * N + i], e = uu[7 * N + i];
        F[0 * N + i] = mx;
        F[1 * N + i] = my;
        F[2 * N + i] = mz;
        F[3 * N + i] = mx * ux - Bx * Bx + ptot;
        F[4 * N + i] = my * ux - By * Bx;
        F[5 * N + i] = mz * ux - Bz * Bx;
        F[6 * N + i] = mx * uy - Bx * By;
        F[7 * N + i] = my * uy - By * By + ptot;
        F[8 * N + i] = mz * uy - Bz * By;
        F[9 * N + i] = mx * uz - Bx * Bz;
        F[10 * N + i] = my * uz - By * Bz;
        F[11 * N + i] = mz * uz - Bz * Bz + ptot;
        F[12 * N + i] = uz * By - uy * Bz;
        F[13 * N + i] = ux * Bz - uz * Bx;
        F[14 * N + i] = uy * Bx - ux * By;
        F[15 * N + i] = (e + ptot) * ux - udotb * Bx;
        F[16 * N + i] = (e + ptot) * uy - udotb * By;
        F[17 * N + i] = (e + ptot) * uz - udotb * Bz;
        if (aeb) {
            F[18 * N + i] = -(2 * g / (g - 1) * p / tau) - (2.0 * Bx * Bx + By * By + Bz * Bz) / tau
                            - (mx * ux + 2 * my * uy + 2 * mz * uz) / tau;
        }
        if (hall) {
            double Jx = J[i], Jy = J[N + i], Jz = J[2 * N + i];
            F[12 * N + i] = F[12 * N + i] + di / rho * (Jy * Bz - Jz * By);
            F[13 * N + i] = F[13 * N + i] + di / rho * (Jz * Bx - Jx * Bz);
            F[14 * N + i] = F[14 * N + i] + di / rho * (Jx * By - Jy * Bx);
        }
    }
}

/* transform_flux_real_to_fourier (mhdrhs.f90:128-172) */
static void transform_flux(oracle* o) {
    int nf = o->p.if_aeb ? 19 : 18;
    oracle_fft_r2c(o, nf, o->flux, (double*)o->fhat);
}

static void update_prim(oracle* o);

/* out = in * (e^{i pi mx/nx} e^{i pi my/ny} e^{i pi mz/nz})^sign, nf fields */
static void phase_shift(const oracle* o, const cx* in, cx* out, int nf, int sign) {
    const size_t M = o->M;
#pragma omp parallel for collapse(2) num_threads(nthreads())
    for (int iz = 0; iz < o->nz; ++iz)
        for (int iy = 0; iy < o->ny; ++iy)
            for (int ix = 0; ix < o->nxh; ++ix) {
                cx ph = cx_mul(cx_mul(o->phx[ix], o->phy[iy]), o->phz[iz]);
                if (sign < 0) ph.im = -ph.im;
                size_t s = SDX3(o, ix, iy, iz);
                for (int v = 0; v < nf; ++v) out[v * M + s] = cx_mul(in[v * M + s], ph);
            }
}

/* Phase-shift dealiasing (Patterson & Orszag 1971; no reference counterpart):
 * the flux spectra are the average of the unshifted evaluation (already in
 * fhat) and the evaluation on the grid shifted by half a cell along every
 * axis, brought back with the conjugate phase. With the sqrt(2)/3 ellipsoid
 * this removes the aliasing of every quadratic product. */
static void shifted_flux(oracle* o) {
    const size_t M = o->M;
    const int nf = o->p.if_aeb ? 19 : 18;
    cx* scratch = (cx*)malloc(sizeof(cx) * M);
    phase_shift(o, o->uhat, o->uhat_s, 8, +1);
    for (int v = 0; v < 8; ++v) inv3d(o, o->uhat_s + (size_t)v * M, o->uu_s + (size_t)v * o->N, scratch);
    free(scratch);
    double* t;
    cx* c;
    t = o->uu; o->uu = o->uu_s; o->uu_s = t;
    t = o->prim; o->prim = o->prim_s; o->prim_s = t;
    c = o->uhat; o->uhat = o->uhat_s; o->uhat_s = c;
    c = o->fhat; o->fhat = o->fhat_s; o->fhat_s = c;   /* fhat_s: the unshifted spectra */
    update_prim(o);
    calc_flux(o);
    transform_flux(o);
    t = o->uu; o->uu = o->uu_s; o->uu_s = t;
    t = o->prim; o->prim = o->prim_s; o->prim_s = t;
    c = o->uhat; o->uhat = o->uhat_s; o->uhat_s = c;
    phase_shift(o, o->fhat, o->fhat, nf, -1);
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (size_t i = 0; i < (size_t)nf * M; ++i) o->fhat[i] = cx_scale(cx_add(o->fhat_s[i], o->fhat[i]), 0.5);
}

/* calc_rhs (mhdrhs.f90:174-279) */
static void calc_rhs(oracle* o) {
    const size_t M = o->M;
    const cx* F = o->fhat;
    const cx* U = o->uhat;
    cx* R = o->fnl;
    const double tau = o->tau, nu = o->p.viscosity, eta = o->p.resistivity;
#pragma omp parallel for collapse(2) num_threads(nthreads())
    for (int iz = 0; iz < o->nz; ++iz)
        for (int iy = 0; iy < o->ny; ++iy)
            for (int ix = 0; ix < o->nxh; ++ix) {
                cx kx, ky, kz;
                kvec(o, ix, iy, iz, &kx, &ky, &kz);
                size_t s = SDX3(o, ix, iy, iz);
#define FF(n) F[(size_t)(n - 1) * M + s]
#define UU(n) U[(size_t)(n - 1) * M + s]
#define RR(n) R[(size_t)(n - 1) * M + s]
                cx t;
                t = cx_add(cx_add(cx_mul(kx, FF(1)), cx_mul(ky, FF(2))), cx_mul(kz, FF(3)));
                RR(1) = cx_scale(t, -1.0);
                t = cx_add(cx_add(cx_mul(kx, FF(4)), cx_mul(ky, FF(5))), cx_mul(kz, FF(6)));
                RR(2) = cx_scale(t, -1.0);
                t = cx_add(cx_add(cx_mul(kx, FF(7)), cx_mul(ky, FF(8))), cx_mul(kz, FF(9)));
                RR(3) = cx_scale(t, -1.0);
                t = cx_add(cx_add(cx_mul(kx, FF(10)), cx_mul(ky, FF(11))), cx_mul(kz, FF(12)));
                RR(4) = cx_scale(t, -1.0);
                RR(5) = cx_sub(cx_mul(kz, FF(14)), cx_mul(ky, FF(15)));
                RR(6) = cx_sub(cx_mul(kx, FF(15)), cx_mul(kz, FF(13)));
                RR(7) = cx_sub(cx_mul(ky, FF(13)), cx_mul(kx, FF(14)));
                t = cx_add(cx_add(cx_mul(kx, FF(16)), cx_mul(ky, FF(17))), cx_mul(kz, FF(18)));
                RR(8) = cx_scale(t, -1.0);
                if (o->p.if_aeb) {
                    RR(1) = cx_sub(RR(1), cx_divr(cx_scale(UU(1), 2.0), tau));
                    RR(2) = cx_sub(RR(2), cx_divr(cx_scale(UU(2), 2.0), tau));
                    RR(3) = cx_sub(RR(3), cx_divr(cx_scale(UU(3), 3.0), tau));
                    RR(4) = cx_sub(RR(4), cx_divr(cx_scale(UU(4), 3.0), tau));
                    RR(5) = cx_sub(RR(5), cx_divr(cx_scale(UU(5), 2.0), tau));
                    RR(6) = cx_sub(RR(6), cx_divr(cx_scale(UU(6), 1.0), tau));
                    RR(7) = cx_sub(RR(7), cx_divr(cx_scale(UU(7), 1.0), tau));
                    RR(8) = cx_add(RR(8), FF(19));
                }
                double k2 = o->ksq[s];
                if (o->p.if_visc && o->p.if_visc_exp) {
                    for (int v = 2; v <= 4; ++v) RR(v) = cx_sub(RR(v), cx_scale(cx_scale(UU(v), nu), k2));
                }
                if (o->p.if_resis && o->p.if_resis_exp) {
                    if (o->p.if_conserve_background && ix == 0 && iz == 0) continue;
                    for (int v = 5; v <= 7; ++v) RR(v) = cx_sub(RR(v), cx_scale(cx_scale(UU(v), eta), k2));
                }
#undef FF
#undef UU
#undef RR
            }
}

/* rkt_init (rktmod.f90:15-32) */
static void rkt_init(oracle* o, double dt) {
    memset(o->fnl_rk, 0, sizeof(cx) * 8 * o->M);
    o->cc1[0] = 8.0 / 15.0 * dt; o->dd1[0] = 0.0;
    o->cc1[1] = 5.0 / 12.0 * dt; o->dd1[1] = -17. / 60. * dt;
    o->cc1[2] = 0.75 * dt;       o->dd1[2] = -5. / 12. * dt;
    o->tstep[0] = 8. / 15. * dt;
    o->tstep[1] = 2. / 15. * dt;
    o->tstep[2] = 1. / 3. * dt;
}

/* rkt (rktmod.f90:34-62) */
static void rkt(oracle* o, int irk) {
    const size_t M = o->M;
    const double c = o->cc1[irk], d = o->dd1[irk], ts = o->tstep[irk];
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (size_t i = 0; i < 8 * M; ++i) {
        o->uhat[i] = cx_add(cx_add(cx_scale(o->fnl[i], c), cx_scale(o->fnl_rk[i], d)), o->uhat[i]);
        o->fnl_rk[i] = o->fnl[i];
    }
    if (o->p.if_visc && !o->p.if_visc_exp) {
        for (int v = 1; v <= 3; ++v)
#pragma omp parallel for num_threads(nthreads()) schedule(static)
            for (size_t s = 0; s < M; ++s)
                o->uhat[v * M + s] = cx_divr(o->uhat[v * M + s], ts * o->ksq[s] * o->p.viscosity + 1.0);
    }
    if (o->p.if_resis && !o->p.if_resis_exp) {
        for (int v = 4; v <= 6; ++v)
#pragma omp parallel for num_threads(nthreads()) schedule(static)
            for (size_t s = 0; s < M; ++s)
                o->uhat[v * M + s] = cx_divr(o->uhat[v * M + s], ts * o->ksq[s] * o->p.resistivity + 1.0);
    }
}

/* dealias (dealiasing.f90:70-113) on an arbitrary (nxh,ny,nz,8) array */
static void dealias_array(const oracle* o, cx* u) {
    const size_t M = o->M;
    const double pi = o->pi;
    if (o->p.dealias_option == 1) {
#pragma omp parallel for collapse(2) num_threads(nthreads())
        for (int iz = 0; iz < o->nz; ++iz)
            for (int iy = 0; iy < o->ny; ++iy)
                for (int ix = 0; ix < o->nxh; ++ix) {
                    double a = o->wx[ix] * o->p.Lx / (2 * pi * o->nx);
                    double b = o->wy[iy] * o->p.Ly / (2 * pi * o->ny);
                    double c = o->wz[iz] * o->p.Lz / (2 * pi * o->nz);
                    double r = sqrt(a * a + b * b + c * c);
                    if (r > 1. / 3.) {
                        size_t s = SDX3(o, ix, iy, iz);
                        for (int v = 0; v < 8; ++v) { u[v * M + s].re = 0.0; u[v * M + s].im = 0.0; }
                    }
                }
    } else if (o->p.dealias_option == 3) {
        /* phase-shift dealiasing keeps the open ellipsoid (mx/nx)^2 + (my/ny)^2 + (mz/nz)^2 < 2/9
         * (no reference counterpart), tested exactly in integers: modes on the boundary can
         * alias in two directions at once when an n is a multiple of 3 */
        const long long X = o->nx, Y = o->ny, Z = o->nz;
#pragma omp parallel for collapse(2) num_threads(nthreads())
        for (int iz = 0; iz < o->nz; ++iz)
            for (int iy = 0; iy < o->ny; ++iy)
                for (int ix = 0; ix < o->nxh; ++ix) {
                    const long long mx = ix, my = iy <= o->ny / 2 ? iy : iy - o->ny;
                    const long long mz = iz <= o->nz / 2 ? iz : iz - o->nz;
                    if (9 * (mx * mx * Y * Y * Z * Z + my * my * X * X * Z * Z + mz * mz * X * X * Y * Y) >=
                        2 * X * X * Y * Y * Z * Z) {
                        size_t s = SDX3(o, ix, iy, iz);
                        for (int v = 0; v < 8; ++v) { u[v * M + s].re = 0.0; u[v * M + s].im = 0.0; }
                    }
                }
    } else if (o->p.dealias_option == 2) {
        for (int v = 0; v < 8; ++v)
#pragma omp parallel for collapse(2) num_threads(nthreads())
            for (int iz = 0; iz < o->nz; ++iz)
                for (int iy = 0; iy < o->ny; ++iy)
                    for (int ix = 0; ix < o->nxh; ++ix) {
                        size_t s = v * M + SDX3(o, ix, iy, iz);
                        u[s] = cx_scale(cx_scale(cx_scale(u[s], o->filtx[ix]), o->filty[iy]), o->filtz[iz]);
                    }
    }
}

/* update_uu_prim_from_uu (mhdrhs.f90:282-294) */
static void update_prim(oracle* o) {
    const size_t N = o->N;
    const double g = o->p.adiabatic_index;
    double* uu = o->uu;
    double* pr = o->prim;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (size_t i = 0; i < N; ++i) {
        pr[i] = uu[N + i] / uu[i];
        pr[N + i] = uu[2 * N + i] / uu[i];
        pr[2 * N + i] = uu[3 * N + i] / uu[i];
        pr[3 * N + i] = (uu[7 * N + i] - 0.5 * (uu[N + i] * pr[i] + uu[2 * N + i] * pr[N + i] +
                                                 uu[3 * N + i] * pr[2 * N + i] + uu[4 * N + i] * uu[4 * N + i] +
                                                 uu[5 * N + i] * uu[5 * N + i] + uu[6 * N + i] * uu[6 * N + i])) *
                        (g - 1);
    }
}

/* vardt reduction (mhd.f90:350-419): minimum over points before *cfl */
static double cfl_dtmin(const oracle* o) {
    const size_t N = o->N;
    const double g = o->p.adiabatic_index, di = o->p.ion_inertial_length;
    const double dx = o->dx, dy = o->dy, dz = o->dz;
    double dmin_xyz = dx < dy ? dx : dy;
    if (dz < dmin_xyz) dmin_xyz = dz;
    const double rr = o->radius / o->p.radius0;
    const double* uu = o->uu;
    const double* pr = o->prim;
    const int hall = o->p.if_hall;
    double dtmin = -1;
    /* serial scan in memory order, as the reference loop (mhd.f90:352-416) */
    for (size_t i = 0; i < N; ++i) {
        double rho = uu[i];
        double cs2 = g * pr[3 * N + i] / rho;
        double cax = uu[4 * N + i] / sqrt(rho), cay = uu[5 * N + i] / sqrt(rho), caz = uu[6 * N + i] / sqrt(rho);
        double ca2 = cax * cax + cay * cay + caz * caz;
        double cms2 = cs2 + ca2;
        double cnx = sqrt(fmax(cms2 * cms2 - 4 * cs2 * cax * cax, 0.));
        double cny = sqrt(fmax(cms2 * cms2 - 4 * cs2 * cay * cay, 0.));
        double cnz = sqrt(fmax(cms2 * cms2 - 4 * cs2 * caz * caz, 0.));
        double cfx = sqrt(cms2 + cnx) / sqrt(2.), cfy = sqrt(cms2 + cny) / sqrt(2.), cfz = sqrt(cms2 + cnz) / sqrt(2.);
        double csx = sqrt(fmax(cms2 - cnx, 0.)) / sqrt(2.);
        double csy = sqrt(fmax(cms2 - cny, 0.)) / sqrt(2.);
        double csz = sqrt(fmax(cms2 - cnz, 0.)) / sqrt(2.);
        double ux = pr[i], uy = pr[N + i], uz = pr[2 * N + i];
#define CMAX(u, cf, cs, ca)                                                                          \
    fmax(fmax(fmax(fabs(u + cf), fabs(u + cs)), fmax(fabs(u + ca), fabs(u - cf))), \
         fmax(fmax(fabs(u - cs), fabs(u - ca)), fabs(u)))
        double cmx = CMAX(ux, cfx, csx, cax);
        double cmy = CMAX(uy, cfy, csy, cay);
        double cmz = CMAX(uz, cfz, csz, caz);
#undef CMAX
        if (hall) {
            double bmax = uu[4 * N + i];
            if (uu[5 * N + i] > bmax) bmax = uu[5 * N + i];
            if (uu[6 * N + i] > bmax) bmax = uu[6 * N + i];
            double ch = di / rho * bmax / dmin_xyz;
            cmx = fmax(cmx, ch); cmy = fmax(cmy, ch); cmz = fmax(cmz, ch);
        }
        double dtx = dx / cmx, dty = dy / cmy * rr, dtz = dz / cmz * rr;
        double d = dtx < dty ? dtx : dty;
        if (dtz < d) d = dtz;
        if (dtmin < 0 || d < dtmin) dtmin = d;
    }
    return dtmin;
}

/* vardt tail (mhd.f90:422-428) */
static void vardt(oracle* o) {
    double dtmin;
    if (o->fixed_dt > 0) {
        o->dt = o->fixed_dt;
    } else {
        dtmin = cfl_dtmin(o) * o->p.cfl;
        if (o->dt < 0.98 * dtmin || o->dt > 1.02 * dtmin) o->dt = dtmin;
    }
    rkt_init(o, o->dt);
}

/* evolve (mhd.f90:298-326) */
static void evolve(oracle* o) {
    cx* scratch = (cx*)malloc(sizeof(cx) * o->M);
    for (int irk = 0; irk < 3; ++irk) {
        calc_flux(o);
        transform_flux(o);
        if (o->p.dealias_option == 3) shifted_flux(o);
        calc_rhs(o);
        rkt(o, irk);
        dealias_array(o, o->uhat);
        for (int v = 0; v < 8; ++v) inv3d(o, o->uhat + (size_t)v * o->M, o->uu + (size_t)v * o->N, scratch);
        update_prim(o);
    }
    free(scratch);
}

/* ------------------------------------------------------------------------ */
/* Entry points                                                             */
/* ------------------------------------------------------------------------ */

/* initial_calc_conserve_variable (mhdinit.f90:1038-1056): on entry uu holds
 * (rho,ux,uy,uz,Bx,By,Bz,p); on exit the conserved state, uu_prim = (u,p). */
void oracle_cons_from_prim(const laps_params* p, int64_t N, double* uu, double* pr) {
    const double g = p->adiabatic_index;
    for (int64_t i = 0; i < N; ++i) {
        pr[i] = uu[N + i];
        pr[N + i] = uu[2 * N + i];
        pr[2 * N + i] = uu[3 * N + i];
        pr[3 * N + i] = uu[7 * N + i];
        uu[N + i] = uu[i] * pr[i];
        uu[2 * N + i] = uu[i] * pr[N + i];
        uu[3 * N + i] = uu[i] * pr[2 * N + i];
        uu[7 * N + i] = pr[3 * N + i] / (g - 1) +
                        0.5 * (uu[i] * (pr[i] * pr[i] + pr[N + i] * pr[N + i] + pr[2 * N + i] * pr[2 * N + i]) +
                               uu[4 * N + i] * uu[4 * N + i] + uu[5 * N + i] * uu[5 * N + i] +
                               uu[6 * N + i] * uu[6 * N + i]);
    }
}

/* mhd.f90:121-136: transform_uu_real_to_fourier, dt = 0, vardt */
int oracle_set_state(oracle* o, const double* uu, const double* pr) {
    memcpy(o->uu, uu, sizeof(double) * 8 * o->N);
    memcpy(o->prim, pr, sizeof(double) * 4 * o->N);
    oracle_fft_r2c(o, 8, o->uu, (double*)o->uhat);
    o->time = 0.0;
    o->dt_last = 0.0;
    o->istep = 0;
    o->status = LAPS_RUNNING;
    o->dt = 0.0;
    vardt(o);
    return 0;
}

/* band-limited state given in spectral space: uu = IFFT(uhat), prim from uu */
int oracle_set_spectral(oracle* o, const double* uhat) {
    memcpy(o->uhat, uhat, sizeof(cx) * 8 * o->M);
    oracle_fft_c2r(o, 8, (const double*)o->uhat, o->uu);
    update_prim(o);
    o->time = 0.0;
    o->dt_last = 0.0;
    o->istep = 0;
    o->status = LAPS_RUNNING;
    o->dt = 0.0;
    vardt(o);
    return 0;
}

int oracle_set_time(oracle* o, double time, double dt_prev) {
    o->time = time;
    evolve_radius(o, time);
    o->dt = dt_prev;
    vardt(o);
    return 0;
}

int oracle_set_fixed_dt(oracle* o, double dt) {
    o->fixed_dt = dt;
    vardt(o);
    return 0;
}

/* main-loop body (mhd.f90:216-285) */
int oracle_step(oracle* o, int nsteps) {
    for (int s = 0; s < nsteps; ++s) {
        if (o->status != LAPS_RUNNING) return 0;
        if (o->dt < 0.00000001) { o->status = LAPS_DT_STOP; return 0; }
        double dt_used = o->dt;
        evolve(o);
        o->time = o->time + dt_used;
        o->istep += 1;
        evolve_radius(o, o->time);
        vardt(o);
        o->dt_last = dt_used;
    }
    return 0;
}

int oracle_get_scalars(oracle* o, double* time, double* dt, double* radius, int64_t* istep, int32_t* status) {
    if (time) *time = o->time;
    if (dt) *dt = o->dt_last;
    if (radius) *radius = o->radius;
    if (istep) *istep = o->istep;
    if (status) *status = o->status;
    return 0;
}

int oracle_get_state(oracle* o, double* uu, double* pr) {
    if (uu) memcpy(uu, o->uu, sizeof(double) * 8 * o->N);
    if (pr) memcpy(pr, o->prim, sizeof(double) * 4 * o->N);
    return 0;
}

int oracle_get_spectral(oracle* o, double* uhat) {
    memcpy(uhat, o->uhat, sizeof(cx) * 8 * o->M);
    return 0;
}

/* vardt without side effects: raw minimum and the hysteresis outcome */
int oracle_cfl(oracle* o, double* dtmin_raw, double* dt_next) {
    double raw = cfl_dtmin(o);
    double dtmin = raw * o->p.cfl;
    double dt = o->dt_last;
    if (dt < 0.98 * dtmin || dt > 1.02 * dtmin) dt = dtmin;
    if (dtmin_raw) *dtmin_raw = raw;
    if (dt_next) *dt_next = dt;
    return 0;
}

/* calc_rms (mhdrms.f90:53-126), B0_ave = 0 */
int oracle_moments(oracle* o, double ave[8], double var[8], double rho_u2[3]) {
    const size_t N = o->N;
    const double* uu = o->uu;
    const double* pr = o->prim;
    double s[8] = {0}, q[8] = {0};
    for (size_t i = 0; i < N; ++i) {
        double v[8] = {uu[i], pr[i], pr[N + i], pr[2 * N + i], uu[4 * N + i], uu[5 * N + i], uu[6 * N + i], pr[3 * N + i]};
        for (int k = 0; k < 8; ++k) { s[k] += v[k]; q[k] += v[k] * v[k]; }
    }
    for (int k = 0; k < 8; ++k) {
        ave[k] = s[k] / (double)N;
        double sq = q[k] / (double)N;
        var[k] = sq - ave[k] * ave[k];
    }
    for (int iv = 0; iv < 3; ++iv) {
        double r = 0.0;
        for (size_t i = 0; i < N; ++i) {
            double d = pr[iv * N + i] - ave[1 + iv];
            r += uu[i] * (d * d);
        }
        rho_u2[iv] = r / (double)N;
    }
    return 0;
}

/* calc_max_divB (mhd.f90:522-570) */
int oracle_max_divb(oracle* o, double* out) {
    double mx = 0.0;
    const size_t M = o->M;
    for (int iz = 0; iz < o->nz; ++iz)
        for (int iy = 0; iy < o->ny; ++iy)
            for (int ix = 0; ix < o->nxh; ++ix) {
                cx kx, ky, kz;
                kvec(o, ix, iy, iz, &kx, &ky, &kz);
                size_t s = SDX3(o, ix, iy, iz);
                cx d = cx_add(cx_add(cx_mul(kx, o->uhat[4 * M + s]), cx_mul(ky, o->uhat[5 * M + s])),
                              cx_mul(kz, o->uhat[6 * M + s]));
                double a = hypot(d.re, d.im);
                if (a > mx) mx = a;
            }
    *out = mx;
    return 0;
}

/* calc_flux + transform_flux_real_to_fourier + calc_rhs on the current state */
int oracle_rhs(oracle* o, double* fnl_out) {
    calc_flux(o);
    transform_flux(o);
    if (o->p.dealias_option == 3) shifted_flux(o);
    calc_rhs(o);
    memcpy(fnl_out, o->fnl, sizeof(cx) * 8 * o->M);
    return 0;
}

int oracle_flux(oracle* o, double* flux_out) {
    calc_flux(o);
    memcpy(flux_out, o->flux, sizeof(double) * 19 * o->N);
    return 0;
}

int oracle_dealias_apply(oracle* o, double* uhat) {
    dealias_array(o, (cx*)uhat);
    return 0;
}

int oracle_filters(oracle* o, double* fx, double* fy, double* fz) {
    memcpy(fx, o->filtx, sizeof(double) * o->nxh);
    memcpy(fy, o->filty, sizeof(double) * o->ny);
    memcpy(fz, o->filtz, sizeof(double) * o->nz);
    return 0;
}

int oracle_ksquare(oracle* o, double* ksq) {
    memcpy(ksq, o->ksq, sizeof(double) * o->M);
    return 0;
}
