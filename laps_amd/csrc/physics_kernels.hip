// physics_kernels.hip -- pointwise kernels of the LAPS step.
//
//  k_curl     : Hall current J_hat = ik x B_hat      (mhdrhs.f90:313-339)
//  k_rhs_rk   : spectral divergence/curl, AEB damping, explicit nu/eta, the
//               low-storage RK3 update with implicit diffusion, dealiasing
//               (mhdrhs.f90:174-279; rktmod.f90:34-62; dealiasing.f90:70-113)
//  k_ctl_*    : the scalar bookkeeping of vardt / evolve_radius on the device
//               (mhd.f90:216, :246-248, :418-428; rktmod.f90:15-32;
//               AEBmod.f90:46-73)
//  reductions : CFL over a real state, moments (mhdrms.f90:53-126) and
//               max |k.B| (mhd.f90:522-570), all in FP64.
#include "internal.hpp"
#include "physics_dev.hpp"

namespace laps {

static constexpr int kThreads = 256;

static inline unsigned grid_for(long long n, int threads = kThreads, long long cap = 1 << 20) {
    long long g = (n + threads - 1) / threads;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// spectral derivative factor i*k (imaginary part) with AEB / corotating scaling;
// evaluation order of mhdrhs.f90:192-204: (w * radius0) / radius.
struct KVec {
    double kx, ky, kz;
};
__device__ __forceinline__ KVec kvec(double wx, double wy, double wz, const PhysParams& pp, double radius) {
    KVec k;
    k.kz = (wz * pp.radius0) / radius;
    k.ky = (wy * pp.radius0) / radius;
    k.kx = wx;
    if (pp.corot_k) {
        k.kx = wx * pp.cos_c + wy * pp.sin_c;
        k.ky = ((-wx * pp.sin_c + wy * pp.cos_c) * pp.radius0) / radius;
    }
    return k;
}

// (i k) * z for real k
__device__ __forceinline__ cpx<double> ik(double k, cpx<double> z) { return {-k * z.y, k * z.x}; }

// k_square: grid_initialize form until the first evolve_radius, then
// update_ksquare (mhdinit.f90:114-122; AEBmod.f90:87-125)
__device__ __forceinline__ double ksquare(double kx, double ky, double kz, const PhysParams& pp, const Ctl* ctl) {
    if (!ctl->ksq_updated) return kx * kx + ky * ky + kz * kz;
    const double R0 = pp.radius0, R = ctl->radius;
    if (pp.corot_ksq) {
        const double c = pp.cos_c, s = pp.sin_c, r0r = R0 / R;
        const double a = s * R0 / R, b = c * R0 / R, z = kz * R0 / R;
        return kx * kx * (c * c + a * a) + ky * ky * (s * s + b * b) + kx * ky * 2 * c * s * (1 - r0r * r0r) + z * z;
    }
    const double a = ky * R0 / R, b = kz * R0 / R;
    return kx * kx + a * a + b * b;
}

// storage index -> (kx, stored ky row, stored kz row); false for row padding
__device__ __forceinline__ bool box_index(long long idx, const Box& b, int& kx, int& cy, int& cz) {
    kx = (int)(idx % b.KXP);
    const long long r = idx / b.KXP;
    cy = (int)(r % b.CY);
    cz = (int)(r / b.CY);
    return kx < b.KX;
}

// ------------------------------------------------------------------ curl
__global__ void k_curl(const cpx<double>* __restrict__ U, cpx<double>* __restrict__ J, long long fs, Box b,
                       BoxTables t, PhysParams pp, const Ctl* __restrict__ ctl) {
    const long long n = b.modes();
    const double radius = ctl->radius;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int kx, cy, cz;
        if (!box_index(i, b, kx, cy, cz)) continue;
        const KVec k = kvec(t.wx[kx], t.wy[cy], t.wz[cz], pp, radius);
        const cpx<double> bx = U[4 * fs + i], by = U[5 * fs + i], bz = U[6 * fs + i];
        J[i] = ik(k.ky, bz) - ik(k.kz, by);
        J[fs + i] = ik(k.kz, bx) - ik(k.kx, bz);
        J[2 * fs + i] = ik(k.kx, by) - ik(k.ky, bx);
    }
}

hipError_t launch_curl(const cpx<double>* B, cpx<double>* J, long long fs, Box b, BoxTables t, const PhysParams& pp,
                       const Ctl* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_curl, dim3(grid_for(b.modes())), dim3(kThreads), 0, s, B, J, fs, b, t, pp, ctl);
    return hipGetLastError();
}

// block-wide minimum of non-negative doubles, folded into *bits with an
// unsigned 64-bit atomic (the IEEE order of non-negative doubles is the
// integer order of their bit patterns; NaN sorts above +Inf and is skipped)
__device__ __forceinline__ void block_min_atomic(double v, unsigned long long* bits) {
    __shared__ double red[kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int i = 1; i < (int)(blockDim.x / 64); ++i) m = fmin(m, red[i]);
        atomicMin(bits, (unsigned long long)__double_as_longlong(m));
    }
}

// ------------------------------------------------------------------ RHS + RK3 + dealias
template <typename T, int OPT>
__global__ __launch_bounds__(kThreads) void k_rhs_rk(const cpx<T>* __restrict__ E, cpx<double>* __restrict__ U,
                                                     cpx<double>* __restrict__ Fp, Box b, BoxTables t, int irk,
                                                     PhysParams pp, const Ctl* __restrict__ ctl) {
    if (ctl->status != 0) return;
    const long long n = b.modes();
    const double radius = ctl->radius, tau = ctl->tau;
    const double c = ctl->cc[irk], d = ctl->dd[irk], ts = ctl->ts[irk];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int kx, cy, cz;
        if (!box_index(i, b, kx, cy, cz)) continue;
        const double wx = t.wx[kx], wy = t.wy[cy], wz = t.wz[cz];
        const KVec k = kvec(wx, wy, wz, pp, radius);
        auto Fv = [&](int f) -> cpx<double> {
            const cpx<T> v = E[f * n + i];
            return {double(v.x), double(v.y)};
        };
        cpx<double> r[8];
        {
            cpx<double> s = ik(k.kx, Fv(0)) + ik(k.ky, Fv(1));
            s = s + ik(k.kz, Fv(2));
            r[0] = {-s.x, -s.y};
            s = ik(k.kx, Fv(3)) + ik(k.ky, Fv(4));
            s = s + ik(k.kz, Fv(5));
            r[1] = {-s.x, -s.y};
            s = ik(k.kx, Fv(6)) + ik(k.ky, Fv(7));
            s = s + ik(k.kz, Fv(8));
            r[2] = {-s.x, -s.y};
            s = ik(k.kx, Fv(9)) + ik(k.ky, Fv(10));
            s = s + ik(k.kz, Fv(11));
            r[3] = {-s.x, -s.y};
            const cpx<double> e1 = Fv(12), e2 = Fv(13), e3 = Fv(14);
            r[4] = ik(k.kz, e2) - ik(k.ky, e3);
            r[5] = ik(k.kx, e3) - ik(k.kz, e1);
            r[6] = ik(k.ky, e1) - ik(k.kx, e2);
            s = ik(k.kx, Fv(15)) + ik(k.ky, Fv(16));
            s = s + ik(k.kz, Fv(17));
            r[7] = {-s.x, -s.y};
        }
        cpx<double> u[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) u[v] = U[v * n + i];
        if (pp.aeb) {
            const double f[7] = {2.0, 2.0, 3.0, 3.0, 2.0, 1.0, 1.0};
#pragma unroll
            for (int v = 0; v < 7; ++v) r[v] = r[v] - cpx<double>{(u[v].x * f[v]) / tau, (u[v].y * f[v]) / tau};
            r[7] = r[7] + Fv(18);
        }
        const double k2 = ksquare(wx, wy, wz, pp, ctl);
        if (pp.visc && pp.visc_exp) {
#pragma unroll
            for (int v = 1; v <= 3; ++v) r[v] = r[v] - cpx<double>{u[v].x * pp.nu * k2, u[v].y * pp.nu * k2};
        }
        if (pp.resis && pp.resis_exp && !(pp.conserve_bg && kx == 0 && t.mz[cz] == 0)) {
#pragma unroll
            for (int v = 4; v <= 6; ++v) r[v] = r[v] - cpx<double>{u[v].x * pp.eta * k2, u[v].y * pp.eta * k2};
        }
        // rkt (rktmod.f90:40-61)
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            cpx<double> un;
            if (irk == 0) {
                // fnl_rk was zeroed by rkt_init (rktmod.f90:22) and d_1 = 0
                un = cpx<double>{c * r[v].x + 0.0, c * r[v].y + 0.0} + u[v];
            } else {
                const cpx<double> fp = Fp[v * n + i];
                un = cpx<double>{c * r[v].x + d * fp.x, c * r[v].y + d * fp.y} + u[v];
            }
            Fp[v * n + i] = r[v];
            u[v] = un;
        }
        if (pp.visc && !pp.visc_exp) {
            const double den = ts * k2 * pp.nu + 1.0;
#pragma unroll
            for (int v = 1; v <= 3; ++v) u[v] = {u[v].x / den, u[v].y / den};
        }
        if (pp.resis && !pp.resis_exp) {
            const double den = ts * k2 * pp.eta + 1.0;
#pragma unroll
            for (int v = 4; v <= 6; ++v) u[v] = {u[v].x / den, u[v].y / den};
        }
        // dealias (dealiasing.f90:152-175)
        if constexpr (OPT == 1) {
            const int lim = t.kzlim[(long long)cy * b.KX + kx];
            const int mz = t.mz[cz];
            if (mz > lim || -mz > lim) {
#pragma unroll
                for (int v = 0; v < 8; ++v) u[v] = {0.0, 0.0};
            }
        } else if constexpr (OPT == 2) {
            const double fx = t.fx[kx], fy = t.fy[cy], fz = t.fz[cz];
#pragma unroll
            for (int v = 0; v < 8; ++v) u[v] = {((u[v].x * fx) * fy) * fz, ((u[v].y * fx) * fy) * fz};
        }
#pragma unroll
        for (int v = 0; v < 8; ++v) U[v * n + i] = u[v];
    }
}

template <typename T>
hipError_t launch_rhs_rk(const cpx<T>* E, cpx<double>* U, cpx<double>* Fp, Box b, BoxTables t, int irk,
                         const PhysParams& pp, const Ctl* ctl, hipStream_t s) {
    const unsigned g = grid_for(b.modes());
    if (pp.dealias == 1)
        hipLaunchKernelGGL((k_rhs_rk<T, 1>), dim3(g), dim3(kThreads), 0, s, E, U, Fp, b, t, irk, pp, ctl);
    else if (pp.dealias == 2)
        hipLaunchKernelGGL((k_rhs_rk<T, 2>), dim3(g), dim3(kThreads), 0, s, E, U, Fp, b, t, irk, pp, ctl);
    else
        hipLaunchKernelGGL((k_rhs_rk<T, 0>), dim3(g), dim3(kThreads), 0, s, E, U, Fp, b, t, irk, pp, ctl);
    return hipGetLastError();
}

// ------------------------------------------------------------------ control block
__global__ void k_ctl_begin(Ctl* ctl) {
    ctl->dtmin_bits = 0x7FF0000000000000ULL;  // +Inf
}

// vardt tail + rkt_init + the dt < 1e-8 stop test (mhd.f90:216, :418-428; rktmod.f90:15-32)
__global__ void k_ctl_finalize(Ctl* ctl, PhysParams pp) {
    if (ctl->status != 0) return;
    double dt;
    if (pp.fixed_dt > 0) {
        dt = pp.fixed_dt;
    } else {
        const double dtmin = __longlong_as_double((long long)ctl->dtmin_bits) * pp.cfl;
        dt = ctl->dt_prev;
        if (dt < 0.98 * dtmin || dt > 1.02 * dtmin) dt = dtmin;
    }
    ctl->dt_prev = dt;
    if (dt < 0.00000001) {
        ctl->status = 1;
        return;
    }
    ctl->dt = dt;
    ctl->cc[0] = 8.0 / 15.0 * dt;
    ctl->dd[0] = 0.0;
    ctl->cc[1] = 5.0 / 12.0 * dt;
    ctl->dd[1] = -17. / 60. * dt;
    ctl->cc[2] = 0.75 * dt;
    ctl->dd[2] = -5. / 12. * dt;
    ctl->ts[0] = 8. / 15. * dt;
    ctl->ts[1] = 2. / 15. * dt;
    ctl->ts[2] = 1. / 3. * dt;
}

// time += dt, istep += 1, evolve_radius(time) (mhd.f90:246-248; AEBmod.f90:56-73)
__global__ void k_ctl_end(Ctl* ctl, PhysParams pp) {
    if (ctl->status != 0) return;
    ctl->time = ctl->time + ctl->dt;
    ctl->istep += 1;
    ctl->radius = pp.radius0 + pp.Ur * ctl->time;
    ctl->tau = ctl->radius / pp.Ur;
    ctl->r0r = pp.radius0 / ctl->radius;
    ctl->ksq_updated = 1;
}

hipError_t launch_ctl_begin(Ctl* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(1), 0, s, ctl);
    return hipGetLastError();
}
hipError_t launch_ctl_finalize(Ctl* ctl, const PhysParams& pp, hipStream_t s) {
    hipLaunchKernelGGL(k_ctl_finalize, dim3(1), dim3(1), 0, s, ctl, pp);
    return hipGetLastError();
}
hipError_t launch_ctl_end(Ctl* ctl, const PhysParams& pp, hipStream_t s) {
    hipLaunchKernelGGL(k_ctl_end, dim3(1), dim3(1), 0, s, ctl, pp);
    return hipGetLastError();
}

// ------------------------------------------------------------------ primitives (FP64)
template <typename T>
__global__ void k_prim(const T* __restrict__ R, double* __restrict__ P, long long N, double g) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        const double rho = R[i], mx = R[N + i], my = R[2 * N + i], mz = R[3 * N + i];
        const double bx = R[4 * N + i], by = R[5 * N + i], bz = R[6 * N + i], e = R[7 * N + i];
        const double ux = mx / rho, uy = my / rho, uz = mz / rho;
        P[i] = ux;
        P[N + i] = uy;
        P[2 * N + i] = uz;
        P[3 * N + i] = (e - 0.5 * (mx * ux + my * uy + mz * uz + bx * bx + by * by + bz * bz)) * (g - 1);
    }
}

template <typename T>
hipError_t launch_prim(const T* R, double* P, long long npts, double gamma, hipStream_t s) {
    hipLaunchKernelGGL((k_prim<T>), dim3(grid_for(npts)), dim3(kThreads), 0, s, R, P, npts, gamma);
    return hipGetLastError();
}
template hipError_t launch_prim<double>(const double*, double*, long long, double, hipStream_t);

// ------------------------------------------------------------------ CFL on a real state
__global__ __launch_bounds__(kThreads) void k_cfl(const double* __restrict__ U, const double* __restrict__ P,
                                                  long long N, PhysParams pp, Ctl* __restrict__ ctl) {
    const double rr = ctl->radius / pp.radius0;
    double dmin = pp.dx < pp.dy ? pp.dx : pp.dy;
    if (pp.dz < dmin) dmin = pp.dz;
    double dtloc = __longlong_as_double(0x7FF0000000000000LL);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        const double d = cfl_point(U[i], U[4 * N + i], U[5 * N + i], U[6 * N + i], P[3 * N + i], P[i], P[N + i],
                                   P[2 * N + i], pp, rr, dmin);
        dtloc = fmin(dtloc, d);
    }
    block_min_atomic(dtloc, &ctl->dtmin_bits);
}

hipError_t launch_cfl(const double* U8, const double* P4, long long npts, const PhysParams& pp, Ctl* ctl,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_cfl, dim3(grid_for(npts, kThreads, 8192)), dim3(kThreads), 0, s, U8, P4, npts, pp, ctl);
    return hipGetLastError();
}

// ------------------------------------------------------------------ moments (mhdrms.f90:73-121)
// pass 1: per-block sums of f and f^2 for f = (rho,ux,uy,uz,Bx,By,Bz,p)
__global__ __launch_bounds__(kThreads) void k_moments(const double* __restrict__ U, const double* __restrict__ P,
                                                      long long N, double* __restrict__ part) {
    double s[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) s[k] = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        const double v[8] = {U[i], P[i], P[N + i], P[2 * N + i], U[4 * N + i], U[5 * N + i], U[6 * N + i], P[3 * N + i]};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            s[k] += v[k];
            s[8 + k] += v[k] * v[k];
        }
    }
    __shared__ double red[16][kThreads];
#pragma unroll
    for (int k = 0; k < 16; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 16; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 16) part[blockIdx.x * 16 + threadIdx.x] = red[threadIdx.x][0];
}

// pass 2: per-block sums of rho (u_i - <u_i>)^2
__global__ __launch_bounds__(kThreads) void k_moments2(const double* __restrict__ U, const double* __restrict__ P,
                                                       long long N, const double* __restrict__ ave,
                                                       double* __restrict__ part) {
    double s[3] = {0.0, 0.0, 0.0};
    const double a1 = ave[1], a2 = ave[2], a3 = ave[3];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        const double rho = U[i];
        const double d1 = P[i] - a1, d2 = P[N + i] - a2, d3 = P[2 * N + i] - a3;
        s[0] += rho * (d1 * d1);
        s[1] += rho * (d2 * d2);
        s[2] += rho * (d3 * d3);
    }
    __shared__ double red[3][kThreads];
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

hipError_t launch_moments(const double* U8, const double* P4, long long npts, double* partial, int nblk,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_moments, dim3(nblk), dim3(kThreads), 0, s, U8, P4, npts, partial);
    return hipGetLastError();
}
hipError_t launch_moments2(const double* U8, const double* P4, long long npts, const double* ave, double* partial,
                           int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_moments2, dim3(nblk), dim3(kThreads), 0, s, U8, P4, npts, ave, partial);
    return hipGetLastError();
}

// ------------------------------------------------------------------ max |k.B| (mhd.f90:541-564)
__global__ __launch_bounds__(kThreads) void k_divb(const cpx<double>* __restrict__ U, long long fs, Box b,
                                                   BoxTables t, PhysParams pp, const Ctl* __restrict__ ctl,
                                                   unsigned long long* out) {
    const long long n = b.modes();
    const double radius = ctl->radius;
    double m = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int kx, cy, cz;
        if (!box_index(i, b, kx, cy, cz)) continue;
        const KVec k = kvec(t.wx[kx], t.wy[cy], t.wz[cz], pp, radius);
        cpx<double> d = ik(k.kx, U[4 * fs + i]) + ik(k.ky, U[5 * fs + i]);
        d = d + ik(k.kz, U[6 * fs + i]);
        m = fmax(m, hypot(d.x, d.y));
    }
    __shared__ double red[kThreads / 64];
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int i = 1; i < kThreads / 64; ++i) v = fmax(v, red[i]);
        atomicMax(out, (unsigned long long)__double_as_longlong(v));
    }
}

hipError_t launch_divb(const cpx<double>* U, long long fs, Box b, BoxTables t, const PhysParams& pp, const Ctl* ctl,
                       unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(k_divb, dim3(grid_for(b.modes(), kThreads, 4096)), dim3(kThreads), 0, s, U, fs, b, t, pp, ctl,
                       out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ full spectrum -> kept box
__global__ void k_gather_box(const cpx<double>* __restrict__ full, cpx<double>* __restrict__ box, int nf,
                             long long fstride, long long bstride, int nxh, int ny, int nz, Box b) {
    const long long n = b.modes();
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n * nf;
         i += (long long)gridDim.x * blockDim.x) {
        const int f = (int)(i / n);
        const long long m = i % n;
        int kx, cy, cz;
        if (!box_index(m, b, kx, cy, cz)) continue;
        const int iy = (b.CY == ny || cy <= (b.CY - 1) / 2) ? cy : ny - (b.CY - cy);
        const int iz = (b.CZ == nz || cz <= (b.CZ - 1) / 2) ? cz : nz - (b.CZ - cz);
        box[f * bstride + m] = full[f * fstride + ((long long)iz * ny + iy) * nxh + kx];
    }
}

hipError_t launch_gather_box(const cpx<double>* full, cpx<double>* box, int nf, long long full_stride,
                             long long box_stride, int nxh, int ny, int nz, Box b, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_box, dim3(grid_for(b.modes() * nf)), dim3(kThreads), 0, s, full, box, nf, full_stride,
                       box_stride, nxh, ny, nz, b);
    return hipGetLastError();
}

template <typename TI, typename TO>
__global__ void k_convert(const TI* __restrict__ in, TO* __restrict__ out, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = TO(in[i]);
}
template <typename TI, typename TO>
hipError_t launch_convert(const TI* in, TO* out, long long n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<TI, TO>), dim3(grid_for(n)), dim3(kThreads), 0, s, in, out, n);
    return hipGetLastError();
}
template hipError_t launch_convert<double, float>(const double*, float*, long long, hipStream_t);
template hipError_t launch_convert<float, double>(const float*, double*, long long, hipStream_t);

template hipError_t launch_rhs_rk<double>(const cpx<double>*, cpx<double>*, cpx<double>*, Box, BoxTables, int,
                                          const PhysParams&, const Ctl*, hipStream_t);
template hipError_t launch_rhs_rk<float>(const cpx<float>*, cpx<double>*, cpx<double>*, Box, BoxTables, int,
                                         const PhysParams&, const Ctl*, hipStream_t);

}  // namespace laps
