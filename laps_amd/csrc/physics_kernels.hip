// physics_kernels.hip -- pointwise kernels of the LAPS step.
//
//  k_curl     : Hall current J_hat = ik x B_hat      (mhdrhs.f90:313-339)
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

// storage index -> (kx, stored ky row, stored kz row); false for row padding
__device__ __forceinline__ bool box_index(long long idx, const Box& b, int& kx, int& cy, int& cz) {
    kx = (int)(idx % b.KXP);
    const long long r = idx / b.KXP;
    cy = (int)(r % b.CY);
    cz = (int)(r / b.CY);
    return kx < b.KX;
}

// ------------------------------------------------------------------ curl
// J = i k' x B on the modes of box b (FP64)
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

// ------------------------------------------------------------------ control block
__global__ void k_ctl_begin(Ctl* ctl) {
    ctl += blockIdx.x;  // one block per batched box
    ctl->dtmin_slot[threadIdx.x] = 0x7FF0000000000000ULL;  // +Inf
    if (threadIdx.x == 0) ctl->dtmin_bits = 0x7FF0000000000000ULL;
}

// vardt tail + rkt_init + the dt < 1e-8 stop test (mhd.f90:216, :418-428; rktmod.f90:15-32)
__global__ void k_ctl_finalize(Ctl* ctl, PhysParams pp) {
    ctl += blockIdx.x;  // one block per batched box
    if (ctl->status != 0) return;
    double dt;
    if (pp.fixed_dt > 0) {
        dt = pp.fixed_dt;
    } else {
        unsigned long long m = ctl->dtmin_bits;
        for (int i = 0; i < kCflSlots; ++i) m = ctl->dtmin_slot[i] < m ? ctl->dtmin_slot[i] : m;
        ctl->dtmin_bits = m;
        const double dtmin = __longlong_as_double((long long)m) * pp.cfl;
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
    ctl += blockIdx.x;  // one block per batched box
    if (ctl->status != 0) return;
    ctl->time = ctl->time + ctl->dt;
    ctl->istep += 1;
    ctl->radius = pp.radius0 + pp.Ur * ctl->time;
    ctl->tau = ctl->radius / pp.Ur;
    ctl->r0r = pp.radius0 / ctl->radius;
    ctl->ksq_updated = 1;
}

hipError_t launch_ctl_begin(Ctl* ctl, hipStream_t s, int nbox) {
    hipLaunchKernelGGL(k_ctl_begin, dim3(nbox > 1 ? nbox : 1), dim3(kCflSlots), 0, s, ctl);
    return hipGetLastError();
}
hipError_t launch_ctl_finalize(Ctl* ctl, const PhysParams& pp, hipStream_t s, int nbox) {
    hipLaunchKernelGGL(k_ctl_finalize, dim3(nbox > 1 ? nbox : 1), dim3(1), 0, s, ctl, pp);
    return hipGetLastError();
}
hipError_t launch_ctl_end(Ctl* ctl, const PhysParams& pp, hipStream_t s, int nbox) {
    hipLaunchKernelGGL(k_ctl_end, dim3(nbox > 1 ? nbox : 1), dim3(1), 0, s, ctl, pp);
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
        const double d = cfl_point<double>(U[i], U[4 * N + i], U[5 * N + i], U[6 * N + i], P[3 * N + i], P[i], P[N + i],
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

// ------------------------------------------------------------------ budgets
// Per-block sums over the real state: e, (m . u)/2, |B|^2/2, p/(gamma-1), u . B.
__global__ __launch_bounds__(kThreads) void k_budget_real(const double* __restrict__ U, const double* __restrict__ P,
                                                          long long N, double gm1, double* __restrict__ part) {
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        const double mx = U[N + i], my = U[2 * N + i], mz = U[3 * N + i];
        const double bx = U[4 * N + i], by = U[5 * N + i], bz = U[6 * N + i];
        const double ux = P[i], uy = P[N + i], uz = P[2 * N + i];
        s[0] += U[7 * N + i];
        s[1] += 0.5 * (mx * ux + my * uy + mz * uz);
        s[2] += 0.5 * (bx * bx + by * by + bz * bz);
        s[3] += P[3 * N + i] / gm1;
        s[4] += ux * bx + uy * by + uz * bz;
    }
    __shared__ double red[5][kThreads];
#pragma unroll
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5) part[blockIdx.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// Per-block sums over the stored half spectrum of w_k Re(conj(A_k) . B_k) and
// w_k Re(conj(J_k) . B_k), with J = i k' x B and the Coulomb-gauge potential
// A = i k' x B / |k'|^2 (k' with the AEB / corotating scaling, as in curl);
// k = 0 is left out (A is undefined there) and w_k = 2 for 0 < kx < nx/2, else 1.
__global__ __launch_bounds__(kThreads) void k_budget_spec(const cpx<double>* __restrict__ U, long long fs, Box b,
                                                          BoxTables t, PhysParams pp, const Ctl* __restrict__ ctl,
                                                          int nx, double* __restrict__ part) {
    const long long n = b.modes();
    const double radius = ctl->radius;
    double hm = 0.0, hj = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        int kx, cy, cz;
        if (!box_index(i, b, kx, cy, cz)) continue;
        const KVec k = kvec(t.wx[kx], t.wy[cy], t.wz[cz], pp, radius);
        const double k2 = k.kx * k.kx + k.ky * k.ky + k.kz * k.kz;
        if (k2 == 0.0) continue;
        const cpx<double> bx = U[4 * fs + i], by = U[5 * fs + i], bz = U[6 * fs + i];
        // J = i k x B
        const cpx<double> jx = ik(k.ky, bz) - ik(k.kz, by);
        const cpx<double> jy = ik(k.kz, bx) - ik(k.kx, bz);
        const cpx<double> jz = ik(k.kx, by) - ik(k.ky, bx);
        const double jb = jx.x * bx.x + jx.y * bx.y + jy.x * by.x + jy.y * by.y + jz.x * bz.x + jz.y * bz.y;
        const double w = (kx == 0 || 2 * kx == nx) ? 1.0 : 2.0;
        hj += w * jb;
        hm += w * jb / k2;
    }
    __shared__ double red[2][kThreads];
    red[0][threadIdx.x] = hm;
    red[1][threadIdx.x] = hj;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 2; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 2) part[blockIdx.x * 2 + threadIdx.x] = red[threadIdx.x][0];
}

hipError_t launch_budget_real(const double* U8, const double* P4, long long npts, double gamma, double* partial,
                              int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_budget_real, dim3(nblk), dim3(kThreads), 0, s, U8, P4, npts, gamma - 1.0, partial);
    return hipGetLastError();
}
hipError_t launch_budget_spec(const cpx<double>* U, long long fs, Box b, BoxTables t, const PhysParams& pp,
                              const Ctl* ctl, int nx, double* partial, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(k_budget_spec, dim3(nblk), dim3(kThreads), 0, s, U, fs, b, t, pp, ctl, nx, partial);
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


}  // namespace laps
