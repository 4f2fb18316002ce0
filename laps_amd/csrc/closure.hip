// closure.hip -- learned subgrid closure: features, network inference and
// training, and its right-hand side.
//
// No counterpart in the reference (its fluxes are the resolved ones only,
// mhdrhs.f90:60-119). On a coarse grid the filtered equations carry
// unresolved flux terms: the subgrid Reynolds stress R_ij, Maxwell stress
// M_ij, electromotive force and energy flux. A small network maps the resolved
// state at a point and its 3^3-stencil gradients to those 18 terms; the step
// adds the divergence of the momentum-flux and energy-flux corrections and the
// curl of the EMF correction to the right-hand side.
//
//   features (28): rho, u (3), B (3), then d/dx, d/dy, d/dz of rho, ux, uy,
//                  uz, Bx, By, Bz (central differences, periodic), each
//                  normalised as (f - mean) / std
//   network:       28 -> 64 (tanh) -> 64 (tanh) -> 18, FP32
//   outputs (18):  R xx, xy, xz, yy, yz, zz; M xx, xy, xz, yy, yz, zz;
//                  EMF x, y, z; energy flux x, y, z;
//                  physical value = output * ystd
//
// Parameter vector (FP32): W1[28][64], b1[64], W2[64][64], b2[64],
// W3[64][18], b3[18], then xmean[28], xstd[28], ystd[18].
#include "internal.hpp"

namespace laps {

namespace {

constexpr int kThreadsCl = 256;

__device__ __forceinline__ float tanh_f(float x) { return tanhf(x); }

}  // namespace

// ---------------------------------------------------------------- features

__global__ void k_cl_features(const double* __restrict__ S, int nx, int ny, int nz, double dx, double dy, double dz,
                              const float* __restrict__ norm, float* __restrict__ X) {
    const long long N = (long long)nx * ny * nz;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int ix = (int)(i % nx);
    const int iy = (int)((i / nx) % ny);
    const int iz = (int)(i / ((long long)nx * ny));
    const long long xp = i - ix + (ix + 1 == nx ? 0 : ix + 1), xm = i - ix + (ix == 0 ? nx - 1 : ix - 1);
    const long long sy = nx, sz = (long long)nx * ny;
    const long long yp = i + (iy + 1 == ny ? -(long long)(ny - 1) * sy : sy);
    const long long ym = i + (iy == 0 ? (long long)(ny - 1) * sy : -sy);
    const long long zp = i + (iz + 1 == nz ? -(long long)(nz - 1) * sz : sz);
    const long long zm = i + (iz == 0 ? (long long)(nz - 1) * sz : -sz);
    // fields: rho = S[0], u = S[8..10] (primitives), B = S[4..6]
    const int fld[7] = {0, 8, 9, 10, 4, 5, 6};
    float f[28];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const double* g = S + fld[q] * N;
        f[q] = (float)g[i];
        f[7 + 3 * q] = (float)((g[xp] - g[xm]) / (2.0 * dx));
        f[8 + 3 * q] = (float)((g[yp] - g[ym]) / (2.0 * dy));
        f[9 + 3 * q] = (float)((g[zp] - g[zm]) / (2.0 * dz));
    }
    const float* mean = norm;
    const float* sd = norm + kClIn;
#pragma unroll
    for (int q = 0; q < kClIn; ++q) X[i * kClIn + q] = (f[q] - mean[q]) / sd[q];
}

// ---------------------------------------------------------------- network

// One work-item per sample; the parameters are staged in LDS. H1, H2 (post-
// activation) are written when non-null, for the backward pass.
__global__ __launch_bounds__(kThreadsCl) void k_mlp_fwd(const float* __restrict__ X, long long n,
                                                       const float* __restrict__ P, float* __restrict__ Y,
                                                       float* __restrict__ H1, float* __restrict__ H2) {
    __shared__ float sp[kClNParam];
    for (int t = threadIdx.x; t < kClNParam; t += blockDim.x) sp[t] = P[t];
    __syncthreads();
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* W1 = sp;
    const float* b1 = W1 + kClIn * kClH;
    const float* W2 = b1 + kClH;
    const float* b2 = W2 + kClH * kClH;
    const float* W3 = b2 + kClH;
    const float* b3 = W3 + kClH * kClOut;
    float x[kClIn];
#pragma unroll
    for (int q = 0; q < kClIn; ++q) x[q] = X[i * kClIn + q];
    float h1[kClH];
#pragma unroll
    for (int o = 0; o < kClH; ++o) {
        float s = b1[o];
#pragma unroll
        for (int q = 0; q < kClIn; ++q) s = fmaf(x[q], W1[q * kClH + o], s);
        h1[o] = tanh_f(s);
    }
    if (H1) {
#pragma unroll
        for (int o = 0; o < kClH; ++o) H1[i * kClH + o] = h1[o];
    }
    float h2[kClH];
#pragma unroll
    for (int o = 0; o < kClH; ++o) {
        float s = b2[o];
#pragma unroll
        for (int q = 0; q < kClH; ++q) s = fmaf(h1[q], W2[q * kClH + o], s);
        h2[o] = tanh_f(s);
    }
    if (H2) {
#pragma unroll
        for (int o = 0; o < kClH; ++o) H2[i * kClH + o] = h2[o];
    }
#pragma unroll
    for (int o = 0; o < kClOut; ++o) {
        float s = b3[o];
#pragma unroll
        for (int q = 0; q < kClH; ++q) s = fmaf(h2[q], W3[q * kClOut + o], s);
        Y[i * kClOut + o] = s;
    }
}

// Mean-squared-error loss L = sum (Y - T)^2 / (n * 18): per-sample deltas
// dY = dL/dY, dZ2 = dL/d(pre-activation 2), dZ1 = dL/d(pre-activation 1),
// and the per-sample squared error for the loss.
__global__ __launch_bounds__(kThreadsCl) void k_mlp_delta(const float* __restrict__ Y, const float* __restrict__ T,
                                                         const float* __restrict__ H1, const float* __restrict__ H2,
                                                         long long n, const float* __restrict__ P,
                                                         float* __restrict__ dY, float* __restrict__ dZ2,
                                                         float* __restrict__ dZ1, float* __restrict__ err) {
    __shared__ float sp[kClNParam];
    for (int t = threadIdx.x; t < kClNParam; t += blockDim.x) sp[t] = P[t];
    __syncthreads();
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* W2 = sp + kClIn * kClH + kClH;
    const float* W3 = W2 + kClH * kClH + kClH;
    const float sc = 2.0f / (float)((double)n * kClOut);
    float dy[kClOut];
    float e = 0.f;
#pragma unroll
    for (int o = 0; o < kClOut; ++o) {
        const float r = Y[i * kClOut + o] - T[i * kClOut + o];
        e = fmaf(r, r, e);
        dy[o] = sc * r;
        dY[i * kClOut + o] = dy[o];
    }
    err[i] = e;
    float d2[kClH];
#pragma unroll
    for (int q = 0; q < kClH; ++q) {
        float s = 0.f;
#pragma unroll
        for (int o = 0; o < kClOut; ++o) s = fmaf(dy[o], W3[q * kClOut + o], s);
        const float h = H2[i * kClH + q];
        d2[q] = s * (1.f - h * h);
        dZ2[i * kClH + q] = d2[q];
    }
#pragma unroll
    for (int q = 0; q < kClH; ++q) {
        float s = 0.f;
#pragma unroll
        for (int o = 0; o < kClH; ++o) s = fmaf(d2[o], W2[q * kClH + o], s);
        const float h = H1[i * kClH + q];
        dZ1[i * kClH + q] = s * (1.f - h * h);
    }
}

// Weight gradients of one layer: gW[a][o] = sum_i A[i][a] D[i][o] and
// gb[o] = sum_i D[i][o], one work-item per output, in a fixed order (the
// result does not depend on the launch).
__global__ void k_mlp_wgrad(const float* __restrict__ A, const float* __restrict__ D, long long n, int na, int no,
                            float* __restrict__ gW, float* __restrict__ gb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (na + 1) * no) return;
    const int a = t / no, o = t % no;
    float s = 0.f;
    if (a < na) {
        for (long long i = 0; i < n; ++i) s = fmaf(A[i * na + a], D[i * no + o], s);
        gW[a * no + o] = s;
    } else {
        for (long long i = 0; i < n; ++i) s += D[i * no + o];
        gb[o] = s;
    }
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1t, float b2t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = G[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
    P[i] -= lr * mh / (sqrtf(vh) + eps);
}

// Sum of n floats into out[0] (one block; fixed order per launch shape).
__global__ void k_sum(const float* __restrict__ x, long long n, float* __restrict__ out) {
    __shared__ double s[kThreadsCl];
    double a = 0.0;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) a += x[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)s[0];
}

// ---------------------------------------------------------------- closure terms

// Network outputs (normalised) -> the 12 real correction fields of the step:
// C[0..5] = momentum flux tau_ij = R_ij - M_ij + delta_ij q (xx, xy, xz, yy,
// yz, zz), C[6..8] = EMF correction, C[9..11] = energy-flux correction. q is
// the subgrid part of the total pressure p + B^2/2 with
// p = (gamma-1)(e - rho u.u/2 - B^2/2): q = -(gamma-1)/2 (tr R + tr M) + tr M / 2.
__global__ void k_cl_terms(const float* __restrict__ Y, long long N, const float* __restrict__ ystd, double gamma,
                           double* __restrict__ C) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= N) return;
    double o[kClOut];
#pragma unroll
    for (int q = 0; q < kClOut; ++q) o[q] = (double)Y[i * kClOut + q] * (double)ystd[q];
    const double trR = o[0] + o[3] + o[5], trM = o[6] + o[9] + o[11];
    const double q = -0.5 * (gamma - 1.0) * (trR + trM) + 0.5 * trM;
#pragma unroll
    for (int c = 0; c < 6; ++c) C[c * N + i] = o[c] - o[6 + c] + ((c == 0 || c == 3 || c == 5) ? q : 0.0);
#pragma unroll
    for (int c = 0; c < 6; ++c) C[(6 + c) * N + i] = o[12 + c];
}

// Spectral right-hand side of the correction fields (full spectra Ch, FP64,
// layout [12][nz][ny][nxh]): momentum_i -= i k'_j tau_ij (as calc_rhs does for
// the resolved momentum flux, mhdrhs.f90:206-222), B -= i k' x E (the
// induction term of the electric field E, mhdrhs.f90:224-233) and e -= i k'.F
// (the energy flux, mhdrhs.f90:234-236). Out: [7][M], momentum x, y, z, B x,
// y, z, energy.
__global__ void k_cl_rhs(const cpx<double>* __restrict__ Ch, long long M, int nxh, int ny,
                         const double* __restrict__ wx, const double* __restrict__ wy, const double* __restrict__ wz,
                         cpx<double>* __restrict__ R) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int kx = (int)(i % nxh);
    const int iy = (int)((i / nxh) % ny);
    const int iz = (int)(i / ((long long)nxh * ny));
    const double k[3] = {wx[kx], wy[iy], wz[iz]};
    auto t = [&](int a, int b) {
        const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        return Ch[idx[a][b] * M + i];
    };
    // -i * (k . v) for complex v: -i (x + iy) = y - ix
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const cpx<double> v = t(a, b);
            sx += k[b] * v.x;
            sy += k[b] * v.y;
        }
        R[a * M + i] = cpx<double>{sy, -sx};
    }
    const cpx<double> E[3] = {Ch[6 * M + i], Ch[7 * M + i], Ch[8 * M + i]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        const double sx = k[b] * E[c].x - k[c] * E[b].x, sy = k[b] * E[c].y - k[c] * E[b].y;
        R[(3 + a) * M + i] = cpx<double>{sy, -sx};
    }
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const cpx<double> v = Ch[(9 + b) * M + i];
        sx += k[b] * v.x;
        sy += k[b] * v.y;
    }
    R[6 * M + i] = cpx<double>{sy, -sx};
}

// ---------------------------------------------------------------- training data

// Box filter of a fine state and of its flux products, one work-item per coarse cell (mean over
// r^3 fine cells): out[q][coarse cell], q = 0..7 rho, m x y z, B x y z, e; 8..13 m_i u_j and
// 14..19 B_i B_j (xx, xy, xz, yy, yz, zz); 20..22 E = -u x B + (d_i/rho) J x B; 23..25 the
// energy flux (e + p + B^2/2) u - (u.B) B (mhdrhs.f90:60-119). S = [uu(8), uu_prim(4)] and J
// (3 fields) are the fine real fields, FP64.
__global__ void k_cl_filter(const double* __restrict__ S, const double* __restrict__ J, int nx, int ny, int nz, int r,
                            double di, double* __restrict__ out) {
    const int cx = nx / r, cy = ny / r, cz = nz / r;
    const long long Nc = (long long)cx * cy * cz;
    const long long ic = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (ic >= Nc) return;
    const int x0 = (int)(ic % cx) * r, y0 = (int)((ic / cx) % cy) * r, z0 = (int)(ic / ((long long)cx * cy)) * r;
    const long long N = (long long)nx * ny * nz;
    double acc[kClFilt];
#pragma unroll
    for (int q = 0; q < kClFilt; ++q) acc[q] = 0.0;
    const int pr[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int dz = 0; dz < r; ++dz)
        for (int dy = 0; dy < r; ++dy)
            for (int dx = 0; dx < r; ++dx) {
                const long long i = ((long long)(z0 + dz) * ny + (y0 + dy)) * nx + (x0 + dx);
                double w[8], u[3], b[3], j[3];
#pragma unroll
                for (int q = 0; q < 8; ++q) w[q] = S[q * N + i];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    u[a] = S[(8 + a) * N + i];
                    b[a] = w[4 + a];
                    j[a] = J[a * N + i];
                }
                const double p = S[11 * N + i], rho = w[0];
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[q] += w[q];
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    acc[8 + c] += w[1 + pr[c][0]] * u[pr[c][1]];
                    acc[14 + c] += b[pr[c][0]] * b[pr[c][1]];
                }
                const double h = di / rho;
                const double ub = u[0] * b[0] + u[1] * b[1] + u[2] * b[2];
                const double et = w[7] + p + 0.5 * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int s = (a + 1) % 3, t = (a + 2) % 3;
                    acc[20 + a] += -(u[s] * b[t] - u[t] * b[s]) + h * (j[s] * b[t] - j[t] * b[s]);
                    acc[23 + a] += et * u[a] - ub * b[a];
                }
            }
    const double inv = 1.0 / ((double)r * r * r);
#pragma unroll
    for (int q = 0; q < kClFilt; ++q) out[q * Nc + ic] = acc[q] * inv;
}

// One field q of the 26 of k_cl_filter (same numbering and formulas) at every fine point, not
// averaged: the input of the spectral (sharp cut-off) filter, one transform per field. Only the
// inputs field q needs are read.
__global__ void k_cl_point(const double* __restrict__ S, const double* __restrict__ J, long long N, int q, double di,
                           double* __restrict__ out) {
    const int pr[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        double v;
        if (q < 8) {
            v = S[q * N + i];
        } else if (q < 14) {
            v = S[(1 + pr[q - 8][0]) * N + i] * S[(8 + pr[q - 8][1]) * N + i];            // m_i u_j
        } else if (q < 20) {
            v = S[(4 + pr[q - 14][0]) * N + i] * S[(4 + pr[q - 14][1]) * N + i];          // B_i B_j
        } else if (q < 23) {
            const int a = q - 20, s1 = (a + 1) % 3, t1 = (a + 2) % 3;
            const double us = S[(8 + s1) * N + i], ut = S[(8 + t1) * N + i];
            const double bs = S[(4 + s1) * N + i], bt = S[(4 + t1) * N + i];
            const double h = di / S[i];
            v = -(us * bt - ut * bs) + h * (J[s1 * N + i] * bt - J[t1 * N + i] * bs);   // E = -u x B + (d_i/rho) J x B
        } else {
            const int a = q - 23;
            double u[3], b[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                u[c] = S[(8 + c) * N + i];
                b[c] = S[(4 + c) * N + i];
            }
            const double ub = u[0] * b[0] + u[1] * b[1] + u[2] * b[2];
            const double et = S[7 * N + i] + S[11 * N + i] + 0.5 * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
            v = et * u[a] - ub * b[a];                                                   // energy flux
        }
        out[i] = v;
    }
}

// Sharp spectral cut-off between a fine and a coarse full spectrum (layout [z][y][KXP] with kx
// fastest): the modes with kx < nxc/2, |my| < nyc/2 and |mz| < nzc/2 (signed; the coarse
// Nyquist rows are dropped, so the cut is the same for +k and -k). dir > 0: gather into the
// coarse spectrum C (zero outside the cut); dir < 0: scatter C into the fine spectrum F (the
// caller zeroes F first).
__global__ void k_spec_cut(cpx<double>* __restrict__ F, int nyf, int nzf, int kxpf, cpx<double>* __restrict__ C,
                           int nxc, int nyc, int nzc, int kxpc, int dir) {
    const long long Mc = (long long)kxpc * nyc * nzc;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < Mc; i += (long long)gridDim.x * blockDim.x) {
        const int kx = (int)(i % kxpc);
        const int ky = (int)((i / kxpc) % nyc);
        const int kz = (int)(i / ((long long)kxpc * nyc));
        const int my = ky <= nyc / 2 ? ky : ky - nyc;
        const int mz = kz <= nzc / 2 ? kz : kz - nzc;
        const bool kept = kx < nxc / 2 && 2 * my < nyc && -2 * my < nyc && 2 * mz < nzc && -2 * mz < nzc;
        const long long f =
            kept ? ((long long)(mz >= 0 ? mz : nzf + mz) * nyf + (my >= 0 ? my : nyf + my)) * kxpf + kx : 0;
        if (dir > 0) C[i] = kept ? F[f] : cpx<double>{0.0, 0.0};
        else if (kept) F[f] = C[i];
    }
}

// ---------------------------------------------------------------- launchers

static unsigned blocks_for(long long n) { return (unsigned)((n + kThreadsCl - 1) / kThreadsCl); }

hipError_t launch_cl_features(const double* S, int nx, int ny, int nz, double dx, double dy, double dz,
                              const float* norm, float* X, hipStream_t s) {
    const long long N = (long long)nx * ny * nz;
    hipLaunchKernelGGL(k_cl_features, dim3(blocks_for(N)), dim3(kThreadsCl), 0, s, S, nx, ny, nz, dx, dy, dz, norm,
                       X);
    return hipGetLastError();
}

hipError_t launch_cl_point(const double* S, const double* J, long long N, int q, double di, double* out,
                           hipStream_t s) {
    const long long nb = blocks_for(N) < 65536 ? blocks_for(N) : 65536;
    hipLaunchKernelGGL(k_cl_point, dim3((unsigned)nb), dim3(kThreadsCl), 0, s, S, J, N, q, di, out);
    return hipGetLastError();
}

hipError_t launch_spec_cut(cpx<double>* F, int nyf, int nzf, int kxpf, cpx<double>* C, int nxc, int nyc, int nzc,
                           int kxpc, int dir, hipStream_t s) {
    const long long Mc = (long long)kxpc * nyc * nzc;
    const long long nb = blocks_for(Mc) < 65536 ? blocks_for(Mc) : 65536;
    hipLaunchKernelGGL(k_spec_cut, dim3((unsigned)nb), dim3(kThreadsCl), 0, s, F, nyf, nzf, kxpf, C, nxc, nyc, nzc,
                       kxpc, dir);
    return hipGetLastError();
}

hipError_t launch_cl_filter(const double* S, const double* J, int nx, int ny, int nz, int r, double di, double* out,
                            hipStream_t s) {
    const long long Nc = (long long)(nx / r) * (ny / r) * (nz / r);
    hipLaunchKernelGGL(k_cl_filter, dim3(blocks_for(Nc)), dim3(kThreadsCl), 0, s, S, J, nx, ny, nz, r, di, out);
    return hipGetLastError();
}

hipError_t launch_mlp_fwd(const float* X, long long n, const float* P, float* Y, float* H1, float* H2,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks_for(n)), dim3(kThreadsCl), 0, s, X, n, P, Y, H1, H2);
    return hipGetLastError();
}

// Gradient of the loss over n samples into G (kClNParam), and the summed
// squared error into loss[0]; scratch holds n * (2 * kClH + kClOut + 1) floats.
hipError_t launch_mlp_grad(const float* X, const float* T, long long n, const float* P, const float* Y,
                           const float* H1, const float* H2, float* scratch, float* G, float* loss, hipStream_t s) {
    float* dY = scratch;
    float* dZ2 = dY + n * kClOut;
    float* dZ1 = dZ2 + n * kClH;
    float* err = dZ1 + n * kClH;
    hipLaunchKernelGGL(k_mlp_delta, dim3(blocks_for(n)), dim3(kThreadsCl), 0, s, Y, T, H1, H2, n, P, dY, dZ2, dZ1,
                       err);
    float* gW1 = G;
    float* gb1 = gW1 + kClIn * kClH;
    float* gW2 = gb1 + kClH;
    float* gb2 = gW2 + kClH * kClH;
    float* gW3 = gb2 + kClH;
    float* gb3 = gW3 + kClH * kClOut;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(blocks_for((kClIn + 1) * kClH)), dim3(kThreadsCl), 0, s, X, dZ1, n, kClIn,
                       kClH, gW1, gb1);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(blocks_for((kClH + 1) * kClH)), dim3(kThreadsCl), 0, s, H1, dZ2, n, kClH,
                       kClH, gW2, gb2);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(blocks_for((kClH + 1) * kClOut)), dim3(kThreadsCl), 0, s, H2, dY, n, kClH,
                       kClOut, gW3, gb3);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(kThreadsCl), 0, s, err, n, loss);
    return hipGetLastError();
}

// mean squared error from its sum over n samples of kClOut outputs, formed as the host side of
// laps_mlp_grad forms it
__global__ void k_loss_mean(const float* __restrict__ sum, long long n, float* __restrict__ out) {
    if (threadIdx.x == 0) out[0] = sum[0] / (float)((double)n * kClOut);
}

hipError_t launch_loss_mean(const float* sum, long long n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_loss_mean, dim3(1), dim3(64), 0, s, sum, n, out);
    return hipGetLastError();
}

hipError_t launch_adam(float* P, const float* G, float* m, float* v, int n, float lr, float b1t, float b2t,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(kThreadsCl), 0, s, P, G, m, v, n, lr, b1t, b2t);
    return hipGetLastError();
}

hipError_t launch_cl_terms(const float* Y, long long N, const float* ystd, double gamma, double* C, hipStream_t s) {
    hipLaunchKernelGGL(k_cl_terms, dim3(blocks_for(N)), dim3(kThreadsCl), 0, s, Y, N, ystd, gamma, C);
    return hipGetLastError();
}

hipError_t launch_cl_rhs(const cpx<double>* Ch, long long M, int nxh, int ny, const double* wx, const double* wy,
                         const double* wz, cpx<double>* R, hipStream_t s) {
    hipLaunchKernelGGL(k_cl_rhs, dim3(blocks_for(M)), dim3(kThreadsCl), 0, s, Ch, M, nxh, ny, wx, wy, wz, R);
    return hipGetLastError();
}

}  // namespace laps
