// xfused.hip -- the x pass of a substage in one kernel: x C2R of the state
// (and Hall current), primitives, the 18(+1) nonlinear fluxes, the CFL
// minimum, and x R2C of the fluxes. The real-space fields never reach HBM.
//
// Reference: calc_flux (mhdrhs.f90:21-124), update_uu_prim_from_uu
// (:282-294), the x legs of transform_uu_fourier_to_real /
// transform_flux_real_to_fourier (fftw.f90:94-100; mhdrhs.f90:143-149) and
// the vardt loop (mhd.f90:352-416).
//
// Layout on the lanes: a row of N points is owned by LR = N/4 lanes; lane l
// holds the points x = l + r*LR, r = 0..3, of every field. Two fields share
// one complex transform (Z = A + iB), so all 11 inputs of a point sit in the
// registers of one lane and the physics needs no exchange. The transform is
// a Stockham FFT with radix-4 stages (plus one radix-2 stage when N = 2*4^k)
// whose inputs and outputs both use that layout; stages exchange through a
// padded LDS line.
#include "internal.hpp"
#include "physics_dev.hpp"

namespace laps {

template <typename T>
struct LanePad {
    static constexpr int SH = sizeof(T) == 8 ? 4 : 5;  // one pad slot per 256 bytes
    __device__ static constexpr int at(int p) { return p + (p >> SH); }
    static constexpr int line(int n) { return n + (n >> SH) + 1; }
};

template <int N>
struct LaneFft {
    static constexpr int LR = N / 4;                     // lanes per row
    static constexpr int NT = LR > 64 ? LR : 64;         // threads per block: one wave, or one row
    static constexpr bool HAS2 = (N == 32 || N == 128 || N == 512);
};

// radix-4 Stockham stages of span NS, NS*4, ... up to N; v[r] is position
// lane + r*LR when read from the line and after the last stage.
template <int N, int NS, int DIR, typename T>
__device__ __forceinline__ void r4_stages(cpx<T> (&v)[4], cpx<T>* line, const cpx<T>* __restrict__ tw, int lane,
                                          bool from_lds) {
    using P = LanePad<T>;
    constexpr int LR = LaneFft<N>::LR;
    if (from_lds) {
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = *(line + P::at(lane + r * LR));
    }
    const int k = lane % NS;
    if constexpr (NS > 1) {
        // W^k from the table, W^2k and W^3k by multiplication (one LDS read per stage)
        constexpr int STEP = N / (NS * 4);
        const cpx<T> w1 = twiddle<DIR>(tw, k * STEP);
        const cpx<T> w2 = cmul(w1, w1);
        const cpx<T> w3 = cmul(w2, w1);
        v[1] = cmul(v[1], w1);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], w3);
    }
    Dft<4, DIR, T>::run(v);
    if constexpr (NS * 4 < N) {
        __syncthreads();
        const int base = (lane / NS) * NS * 4 + k;
#pragma unroll
        for (int r = 0; r < 4; ++r) *(line + P::at(base + r * NS)) = v[r];
        __syncthreads();
        r4_stages<N, NS * 4, DIR, T>(v, line, tw, lane, true);
    }
}

// In-place transform of one row: v[r] holds position lane + r*LR on entry and
// on exit. `line` is this row's LDS line. All threads of the block call it.
template <int N, int DIR, typename T>
__device__ __forceinline__ void lane_fft(cpx<T> (&v)[4], cpx<T>* line, const cpx<T>* __restrict__ tw, int lane) {
    using P = LanePad<T>;
    constexpr int LR = LaneFft<N>::LR;
    if constexpr (LaneFft<N>::HAS2) {
        // radix-2 stage of span 1: butterfly j = lane pairs r = 0,2; j = lane + LR pairs r = 1,3;
        // outputs of butterfly j go to 2j and 2j + 1
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const cpx<T> a = v[h], b = v[h + 2];
            v[h] = a + b;
            v[h + 2] = a - b;
        }
        __syncthreads();
        *(line + P::at(2 * lane)) = v[0];
        *(line + P::at(2 * lane + 1)) = v[2];
        *(line + P::at(2 * (lane + LR))) = v[1];
        *(line + P::at(2 * (lane + LR) + 1)) = v[3];
        __syncthreads();
        r4_stages<N, 2, DIR, T>(v, line, tw, lane, true);
    } else {
        r4_stages<N, 1, DIR, T>(v, line, tw, lane, false);
    }
}

// Z_p of the paired C2R (Z = A + iB with Hermitian extension; only the real
// parts of the DC and Nyquist coefficients enter, as in FFTW's c2r).
template <int N, typename T>
__device__ __forceinline__ cpx<T> c2r_input(const cpx<T>* __restrict__ A, const cpx<T>* __restrict__ B, int kx,
                                            int p) {
    const cpx<T> zero{T(0), T(0)};
    if (p == 0) return cpx<T>{A[0].x, B ? B[0].x : T(0)};
    if (p < N / 2) {
        if (p >= kx) return zero;
        const cpx<T> a = A[p], b = B ? B[p] : zero;
        return cpx<T>{a.x - b.y, a.y + b.x};
    }
    if (p == N / 2) {
        if (p >= kx) return zero;
        return cpx<T>{A[p].x, B ? B[p].x : T(0)};
    }
    const int m = N - p;
    if (m >= kx) return zero;
    const cpx<T> a = A[m], b = B ? B[m] : zero;
    return cpx<T>{a.x + b.y, b.x - a.y};
}

template <int N, typename T>
__device__ __forceinline__ void c2r_pair(const cpx<T>* A, const cpx<T>* B, int kx, cpx<T>* line,
                                         const cpx<T>* tw, int lane, T (&a)[4], T (&b)[4]) {
    constexpr int LR = LaneFft<N>::LR;
    cpx<T> v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = c2r_input<N, T>(A, B, kx, lane + r * LR);
    lane_fft<N, +1, T>(v, line, tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        a[r] = v[r].x;
        b[r] = v[r].y;
    }
}

// forward transform of Z = fa + i fb and the split into the two half spectra
template <int N, typename T>
__device__ __forceinline__ void r2c_pair(const T (&fa)[4], const T (&fb)[4], cpx<T>* oa, cpx<T>* ob, int kx,
                                         cpx<T>* line, const cpx<T>* tw, int lane) {
    using P = LanePad<T>;
    constexpr int LR = LaneFft<N>::LR;
    cpx<T> v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = cpx<T>{fa[r], fb[r]};
    lane_fft<N, -1, T>(v, line, tw, lane);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) *(line + P::at(lane + r * LR)) = v[r];
    __syncthreads();
    const T h = T(0.5);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = lane + r * LR;
        if (k < kx) {
            const cpx<T> zk = *(line + P::at(k));
            const cpx<T> zn = *(line + P::at((N - k) & (N - 1)));
            oa[k] = cpx<T>{h * (zk.x + zn.x), h * (zk.y - zn.y)};
            if (ob) ob[k] = cpx<T>{h * (zk.y + zn.y), h * (zn.x - zk.x)};
        }
    }
}

template <typename T, int N, bool FIRST, bool CFL, bool HALL, bool AEB>
__global__ __launch_bounds__(LaneFft<N>::NT) void k_xfused(XFArgs a, const cpx<T>* __restrict__ twg, PhysParams pp,
                                                Ctl* __restrict__ ctl) {
    constexpr int LR = LaneFft<N>::LR;
    constexpr int NT = LaneFft<N>::NT;
    constexpr int RPB = NT / LR;  // rows per block
    constexpr int LS = LanePad<T>::line(N);
    __shared__ cpx<T> sh[RPB * LS];
    __shared__ cpx<T> stw[N];
    for (int i = threadIdx.x; i < N; i += NT) stw[i] = twg[i];
    __syncthreads();
    const cpx<T>* tw = stw;
    const int lane = threadIdx.x % LR;
    const int rib = threadIdx.x / LR;
    long long row = (long long)blockIdx.x * RPB + rib;
    const bool valid = row < a.nrows;
    if (!valid) row = a.nrows - 1;  // computes a duplicate row, never stored
    cpx<T>* line = sh + rib * LS;

    // ---- inputs: rho, m(3), B(3), e [, J(3)] at x = lane + r*LR
    T rho[4], mx[4], my[4], mz[4], bx[4], by[4], bz[4], e[4], jx[4], jy[4], jz[4];
    T ux[4], uy[4], uz[4], p[4];
    if constexpr (FIRST) {
        const long long N3 = a.npts;
        const double* S = a.S + row * N;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int x = lane + r * LR;
            rho[r] = T(S[x]); mx[r] = T(S[N3 + x]); my[r] = T(S[2 * N3 + x]); mz[r] = T(S[3 * N3 + x]);
            bx[r] = T(S[4 * N3 + x]); by[r] = T(S[5 * N3 + x]); bz[r] = T(S[6 * N3 + x]); e[r] = T(S[7 * N3 + x]);
            ux[r] = T(S[8 * N3 + x]); uy[r] = T(S[9 * N3 + x]); uz[r] = T(S[10 * N3 + x]); p[r] = T(S[11 * N3 + x]);
        }
    } else {
        const cpx<T>* st = (const cpx<T>*)a.st + row * a.rs_s;
        const long long fs = a.fs_s;
        c2r_pair<N, T>(st, st + fs, a.kx_s, line, tw, lane, rho, mx);
        c2r_pair<N, T>(st + 2 * fs, st + 3 * fs, a.kx_s, line, tw, lane, my, mz);
        c2r_pair<N, T>(st + 4 * fs, st + 5 * fs, a.kx_s, line, tw, lane, bx, by);
        c2r_pair<N, T>(st + 6 * fs, st + 7 * fs, a.kx_s, line, tw, lane, bz, e);
        const T g1 = T(pp.gamma) - T(1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const T ir = T(1) / rho[r];
            ux[r] = mx[r] * ir;
            uy[r] = my[r] * ir;
            uz[r] = mz[r] * ir;
            p[r] = (e[r] - T(0.5) * (mx[r] * ux[r] + my[r] * uy[r] + mz[r] * uz[r] + bx[r] * bx[r] + by[r] * by[r] +
                                     bz[r] * bz[r])) * g1;
        }
    }
    // ---- CFL minimum of the first substage (vardt, mhd.f90:352-416)
    if constexpr (CFL) {
        const double rr = ctl->radius / pp.radius0;
        double dmin = pp.dx < pp.dy ? pp.dx : pp.dy;
        if (pp.dz < dmin) dmin = pp.dz;
        double dtl = __longlong_as_double(0x7FF0000000000000LL);
        if (valid) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double d = cfl_point(double(rho[r]), double(bx[r]), double(by[r]), double(bz[r]),
                                               double(p[r]), double(ux[r]), double(uy[r]), double(uz[r]), pp, rr,
                                               dmin);
                dtl = fmin(dtl, d);
            }
        }
        for (int o = 32; o > 0; o >>= 1) dtl = fmin(dtl, __shfl_xor(dtl, o));
        if ((threadIdx.x & 63) == 0) atomicMin(&ctl->dtmin_bits, (unsigned long long)__double_as_longlong(dtl));
    }

    if constexpr (HALL) {
        const cpx<T>* jb = (const cpx<T>*)a.jb + row * a.rs_j;
        const long long fs = a.fs_j;
        c2r_pair<N, T>(jb, jb + fs, a.kx_j, line, tw, lane, jx, jy);
        T dummy[4];
        c2r_pair<N, T>(jb + 2 * fs, nullptr, a.kx_j, line, tw, lane, jz, dummy);
    }

    // ---- fluxes (mhdrhs.f90:60-119), kept as point values
    T ptot[4], udb[4], ep[4], ex[4], ey[4], ez[4], f19[4];
    const T tau = T(ctl->tau), g = T(pp.gamma);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        ptot[r] = p[r] + T(0.5) * (bx[r] * bx[r] + by[r] * by[r] + bz[r] * bz[r]);
        udb[r] = ux[r] * bx[r] + uy[r] * by[r] + uz[r] * bz[r];
        ep[r] = e[r] + ptot[r];
        ex[r] = uz[r] * by[r] - uy[r] * bz[r];
        ey[r] = ux[r] * bz[r] - uz[r] * bx[r];
        ez[r] = uy[r] * bx[r] - ux[r] * by[r];
        if constexpr (HALL) {
            const T h = T(pp.di) / rho[r];
            ex[r] = ex[r] + h * (jy[r] * bz[r] - jz[r] * by[r]);
            ey[r] = ey[r] + h * (jz[r] * bx[r] - jx[r] * bz[r]);
            ez[r] = ez[r] + h * (jx[r] * by[r] - jy[r] * bx[r]);
        }
        if constexpr (AEB) {
            f19[r] = -(T(2) * g / (g - T(1)) * p[r] / tau) - (T(2) * bx[r] * bx[r] + by[r] * by[r] + bz[r] * bz[r]) / tau -
                     (mx[r] * ux[r] + T(2) * my[r] * uy[r] + T(2) * mz[r] * uz[r]) / tau;
        }
    }

    cpx<T>* out = (cpx<T>*)a.out + row * a.rs_o;
    const long long fo = a.fs_o;
    const int kxo = valid ? a.kx_o : 0;
    T fa[4], fb[4];
    // pairs ordered so that the inputs used by few fluxes die first
    if constexpr (AEB) {
#pragma unroll
        for (int r = 0; r < 4; ++r) fb[r] = T(0);
        r2c_pair<N, T>(f19, fb, out + 18 * fo, nullptr, kxo, line, tw, lane);
    }
    // (F13, F14), (F15, F16), (F17, F18)
    r2c_pair<N, T>(ex, ey, out + 12 * fo, out + 13 * fo, kxo, line, tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) { fa[r] = ez[r]; fb[r] = ep[r] * ux[r] - udb[r] * bx[r]; }
    r2c_pair<N, T>(fa, fb, out + 14 * fo, out + 15 * fo, kxo, line, tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) { fa[r] = ep[r] * uy[r] - udb[r] * by[r]; fb[r] = ep[r] * uz[r] - udb[r] * bz[r]; }
    r2c_pair<N, T>(fa, fb, out + 16 * fo, out + 17 * fo, kxo, line, tw, lane);
    // (F4, F5), (F6, F7), (F8, F9), (F10, F11), (F12, F1), (F2, F3)
#pragma unroll
    for (int r = 0; r < 4; ++r) { fa[r] = mx[r] * ux[r] - bx[r] * bx[r] + ptot[r]; fb[r] = my[r] * ux[r] - by[r] * bx[r]; }
    r2c_pair<N, T>(fa, fb, out + 3 * fo, out + 4 * fo, kxo, line, tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) { fa[r] = mz[r] * ux[r] - bz[r] * bx[r]; fb[r] = mx[r] * uy[r] - bx[r] * by[r]; }
    r2c_pair<N, T>(fa, fb, out + 5 * fo, out + 6 * fo, kxo, line, tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) { fa[r] = my[r] * uy[r] - by[r] * by[r] + ptot[r]; fb[r] = mz[r] * uy[r] - bz[r] * by[r]; }
    r2c_pair<N, T>(fa, fb, out + 7 * fo, out + 8 * fo, kxo, line, tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) { fa[r] = mx[r] * uz[r] - bx[r] * bz[r]; fb[r] = my[r] * uz[r] - by[r] * bz[r]; }
    r2c_pair<N, T>(fa, fb, out + 9 * fo, out + 10 * fo, kxo, line, tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) { fa[r] = mz[r] * uz[r] - bz[r] * bz[r] + ptot[r]; fb[r] = mx[r]; }
    r2c_pair<N, T>(fa, fb, out + 11 * fo, out, kxo, line, tw, lane);
    r2c_pair<N, T>(my, mz, out + fo, out + 2 * fo, kxo, line, tw, lane);
}

template <typename T, int N, bool FIRST, bool CFL, bool HALL>
static void xf_aeb(bool aeb, const XFArgs& a, const cpx<T>* tw, const PhysParams& pp, Ctl* ctl, hipStream_t s) {
    constexpr int NT = LaneFft<N>::NT;
    constexpr int RPB = NT / LaneFft<N>::LR;
    const unsigned g = (unsigned)((a.nrows + RPB - 1) / RPB);
    if (aeb)
        hipLaunchKernelGGL((k_xfused<T, N, FIRST, CFL, HALL, true>), dim3(g), dim3(NT), 0, s, a, tw, pp, ctl);
    else
        hipLaunchKernelGGL((k_xfused<T, N, FIRST, CFL, HALL, false>), dim3(g), dim3(NT), 0, s, a, tw, pp, ctl);
}

template <typename T, int N, bool FIRST, bool CFL>
static void xf_hall(bool hall, bool aeb, const XFArgs& a, const cpx<T>* tw, const PhysParams& pp, Ctl* ctl,
                    hipStream_t s) {
    if (hall) xf_aeb<T, N, FIRST, CFL, true>(aeb, a, tw, pp, ctl, s);
    else xf_aeb<T, N, FIRST, CFL, false>(aeb, a, tw, pp, ctl, s);
}

template <typename T, int N>
static void xf_n(bool first, bool cfl, const XFArgs& a, const cpx<T>* tw, const PhysParams& pp, Ctl* ctl,
                 hipStream_t s) {
    const bool hall = pp.hall, aeb = pp.aeb;
    if (first) {
        if (cfl) xf_hall<T, N, true, true>(hall, aeb, a, tw, pp, ctl, s);
        else xf_hall<T, N, true, false>(hall, aeb, a, tw, pp, ctl, s);
    } else {
        if (cfl) xf_hall<T, N, false, true>(hall, aeb, a, tw, pp, ctl, s);
        else xf_hall<T, N, false, false>(hall, aeb, a, tw, pp, ctl, s);
    }
}

template <typename T>
hipError_t launch_xfused(int nx, bool first, bool cfl, const XFArgs& a, const cpx<T>* tw, const PhysParams& pp,
                         Ctl* ctl, hipStream_t s) {
    switch (nx) {
        case 16: xf_n<T, 16>(first, cfl, a, tw, pp, ctl, s); break;
        case 32: xf_n<T, 32>(first, cfl, a, tw, pp, ctl, s); break;
        case 64: xf_n<T, 64>(first, cfl, a, tw, pp, ctl, s); break;
        case 128: xf_n<T, 128>(first, cfl, a, tw, pp, ctl, s); break;
        case 256: xf_n<T, 256>(first, cfl, a, tw, pp, ctl, s); break;
        case 512: xf_n<T, 512>(first, cfl, a, tw, pp, ctl, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_xfused<double>(int, bool, bool, const XFArgs&, const cpx<double>*, const PhysParams&, Ctl*,
                                          hipStream_t);
template hipError_t launch_xfused<float>(int, bool, bool, const XFArgs&, const cpx<float>*, const PhysParams&, Ctl*,
                                         hipStream_t);

}  // namespace laps
