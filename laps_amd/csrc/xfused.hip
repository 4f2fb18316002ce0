// xfused.hip -- the x pass of a substage in one kernel: x C2R of the state
// (and Hall current), primitives, the 15(+1) nonlinear fluxes, the CFL
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

// LDS slot of line position p: an XOR swizzle inside each 128-byte group
// (M + 1 complex values). Lanes are served in groups that together cover
// 128 bytes (8 lanes of 16-byte values, 16 lanes of 8-byte values); with
// the swizzle, the contiguous reads (lane + r*LR) and the strided stage
// writes ((lane/NS)*4NS + lane%NS + r*NS, NS = 1, 2, 4, 16, ...) of one
// group land in distinct banks. Only the mirrored read of the R2C split
// pairs up banks. No padding, so a line holds exactly N values.
template <typename T>
struct LanePad {
    static constexpr int M = 128 / (2 * (int)sizeof(T)) - 1;
    __device__ static constexpr int at(int p) { return p ^ ((p >> 2) & M); }
    static constexpr int line(int n) { return n; }
};

template <int N>
struct LaneFft {
    static constexpr int LR = N / 4;                     // lanes per row
    static constexpr int NT = LR > 64 ? LR : 64;         // threads per block: one wave, or one row
    static constexpr bool HAS2 = (N == 32 || N == 128 || N == 512);
};

// radix-4 Stockham stages of span NS, NS*4, ... up to N, applied to K
// independent transforms at once (each with its own LDS line), so that one
// transform's LDS round trip overlaps the other's arithmetic. v[t][r] is
// position lane + r*LR of transform t when read from the line and after the
// last stage.
template <int N, int NS, int DIR, int K, typename T>
__device__ __forceinline__ void r4_stages(cpx<T> (&v)[K][4], cpx<T>* const (&line)[K], const cpx<T>* __restrict__ tw,
                                          int lane, bool from_lds) {
    using P = LanePad<T>;
    constexpr int LR = LaneFft<N>::LR;
    if (from_lds) {
#pragma unroll
        for (int t = 0; t < K; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) v[t][r] = *(line[t] + P::at(lane + r * LR));
    }
    const int k = lane % NS;
    if constexpr (NS > 1) {
        // W^k from the table, W^2k and W^3k by multiplication (one LDS read per stage)
        constexpr int STEP = N / (NS * 4);
        const cpx<T> w1 = twiddle<DIR>(tw, k * STEP);
        const cpx<T> w2 = cmul(w1, w1);
        const cpx<T> w3 = cmul(w2, w1);
#pragma unroll
        for (int t = 0; t < K; ++t) {
            v[t][1] = cmul(v[t][1], w1);
            v[t][2] = cmul(v[t][2], w2);
            v[t][3] = cmul(v[t][3], w3);
        }
    }
#pragma unroll
    for (int t = 0; t < K; ++t) Dft<4, DIR, T>::run(v[t]);
    if constexpr (NS * 4 < N) {
        __syncthreads();
        const int base = (lane / NS) * NS * 4 + k;
#pragma unroll
        for (int t = 0; t < K; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) *(line[t] + P::at(base + r * NS)) = v[t][r];
        __syncthreads();
        r4_stages<N, NS * 4, DIR, K, T>(v, line, tw, lane, true);
    }
}

// In-place transform of K rows of data: v[t][r] holds position lane + r*LR
// on entry and on exit. line[t] is transform t's LDS line. All threads of the
// block call it.
// nx = 96 = 4 * 6 * 4: the first and last stages are radix 4 (24 butterflies, one per lane,
// so that the lane layout of the input and output is that of the other lengths); the middle
// radix-6 stage (span 4) runs on lanes 0..15 with its operands through LDS.
template <int DIR, int K, typename T>
__device__ __forceinline__ void lane_fft_96(cpx<T> (&v)[K][4], cpx<T>* const (&line)[K],
                                            const cpx<T>* __restrict__ tw, int lane) {
    using P = LanePad<T>;
    constexpr int N = 96, LR = 24;
#pragma unroll
    for (int t = 0; t < K; ++t) Dft<4, DIR, T>::run(v[t]);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < K; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) *(line[t] + P::at(4 * lane + r)) = v[t][r];
    __syncthreads();
    const int k = lane % 4;  // span 4, twiddle step N / (4 * 6) = 4
    cpx<T> u[K][6];
    if (lane < 16) {
#pragma unroll
        for (int t = 0; t < K; ++t) {
#pragma unroll
            for (int r = 0; r < 6; ++r) u[t][r] = *(line[t] + P::at(lane + 16 * r));
#pragma unroll
            for (int r = 1; r < 6; ++r) u[t][r] = cmul(u[t][r], twiddle<DIR>(tw, 4 * r * k));
            Dft<6, DIR, T>::run(u[t]);
        }
    }
    __syncthreads();  // every read of the line before any write
    if (lane < 16) {
        const int base = (lane / 4) * 24 + k;
#pragma unroll
        for (int t = 0; t < K; ++t)
#pragma unroll
            for (int r = 0; r < 6; ++r) *(line[t] + P::at(base + 4 * r)) = u[t][r];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < K; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) v[t][r] = *(line[t] + P::at(lane + r * LR));
        // span 24: k = lane, twiddle step N / (24 * 4) = 1
#pragma unroll
        for (int r = 1; r < 4; ++r) v[t][r] = cmul(v[t][r], twiddle<DIR>(tw, r * lane));
        Dft<4, DIR, T>::run(v[t]);
    }
    static_assert(N == 4 * LR, "lane layout");
}

template <int N, int DIR, int K, typename T>
__device__ __forceinline__ void lane_fft(cpx<T> (&v)[K][4], cpx<T>* const (&line)[K], const cpx<T>* __restrict__ tw,
                                         int lane) {
    using P = LanePad<T>;
    constexpr int LR = LaneFft<N>::LR;
    if constexpr (N == 96) {
        lane_fft_96<DIR, K, T>(v, line, tw, lane);
    } else if constexpr (LaneFft<N>::HAS2) {
        // radix-2 stage of span 1: butterfly j = lane pairs r = 0,2; j = lane + LR pairs r = 1,3;
        // outputs of butterfly j go to 2j and 2j + 1
#pragma unroll
        for (int t = 0; t < K; ++t)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const cpx<T> a = v[t][h], b = v[t][h + 2];
                v[t][h] = a + b;
                v[t][h + 2] = a - b;
            }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < K; ++t) {
            *(line[t] + P::at(2 * lane)) = v[t][0];
            *(line[t] + P::at(2 * lane + 1)) = v[t][2];
            *(line[t] + P::at(2 * (lane + LR))) = v[t][1];
            *(line[t] + P::at(2 * (lane + LR) + 1)) = v[t][3];
        }
        __syncthreads();
        r4_stages<N, 2, DIR, K, T>(v, line, tw, lane, true);
    } else {
        r4_stages<N, 1, DIR, K, T>(v, line, tw, lane, false);
    }
}

// Raw spectral values that lane `lane` needs for two paired C2Rs (A0 + iB0
// and A1 + iB1): for position p = lane + r*LR, the coefficients at mode
// m = min(p, N - p), loaded at the clamped index (m < kx ? m : 0) with no
// branch around the loads. c2r_load only issues the loads, so a kernel can
// start the next pair's loads before it transforms the current one.
template <typename T>
struct C2RRaw {
    cpx<T> a[2][4], b[2][4];
};

// HAS_B1 = false: the second transform has no B (its imaginary rows are zero)
template <int N, bool HAS_B1, typename T>
__device__ __forceinline__ void c2r_load(const cpx<T>* __restrict__ A0, const cpx<T>* __restrict__ B0,
                                         const cpx<T>* __restrict__ A1, const cpx<T>* __restrict__ B1, int kx,
                                         int lane, C2RRaw<T>& q) {
    constexpr int LR = LaneFft<N>::LR;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int p = lane + r * LR;
        const int m = r < 2 ? p : N - p;  // r < 2: p < N/2; r >= 2: p >= N/2
        const int mc = m < kx ? m : 0;
        q.a[0][r] = A0[mc];
        q.b[0][r] = B0[mc];
        q.a[1][r] = A1[mc];
        q.b[1][r] = HAS_B1 ? B1[mc] : cpx<T>{T(0), T(0)};
    }
}

// Z_p of the paired C2R (Z = A + iB with Hermitian extension; only the real
// parts of the DC and Nyquist coefficients enter, as in FFTW's c2r).
template <int N, typename T>
__device__ __forceinline__ void c2r_combine(const cpx<T> (&a)[4], const cpx<T> (&b)[4], int kx, int lane,
                                            cpx<T> (&v)[4]) {
    constexpr int LR = LaneFft<N>::LR;
    const cpx<T> zero{T(0), T(0)};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int p = lane + r * LR;
        const int m = r < 2 ? p : N - p;
        cpx<T> z;
        if (p == 0 || p == N / 2) z = cpx<T>{a[r].x, b[r].x};
        else if (r < 2) z = cpx<T>{a[r].x - b[r].y, a[r].y + b[r].x};
        else z = cpx<T>{a[r].x + b[r].y, b[r].x - a[r].y};
        v[r] = m < kx ? z : zero;
    }
}

// The two paired C2Rs of q at once: real rows (a0, b0) and (a1, b1).
template <int N, typename T>
__device__ __forceinline__ void c2r_pair2(const C2RRaw<T>& q, int kx, cpx<T>* const (&line)[2], const cpx<T>* tw,
                                          int lane, T (&a0)[4], T (&b0)[4], T (&a1)[4], T (&b1)[4]) {
    cpx<T> v[2][4];
    c2r_combine<N, T>(q.a[0], q.b[0], kx, lane, v[0]);
    c2r_combine<N, T>(q.a[1], q.b[1], kx, lane, v[1]);
    lane_fft<N, +1, 2, T>(v, line, tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        a0[r] = v[0][r].x;
        b0[r] = v[0][r].y;
        a1[r] = v[1][r].x;
        b1[r] = v[1][r].y;
    }
}

// Split of the forward transform Z = FFT(fa + i fb) into the two half spectra
// oa, ob (ob may be null). Only k < kx <= N/2 + 1 is stored, so the mirrors
// N - k that the split needs are 0 (k = 0, its own mirror) or lie in [N/2, N):
// registers r = 2, 3. Each lane writes just those two values, at the slots its
// own last stage read, so the write needs no barrier; Z_k itself comes from
// the registers. The caller puts a barrier between r2c_mirror and r2c_split.
template <int N, typename T>
__device__ __forceinline__ void r2c_mirror(const cpx<T> (&v)[4], cpx<T>* line, int lane) {
    using P = LanePad<T>;
    constexpr int LR = LaneFft<N>::LR;
    *(line + P::at(lane + 2 * LR)) = v[2];
    *(line + P::at(lane + 3 * LR)) = v[3];
}

template <int N, typename T>
__device__ __forceinline__ void r2c_split(const cpx<T> (&v)[4], const cpx<T>* line, int lane, int kx, cpx<T>* oa,
                                          cpx<T>* ob) {
    using P = LanePad<T>;
    constexpr int LR = LaneFft<N>::LR;
    const T h = T(0.5);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int k = lane + r * LR;
        if (k < kx) {
            const cpx<T> zk = v[r];
            const cpx<T> zn = k == 0 ? zk : *(line + P::at(N - k));
            store_stream(oa + k, cpx<T>{h * (zk.x + zn.x), h * (zk.y - zn.y)});
            if (ob) store_stream(ob + k, cpx<T>{h * (zk.y + zn.y), h * (zn.x - zk.x)});
        }
    }
}

// forward transform of one pair (fa, fb) -> (oa, ob)
template <int N, typename T>
__device__ __forceinline__ void r2c_pair(const T (&fa)[4], const T (&fb)[4], cpx<T>* oa, cpx<T>* ob, int kx,
                                         cpx<T>* const (&line)[1], const cpx<T>* tw, int lane) {
    cpx<T> v[1][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[0][r] = cpx<T>{fa[r], fb[r]};
    lane_fft<N, -1, 1, T>(v, line, tw, lane);
    r2c_mirror<N, T>(v[0], line[0], lane);
    __syncthreads();
    r2c_split<N, T>(v[0], line[0], lane, kx, oa, ob);
}

// two forward transforms at once: (fa0, fb0) -> (oa0, ob0), (fa1, fb1) -> (oa1, ob1)
template <int N, typename T>
__device__ __forceinline__ void r2c_pair2(const T (&fa0)[4], const T (&fb0)[4], cpx<T>* oa0, cpx<T>* ob0,
                                          const T (&fa1)[4], const T (&fb1)[4], cpx<T>* oa1, cpx<T>* ob1, int kx,
                                          cpx<T>* const (&line)[2], const cpx<T>* tw, int lane) {
    cpx<T> v[2][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        v[0][r] = cpx<T>{fa0[r], fb0[r]};
        v[1][r] = cpx<T>{fa1[r], fb1[r]};
    }
    lane_fft<N, -1, 2, T>(v, line, tw, lane);
    r2c_mirror<N, T>(v[0], line[0], lane);
    r2c_mirror<N, T>(v[1], line[1], lane);
    __syncthreads();
    r2c_split<N, T>(v[0], line[0], lane, kx, oa0, ob0);
    r2c_split<N, T>(v[1], line[1], lane, kx, oa1, ob1);
}

// waves per SIMD the register allocation must allow: at 512 the FP64 CFL variants need a few
// more than 256 VGPRs otherwise, one wave per SIMD
template <typename T, int N>
constexpr int xf_min_waves() { return (sizeof(T) == 8 && N == 512) ? 2 : 1; }

template <typename T, int N, bool FIRST, bool CFL, bool HALL, bool AEB>
__global__ __launch_bounds__(LaneFft<N>::NT, (xf_min_waves<T, N>())) void k_xfused(XFArgs a, const cpx<T>* __restrict__ twg, PhysParams pp,
                                                Ctl* __restrict__ ctl) {
    constexpr int LR = LaneFft<N>::LR;
    constexpr int NT = LaneFft<N>::NT;
    constexpr int RPB = NT / LR;  // rows per block
    // row slots: at nx = 96 a wave holds 2 rows of 24 lanes and 16 lanes more, which run a
    // dummy row slot of their own (own LDS lines, no stores)
    constexpr int SLOTS = (NT + LR - 1) / LR;
    constexpr int LS = LanePad<T>::line(N);
    __shared__ cpx<T> sh[2 * SLOTS * LS];
    __shared__ cpx<T> stw[N];
    const cpx<T>* tw = stw;
    const int lane = threadIdx.x % LR;
    const int rib = threadIdx.x / LR;
    // batched boxes: one per grid row (offsets and control block of box blockIdx.y)
    ctl += blockIdx.y;
    a.st = (const cpx<T>*)a.st + blockIdx.y * a.st_bx;
    a.jb = (const cpx<T>*)a.jb + blockIdx.y * a.jb_bx;
    a.out = (cpx<T>*)a.out + blockIdx.y * a.out_bx;
    // two chains: the blocks of chain 1 follow those of chain 0
    long long blk = blockIdx.x;
    const long long gpc = (a.nrows + RPB - 1) / RPB;
    const int ch = (a.nch == 2 && blk >= gpc) ? 1 : 0;
    blk -= ch * gpc;
    long long row = blk * RPB + (rib < RPB ? rib : 0);
    const bool valid = rib < RPB && row < a.nrows;
    if (row >= a.nrows) row = a.nrows - 1;  // computes a duplicate row, never stored
    [[maybe_unused]] cpx<T>* const line1[1] = {sh + 2 * rib * LS};
    cpx<T>* const line2[2] = {sh + 2 * rib * LS, sh + (2 * rib + 1) * LS};
    // the twiddle table is fetched first and written to LDS once the row's first
    // loads are in flight, so the wait for it does not wait for them
    LdsTable<cpx<T>, N, NT> twt;
    twt.fetch(twg);
    auto stage_twiddles = [&]() {
        twt.put(stw);
        __syncthreads();
    };

    // ---- inputs: rho, m(3), B(3), e [, J(3)] at x = lane + r*LR. Each group
    // of two paired C2Rs has its loads issued before the previous group's
    // transforms run, so HBM latency overlaps the FFT work.
    T rho[4], mx[4], my[4], mz[4], bx[4], by[4], bz[4], e[4], jx[4], jy[4], jz[4];
    T ux[4], uy[4], uz[4], p[4];
    C2RRaw<T> qa, qb;
    const cpx<T>* jb = (const cpx<T>*)a.jb + ch * a.jb_ch + row * a.rs_j;
    const long long fsj = a.fs_j;
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
        if constexpr (HALL) c2r_load<N, false, T>(jb, jb + fsj, jb + 2 * fsj, nullptr, a.kx_j, lane, qa);
        stage_twiddles();
    } else {
        const cpx<T>* st = (const cpx<T>*)a.st + ch * a.st_ch + row * a.rs_s;
        const long long fs = a.fs_s;
        c2r_load<N, true, T>(st, st + fs, st + 2 * fs, st + 3 * fs, a.kx_s, lane, qa);
        stage_twiddles();
        c2r_load<N, true, T>(st + 4 * fs, st + 5 * fs, st + 6 * fs, st + 7 * fs, a.kx_s, lane, qb);
        c2r_pair2<N, T>(qa, a.kx_s, line2, tw, lane, rho, mx, my, mz);
        if constexpr (HALL) c2r_load<N, false, T>(jb, jb + fsj, jb + 2 * fsj, nullptr, a.kx_j, lane, qa);
        c2r_pair2<N, T>(qb, a.kx_s, line2, tw, lane, bx, by, bz, e);
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
    // ---- CFL minimum of the first substage (vardt, mhd.f90:352-416), in the
    // arithmetic of the pass (FP64, or FP32 in the mixed-precision step)
    if (CFL && ch == 0) {
        const T rr = T(ctl->radius / pp.radius0);
        double dmin = pp.dx < pp.dy ? pp.dx : pp.dy;
        if (pp.dz < dmin) dmin = pp.dz;
        double dtl = __longlong_as_double(0x7FF0000000000000LL);
        if (valid) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const T d = cfl_point<T>(rho[r], bx[r], by[r], bz[r], p[r], ux[r], uy[r], uz[r], pp, rr, T(dmin));
                dtl = fmin(dtl, double(d));
            }
        }
        for (int o = 32; o > 0; o >>= 1) dtl = fmin(dtl, __shfl_xor(dtl, o));
        if ((threadIdx.x & 63) == 0)
            atomicMin(&ctl->dtmin_slot[(blockIdx.x * (NT / 64) + threadIdx.x / 64) % kCflSlots],
                      (unsigned long long)__double_as_longlong(dtl));
    }

    if constexpr (HALL) {
        T dummy[4];
        c2r_pair2<N, T>(qa, a.kx_j, line2, tw, lane, jx, jy, jz, dummy);
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

    cpx<T>* out = (cpx<T>*)a.out + ch * a.out_ch + row * a.rs_o;
    const long long fo = a.fs_o;
    const int kxo = valid ? a.kx_o : 0;
    // Output slots: 0-2 mass flux (rho u), 3-8 the symmetric momentum flux tensor
    // M_ij = (rho u_i) u_j - B_i B_j + ptot d_ij as xx, xy, xz, yy, yz, zz, 9-11 the EMF,
    // 12-14 the energy flux, 15 the expanding-box energy source. The reference's 18
    // fluxes (mhdrhs.f90:60-119) hold both m_j u_i and m_i u_j off the diagonal; they are
    // equal up to rounding, so each is formed once, with the reference's expression of
    // its first-row use (xy: my ux - By Bx, xz: mz ux - Bz Bx, yz: mz uy - Bz By).
    // Two transforms at a time; pairs ordered so that the inputs used by few fluxes die first.
    auto O = [&](int f) { return out + f * fo; };
    T f0[4], f1[4], f2[4], f3[4];
    // (E_x, E_y) with (E_z, energy flux x)
#pragma unroll
    for (int r = 0; r < 4; ++r) { f0[r] = ez[r]; f1[r] = ep[r] * ux[r] - udb[r] * bx[r]; }
    r2c_pair2<N, T>(ex, ey, O(9), O(10), f0, f1, O(11), O(12), kxo, line2, tw, lane);
    // (energy flux y, z) with (M_xx, M_xy)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        f0[r] = ep[r] * uy[r] - udb[r] * by[r]; f1[r] = ep[r] * uz[r] - udb[r] * bz[r];
        f2[r] = mx[r] * ux[r] - bx[r] * bx[r] + ptot[r]; f3[r] = my[r] * ux[r] - by[r] * bx[r];
    }
    r2c_pair2<N, T>(f0, f1, O(13), O(14), f2, f3, O(3), O(4), kxo, line2, tw, lane);
    // (M_xz, M_yy) with (M_yz, M_zz)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        f0[r] = mz[r] * ux[r] - bz[r] * bx[r]; f1[r] = my[r] * uy[r] - by[r] * by[r] + ptot[r];
        f2[r] = mz[r] * uy[r] - bz[r] * by[r]; f3[r] = mz[r] * uz[r] - bz[r] * bz[r] + ptot[r];
    }
    r2c_pair2<N, T>(f0, f1, O(5), O(6), f2, f3, O(7), O(8), kxo, line2, tw, lane);
    // (mass flux x, y) with (mass flux z, expanding-box source or nothing)
    if constexpr (AEB) {
        r2c_pair2<N, T>(mx, my, O(0), O(1), mz, f19, O(2), O(15), kxo, line2, tw, lane);
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) f0[r] = T(0);
        r2c_pair2<N, T>(mx, my, O(0), O(1), mz, f0, O(2), nullptr, kxo, line2, tw, lane);
    }
}

template <typename T, int N, bool FIRST, bool CFL, bool HALL>
static void xf_aeb(bool aeb, const XFArgs& a, const cpx<T>* tw, const PhysParams& pp, Ctl* ctl, hipStream_t s) {
    constexpr int NT = LaneFft<N>::NT;
    constexpr int RPB = NT / LaneFft<N>::LR;
    const unsigned g = (unsigned)((a.nrows + RPB - 1) / RPB * (a.nch == 2 ? 2 : 1));
    const dim3 grid(g, a.nbox > 1 ? a.nbox : 1);
    if (aeb)
        hipLaunchKernelGGL((k_xfused<T, N, FIRST, CFL, HALL, true>), grid, dim3(NT), 0, s, a, tw, pp, ctl);
    else
        hipLaunchKernelGGL((k_xfused<T, N, FIRST, CFL, HALL, false>), grid, dim3(NT), 0, s, a, tw, pp, ctl);
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
        case 96: xf_n<T, 96>(first, cfl, a, tw, pp, ctl, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_xfused<double>(int, bool, bool, const XFArgs&, const cpx<double>*, const PhysParams&, Ctl*,
                                          hipStream_t);
template hipError_t launch_xfused<float>(int, bool, bool, const XFArgs&, const cpx<float>*, const PhysParams&, Ctl*,
                                         hipStream_t);

}  // namespace laps
