// fft_dev.hpp -- device-side FFT building blocks for gfx950 (CDNA4).
//
// A 1D transform of length N is factored into 2-3 Stockham stages of radix
// R_s (register butterflies). One work-item owns butterfly j of a line; the
// first stage reads its operands straight from HBM (positions j + r*N/R0),
// the last stage leaves its results in registers at positions j + r*N/RL,
// and the stages in between exchange data through LDS. A 256-point line
// therefore costs one LDS round trip. Twiddles W_N^t come from a table
// computed in FP64 on the host (rounded once for FP32 kernels).
//
// Sign convention (fftw.f90:11): forward DIR = -1, inverse DIR = +1,
// both unnormalised.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace laps {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }

// one complex value stored with the non-temporal hint (a pass's output is read by the next
// pass, after more data than the caches hold has gone through them)
template <typename T>
__device__ __forceinline__ void store_stream(cpx<T>* p, cpx<T> v) {
    typedef T V2 __attribute__((ext_vector_type(2)));
    const V2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<V2*>(p));
}

// one complex value loaded with the non-temporal hint (read once by its pass)
template <typename T>
__device__ __forceinline__ cpx<T> load_stream(const cpx<T>* p) {
    typedef T V2 __attribute__((ext_vector_type(2)));
    const V2 w = __builtin_nontemporal_load(reinterpret_cast<const V2*>(p));
    return cpx<T>{w.x, w.y};
}

// multiply by exp(DIR * i pi / 2): DIR=-1 -> -i, DIR=+1 -> +i
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> rotq(cpx<T> a) {
    if constexpr (DIR < 0) return {a.y, -a.x};
    else return {-a.y, a.x};
}

// cos / sin of 2*pi*m/16, m = 0..15
__device__ constexpr double kCos16[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
__device__ constexpr double kSin16[16] = {
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

// cos / sin of 2*pi*m/12, m = 0..11
__device__ constexpr double kCos12[12] = {
    1.0, 0.86602540378443864676, 0.5, 0.0, -0.5, -0.86602540378443864676,
    -1.0, -0.86602540378443864676, -0.5, 0.0, 0.5, 0.86602540378443864676};
__device__ constexpr double kSin12[12] = {
    0.0, 0.5, 0.86602540378443864676, 1.0, 0.86602540378443864676, 0.5,
    0.0, -0.5, -0.86602540378443864676, -1.0, -0.86602540378443864676, -0.5};

// W_R^m = exp(DIR * 2 pi i m / R) for R | 16 or R | 12
template <int R, int DIR, typename T>
__device__ __forceinline__ cpx<T> wconst(int m) {
    if constexpr (16 % R == 0) {
        const int i = (m * (16 / R)) & 15;
        return {T(kCos16[i]), T(DIR * kSin16[i])};
    } else {
        static_assert(12 % R == 0, "radix must divide 16 or 12");
        const int i = (m * (12 / R)) % 12;
        return {T(kCos12[i]), T(DIR * kSin12[i])};
    }
}

template <int R, int DIR, typename T>
struct Dft;

template <int DIR, typename T>
struct Dft<2, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* v) {
        cpx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

// X_1,2 = a - (b + c)/2 -+ DIR i (sqrt(3)/2)(b - c) (DIR = -1: the forward roots)
template <int DIR, typename T>
struct Dft<3, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* v) {
        const cpx<T> s = v[1] + v[2];
        const cpx<T> h = v[0] - cscale(s, T(0.5));
        const cpx<T> d = cscale(rotq<DIR>(v[1] - v[2]), T(0.86602540378443864676));
        v[0] = v[0] + s;
        v[1] = h + d;
        v[2] = h - d;
    }
};

template <int DIR, typename T>
struct Dft<4, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* v) {
        cpx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
        cpx<T> t2 = v[1] + v[3], t3 = rotq<DIR>(v[1] - v[3]);
        v[0] = t0 + t2;
        v[2] = t0 - t2;
        v[1] = t1 + t3;
        v[3] = t1 - t3;
    }
};

// R = A*B in registers: n = B*n1 + n2, k = k1 + A*k2.
template <int A, int B, int DIR, typename T>
__device__ __forceinline__ void dft_split(cpx<T>* v) {
    constexpr int R = A * B;
    cpx<T> y[R];  // y[n2*A + k1]
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
        cpx<T> t[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) t[n1] = v[B * n1 + n2];
        Dft<A, DIR, T>::run(t);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            cpx<T> w = (n2 * k1 == 0) ? cpx<T>{T(1), T(0)} : wconst<R, DIR, T>(n2 * k1);
            y[n2 * A + k1] = (n2 * k1 == 0) ? t[k1] : cmul(t[k1], w);
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        cpx<T> t[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) t[n2] = y[n2 * A + k1];
        Dft<B, DIR, T>::run(t);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
    }
}

template <int DIR, typename T>
struct Dft<8, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* v) { dft_split<2, 4, DIR, T>(v); }
};

template <int DIR, typename T>
struct Dft<16, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* v) { dft_split<4, 4, DIR, T>(v); }
};

template <int DIR, typename T>
struct Dft<6, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* v) { dft_split<2, 3, DIR, T>(v); }
};

template <int DIR, typename T>
struct Dft<12, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* v) { dft_split<4, 3, DIR, T>(v); }
};

// Stage plans: radices R0, R1[, R2]. TPL = work-items per line (max butterflies).
// Plan: two stages where possible (one LDS exchange, 16 values per work-item at 256).
template <int N>
struct Plan;
template <> struct Plan<16>  { static constexpr int NST = 2, R0 = 4,  R1 = 4,  R2 = 1; };
template <> struct Plan<32>  { static constexpr int NST = 2, R0 = 8,  R1 = 4,  R2 = 1; };
template <> struct Plan<64>  { static constexpr int NST = 2, R0 = 8,  R1 = 8,  R2 = 1; };
template <> struct Plan<128> { static constexpr int NST = 2, R0 = 16, R1 = 8,  R2 = 1; };
template <> struct Plan<256> { static constexpr int NST = 2, R0 = 16, R1 = 16, R2 = 1; };
template <> struct Plan<512> { static constexpr int NST = 3, R0 = 8,  R1 = 8,  R2 = 8; };
template <> struct Plan<96>  { static constexpr int NST = 2, R0 = 16, R1 = 6,  R2 = 1; };

// PlanLow: radices <= 8, last stage radix 4 where possible: fewer values per
// work-item (more work-items per line) for kernels that keep a transform
// result live across a second transform.
template <int N>
struct PlanLow;
template <> struct PlanLow<16>  { static constexpr int NST = 2, R0 = 4, R1 = 4, R2 = 1; };
template <> struct PlanLow<32>  { static constexpr int NST = 2, R0 = 8, R1 = 4, R2 = 1; };
template <> struct PlanLow<64>  { static constexpr int NST = 3, R0 = 4, R1 = 4, R2 = 4; };
template <> struct PlanLow<128> { static constexpr int NST = 3, R0 = 8, R1 = 4, R2 = 4; };
template <> struct PlanLow<256> { static constexpr int NST = 3, R0 = 8, R1 = 8, R2 = 4; };
template <> struct PlanLow<512> { static constexpr int NST = 3, R0 = 8, R1 = 8, R2 = 8; };
template <> struct PlanLow<96>  { static constexpr int NST = 3, R0 = 4, R1 = 4, R2 = 6; };

// PlanLowW: the radices of PlanLow<256> (8·8·4), with two radix-4 butterflies per work-item in
// the last stage, so that 32 work-items per line are busy in every stage (PlanLow's last stage
// needs 64 and leaves half of them idle in the first two); 8 outputs per work-item.
template <int N>
struct PlanLowW;
template <> struct PlanLowW<256> { static constexpr int NST = 3, R0 = 8, R1 = 8, R2 = 4, S2 = 2; };

// last-stage butterflies per work-item (1 unless the plan says otherwise)
template <class P, class = void>
struct LastSplit { static constexpr int S = 1; };
template <class P>
struct LastSplit<P, std::void_t<decltype(P::S2)>> { static constexpr int S = P::S2; };

template <int N, class P = Plan<N>>
struct PlanInfo {
    static constexpr int S2 = LastSplit<P>::S;
    static constexpr int B0 = N / P::R0;
    static constexpr int B1 = N / P::R1;
    static constexpr int B2 = P::NST == 3 ? N / P::R2 / S2 : 0;   // work-items of the last stage
    static constexpr int TPL = B0 > B1 ? (B0 > B2 ? B0 : B2) : (B1 > B2 ? B1 : B2);
    static constexpr int RL2 = P::R2 * S2;
    static constexpr int RMAX = P::R0 > P::R1 ? (P::R0 > RL2 ? P::R0 : RL2) : (P::R1 > RL2 ? P::R1 : RL2);
    static constexpr int RLAST = P::NST == 3 ? RL2 : P::R1;
    static constexpr int BLAST = N / RLAST;
};

// twiddle W_N^t (table holds the forward roots exp(-2 pi i t/N))
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T>* __restrict__ tw, int t) {
    cpx<T> w = tw[t];
    if constexpr (DIR > 0) w.y = -w.y;
    return w;
}

// One radix-R butterfly of the Stockham stage with span NS, butterfly j.
template <int N, int R, int NS, int DIR, typename T>
__device__ __forceinline__ void stage_bfly(cpx<T>* v, int j, const cpx<T>* __restrict__ tw) {
    if constexpr (NS > 1) {
        const int k = j % NS;
        constexpr int STEP = N / (NS * R);
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<DIR>(tw, r * k * STEP));
    }
    Dft<R, DIR, T>::run(v);
}

template <int NS, int R>
__device__ __forceinline__ int stage_out(int j, int r) {
    return (j / NS) * NS * R + (j % NS) + r * NS;
}

struct NoPre {
    __device__ void operator()() const {}
};

// A table of N values copied from HBM into LDS by the NT work-items of a block
// in two steps: fetch() issues the loads into registers, put() writes LDS. A
// kernel calls fetch() first and put() after issuing its own first loads, so
// the table's latency hides behind them; a barrier must follow put().
template <typename V, int N, int NT>
struct LdsTable {
    static constexpr int PER = (N + NT - 1) / NT;
    V r[PER];
    __device__ __forceinline__ void fetch(const V* __restrict__ g) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int t = (int)threadIdx.x + i * NT;
            r[i] = g[t < N ? t : N - 1];  // every element assigned: the array stays in registers
        }
    }
    __device__ __forceinline__ void put(V* s) const {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int t = (int)threadIdx.x + i * NT;
            if (N % NT == 0 || t < N) s[t] = r[i];
        }
    }
};

// Stages of a line transform whose first-stage inputs are already in registers: v[r] holds
// input position j + r*B0 (r < R0) for j < B0. `after0()` runs on every work-item once the
// first stage has written its results to LDS and before the first barrier: a kernel can issue
// the loads of its next transform there, so that they are in flight during this one's LDS
// stages. On return, if j < BLAST, v[r] holds output position j + r*BLAST (r < RLAST). Every
// work-item of the block must call this. `tw` may point to LDS or HBM.
template <int N, int DIR, typename T, class P = Plan<N>, class Lds, class After0 = NoPre>
__device__ __forceinline__ bool line_fft_stages(int j, const cpx<T>* __restrict__ tw, Lds lds, cpx<T>* v,
                                                After0 after0 = {}) {
    using I = PlanInfo<N, P>;
    // stage 0 (span 1): registers -> LDS
    if (j < I::B0) {
        stage_bfly<N, P::R0, 1, DIR, T>(v, j, tw);
#pragma unroll
        for (int r = 0; r < P::R0; ++r) *lds(stage_out<1, P::R0>(j, r)) = v[r];
    }
    after0();
    __syncthreads();
    if constexpr (P::NST == 3) {
        constexpr int NS1 = P::R0;
        if (j < I::B1) {
#pragma unroll
            for (int r = 0; r < P::R1; ++r) v[r] = *lds(j + r * I::B1);
            stage_bfly<N, P::R1, NS1, DIR, T>(v, j, tw);
        }
        __syncthreads();
        if (j < I::B1) {
#pragma unroll
            for (int r = 0; r < P::R1; ++r) *lds(stage_out<NS1, P::R1>(j, r)) = v[r];
        }
        __syncthreads();
        constexpr int NS2 = P::R0 * P::R1;
        if constexpr (I::S2 == 1) {
            if (j < I::B2) {
#pragma unroll
                for (int r = 0; r < P::R2; ++r) v[r] = *lds(j + r * I::B2);
                stage_bfly<N, P::R2, NS2, DIR, T>(v, j, tw);
            }
        } else {
            // butterflies j + h*B2 (h < S2); element r of butterfly jj is position jj + r*N/R2 =
            // j + (h + S2*r)*B2, kept in v[h + S2*r]: output q at j + q*BLAST as for the others
            if (j < I::B2) {
#pragma unroll
                for (int h = 0; h < I::S2; ++h) {
                    const int jj = j + h * I::B2;
                    cpx<T> u[P::R2];
#pragma unroll
                    for (int r = 0; r < P::R2; ++r) u[r] = *lds(jj + r * (N / P::R2));
                    stage_bfly<N, P::R2, NS2, DIR, T>(u, jj, tw);
#pragma unroll
                    for (int r = 0; r < P::R2; ++r) v[h + I::S2 * r] = u[r];
                }
            }
        }
        return j < I::B2;
    } else {
        constexpr int NS1 = P::R0;
        if (j < I::B1) {
#pragma unroll
            for (int r = 0; r < P::R1; ++r) v[r] = *lds(j + r * I::B1);
            stage_bfly<N, P::R1, NS1, DIR, T>(v, j, tw);
        }
        return j < I::B1;
    }
}

// Full line transform up to the last stage's butterfly. `load(p)` supplies
// input position p and `lds(p)` is the LDS address of position p of this
// work-item's line. `pre()` runs on every work-item after the first stage's
// loads are issued and before the first barrier. On return, if j < BLAST, v[r]
// holds output position j + r*BLAST (r < RLAST). Every work-item of the block
// must call this. `tw` may point to LDS or HBM.
template <int N, int DIR, typename T, class P = Plan<N>, class Load, class Lds, class Pre = NoPre>
__device__ __forceinline__ bool line_fft_regs(int j, const cpx<T>* __restrict__ tw, Load load, Lds lds,
                                              cpx<T>* v, Pre pre = {}) {
    using I = PlanInfo<N, P>;
    if (j < I::B0) {
#pragma unroll
        for (int r = 0; r < P::R0; ++r) v[r] = load(j + r * I::B0);
    }
    pre();
    return line_fft_stages<N, DIR, T, P>(j, tw, lds, v);
}

// Full line transform; `store(p, v)` consumes output position p.
template <int N, int DIR, typename T, class P = Plan<N>, class Load, class Lds, class Store, class Pre = NoPre>
__device__ __forceinline__ void line_fft(int j, const cpx<T>* __restrict__ tw, Load load, Lds lds, Store store,
                                         Pre pre = {}) {
    using I = PlanInfo<N, P>;
    cpx<T> v[I::RMAX];
    if (line_fft_regs<N, DIR, T, P>(j, tw, load, lds, v, pre)) {
#pragma unroll
        for (int r = 0; r < I::RLAST; ++r) store(j + r * I::BLAST, v[r]);
    }
}

// Stored row s of a line kept in compact form: cnt == N is the natural FFT
// order; cnt = 2h+1 < N stores indices 0..h then N-h..N-1 (signed -h..-1).
__device__ __forceinline__ int pos_to_stored(int p, int n, int cnt) {
    if (cnt == n) return p;
    const int h = (cnt - 1) >> 1;
    if (p <= h) return p;
    if (p >= n - h) return cnt - (n - p);
    return -1;
}

}  // namespace laps
