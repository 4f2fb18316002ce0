// fft_kernels.hip -- batched 1D transforms of the 3D R2C/C2R pipeline.
//
// These replace the FFTW pencil loops and the MPI transposes of the
// reference (fftw.f90:42-222, parallel.f90:218-324): on one GPU the data
// never leaves HBM, and each pass reads its lines in whatever axis order the
// layout gives, coalesced over a tile of adjacent lines.
//
//  k_strided  : C2C along y or z. A work-group owns TX adjacent lines
//               (columns, stride 1 in memory) of one (batch1, batch2) slab;
//               adjacent work-items touch adjacent columns, so each HBM
//               access of a row of the tile is one 128-byte segment.
//               Input and output rows may be stored pruned (kept ky/kz only).
//               The z inverse of a Hall state forms the current J = ik' x B
//               from the B fields as it loads them (no J is stored).
//  k_x_c2r    : x C2R of two rows at once (Z = A + iB, one complex line).
//  k_x_r2c    : x R2C of two rows at once, split with the conjugate trick.
#include "internal.hpp"

namespace laps {

template <typename T, int N, class P = Plan<N>, int W = 0>
struct StridedCfg {
    static constexpr int TPL = PlanInfo<N, P>::TPL;
    // 128-byte rows at least (8 FP64 / 16 FP32 columns at 512, whose three-stage plan has 64
    // work-items per line: 512 or 1024 per block); 256-byte rows below 512
    static constexpr int TX0 = (N >= 512) ? (sizeof(T) == 8 ? 8 : 16) : ((sizeof(T) == 4 && TPL <= 16) ? 32 : 16);
    // W = 1: the other tile width the plan autotuner tries: half as many columns (smaller blocks,
    // more of them per CU) where the default has 16 or more, else twice as many
    static constexpr int TX = W == 0 ? TX0 : (TX0 >= 16 ? TX0 / 2 : 2 * TX0);
    static constexpr int NT = TX * TPL;
    static constexpr int LS = N + 1;  // odd line stride: adjacent lines hit different banks
    static_assert(64 % TX == 0 && (NT % 64 == 0 || 64 % NT == 0), "every wave holds whole rows of the tile");
};

// DUAL: both chains of a slab in one work-group (StridedArgs::dual), a separate instantiation so
// that the second line's registers do not weigh on the other launches of the kernel
template <typename TI, typename T, int N, int DIR, class P, int W, bool DUAL>
__global__ __launch_bounds__((StridedCfg<T, N, P, W>::NT)) void k_strided(const cpx<TI>* __restrict__ in,
                                                                     cpx<T>* __restrict__ out, StridedArgs a,
                                                                     const cpx<T>* __restrict__ tw, T scale,
                                                                     int ntile, long long nrest) {
    using C = StridedCfg<T, N, P, W>;
    __shared__ cpx<T> sh[C::TX * C::LS];
    __shared__ cpx<T> stw[N];
    long long bid = blockIdx.x;
    in += blockIdx.y * a.in_bx;    // batched boxes: one per grid row
    out += blockIdx.y * a.out_bx;
    const cpx<double>* ph = a.ph;
    if (a.nch == 2 && !DUAL) {
        int ch;
        if (a.pair) {
            // blocks b and b + 8 run the same slab of chains 0 and 1: work-groups are dealt
            // round-robin over the 8 XCDs, so both land on one XCD and the second reads the
            // input through that XCD's L2. The grid is padded to whole groups of 16.
            ch = (int)((bid >> 3) & 1);
            bid = ((bid >> 4) << 3) | (bid & 7);
            if (bid >= nrest) return;  // whole work-group, before any barrier
        } else {
            ch = bid >= nrest ? 1 : 0;  // the blocks of chain 1 follow those of chain 0
            bid -= ch * nrest;
        }
        in += ch * a.in_ch;
        out += ch * a.out_ch;
        if (ch == 0) ph = nullptr;
    }
    if (a.curl) {
        // J formed here: the nb2 field blocks of a slab (tile, b1) run close together on one
        // XCD, so the J blocks' reads of the B fields hit that XCD's L2. Work-groups are dealt
        // round-robin over the 8 XCDs; slab s, field f is block ((s / 8) nb2 + f) 8 + s % 8
        // (the grid is padded to whole groups of 8 slabs).
        const long long ns = (long long)ntile * a.nb1;
        const long long g = bid >> 3;
        const long long sl = (g / a.nb2) * 8 + (bid & 7);
        if (sl >= ns) return;  // whole work-group, before any barrier
        bid = (g % a.nb2) * ns + sl;
    }
    LdsTable<cpx<T>, N, C::NT> twt;
    twt.fetch(tw);
    const int t = threadIdx.x;
    const int l = t % C::TX, j = t / C::TX;
    const int tile = (int)(bid % ntile);
    bid /= ntile;
    const int b1 = (int)(bid % a.nb1);
    const int b2 = (int)(bid / a.nb1);
    const int col = tile * C::TX + l;
    const bool valid = col < a.ncol;
    const cpx<TI>* ib = in + b1 * a.in_b1 + b2 * a.in_b2 + (valid ? col : 0);
    cpx<T>* ob = out + b1 * a.out_b1 + b2 * a.out_b2 + (valid ? col : 0);
    cpx<T>* line = sh + l * C::LS;
    const int in_cnt = a.in_cnt, out_cnt = a.out_cnt;
    const long long in_ls = a.in_ls, out_ls = a.out_ls;
    // grid shift: the column and slab factors are fixed per work-item
    const cpx<double> wcb = ph ? cmul(a.phc[valid ? col : 0], a.phb[b1]) : cpx<double>{1.0, 0.0};
    // z inverse: a row of the mask with no kept mode is zero in and out
    bool dead = false;
    if (a.row_role == 1) {
        const int my = (a.cy_cnt == a.mY) ? (b1 <= a.cy_cnt / 2 ? b1 : b1 - a.cy_cnt)
                                          : (b1 <= (a.cy_cnt - 1) / 2 ? b1 : b1 - a.cy_cnt);
        dead = mask_row_dead(a.mask, a.mX, a.mY, a.mZ, col, my);
        // every wave holds all TX columns of the tile (TX divides 64): a tile with no live row
        // is skipped whole, before any barrier (about a quarter of the tiles of a mask)
        if (__ballot(valid && !dead) == 0) return;
    }
    // Hall current from B (curl): component jc of J = i k' x B, from fields 4 + (jc+1)%3 and
    // 4 + (jc+2)%3 of this slab's column (the evaluation of k_curl, mhdrhs.f90:313-339)
    const int jc = (a.curl && b2 >= 8) ? b2 - 8 : -1;
    const cpx<TI>* ibs = ib;
    const cpx<TI>* ibt = ib;
    double kxy[2] = {0.0, 0.0};  // the k' components of index (jc+1)%3 and (jc+2)%3 that do not depend on kz
    double rz = 1.0;             // R0/R
    int sz = 0;                  // which of the two is kz' (0: none)
    if (jc >= 0) {
        const int c1 = (jc + 1) % 3, c2 = (jc + 2) % 3;
        ibs = ib + (long long)(4 + c1 - b2) * a.in_b2;
        ibt = ib + (long long)(4 + c2 - b2) * a.in_b2;
        rz = a.k_r0 / (a.kctl + blockIdx.y)->radius;
        const double wx = a.kwx[valid ? col : 0], wy = a.kwy[b1];
        double kx = wx, ky = wy * rz;
        if (a.k_corot) {
            kx = wx * a.k_cos + wy * a.k_sin;
            ky = (-wx * a.k_sin + wy * a.k_cos) * rz;
        }
        const double kk[3] = {kx, ky, 0.0};
        kxy[0] = kk[c1];
        kxy[1] = kk[c2];
        sz = c1 == 2 ? 1 : (c2 == 2 ? 2 : 0);
    }
    // J_jc = i (k_c1 B_c2 - k_c2 B_c1), kz' = wz R0/R of stored z row s (the reference's
    // (w R0)/R within an ulp)
    auto curl_at = [&](int s) -> cpx<T> {
        const double kz = a.kwz[s] * rz;
        const double k1 = sz == 1 ? kz : kxy[0], k2 = sz == 2 ? kz : kxy[1];
        const cpx<TI> b1v = ibs[s * in_ls], b2v = ibt[s * in_ls];
        const cpx<double> u{double(b2v.x) * k1 - double(b1v.x) * k2, double(b2v.y) * k1 - double(b1v.y) * k2};
        return cpx<T>{T(-u.y), T(u.x)};
    };
    // every input element is read once, except the B fields that the J blocks read again
    const bool once = !a.curl || b2 < 4 || b2 == 7;
    auto rd = [&](int s) -> cpx<TI> { return once ? load_stream(ib + s * in_ls) : ib[s * in_ls]; };
    auto load = [&](int p) -> cpx<T> {
        const int s = pos_to_stored(p, N, in_cnt);
        if (!valid || s < 0 || dead) return cpx<T>{T(0), T(0)};
        if (a.row_role == 2 && mask_row_dead(a.mask, a.mX, a.mY, a.mZ, col, p <= N / 2 ? p : p - N))
            return cpx<T>{T(0), T(0)};
        if (jc >= 0) {
            const cpx<T> v = curl_at(s);
            if (ph) {
                const cpx<double> w = cmul(ph[p], wcb);
                return cmul(v, cpx<T>{T(w.x), T(w.y)});
            }
            return v;
        }
        const cpx<TI> v = rd(s);
        if (ph) {
            const cpx<double> w = cmul(ph[p], wcb);
            return cmul(cpx<T>{T(v.x), T(v.y)}, cpx<T>{T(w.x), T(w.y)});
        }
        return cpx<T>{T(v.x), T(v.y)};
    };
    auto lds = [&](int p) { return line + p; };
    auto store = [&](int p, cpx<T> v) {
        const int s = pos_to_stored(p, N, out_cnt);
        if (valid && s >= 0 && !dead) store_stream(ob + s * out_ls, cpx<T>{v.x * scale, v.y * scale});
    };
    cpx<T>* const stp = stw;
    if constexpr (DUAL) {
        // both chains: chain 0 from the input as loaded, chain 1 from the same values times the
        // grid shift, so the input is read and waited for once per slab
        using I = PlanInfo<N, P>;
        cpx<T> v[I::RMAX], w[I::RMAX];
        if (j < I::B0) {
#pragma unroll
            for (int r = 0; r < P::R0; ++r) {
                const int p = j + r * I::B0;
                const int s = pos_to_stored(p, N, in_cnt);
                cpx<T> x{T(0), T(0)};
                if (valid && s >= 0 && !dead) {
                    if (jc >= 0) {
                        x = curl_at(s);
                    } else {
                        const cpx<TI> q = rd(s);
                        x = cpx<T>{T(q.x), T(q.y)};
                    }
                }
                v[r] = x;
                const cpx<double> sh1 = cmul(ph[p], wcb);
                w[r] = cmul(x, cpx<T>{T(sh1.x), T(sh1.y)});
            }
        }
        twt.put(stp);
        if (line_fft_stages<N, DIR, T, P>(j, stw, lds, v)) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) store(j + r * I::BLAST, v[r]);
        }
        __syncthreads();  // every read of the LDS line before the second transform writes it
        ob += a.out_ch;
        if (line_fft_stages<N, DIR, T, P>(j, stw, lds, w)) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) store(j + r * I::BLAST, w[r]);
        }
        return;
    }
    line_fft<N, DIR, T, P>(j, stw, load, lds, store, [twt, stp]() { twt.put(stp); });
}

template <typename T, int N>
struct XCfg {
    static constexpr int TPL = PlanInfo<N>::TPL;  // default plan
    static constexpr int LP = TPL >= 128 ? 1 : 128 / TPL;  // line pairs per work-group
    static constexpr int NT = LP * TPL;
    static constexpr int PADSH = sizeof(T) == 8 ? 4 : 5;    // one pad slot per 256 bytes
    static constexpr int LS = N + (N >> PADSH) + 1;
    __device__ static constexpr int pad(int p) { return p + (p >> PADSH); }
};

// x C2R, unnormalised (plan_ifft_x, fftw.f90:31, :94-100). Rows 2q and 2q+1
// are packed as Z_p = A_p + i B_p over the Hermitian extension; only the
// real parts of the DC and Nyquist coefficients are read.
template <typename T, int N>
__global__ __launch_bounds__((XCfg<T, N>::NT)) void k_x_c2r(const cpx<T>* __restrict__ in, T* __restrict__ out,
                                                          XArgs a, const cpx<T>* __restrict__ tw) {
    using C = XCfg<T, N>;
    __shared__ cpx<T> sh[C::LP * C::LS];
    const int l = threadIdx.x / C::TPL, j = threadIdx.x % C::TPL;
    const long long q = (long long)blockIdx.x * C::LP + l;
    const bool valid = 2 * q + 1 < a.nrows;
    const long long qq = valid ? q : 0;
    const cpx<T>* ra = in + (2 * qq) * a.crow;
    const cpx<T>* rb = ra + a.crow;
    T* oa = out + (2 * qq) * a.rrow;
    T* ob = oa + a.rrow;
    const int kx = a.kx;
    cpx<T>* line = sh + l * C::LS;
    auto load = [&](int p) -> cpx<T> {
        if (!valid) return cpx<T>{T(0), T(0)};
        if (p == 0) return cpx<T>{ra[0].x, rb[0].x};
        if (p < N / 2) {
            if (p >= kx) return cpx<T>{T(0), T(0)};
            const cpx<T> A = ra[p], B = rb[p];
            return cpx<T>{A.x - B.y, A.y + B.x};
        }
        if (p == N / 2) {
            if (p >= kx) return cpx<T>{T(0), T(0)};
            return cpx<T>{ra[p].x, rb[p].x};
        }
        const int m = N - p;
        if (m >= kx) return cpx<T>{T(0), T(0)};
        const cpx<T> A = ra[m], B = rb[m];
        return cpx<T>{A.x + B.y, B.x - A.y};
    };
    auto lds = [&](int p) { return line + C::pad(p); };
    auto store = [&](int p, cpx<T> v) {
        if (valid) {
            oa[p] = v.x;
            ob[p] = v.y;
        }
    };
    line_fft<N, +1, T>(j, tw, load, lds, store);
}

// x R2C, unnormalised (plan_fft_x, fftw.f90:27): Z = FFT(a + i b);
// A_k = (Z_k + conj Z_{N-k})/2, B_k = (Z_k - conj Z_{N-k})/(2i), k < kx.
template <typename T, int N>
__global__ __launch_bounds__((XCfg<T, N>::NT)) void k_x_r2c(const T* __restrict__ in, cpx<T>* __restrict__ out,
                                                          XArgs a, const cpx<T>* __restrict__ tw) {
    using C = XCfg<T, N>;
    using I = PlanInfo<N>;
    __shared__ cpx<T> sh[C::LP * C::LS];
    const int l = threadIdx.x / C::TPL, j = threadIdx.x % C::TPL;
    const long long q = (long long)blockIdx.x * C::LP + l;
    const bool valid = 2 * q + 1 < a.nrows;
    const long long qq = valid ? q : 0;
    const T* ia = in + (2 * qq) * a.rrow;
    const T* ib = ia + a.rrow;
    cpx<T>* oa = out + (2 * qq) * a.crow;
    cpx<T>* ob = oa + a.crow;
    cpx<T>* line = sh + l * C::LS;
    auto load = [&](int p) -> cpx<T> {
        if (!valid) return cpx<T>{T(0), T(0)};
        return cpx<T>{ia[p], ib[p]};
    };
    auto lds = [&](int p) { return line + C::pad(p); };
    cpx<T> v[I::RMAX];
    const bool has = line_fft_regs<N, -1, T>(j, tw, load, lds, v);
    __syncthreads();
    if (has) {
#pragma unroll
        for (int r = 0; r < I::RLAST; ++r) *lds(j + r * I::BLAST) = v[r];
    }
    __syncthreads();
    if (valid) {
        const T h = T(0.5);
        for (int k = j; k < a.kx; k += C::TPL) {
            const cpx<T> zk = *lds(k);
            const cpx<T> zn = *lds(k == 0 ? 0 : N - k);
            oa[k] = cpx<T>{h * (zk.x + zn.x), h * (zk.y - zn.y)};
            ob[k] = cpx<T>{h * (zk.y + zn.y), h * (zn.x - zk.x)};
        }
    }
}

// ---------------------------------------------------------------- launchers

template <typename TI, typename T, int N, int DIR, class P, int W>
static hipError_t do_strided_p(const cpx<TI>* in, cpx<T>* out, const StridedArgs& a, const cpx<T>* tw, T scale,
                               hipStream_t s) {
    using C = StridedCfg<T, N, P, W>;
    const int ntile = (a.ncol + C::TX - 1) / C::TX;
    const long long nrest = (long long)ntile * a.nb1 * a.nb2;
    if (nrest <= 0) return hipSuccess;
    if (a.curl && (DIR < 0 || (a.nch == 2 && !a.dual))) return hipErrorInvalidValue;
    const long long nslab = (long long)ntile * a.nb1;
    const long long nblk = a.curl ? (nslab + 7) / 8 * 8 * a.nb2
                                  : ((a.nch == 2 && !a.dual) ? 2 * (a.pair ? (nrest + 7) / 8 * 8 : nrest) : nrest);
    const dim3 grid((unsigned)nblk, a.nbox > 1 ? a.nbox : 1);
    if constexpr (DIR > 0) {
        if (a.dual) {
            if (a.nch != 2 || !a.ph) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_strided<TI, T, N, DIR, P, W, true>), grid, dim3(C::NT), 0, s, in, out, a, tw, scale,
                               ntile, nrest);
            return hipGetLastError();
        }
    } else {
        if (a.dual) return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL((k_strided<TI, T, N, DIR, P, W, false>), grid, dim3(C::NT), 0, s, in, out, a, tw, scale, ntile,
                       nrest);
    return hipGetLastError();
}

// variant: bit 0 the plan (0 Plan, 1 the low-radix plan: PlanLowW at 256, PlanLow otherwise),
// bit 1 the tile width (inverse passes only)
template <typename TI, typename T, int N, int DIR>
static hipError_t do_strided(int variant, const cpx<TI>* in, cpx<T>* out, const StridedArgs& a, const cpx<T>* tw,
                             T scale, hipStream_t s) {
    using Low = std::conditional_t<N == 256, PlanLowW<256>, PlanLow<N>>;
    const bool low = variant & 1;
    if constexpr (DIR > 0) {
        if (variant & 2) {
            if (low) return do_strided_p<TI, T, N, DIR, Low, 1>(in, out, a, tw, scale, s);
            return do_strided_p<TI, T, N, DIR, Plan<N>, 1>(in, out, a, tw, scale, s);
        }
    }
    if (low) return do_strided_p<TI, T, N, DIR, Low, 0>(in, out, a, tw, scale, s);
    return do_strided_p<TI, T, N, DIR, Plan<N>, 0>(in, out, a, tw, scale, s);
}

template <typename TI, typename T>
hipError_t launch_strided(int n, int dir, const cpx<TI>* in, cpx<T>* out, const StridedArgs& a, const cpx<T>* tw,
                          T scale, hipStream_t s, int variant) {
#define LAPS_STRIDED_CASE(NN)                                                              \
    case NN:                                                                               \
        return dir < 0 ? do_strided<TI, T, NN, -1>(variant, in, out, a, tw, scale, s)      \
                       : do_strided<TI, T, NN, +1>(variant, in, out, a, tw, scale, s);
    switch (n) {
        LAPS_STRIDED_CASE(16)
        LAPS_STRIDED_CASE(32)
        LAPS_STRIDED_CASE(64)
        LAPS_STRIDED_CASE(128)
        LAPS_STRIDED_CASE(256)
        LAPS_STRIDED_CASE(512)
        LAPS_STRIDED_CASE(96)
        default:
            return hipErrorInvalidValue;
    }
#undef LAPS_STRIDED_CASE
}

template <typename T, int N>
static hipError_t do_x(bool c2r, const void* in, void* out, const XArgs& a, const cpx<T>* tw, hipStream_t s) {
    using C = XCfg<T, N>;
    const long long npairs = a.nrows / 2;
    const long long nblk = (npairs + C::LP - 1) / C::LP;
    if (nblk <= 0) return hipSuccess;
    if (c2r)
        hipLaunchKernelGGL((k_x_c2r<T, N>), dim3((unsigned)nblk), dim3(C::NT), 0, s, (const cpx<T>*)in, (T*)out, a,
                           tw);
    else
        hipLaunchKernelGGL((k_x_r2c<T, N>), dim3((unsigned)nblk), dim3(C::NT), 0, s, (const T*)in, (cpx<T>*)out, a,
                           tw);
    return hipGetLastError();
}

template <typename T>
static hipError_t dispatch_x(int n, bool c2r, const void* in, void* out, const XArgs& a, const cpx<T>* tw,
                             hipStream_t s) {
    if (a.nrows % 2) return hipErrorInvalidValue;
    switch (n) {
        case 16: return do_x<T, 16>(c2r, in, out, a, tw, s);
        case 32: return do_x<T, 32>(c2r, in, out, a, tw, s);
        case 64: return do_x<T, 64>(c2r, in, out, a, tw, s);
        case 128: return do_x<T, 128>(c2r, in, out, a, tw, s);
        case 256: return do_x<T, 256>(c2r, in, out, a, tw, s);
        case 512: return do_x<T, 512>(c2r, in, out, a, tw, s);
        case 96: return do_x<T, 96>(c2r, in, out, a, tw, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_x_c2r(int n, const cpx<T>* in, T* out, const XArgs& a, const cpx<T>* tw, hipStream_t s) {
    return dispatch_x<T>(n, true, in, out, a, tw, s);
}
template <typename T>
hipError_t launch_x_r2c(int n, const T* in, cpx<T>* out, const XArgs& a, const cpx<T>* tw, hipStream_t s) {
    return dispatch_x<T>(n, false, in, out, a, tw, s);
}

template hipError_t launch_strided<double, double>(int, int, const cpx<double>*, cpx<double>*, const StridedArgs&,
                                                   const cpx<double>*, double, hipStream_t, int);
template hipError_t launch_strided<double, float>(int, int, const cpx<double>*, cpx<float>*, const StridedArgs&,
                                                  const cpx<float>*, float, hipStream_t, int);
template hipError_t launch_strided<float, float>(int, int, const cpx<float>*, cpx<float>*, const StridedArgs&,
                                                 const cpx<float>*, float, hipStream_t, int);
template hipError_t launch_x_c2r<double>(int, const cpx<double>*, double*, const XArgs&, const cpx<double>*,
                                         hipStream_t);
template hipError_t launch_x_c2r<float>(int, const cpx<float>*, float*, const XArgs&, const cpx<float>*,
                                        hipStream_t);
template hipError_t launch_x_r2c<double>(int, const double*, cpx<double>*, const XArgs&, const cpx<double>*,
                                         hipStream_t);
template hipError_t launch_x_r2c<float>(int, const float*, cpx<float>*, const XArgs&, const cpx<float>*,
                                        hipStream_t);

}  // namespace laps
