// fwd_fused.hip -- forward y and z passes of the fluxes with the spectral
// divergence / curl folded in, and the RK3 update folded into the z pass.
//
// Reference: calc_rhs (mhdrhs.f90:174-279), rkt (rktmod.f90:34-62),
// dealias (dealiasing.f90:70-113) and the y/z legs of
// transform_flux_real_to_fourier (fftw.f90:136-180).
//
// fnl = -(kx F_a + ky F_b + kz F_c) (k = i*k') is linear and each derivative
// factor depends only on modes the pass already knows: kx' and ky' on the y
// pass output, kz' on the z pass output. So the y pass of a divergence group
// (F_a, F_b, F_c) writes  G = -i kx' Y(F_a) - i ky' Y(F_b)  and  Y(F_c)
// (three fields in, two out), and the z pass forms fnl = Z(G) - i kz' Z(Y F_c)
// per mode, where the RK3 update, implicit diffusion and dealiasing follow in
// registers. The induction group (E) writes Y(E_x), Y(E_y) and the partial
// curls P5 = -i ky' Y(E_z), P6 = i kx' Y(E_z), P7 = i ky' Y(E_x) - i kx' Y(E_y).
// With corotating AEB, kx' and ky' both depend on (kx, ky); both are known per
// output mode of the y pass.

#include <type_traits>

#include "internal.hpp"
#include "physics_dev.hpp"

namespace laps {

// ZP: the z pass, whose blocks also carry the RHS and RK3 update. The passes use the plan of
// radices <= 8 (8 values per work-item, so a block keeps one transform's result live across
// the next at 4-5 waves per SIMD), except the FP32 y pass at 256, which runs faster on the
// two-stage radix-16x16 plan (one LDS exchange, 16 values per work-item: 2.62 against 3.15
// ms per mixed phase-shift step). On that plan the FP32 z pass ran at 2 waves per SIMD
// (236 registers) and took 2.83 ms; on the low-radix plan 2.74 ms, and 2.58 ms on its
// variant with 32 busy work-items per line and 16 columns (PlanLowW; the FP64 passes on it
// need 150-180 registers, one block per CU, or spill).
// ALT (the y pass at 256 only): the wide 8·8·4 plan with 16 columns. FP32: instead of the
// two-stage plan with 32, for launches of one chain (0.82 -> 0.79 ms per mixed 2/3-rule step;
// with the two chains of phase shift the two-stage plan is faster, 2.47 against 2.69). FP64:
// instead of PlanLow with 8, for launches of two chains, held to 128 registers (92 bytes
// spilled; 4.56 -> 4.16 ms per phase-shift step; with one chain 1.17 -> 1.18, not used)
template <typename T, int N, bool ZP, bool ALT = false>
struct FwdCfg {
    // the FP32 z pass at 256 and the ALT y passes: the 8·8·4 plan with two last-stage
    // butterflies per work-item (32 per line, none idle) and 16 columns, 128-byte rows of D,
    // Fp and Uf in FP32, 256-byte rows of the y pass in FP64
    static constexpr bool WIDE = N == 256 && (ALT || (sizeof(T) == 4 && ZP));
    using P = std::conditional_t<(sizeof(T) == 4 && N == 256 && !ZP && !ALT), Plan<N>,
                                 std::conditional_t<WIDE, PlanLowW<N>, PlanLow<N>>>;
    using I = PlanInfo<N, P>;
    // columns per block: rows of 128 bytes (8 FP64 or 16 FP32 columns), 256 bytes for the FP32
    // y pass on the two-stage plan; the FP32 z pass has 16 columns only on the wide plan (on
    // PlanLow, 1024-work-item blocks, it measured slower at 256), else 8
    static constexpr int TX = (sizeof(T) == 8) ? (WIDE ? 16 : 8) : (WIDE ? 16 : (ZP ? 8 : (N == 256 ? 32 : 16)));
    static constexpr int TPL = I::TPL;
    static constexpr int NT = TX * TPL;
    static constexpr int LS = N + 1;
    // waves per SIMD the register allocation must allow: the FP32 y pass at 256 on the
    // two-stage plan and the FP64 ALT y pass (512-thread blocks, 66-72 KB of LDS) fit two
    // blocks per CU only at <= 128 VGPRs
    static constexpr int MINW = (N == 256 && !ZP && (sizeof(T) == 4 ? !ALT : ALT)) ? 4 : 1;
    static_assert(64 % TX == 0 && (NT % 64 == 0 || 64 % NT == 0), "every wave holds whole rows of the tile");
};

// derivative factors k' (real; the spectral factor is i*k'), mhdrhs.f90:192-204; r0r = R0/R,
// formed once per work-item (the reference's (k*R0)/R per mode, within an ulp)
__device__ __forceinline__ void kprime_xy(double wx, double wy, const PhysParams& pp, double r0r, double& kx,
                                          double& ky) {
    kx = wx;
    ky = wy * r0r;
    if (pp.corot_k) {
        kx = wx * pp.cos_c + wy * pp.sin_c;
        ky = (-wx * pp.sin_c + wy * pp.cos_c) * r0r;
    }
}

template <typename T>
__device__ __forceinline__ cpx<T> mul_i(T k, cpx<T> z) { return {-k * z.y, k * z.x}; }  // (i k) z

// ---------------------------------------------------------------- y pass
template <typename T, int N, bool ALT>
__global__ __launch_bounds__((FwdCfg<T, N, false, ALT>::NT), (FwdCfg<T, N, false, ALT>::MINW)) void k_yfwd_group(const cpx<T>* __restrict__ in,
                                                                 cpx<T>* __restrict__ out, YGArgs a,
                                                                 const cpx<T>* __restrict__ tw,
                                                                 const double* __restrict__ wxt,
                                                                 const double* __restrict__ wyf, PhysParams pp,
                                                                 const Ctl* __restrict__ ctl, int ntile, int nz) {
    using C = FwdCfg<T, N, false, ALT>;
    using I = typename C::I;
    // FP32: the derivative factors come from LDS tables of their y-dependent parts (below)
    constexpr bool TAB = sizeof(T) == 4;
    __shared__ cpx<T> sh[C::TX * C::LS];
    __shared__ cpx<T> stw[N];
    __shared__ double swy[TAB ? 1 : N];
    __shared__ T sax[TAB ? N : 1], say[TAB ? N : 1];
    LdsTable<cpx<T>, N, C::NT> twt;
    LdsTable<double, N, C::NT> wyt;
    twt.fetch(tw);
    wyt.fetch(wyf);
    twt.put(stw);  // visible after the barrier that opens each transform
    if constexpr (!TAB) wyt.put(swy);
    const int l = threadIdx.x % C::TX, j = threadIdx.x / C::TX;
    in += blockIdx.y * a.in_bx;    // batched boxes: one per grid row
    out += blockIdx.y * a.out_bx;
    ctl += blockIdx.y;
    long long bid = blockIdx.x;
    const long long bpc = (long long)ntile * nz * a.ngroups;  // blocks per chain
    if (a.nch == 2 && bid >= bpc) {
        bid -= bpc;
        in += a.in_ch;
        out += a.out_ch;
    }
    const int tile = (int)(bid % ntile);
    bid /= ntile;
    const int z = (int)(bid % nz);
    const int gi = (int)(bid / nz);
    const YGroup G = a.g[gi];
    const int col = tile * C::TX + l;
    const bool valid = col < a.ncol;
    const int kx = valid ? col : 0;
    const double radius = ctl->radius;
    const double r0r = pp.radius0 / radius;  // R0/R of the wave-number scalings
    const double wx = wxt[kx];
    // derivative factors kx', ky' of output position p of this column (mhdrhs.f90:192-204).
    // FP64: the reference's expressions per mode (kprime_xy, R0/R formed once). FP32: kx' = cx + ax[p] and
    // ky' = cy + ay[p], the y-dependent parts tabulated in LDS once per block (the pass
    // rounds the factors to FP32 either way), so no factor is held in registers across
    // transforms and no FP64 quotient is formed per mode.
    T cx = T(0), cy = T(0);
    if constexpr (TAB) {
        const double rr = r0r;
#pragma unroll
        for (int i = 0; i < decltype(wyt)::PER; ++i) {
            const int t = (int)threadIdx.x + i * C::NT;
            if (t < N) {
                const double wy = wyt.r[i];
                sax[t] = pp.corot_k ? T(wy * pp.sin_c) : T(0);
                say[t] = pp.corot_k ? T(wy * pp.cos_c * rr) : T(wy * rr);
            }
        }
        cx = pp.corot_k ? T(wx * pp.cos_c) : T(wx);
        cy = pp.corot_k ? T(-wx * pp.sin_c * rr) : T(0);
    }
    auto kxy = [&](int p, T& kxv, T& kyv) {
        if constexpr (TAB) {
            kxv = cx + sax[p];
            kyv = cy + say[p];
        } else {
            double kxp, kyp;
            kprime_xy(wx, swy[p], pp, r0r, kxp, kyp);
            kxv = T(kxp);
            kyv = T(kyp);
        }
    };
    const int ylim = a.ylim ? a.ylim[kx] : N;
    cpx<T>* line = sh + l * C::LS;
    auto lds = [&](int p) { return line + p; };
    const cpx<T>* ib = in + (long long)z * N * a.rs + kx;
    cpx<T>* ob = out + (long long)z * a.cy * a.rs + kx;

    auto fft_field = [&](int f, cpx<T>* v) {
        const cpx<T>* src = ib + f * a.in_fs;
        auto load = [&](int p) -> cpx<T> { return valid ? load_stream(src + p * a.rs) : cpx<T>{T(0), T(0)}; };
        __syncthreads();  // the LDS line may still be read by the previous transform
        return line_fft_regs<N, -1, T, typename C::P>(j, stw, load, lds, v);
    };
    auto store = [&](int f, int p, cpx<T> val) {
        const int s = pos_to_stored(p, N, a.cy);
        const int my = p <= N / 2 ? p : p - N;
        if (valid && s >= 0 && my <= ylim && -my <= ylim) store_stream(ob + f * a.out_fs + s * a.rs, val);
    };

    cpx<T> va[I::RMAX], vb[I::RMAX];
    if (G.kind == 3) {
        // symmetric momentum flux tensor M = (xx, xy, xz, yy, yz, zz) in in[0..5]:
        // out = -i(kx' Y M_xx + ky' Y M_xy), Y M_xz, -i(kx' Y M_xy + ky' Y M_yy), Y M_yz,
        //       -i(kx' Y M_xz + ky' Y M_yz), Y M_zz  (the x, y, z momentum divergence groups;
        // each off-diagonal component is transformed once and used by both rows)
        auto neg = [](cpx<T> z) { return cpx<T>{-z.x, -z.y}; };
        const bool h = fft_field(G.in[0], va);                                      // xx
        // the tables are visible after the transform's first barrier; FP64 keeps the
        // factors in registers (4 per work-item), FP32 reads them from the tables at each use
        constexpr int NKC = TAB ? 1 : I::RLAST;
        T kxc[NKC], kyc[NKC];
        if constexpr (!TAB) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) kxy(h ? j + r * I::BLAST : 0, kxc[r], kyc[r]);
        }
        auto kxr = [&](int r) -> T {
            if constexpr (TAB) return cx + sax[h ? j + r * I::BLAST : 0];
            else return kxc[r];
        };
        auto kyr = [&](int r) -> T {
            if constexpr (TAB) return cy + say[h ? j + r * I::BLAST : 0];
            else return kyc[r];
        };
#pragma unroll
        for (int r = 0; r < I::RLAST; ++r) va[r] = mul_i(kxr(r), va[r]);
        fft_field(G.in[1], vb);                                                     // xy
        if (h) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) {
                store(G.out[0], j + r * I::BLAST, neg(va[r] + mul_i(kyr(r), vb[r])));
                vb[r] = mul_i(kxr(r), vb[r]);                                       // y row, x part
            }
        }
        fft_field(G.in[3], va);                                                     // yy
        if (h) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) store(G.out[2], j + r * I::BLAST, neg(vb[r] + mul_i(kyr(r), va[r])));
        }
        fft_field(G.in[2], va);                                                     // xz
        if (h) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) {
                store(G.out[1], j + r * I::BLAST, va[r]);
                va[r] = mul_i(kxr(r), va[r]);                                       // z row, x part
            }
        }
        fft_field(G.in[4], vb);                                                     // yz
        if (h) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) {
                store(G.out[3], j + r * I::BLAST, vb[r]);
                store(G.out[4], j + r * I::BLAST, neg(va[r] + mul_i(kyr(r), vb[r])));
            }
        }
        fft_field(G.in[5], va);                                                     // zz
        if (h) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) store(G.out[5], j + r * I::BLAST, va[r]);
        }
        return;
    }
    if (G.kind == 2) {
        if (fft_field(G.in[0], va)) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) store(G.out[0], j + r * I::BLAST, va[r]);
        }
        return;
    }
    const bool has = fft_field(G.in[0], va);
    if (G.kind == 1 && has) {
#pragma unroll
        for (int r = 0; r < I::RLAST; ++r) store(G.out[0], j + r * I::BLAST, va[r]);  // Y(E_x)
    }
    fft_field(G.in[1], vb);
    if (has) {
#pragma unroll
        for (int r = 0; r < I::RLAST; ++r) {
            const int p = j + r * I::BLAST;
            T kx_, ky_;
            kxy(p, kx_, ky_);
            if (G.kind == 0) {
                const cpx<T> gsum = mul_i(kx_, va[r]) + mul_i(ky_, vb[r]);
                store(G.out[0], p, cpx<T>{-gsum.x, -gsum.y});                  // -(i kx' Ya + i ky' Yb)
            } else {
                store(G.out[1], p, vb[r]);                                       // Y(E_y)
                store(G.out[4], p, mul_i(ky_, va[r]) - mul_i(kx_, vb[r]));      // P7
            }
        }
    }
    fft_field(G.in[2], va);
    if (has) {
#pragma unroll
        for (int r = 0; r < I::RLAST; ++r) {
            const int p = j + r * I::BLAST;
            if (G.kind == 0) {
                store(G.out[1], p, va[r]);                                       // Y(F_c)
            } else {
                T kx_, ky_;
                kxy(p, kx_, ky_);
                const cpx<T> t5 = mul_i(ky_, va[r]);
                store(G.out[2], p, cpx<T>{-t5.x, -t5.y});                        // P5 = -i ky' Y(E_z)
                store(G.out[3], p, mul_i(kx_, va[r]));                           // P6 =  i kx' Y(E_z)
            }
        }
    }
}

// ---------------------------------------------------------------- z pass + RK3
// Output variable v is  sum_t  m_t * Z(D[f_t])  with m_t in {1, -i kz', +i kz'}.
template <typename T, int N, int OPT>
__global__ __launch_bounds__((FwdCfg<T, N, true>::NT), (FwdCfg<T, N, true>::MINW)) void k_zfwd_rhs(const cpx<T>* __restrict__ D,
                                                               cpx<double>* __restrict__ U,
                                                               cpx<T>* __restrict__ Fp, ZArgs a,
                                                               const cpx<T>* __restrict__ tw, BoxTables tb,
                                                               const double* __restrict__ wzf, PhysParams pp,
                                                               const Ctl* __restrict__ ctl, int irk, int ntile,
                                                               T scale) {
    using C = FwdCfg<T, N, true>;
    using I = typename C::I;
    // batched boxes: one per grid row
    ctl += blockIdx.y;
    D += blockIdx.y * a.d_bx;
    U += blockIdx.y * a.u_bx;
    Fp += blockIdx.y * a.fp_bx;
    const cpx<T>* const Ds = a.Ds ? (const cpx<T>*)a.Ds + blockIdx.y * a.d_bx : nullptr;
    cpx<float>* const Uf = a.Uf ? a.Uf + blockIdx.y * a.uf_bx : nullptr;
    if (ctl->status != 0) return;  // uniform over the box: the step was refused
    __shared__ cpx<T> sh[C::TX * C::LS];
    __shared__ cpx<T> stw[N];
    __shared__ double swz[N];
    LdsTable<cpx<T>, N, C::NT> twt;
    LdsTable<double, N, C::NT> wzt;
    twt.fetch(tw);
    wzt.fetch(wzf);
    twt.put(stw);  // visible after the barrier that opens each transform
    wzt.put(swz);
    const int l = threadIdx.x % C::TX, j = threadIdx.x / C::TX;
    // block -> (variable v, kx tile, cy): the eight conserved variables of a
    // column tile are independent, one block each
    const int v = (int)(blockIdx.x % 8);
    const int tile = (int)((blockIdx.x / 8) % ntile);
    const int cy = (int)(blockIdx.x / 8 / ntile);
    const int col = tile * C::TX + l;
    const bool valid = col < a.KX;
    const int kx = valid ? col : 0;
    const double radius = ctl->radius, tau = ctl->tau;
    const double r0r = pp.radius0 / radius;  // R0/R of the wave-number scalings
    const double c = ctl->cc[irk], d = ctl->dd[irk], ts = ctl->ts[irk];
    const double wx = tb.wx[kx], wy = tb.wy[cy];
    const int lim = tb.kzlim[(long long)cy * a.KX + kx];
    const bool live = valid && (OPT != 1 || lim >= 0);
    // every wave holds all TX columns of the tile: a tile with no kept mode has nothing to
    // transform, read or write (unless the modes outside the kept set must be cleared)
    if (OPT == 1 && !a.clear_outside && __ballot(live) == 0) return;
    cpx<T>* line = sh + l * C::LS;
    auto lds = [&](int p) { return line + p; };
    const cpx<T>* ib = D + (long long)cy * a.rs + kx;
    const long long zls = (long long)a.CY * a.rs;   // z line stride of D
    const long long mode0 = (long long)cy * a.rs + kx;

    {
        const ZVar V = a.v[v];
        cpx<T> acc[I::RMAX], w[I::RMAX];
        // sum_t m_t Z(base[f_t]) into acc (m_t in {1, -i kz', +i kz'})
        auto terms = [&](const cpx<T>* base) {
            bool h = false;
#pragma unroll
            for (int r = 0; r < I::RMAX; ++r) acc[r] = cpx<T>{T(0), T(0)};
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (t >= V.nterm) break;
                const cpx<T>* src = base + V.t[t].field * a.in_fs;
                // a column with no kept mode (OPT 1) is not read: the y pass did not store it
                auto load = [&](int p) -> cpx<T> { return live ? load_stream(src + p * zls) : cpx<T>{T(0), T(0)}; };
                __syncthreads();
                h = line_fft_regs<N, -1, T, typename C::P>(j, stw, load, lds, w);
                if (h) {
                    const int m = V.t[t].mult;
#pragma unroll
                    for (int r = 0; r < I::RLAST; ++r) {
                        if (m == 0) {
                            acc[r] = acc[r] + w[r];
                        } else {
                            const int p = j + r * I::BLAST;
                            const T kz = T(swz[p] * r0r);
                            const cpx<T> q = mul_i(kz, w[r]);
                            acc[r] = m > 0 ? acc[r] + q : acc[r] - q;
                        }
                    }
                }
            }
            return h;
        };
        bool has;
        if (Ds) {
            // phase-shift dealiasing: fnl = (Z(D) + conj(ph) Z(Ds)) / 2, ph of the output mode
            const cpx<T>* ibs = Ds + (long long)cy * a.rs + kx;
            const cpx<double> wcb = cmul(a.phc[kx], a.phb[cy]);
            has = terms(ibs);
            cpx<T> accs[I::RLAST];
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) {
                const cpx<double> q = cmul(a.ph[has ? j + r * I::BLAST : 0], wcb);
                accs[r] = cmul(acc[r], cpx<T>{T(q.x), T(-q.y)});
            }
            has = terms(ib);
            const T half = T(0.5);
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r)
                acc[r] = cpx<T>{half * (acc[r].x + accs[r].x), half * (acc[r].y + accs[r].y)};
        } else {
            has = terms(ib);
        }
        // stage fnl of this column in the LDS line, then update the kept modes: adjacent
        // work-items take adjacent kx (coalesced U, Fp). A work-item takes its modes CH at a
        // time and issues the state and previous-RHS loads of all CH before updating any, so
        // a group costs one memory latency instead of one per mode.
        __syncthreads();
        if (has) {
#pragma unroll
            for (int r = 0; r < I::RLAST; ++r) *lds(j + r * I::BLAST) = acc[r];
        }
        __syncthreads();
        if (!valid) return;
        constexpr int NIT = (N + C::TPL - 1) / C::TPL;
        constexpr int CH = NIT < 4 ? NIT : 4;
#pragma unroll 1
        for (int i0 = 0; i0 < NIT; i0 += CH) {
            long long mi[CH];
            int czq[CH];
            bool act[CH];
            cpx<double> uq[CH];
            cpx<T> fq[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const int p = j + (i0 + q) * C::TPL;
                const int cz = p < N ? pos_to_stored(p, N, a.CZ) : -1;
                czq[q] = cz;
                mi[q] = (long long)(cz < 0 ? 0 : cz) * a.CY * a.rs + mode0;
                bool k = cz >= 0;
                if constexpr (OPT == 1) {
                    // outside the kept set the new state is zero whatever U and Fp hold there:
                    // neither is read, and Fp there is never read again
                    const int mz = p <= N / 2 ? p : p - N;
                    if (k && (mz > lim || -mz > lim)) {
                        if (a.clear_outside) {
                            U[v * a.ufs + mi[q]] = cpx<double>{0.0, 0.0};
                            if (Uf) Uf[v * a.ufs + mi[q]] = cpx<float>{0.0f, 0.0f};
                        }
                        k = false;
                    }
                }
                act[q] = k;
                uq[q] = k ? load_stream(U + v * a.ufs + mi[q]) : cpx<double>{0.0, 0.0};
                fq[q] = (k && irk != 0) ? load_stream(Fp + v * a.ufs + mi[q]) : cpx<T>{T(0), T(0)};
            }
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                if (!act[q]) continue;
                const int p = j + (i0 + q) * C::TPL;
                const int cz = czq[q];
                const double wz = swz[p];
                const cpx<T> fv = *lds(p);
                cpx<double> fnl{double(fv.x) * double(scale), double(fv.y) * double(scale)};
                if (a.clos && v >= 1) fnl = fnl + a.clos[(v - 1) * a.ufs + mi[q]];
                const cpx<double> u = uq[q];
                if (pp.aeb && v < 7) {
                    const double fac = (v == 2 || v == 3) ? 3.0 : (v >= 5 ? 1.0 : 2.0);
                    fnl = fnl - cpx<double>{(u.x * fac) / tau, (u.y * fac) / tau};
                }
                const double k2 = (pp.visc || pp.resis) ? ksquare(wx, wy, wz, pp, ctl, r0r) : 0.0;
                if (pp.visc && pp.visc_exp && v >= 1 && v <= 3)
                    fnl = fnl - cpx<double>{u.x * pp.nu * k2, u.y * pp.nu * k2};
                if (pp.resis && pp.resis_exp && v >= 4 && v <= 6 && !(pp.conserve_bg && kx == 0 && p == 0))
                    fnl = fnl - cpx<double>{u.x * pp.eta * k2, u.y * pp.eta * k2};
                cpx<double> un;
                if (irk == 0) {
                    un = cpx<double>{c * fnl.x + 0.0, c * fnl.y + 0.0} + u;
                } else {
                    const cpx<double> fp{double(fq[q].x), double(fq[q].y)};
                    un = cpx<double>{c * fnl.x + d * fp.x, c * fnl.y + d * fp.y} + u;
                }
                // the RHS of the last substage is never read: the next step's first substage
                // weighs the previous RHS by dd(0) = 0 and reads none (rktmod.f90:40-42)
                if (irk != 2) store_stream(Fp + v * a.ufs + mi[q], cpx<T>{T(fnl.x), T(fnl.y)});
                if (pp.visc && !pp.visc_exp && v >= 1 && v <= 3) {
                    const double den = ts * k2 * pp.nu + 1.0;
                    const double inv = 1.0 / den;
                    un = {un.x * inv, un.y * inv};
                }
                if (pp.resis && !pp.resis_exp && v >= 4 && v <= 6) {
                    const double den = ts * k2 * pp.eta + 1.0;
                    const double inv = 1.0 / den;
                    un = {un.x * inv, un.y * inv};
                }
                if constexpr (OPT == 2) {
                    const double fx = tb.fx[kx], fy = tb.fy[cy], fz = tb.fz[cz];
                    un = {((un.x * fx) * fy) * fz, ((un.y * fx) * fy) * fz};
                }
                store_stream(U + v * a.ufs + mi[q], un);
                if (Uf) store_stream(Uf + v * a.ufs + mi[q], cpx<float>{float(un.x), float(un.y)});
            }
        }
    }
}

// ---------------------------------------------------------------- launchers

template <typename T, int N, bool ALT>
static hipError_t do_yg_c(const cpx<T>* in, cpx<T>* out, const YGArgs& a, const cpx<T>* tw, const double* wxt,
                          const double* wyf, const PhysParams& pp, const Ctl* ctl, int nz, hipStream_t s) {
    using C = FwdCfg<T, N, false, ALT>;
    const int ntile = (a.ncol + C::TX - 1) / C::TX;
    const long long nblk = (long long)ntile * nz * a.ngroups * (a.nch == 2 ? 2 : 1);
    hipLaunchKernelGGL((k_yfwd_group<T, N, ALT>), dim3((unsigned)nblk, a.nbox > 1 ? a.nbox : 1), dim3(C::NT), 0, s, in,
                       out, a, tw, wxt, wyf, pp, ctl, ntile, nz);
    return hipGetLastError();
}

template <typename T, int N>
static hipError_t do_yg(const cpx<T>* in, cpx<T>* out, const YGArgs& a, const cpx<T>* tw, const double* wxt,
                        const double* wyf, const PhysParams& pp, const Ctl* ctl, int nz, hipStream_t s) {
    if constexpr (N == 256) {
        // the wide plan: FP32 launches of one chain, FP64 launches of two
        if ((a.nch == 2) == (sizeof(T) == 8)) return do_yg_c<T, N, true>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
    }
    return do_yg_c<T, N, false>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
}

template <typename T>
hipError_t launch_yfwd_group(int ny, const cpx<T>* in, cpx<T>* out, const YGArgs& a, const cpx<T>* tw,
                             const double* wxt, const double* wyf, const PhysParams& pp, const Ctl* ctl, int nz,
                             hipStream_t s) {
    switch (ny) {
        case 16: return do_yg<T, 16>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
        case 32: return do_yg<T, 32>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
        case 64: return do_yg<T, 64>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
        case 128: return do_yg<T, 128>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
        case 256: return do_yg<T, 256>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
        case 512: return do_yg<T, 512>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
        case 96: return do_yg<T, 96>(in, out, a, tw, wxt, wyf, pp, ctl, nz, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename T, int N, int OPT>
static hipError_t do_zr(const cpx<T>* D, cpx<double>* U, cpx<T>* Fp, const ZArgs& a, const cpx<T>* tw,
                        const BoxTables& tb, const double* wzf, const PhysParams& pp, const Ctl* ctl, int irk,
                        T scale, hipStream_t s) {
    using C = FwdCfg<T, N, true>;
    const int ntile = (a.KX + C::TX - 1) / C::TX;
    const long long nblk = 8LL * ntile * a.CY;
    hipLaunchKernelGGL((k_zfwd_rhs<T, N, OPT>), dim3((unsigned)nblk, a.nbox > 1 ? a.nbox : 1), dim3(C::NT), 0, s, D, U,
                       Fp, a, tw, tb, wzf, pp, ctl, irk, ntile, scale);
    return hipGetLastError();
}

template <typename T, int N>
static hipError_t do_zr_opt(const cpx<T>* D, cpx<double>* U, cpx<T>* Fp, const ZArgs& a, const cpx<T>* tw,
                            const BoxTables& tb, const double* wzf, const PhysParams& pp, const Ctl* ctl, int irk,
                            T scale, hipStream_t s) {
    if (pp.dealias == 1 || pp.dealias == 3) return do_zr<T, N, 1>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
    if (pp.dealias == 2) return do_zr<T, N, 2>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
    return do_zr<T, N, 0>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
}

template <typename T>
hipError_t launch_zfwd_rhs(int nz, const cpx<T>* D, cpx<double>* U, cpx<T>* Fp, const ZArgs& a,
                           const cpx<T>* tw, const BoxTables& tb, const double* wzf, const PhysParams& pp,
                           const Ctl* ctl, int irk, T scale, hipStream_t s) {
    switch (nz) {
        case 16: return do_zr_opt<T, 16>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
        case 32: return do_zr_opt<T, 32>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
        case 64: return do_zr_opt<T, 64>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
        case 128: return do_zr_opt<T, 128>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
        case 256: return do_zr_opt<T, 256>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
        case 512: return do_zr_opt<T, 512>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
        case 96: return do_zr_opt<T, 96>(D, U, Fp, a, tw, tb, wzf, pp, ctl, irk, scale, s);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_yfwd_group<double>(int, const cpx<double>*, cpx<double>*, const YGArgs&,
                                              const cpx<double>*, const double*, const double*, const PhysParams&,
                                              const Ctl*, int, hipStream_t);
template hipError_t launch_yfwd_group<float>(int, const cpx<float>*, cpx<float>*, const YGArgs&, const cpx<float>*,
                                             const double*, const double*, const PhysParams&, const Ctl*, int,
                                             hipStream_t);
template hipError_t launch_zfwd_rhs<double>(int, const cpx<double>*, cpx<double>*, cpx<double>*, const ZArgs&,
                                            const cpx<double>*, const BoxTables&, const double*, const PhysParams&,
                                            const Ctl*, int, double, hipStream_t);
template hipError_t launch_zfwd_rhs<float>(int, const cpx<float>*, cpx<double>*, cpx<float>*, const ZArgs&,
                                           const cpx<float>*, const BoxTables&, const double*, const PhysParams&,
                                           const Ctl*, int, float, hipStream_t);

}  // namespace laps
