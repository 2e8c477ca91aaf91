// Persistent banded chain propagation on MFMA (gfx950).
//
// One workgroup owns one (state point, sweep) of a linear chain species and runs
// all N-1 propagator steps (child_propagators.jl:93-141 for sweep 0, the parent
// recursion for sweep 1) without leaving the kernel: the running segment sums Q
// stay in LDS, only the per-segment propagators gC / gP are written to HBM for
// the density accumulation.
//
// The bond operator of one step is A[n][n'] = c[n - n'] + c[n + n' + 1]
// (indices mod L = 2 NP) with c truncated to the half-width Be.  On the
// mirror-extended input
//     x_ext[m] = x[m] (0 <= m < NP),  x[-1 - m] (m < 0),  x[2 NP - 1 - m] (m >= NP)
// the Toeplitz and the Hankel part together are one plain convolution
//     y[n] = sum_{|j| <= Be} c[j] x_ext[n - j].
// Writing n = 16 p + r, a 256-output tile is the GEMM
//     D[r][p] = sum_t A[r][t] X[t][p],   A[r][t] = c[r - t + Be],  X[t][p] = x_ext[16 p + t - Be]
// with K = 2 Be + 16: M = 16 output phases, N = 16 position blocks.  A is a
// sliding window of c in LDS, X the extended input in LDS; both feed
// v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64 directly.
//
// Per step and column (a, b) (a: state of the output segment, b: state of the
// input segment):
//     x_ab = exp(-lambda_sj) exp(lng_bond / 2) Q_b trapez_sj            (input)
//     g[outseg][b][a] = |A x_ab| exp(lng_bond / 2) / normative          (output)
//     Q_a (next step) = sum_b g[outseg][b][a]
// The weights W_ab = exp(-lambda) exp(lng/2) trapez do not change along a chain
// whose segments all carry the same states; they are then kept in registers.
#include <cstring>

#include "sdft_launch.h"
#include "sdft_mfma.h"

namespace sdft {

constexpr int CHAIN_THREADS = 512;
// accumulation chains per MFMA unit: one (measured: four chains did not shorten the launch, so the
// registers go to occupancy instead)
constexpr int NACC = 1;
constexpr int CHAIN_WAVES = CHAIN_THREADS / 64;

struct ChainArgs {
    int32_t u, nseg, S, state_off;
    int32_t GF, MB, NB, homog;     // homog: every segment has the same state beads
    int32_t hv;                    // output factors exp(lng/2)/normative tabulated in LDS (homog only)
    int32_t shared_a;              // every (a, b) column of both sweeps applies the same bond operator
    int64_t goff;
    int32_t ck[2][SDFT_FAST_S * SDFT_FAST_S];  // canonical bond kernel of (sweep, a * S + b)
    int32_t npad_max, be_max;                // LDS carve-out (largest point of the launch)
    int32_t xl_need, pad3;                   // largest NP + 2 Be
    int32_t rescale, nsp, nmax, pad;         // per-step rescaling (scale_slot tables)
    unsigned long long* smax;
    const ChainCol* info;                    // [2][N-1][S*S] column table of every step (host-built)
};

__device__ __forceinline__ int chain_phys(int i) { return i + (i >> 4); }  // 1 pad word per 16: conflict-free B reads


constexpr int CHAIN_LDS_SLACK = 64;  // elements after the carve-out: the operand prefetch may read one group past the end

// Register use decides how many 8-wave workgroups share a CU: two need <= 128 VGPRs per lane (at
// least 4 waves per SIMD), which the FP32 variants are held to; the FP64 variants run one per CU
// (held to 128 they spill the input weights to scratch: faster at full load, slower in the polish
// phase where few points are active, so the solve is slower overall).
template <typename T, int EPC>
__global__ void __launch_bounds__(CHAIN_THREADS, sizeof(T) == 4 ? 4 : 1) k_chain(const DevSystem* __restrict__ psys,
                                                         const DevPoint* __restrict__ pts,
                                                         const int32_t* __restrict__ evalflag, int prec, ChainArgs ca,
                                                         const int32_t* __restrict__ seg_states,
                                                         const int32_t* __restrict__ band, const T* __restrict__ kern,
                                                         const T* __restrict__ elam, const T* __restrict__ hl,
                                                         const T* __restrict__ trapez, T* __restrict__ gC,
                                                         T* __restrict__ gP) {
    const int p = blockIdx.x, sw = blockIdx.y;
    if (evalflag[p] != prec) return;
    const DevSystem& sys = *psys;
    const DevPoint& P = pts[p];
    const int NP = P.NP, L = P.L;
    const int B = band[P.kset];
    const int Be = chain_half_width(B);
    const int KS = Be / 2 + 4;                 // k-steps: K / 4 with K = 2 Be + 16
    const int ntile = (NP + 255) >> 8;
    const int XL = ntile * 256 + 2 * Be;       // extended input length of this point
    const int S = ca.S, C = S * S, N = ca.nseg;
    const int KC = 2 * ca.be_max + 32;
    const int XLm = ca.npad_max + 2 * ca.be_max;
    const int XLP = XLm + (XLm >> 4) + 1;
    const int NPQ = ca.npad_max;

    extern __shared__ unsigned char chain_lds[];
    T* cw = reinterpret_cast<T*>(chain_lds);   // [C][KC]
    T* X = cw + C * KC;                        // [C][XLP]
    T* Q = X + C * XLP;                        // [S][NPQ]
    T* HV = Q + S * NPQ;                       // [C][NPQ] exp(lng/2) of the output end / normative (ca.hv)
    __shared__ ChainCol info[2][SDFT_FAST_S * SDFT_FAST_S];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double norm = P.normative[ca.u];

    // operator windows, stored reversed so that the A reads ascend with k:
    // cw[col][ri] = c[j], j = (KC - 1 - ri) - 15 - Be, for |j| <= Be, else 0
    {
        const T* cset = kern + (int64_t)P.kbase * sys.K;
        for (int e = tid; e < C * KC; e += CHAIN_THREADS) {
            const int col = e / KC, ri = e - col * KC;
            const int k = ca.ck[sw][col];
            const int j = (KC - 1 - ri) - 15 - Be;
            T v = T(0);
            if (k >= 0 && j >= -Be && j <= Be) {
                int jj = j % L;
                if (jj < 0) jj += L;
                v = cset[(int64_t)k * L + jj];
            }
            cw[e] = v;
        }
    }

    // input weights W_ab of the extended element mi of each column; with EPC > 0 each thread keeps
    // the weights of its elements (mi = tid + k * CHAIN_THREADS) in registers
    T wreg[SDFT_FAST_S * SDFT_FAST_S][EPC > 0 ? EPC : 1];
    const int64_t eb = P.base * (int64_t)ca.NB;
    const int64_t hb = P.base * (int64_t)(2 * ca.MB);
    auto ext_index = [&](int mi) -> int {  // point index of extended element mi, -1 beyond the mirror
        const int m = mi - Be;
        return m < 0 ? -1 - m : (m < NP ? m : 2 * NP - 1 - m);
    };

    // Column table of the steps (static topology, built on the host): threads tid < C copy the
    // current step's entry into LDS and already load the next one, so no global-memory latency sits
    // between two steps.
    const ChainCol* tab = ca.info + (int64_t)sw * (N - 1) * C;
    const int4* tab4 = reinterpret_cast<const int4*>(tab);  // one 16-byte load per entry, kept in registers
    static_assert(sizeof(ChainCol) == sizeof(int4), "ChainCol is four int32");
    int4 nxt = make_int4(0, 0, 0, 0);
    if (tid < C) nxt = tab4[tid];
    // homogeneous chain: every step has the same columns, so the output factors of the epilogue are
    // read from HBM once here (the first step's barrier orders these writes before their reads)
    if (ca.hv) {
        for (int e = tid; e < C * NP; e += CHAIN_THREADS) {
            const int col = e / NP, n = e - col * NP;
            const ChainCol cc = tab[col];
            HV[col * NPQ + n] = cc.valid ? hl[hb + (int64_t)cc.hl_out * NP + n] / (T)norm : T(0);
        }
    }
    for (int s = 1; s < N; ++s) {
        const int outseg = sw == 0 ? N - 1 - s : s;
        ChainCol* ci = info[s & 1];
        if (tid < C) {
            ci[tid] = ChainCol{nxt.x, nxt.y, nxt.z, nxt.w};
            if (s + 1 < N) nxt = tab4[s * C + tid];
        }
        __syncthreads();  // column info of this step; Q of the previous step complete
        // ---- extended inputs X = W * Q (Q = 1 at the free chain end)
        const bool first = s == 1;
        // per-step rescaling: divide the input by the largest Q of the previous step (recorded for
        // the segment log scales)
        T qscale = T(1);
        if (ca.rescale && !first) {
            __shared__ double redm[CHAIN_WAVES];
            double mx = 0.0;
            for (int a = 0; a < S; ++a)
                for (int n = tid; n < NP; n += CHAIN_THREADS) {
                    const double v = (double)Q[a * NPQ + n];
                    mx = v > mx ? v : mx;
                }
            for (int o = 32; o > 0; o >>= 1) {
                const double w = __shfl_xor(mx, o, 64);
                mx = w > mx ? w : mx;
            }
            if (lane == 0) redm[wave] = mx;
            __syncthreads();
            double m = 0.0;
            for (int w = 0; w < CHAIN_WAVES; ++w) m = redm[w] > m ? redm[w] : m;
            if (tid == 0) {
                const int64_t slot = scale_slot(p, ca.u, sw, s - 1, ca.nsp, ca.nmax);
                ca.smax[slot] = (unsigned long long)__double_as_longlong(m);
            }
            m = (m > 0.0 && m <= 1.7976931348623157e308) ? m : 1.0;  // scale_read's rule
            qscale = (T)(1.0 / m);
        }
#pragma unroll
        for (int col = 0; col < SDFT_FAST_S * SDFT_FAST_S; ++col) {
            if (col < C) {
                const ChainCol cc = ci[col];
                const T* el = elam + eb + (int64_t)cc.sj * NP;
                const T* tz = trapez + eb + (int64_t)cc.sj * NP;
                const T* hh = hl + hb + (int64_t)cc.hl_in * NP;
                const T* qb = Q + (col % S) * NPQ;
                T* xc = X + col * XLP;
                if (EPC > 0) {
                    if (first || !ca.homog) {
#pragma unroll
                        for (int k = 0; k < (EPC > 0 ? EPC : 1); ++k) {
                            const int mi = tid + k * CHAIN_THREADS;
                            const int n = ext_index(mi);
                            wreg[col][k] = (mi < XL && cc.valid && n >= 0) ? el[n] * hh[n] * tz[n] : T(0);
                        }
                    }
#pragma unroll
                    for (int k = 0; k < (EPC > 0 ? EPC : 1); ++k) {
                        const int mi = tid + k * CHAIN_THREADS;
                        if (mi < XL) {
                            const int n = ext_index(mi);
                            const T v = (first || n < 0) ? wreg[col][k] : wreg[col][k] * (qb[n] * qscale);
                            xc[chain_phys(mi)] = v;
                        }
                    }
                } else {
                    for (int mi = tid; mi < XL; mi += CHAIN_THREADS) {
                        const int n = ext_index(mi);
                        T v = (cc.valid && n >= 0) ? el[n] * hh[n] * tz[n] : T(0);
                        if (!first && n >= 0) v *= qb[n] * qscale;
                        xc[chain_phys(mi)] = v;
                    }
                }
            }
        }
        __syncthreads();
        // ---- MFMA tiles: unit = (a, tile), summing the S input states into Q_a
        T* gout = sw == 0 ? gC : gP;
        const int64_t gb = P.gbase + ca.goff * P.gstride;
        const int q = lane & 15, g = lane >> 4;
        for (int unit = wave; unit < S * ntile; unit += CHAIN_WAVES) {
            const int a = unit / ntile, tile = unit - a * ntile;
            const int p0 = tile * 16;
            // epilogue factors exp(lng/2) of the output end, fetched before the products
            T hv[SDFT_FAST_S][4];
            if (!ca.hv) {
#pragma unroll
                for (int b = 0; b < SDFT_FAST_S; ++b) {
                    const ChainCol cc = ci[(a * S + b) % C];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int n = 16 * (p0 + q) + Mfma<T>::row(lane, v);
                        hv[b][v] = (b < S && cc.valid && n < NP) ? hl[hb + (int64_t)cc.hl_out * NP + n] : T(0);
                    }
                }
            }
            T qs[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int b = 0; b < SDFT_FAST_S; ++b) {
                if (b >= S) break;
                const int col = a * S + b;
                const ChainCol cc = ci[col];
                if (!cc.valid) continue;
                // k-step ks = 4 K + k4 (t = 4 ks + g): A[q][t] = c[q - t + Be] = cw[col][KC - 16 - 2 Be - q + g + 4 ks],
                // B = X[col][phys(16 (p0 + q) + g + 4 ks)] = X[col][17 (p0 + q) + g + 17 K + 4 k4]
                const T* ap = cw + col * KC + (KC - 16 - 2 * Be) - q + g;
                const T* xp = X + col * XLP + 17 * (p0 + q) + g;
                // NACC independent accumulation chains (k-step k4 goes to chain k4 % NACC), summed in
                // a fixed order at the end
                typename Mfma<T>::acc_t acc4[NACC];
#pragma unroll
                for (int k4 = 0; k4 < NACC; ++k4)
                    for (int v = 0; v < 4; ++v) acc4[k4][v] = T(0);
                // ping-pong operand buffers: the next group's LDS reads are in flight while
                // the current group runs on the matrix core
                T a0[4], b0[4], a1[4], b1[4];
                auto fetch = [&](int K, T (&av)[4], T (&bv)[4]) {
#pragma unroll
                    for (int k4 = 0; k4 < 4; ++k4) {
                        av[k4] = ap[16 * K + 4 * k4];
                        bv[k4] = xp[17 * K + 4 * k4];
                    }
                };
                const int KG = KS / 4;  // full 4-step groups, then KS - 4 KG single k-steps
                fetch(0, a0, b0);
                int K = 0;
                for (; K + 1 < KG; K += 2) {
                    fetch(K + 1, a1, b1);
                    __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the products
#pragma unroll
                    for (int k4 = 0; k4 < 4; ++k4) acc4[k4 % NACC] = Mfma<T>::mma(a0[k4], b0[k4], acc4[k4 % NACC]);
                    __builtin_amdgcn_sched_barrier(0);
                    fetch(K + 2, a0, b0);  // at most one group past the last: within the LDS carve-out, unused
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int k4 = 0; k4 < 4; ++k4) acc4[k4 % NACC] = Mfma<T>::mma(a1[k4], b1[k4], acc4[k4 % NACC]);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (K < KG) {  // odd group count: the last group, fetched above
#pragma unroll
                    for (int k4 = 0; k4 < 4; ++k4) acc4[k4 % NACC] = Mfma<T>::mma(a0[k4], b0[k4], acc4[k4 % NACC]);
                }
                for (int ks = 4 * KG; ks < KS; ++ks)
                    acc4[0] = Mfma<T>::mma(ap[4 * ks], xp[17 * (ks >> 2) + 4 * (ks & 3)], acc4[0]);
                typename Mfma<T>::acc_t acc = acc4[0];
#pragma unroll
                for (int k4 = 1; k4 < NACC; ++k4)
                    for (int v = 0; v < 4; ++v) acc[v] += acc4[k4][v];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int n = 16 * (p0 + q) + Mfma<T>::row(lane, v);
                    if (n < NP) {
                        const T y = acc[v];
                        const T ay = y < T(0) ? -y : y;
                        const T val = ca.hv ? ay * HV[col * NPQ + n] : ay * hv[b][v] / (T)norm;
                        gout[gb + (((int64_t)outseg * S + b) * S + a) * P.gstride + n] = val;
                        qs[v] += val;
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int n = 16 * (p0 + q) + Mfma<T>::row(lane, v);
                if (n < NP) Q[a * NPQ + n] = qs[v];
            }
        }
        // the next step's first barrier orders these Q writes before their reads
    }
}

// --------------------------------------------------------------------------
// Homogeneous chains (every segment carries the same state beads, so the input weights
// W_ab = exp(-lambda) exp(lng/2) trapez and the output factors are the same at every step): the
// next step's extended inputs X_{s+1}[(a', a)] = W_(a', a) Q_a are written by the epilogue that
// produces Q_a, so no separate staging pass runs between the steps. The per-step rescaling is
// applied at the output instead of the input, g_{s+1} = |A (W Q_s)| hv / max Q_s (the operator is
// linear, so the propagators are the same up to rounding), and the maximum is reduced per wave in
// the epilogue.
// The MFMA row r of a tile holds output phase phi(r): for FP64 (D[4 v + l/16][l % 16]) phi permutes
// the rows so that the four accumulators of a lane are four consecutive outputs, as they are for
// FP32 (D[4 (l/16) + v][l % 16]); the epilogue then stores its propagators as 4-element vectors.
// X holds only the extended positions that valid outputs read (m < NP + 2 Be) and 16 zeros after
// them: a tile column reads 2 Be + 16 positions for all 16 rows, where the operator window is zero,
// so those positions must be finite (0 x NaN would reach a valid output). Reads past that feed only
// tile columns whose outputs all lie beyond NP, which the epilogue discards, and stay inside the
// carve-out.
// --------------------------------------------------------------------------
__host__ __device__ __forceinline__ int round4(int n) { return (n + 3) & ~3; }
__host__ __device__ __forceinline__ int chain_xlp(int xl_need) {
    return round4(xl_need + 16 + ((xl_need + 16) >> 4) + 1);
}

template <typename T> struct Vec4;
template <> struct Vec4<float> {
    static __device__ __forceinline__ void load(const float* p, float (&v)[4]) {
        const float4 x = *reinterpret_cast<const float4*>(p);
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    }
    static __device__ __forceinline__ void store(float* p, const float (&v)[4]) {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    }
};
template <> struct Vec4<double> {
    static __device__ __forceinline__ void load(const double* p, double (&v)[4]) {
        const double2 x = reinterpret_cast<const double2*>(p)[0], y = reinterpret_cast<const double2*>(p)[1];
        v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y;
    }
    static __device__ __forceinline__ void store(double* p, const double (&v)[4]) {
        reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
        reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
    }
};

// --------------------------------------------------------------------------
// Both sweeps of a point in one 16-wave workgroup, offset by half a step (ping-pong): waves 0-7 run
// the child sweep and waves 8-15 the parent sweep, and in every phase between two barriers one half
// runs the matrix products of its step while the other half finishes its previous step (output
// factors, propagator stores, next inputs, maximum). The matrix cores then always have a half with
// products to issue, and each half needs only one X buffer: a half writes its next inputs in the
// phase after its products have read the current ones. Phase ph: half h runs the products of step
// (ph - h) / 2 + 1 when ph - h is even and the finish of step (ph - h + 1) / 2 when it is odd.
// The two halves share a CU, so a point's chain takes 2 N - 1 phases instead of N - 1 steps on two
// CUs: at low load (at most half as many active points as CUs) each workgroup instead runs one
// sweep on waves 0-7 with a double-buffered X and one barrier per step, and the second workgroup
// of the point runs the other sweep. Every wave runs the same barrier sequence.
// One (a, tile) unit per wave, its weights and output factors in registers.
// --------------------------------------------------------------------------
size_t chain_pp_lds_bytes(const ChainPlanHost& ch, size_t elem) {
    const int S = ch.S, C = S * S;
    const int KC = 2 * ch.be_max + 32;
    const int full = ch.npad_max + 2 * ch.be_max + 16;
    const int over = full + (full >> 4) + 1 - chain_xlp(ch.xl_need);
    return (size_t)(2 * round4(C * KC) + 2 * C * chain_xlp(ch.xl_need) + (over > 0 ? over : 0) +
                    CHAIN_LDS_SLACK) * elem;
}

constexpr int PP_THREADS = 1024;

// Largest 64-bit pattern over the wave (every lane active): row prefix maxima by DPP row shifts,
// then the four row maxima read as scalars. Non-negative doubles order as their bit patterns.
template <int CTRL>
__device__ __forceinline__ void dpp_max_step(unsigned& lo, unsigned& hi) {
    const unsigned lo2 = __builtin_amdgcn_update_dpp(0u, lo, CTRL, 0xf, 0xf, false);
    const unsigned hi2 = __builtin_amdgcn_update_dpp(0u, hi, CTRL, 0xf, 0xf, false);
    const unsigned long long v = ((unsigned long long)hi2 << 32) | lo2;
    const unsigned long long w = ((unsigned long long)hi << 32) | lo;
    if (v > w) {
        lo = lo2;
        hi = hi2;
    }
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long u) {
    unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    dpp_max_step<0x111>(lo, hi);   // row_shr:1
    dpp_max_step<0x112>(lo, hi);   // row_shr:2
    dpp_max_step<0x114>(lo, hi);   // row_shr:4
    dpp_max_step<0x118>(lo, hi);   // row_shr:8: lane 15 of each row holds the row maximum
    unsigned long long m = 0ull;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const unsigned long long v =
            ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)hi, 16 * r + 15) << 32) |
            (unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)lo, 16 * r + 15);
        m = v > m ? v : m;
    }
    return m;
}

template <typename T, bool SHA>
__global__ void __launch_bounds__(PP_THREADS) k_chain_pp(
        const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts, const int32_t* __restrict__ evalflag,
        int prec, int npts, int pp_min, ChainArgs ca, const int32_t* __restrict__ band, const T* __restrict__ kern,
        const T* __restrict__ elam, const T* __restrict__ hl, const T* __restrict__ trapez, T* __restrict__ gC,
        T* __restrict__ gP) {
    const int p = blockIdx.x, wg = blockIdx.y;
    if (evalflag[p] != prec) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: scalar branches
    const int half = wave >> 3, hw = wave & 7;
    // regime: the number of points evaluated at this precision
    __shared__ int nact_s;
    if (tid == 0) nact_s = 0;
    __syncthreads();
    {
        int c = 0;
        for (int i = tid; i < npts; i += PP_THREADS) c += evalflag[i] == prec ? 1 : 0;
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0 && c) atomicAdd(&nact_s, c);
    }
    __syncthreads();
    const bool pp = nact_s >= pp_min;
    if (pp && wg == 1) return;                       // the first workgroup runs both sweeps
    const DevSystem& sys = *psys;
    const DevPoint& P = pts[p];
    const int NP = P.NP, L = P.L;
    const int Be = chain_half_width(band[P.kset]);
    const int KS = Be / 2 + 4;
    const int ntile = (NP + 255) >> 8;
    const int XN = NP + 2 * Be;
    const int S = ca.S, C = S * S, N = ca.nseg;
    const int KC = 2 * ca.be_max + 32;
    const int XLR = chain_xlp(ca.xl_need);
    const int nslot = pp ? 2 : 1;                    // sweeps in this workgroup
    const bool active = half < nslot;                // this wave's half has a sweep
    const int sw = pp ? half : wg;                   // its sweep

    extern __shared__ __attribute__((aligned(16))) unsigned char chain_pp_lds[];
    T* cwb = reinterpret_cast<T*>(chain_pp_lds);    // [2][C][KC] operator windows of each half
    T* XB = cwb + 2 * round4(C * KC);               // [2][C][XLR]: one X per half (pp), or the two
                                                    // parities of the single sweep
    __shared__ ChainCol cinfo[2][SDFT_FAST_S * SDFT_FAST_S];
    __shared__ unsigned long long qmax[2][3];       // largest Q of step s in [half][s mod 3]

    const double norm = P.normative[ca.u];
    {
        const T* cset = kern + (int64_t)P.kbase * sys.K;
        for (int e = tid; e < nslot * C * KC; e += PP_THREADS) {
            const int j = e / (C * KC), e2 = e - j * C * KC;
            const int col = e2 / KC, ri = e2 - col * KC;
            const int k = ca.ck[pp ? j : wg][col];
            const int jj0 = (KC - 1 - ri) - 15 - Be;
            T v = T(0);
            if (k >= 0 && jj0 >= -Be && jj0 <= Be) {
                int jj = jj0 % L;
                if (jj < 0) jj += L;
                v = cset[(int64_t)k * L + jj];
            }
            cwb[j * round4(C * KC) + e2] = v;
        }
    }
    if (tid < nslot * C) {
        const int j = tid / C, col = tid - j * C;
        cinfo[j][col] = ca.info[(int64_t)(pp ? j : wg) * (N - 1) * C + col];
    }
    if (tid < 6) qmax[tid / 3][tid % 3] = 0ull;
    __syncthreads();
    const int64_t eb = P.base * (int64_t)ca.NB;
    const int64_t hb = P.base * (int64_t)(2 * ca.MB);
    auto weight = [&](const ChainCol& cc, int n) -> T {
        return cc.valid ? elam[eb + (int64_t)cc.sj * NP + n] * hl[hb + (int64_t)cc.hl_in * NP + n] *
                              trapez[eb + (int64_t)cc.sj * NP + n]
                        : T(0);
    };
    const int slot = pp ? half : 0;
    const ChainCol* ci = cinfo[slot];
    int validmask = 0;
    for (int col = 0; col < C; ++col) validmask |= ci[col].valid ? 1 << col : 0;
    validmask = __builtin_amdgcn_readfirstlane(validmask);
    const T* cw = cwb + slot * round4(C * KC);
    unsigned long long* qm = qmax[slot];
    T* gout = sw == 0 ? gC : gP;
    const int64_t gstride = P.gstride;
    const int64_t gb = P.gbase + ca.goff * gstride;
    const int q = lane & 15, g = lane >> 4;
    const int phq = sizeof(T) == 8 ? 4 * (q & 3) + (q >> 2) : q;
    const bool has_unit = active && hw < S * ntile;
    const int a = has_unit ? hw / ntile : 0, tile = has_unit ? hw - a * ntile : 0;
    const int p0 = tile * 16;
    const int n0 = 16 * (p0 + q) + 4 * g;
    // Each lane's output factors and next-input weights are four consecutive positions of a row;
    // read per lane from global memory, a wave's reads would touch every cache line of its 256
    // positions four times. The rows are instead read once, coalesced, into the X buffers: first
    // the output factor rows e^{lng_i/2} of every column (unpadded, [j][col][n]), then, after the
    // lanes have taken theirs, the step-1 inputs, which are the weights themselves.
    for (int e = tid; e < nslot * C * NP; e += PP_THREADS) {
        const int j = e / (C * NP), e2 = e - j * C * NP;
        const int col = e2 / NP, n = e2 - col * NP;
        XB[(j * C + col) * XLR + n] = hl[hb + (int64_t)cinfo[j][col].hl_out * NP + n];
    }
    __syncthreads();
    T hr[SDFT_FAST_S][4], wr[SDFT_FAST_S][4];
#pragma unroll
    for (int b = 0; b < SDFT_FAST_S; ++b) {
        const int col = (a * S + b) % C;
        const bool vo = ci[col].valid;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const bool in = has_unit && b < S && n0 + v < NP;
            const T h = XB[(slot * C + col) * XLR + (n0 + v < NP ? n0 + v : NP - 1)];
            hr[b][v] = in && vo ? h / (T)norm : T(0);
        }
    }
    __syncthreads();
    // step-1 inputs: X of half j (pp) or the parity-1 buffer (one sweep), and the 16 zeros past XN in
    // both buffers
    for (int e = tid; e < 2 * C * (XN + 16); e += PP_THREADS) {
        const int j = e / (C * (XN + 16)), e2 = e - j * C * (XN + 16);
        const int col = e2 / (XN + 16), mi = e2 - col * (XN + 16);
        T* xb = XB + (j * C + col) * XLR;
        if (mi >= XN) {
            xb[chain_phys(mi)] = T(0);
        } else if (pp || j == 1) {
            const int m = mi - Be;
            const int n = m < 0 ? -1 - m : (m < NP ? m : 2 * NP - 1 - m);
            xb[chain_phys(mi)] = weight(cinfo[pp ? j : 0][col], n);
        }
    }
    __syncthreads();
    {
        // the weight of column (b, a) at position n is the step-1 input at extended position n + Be
        const T* x1 = XB + (pp ? half : 1) * C * XLR;
#pragma unroll
        for (int b = 0; b < SDFT_FAST_S; ++b) {
            const int col = (b * S + a) % C;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const bool in = has_unit && b < S && n0 + v < NP;
                const T w = x1[col * XLR + chain_phys((n0 + v < NP ? n0 + v : NP - 1) + Be)];
                wr[b][v] = in ? w : T(0);
            }
        }
    }
    typename Mfma<T>::acc_t acc[SDFT_FAST_S];
    const int nph = pp ? 2 * N - 1 : N - 1;
    for (int ph = 0; ph < nph; ++ph) {
        __syncthreads();
        int sm = 0, se = 0;   // step whose products / finish this half runs in this phase (0: none)
        if (active) {
            if (pp) {
                const int r = ph - half;
                if (r >= 0) {
                    if ((r & 1) == 0) sm = r / 2 + 1 <= N - 1 ? r / 2 + 1 : 0;
                    else se = (r + 1) / 2 <= N - 1 ? (r + 1) / 2 : 0;
                }
            } else {
                sm = se = ph + 1;
            }
        }
        if (sm) {
            if (hw == 0 && lane == 0) qm[(sm + 1) % 3] = 0ull;
            const T* Xc = XB + (pp ? half : (sm & 1)) * C * XLR;
            if (SHA && has_unit) {
                // one operator window for both input states b: two independent accumulation chains
                // share its reads (k-steps taken two at a time, the next pair prefetched)
#pragma unroll
                for (int b = 0; b < SDFT_FAST_S; ++b)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[b][v] = T(0);
                const T* ap = cw + a * S * KC + (KC - 16 - 2 * Be) - phq + g;
                const T* xp0 = Xc + (a * S) * XLR + 17 * (p0 + q) + g;
                const T* xp1 = xp0 + XLR;
                const bool v0 = (validmask >> (a * S)) & 1, v1 = (validmask >> (a * S + 1)) & 1;
                T a0[2], b0[2], c0[2], a1[2], b1[2], c1[2];
                // k-step ks = 2 H + k: A at ap[4 ks], B at xp[17 (ks / 4) + 4 (ks % 4)]
                auto fetch2 = [&](int H, T (&av)[2], T (&bv)[2], T (&cv)[2]) {
                    const int xo = 17 * (H >> 1) + 8 * (H & 1);
#pragma unroll
                    for (int kk = 0; kk < 2; ++kk) {
                        av[kk] = ap[8 * H + 4 * kk];
                        bv[kk] = xp0[xo + 4 * kk];
                        cv[kk] = xp1[xo + 4 * kk];
                    }
                };
                auto mma2 = [&](const T (&av)[2], const T (&bv)[2], const T (&cv)[2]) {
#pragma unroll
                    for (int kk = 0; kk < 2; ++kk) {
                        acc[0] = Mfma<T>::mma(av[kk], bv[kk], acc[0]);
                        acc[1] = Mfma<T>::mma(av[kk], cv[kk], acc[1]);
                    }
                };
                const int NH = KS >> 1;
                fetch2(0, a0, b0, c0);
                int H = 0;
                for (; H + 1 < NH; H += 2) {
                    fetch2(H + 1, a1, b1, c1);
                    __builtin_amdgcn_sched_barrier(0);
                    mma2(a0, b0, c0);
                    __builtin_amdgcn_sched_barrier(0);
                    fetch2(H + 2, a0, b0, c0);  // at most one pair past the last: within the LDS carve-out, unused
                    __builtin_amdgcn_sched_barrier(0);
                    mma2(a1, b1, c1);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (H < NH) mma2(a0, b0, c0);
                if (KS & 1) {
                    const int ks = KS - 1;
                    const T av = ap[4 * ks];
                    const int xo = 17 * (ks >> 2) + 4 * (ks & 3);
                    acc[0] = Mfma<T>::mma(av, xp0[xo], acc[0]);
                    acc[1] = Mfma<T>::mma(av, xp1[xo], acc[1]);
                }
                // a column without a bond contributes nothing (its window and inputs are zero; keep
                // the finish's rule of skipping it)
                if (!v0)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[0][v] = T(0);
                if (!v1)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[1][v] = T(0);
            } else if (!SHA && has_unit) {
#pragma unroll
                for (int b = 0; b < SDFT_FAST_S; ++b) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[b][v] = T(0);
                    if (b >= S) break;
                    const int col = a * S + b;
                    if (!((validmask >> col) & 1)) continue;
                    const T* ap = cw + col * KC + (KC - 16 - 2 * Be) - phq + g;
                    const T* xp = Xc + col * XLR + 17 * (p0 + q) + g;
                    T a0[4], b0[4], a1[4], b1[4];
                    auto fetch = [&](int K, T (&av)[4], T (&bv)[4]) {
#pragma unroll
                        for (int k4 = 0; k4 < 4; ++k4) {
                            av[k4] = ap[16 * K + 4 * k4];
                            bv[k4] = xp[17 * K + 4 * k4];
                        }
                    };
                    const int KG = KS / 4;
                    fetch(0, a0, b0);
                    int K = 0;
                    for (; K + 1 < KG; K += 2) {
                        fetch(K + 1, a1, b1);
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int k4 = 0; k4 < 4; ++k4) acc[b] = Mfma<T>::mma(a0[k4], b0[k4], acc[b]);
                        __builtin_amdgcn_sched_barrier(0);
                        fetch(K + 2, a0, b0);  // at most one group past the last: within the LDS carve-out, unused
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int k4 = 0; k4 < 4; ++k4) acc[b] = Mfma<T>::mma(a1[k4], b1[k4], acc[b]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    if (K < KG) {
#pragma unroll
                        for (int k4 = 0; k4 < 4; ++k4) acc[b] = Mfma<T>::mma(a0[k4], b0[k4], acc[b]);
                    }
                    for (int ks = 4 * KG; ks < KS; ++ks)
                        acc[b] = Mfma<T>::mma(ap[4 * ks], xp[17 * (ks >> 2) + 4 * (ks & 3)], acc[b]);
                }
            }
        }
        if (se) {
            const int s = se;
            __builtin_amdgcn_s_setprio(1);   // the finish is latency-bound: ahead of the other half's products
            T scale = T(1);
            if (ca.rescale && s > 1) {
                double m = __longlong_as_double((long long)qm[(s - 1) % 3]);
                if (hw == 0 && lane == 0) {
                    const int64_t sl = scale_slot(p, ca.u, sw, s - 1, ca.nsp, ca.nmax);
                    ca.smax[sl] = (unsigned long long)__double_as_longlong(m);
                }
                m = (m > 0.0 && m <= 1.7976931348623157e308) ? m : 1.0;  // scale_read's rule
                scale = (T)(1.0 / m);
            }
            double mx = 0.0;   // largest Q of this lane's outputs (step s < N - 1)
            if (has_unit && n0 < NP) {
                const int outseg = sw == 0 ? N - 1 - s : s;
                T qs[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int b = 0; b < SDFT_FAST_S; ++b) {
                    if (b >= S) break;
                    const int col = a * S + b;
                    if (!((validmask >> col) & 1)) continue;
                    T val[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const T y = acc[b][v];
                        const T ay = y < T(0) ? -y : y;
                        val[v] = n0 + v < NP ? ay * hr[b][v] * scale : T(0);
                        qs[v] += val[v];
                    }
                    // the padding of the propagator rows past NP receives zeros
                    Vec4<T>::store(gout + gb + (((int64_t)outseg * S + b) * S + a) * gstride + n0, val);
                }
                if (s < N - 1) {
                    T* Xn = XB + (pp ? half : ((s + 1) & 1)) * C * XLR;
#pragma unroll
                    for (int a2 = 0; a2 < SDFT_FAST_S; ++a2) {
                        if (a2 >= S) break;
                        T* xn = Xn + (a2 * S + a) * XLR;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int n = n0 + v;
                            if (n < NP) {
                                const T x = wr[a2][v] * qs[v];
                                xn[chain_phys(n + Be)] = x;
                                if (n < Be) xn[chain_phys(Be - 1 - n)] = x;
                                const int mt = 2 * NP - 1 - n + Be;
                                if (mt < XN) xn[chain_phys(mt)] = x;
                            }
                        }
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const double qv = (double)qs[v];
                        mx = qv > mx ? qv : mx;   // qs is zero past NP
                    }
                }
            }
            if (has_unit && ca.rescale && s < N - 1) {   // wave-uniform: every lane takes part
                const unsigned long long wm = wave_max_u64((unsigned long long)__double_as_longlong(mx));
                if (lane == 0) atomicMax(&qm[s % 3], wm);
            }
            __builtin_amdgcn_s_setprio(0);
        }
        // the next phase's barrier orders these writes before their reads
    }
}

size_t chain_lds_bytes(const ChainPlanHost& ch, size_t elem, bool hv_table) {
    const int S = ch.S, C = S * S;
    const int KC = 2 * ch.be_max + 32;
    const int XLm = ch.npad_max + 2 * ch.be_max;
    const int XLP = XLm + (XLm >> 4) + 1;
    const int HV = hv_table ? C * ch.npad_max : 0;
    return (size_t)(C * KC + C * XLP + S * ch.npad_max + HV + CHAIN_LDS_SLACK) * elem;
}

template <typename T, int EPC>
static hipError_t chain_set_lds(size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_chain<T, EPC>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t chain_prepare(const ChainPlanHost& ch) {
    hipError_t e = hipSuccess;
    auto pp_attr = [&](const void* f, size_t bytes) {
        return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    };
    if (ch.pp32 && (e = ch.shared_a ? pp_attr(reinterpret_cast<const void*>(&k_chain_pp<float, true>),
                                              chain_pp_lds_bytes(ch, sizeof(float)))
                                    : pp_attr(reinterpret_cast<const void*>(&k_chain_pp<float, false>),
                                              chain_pp_lds_bytes(ch, sizeof(float)))) != hipSuccess)
        return e;
    if (ch.pp64 && (e = ch.shared_a ? pp_attr(reinterpret_cast<const void*>(&k_chain_pp<double, true>),
                                              chain_pp_lds_bytes(ch, sizeof(double)))
                                    : pp_attr(reinterpret_cast<const void*>(&k_chain_pp<double, false>),
                                              chain_pp_lds_bytes(ch, sizeof(double)))) != hipSuccess)
        return e;
    if (ch.ok32) {
        const size_t b = chain_lds_bytes(ch, sizeof(float), ch.hv32);
        if ((e = chain_set_lds<float, 0>(b)) != hipSuccess) return e;
        if ((e = chain_set_lds<float, 2>(b)) != hipSuccess) return e;
        if ((e = chain_set_lds<float, 3>(b)) != hipSuccess) return e;
        if ((e = chain_set_lds<float, 4>(b)) != hipSuccess) return e;
    }
    if (ch.ok64) {
        const size_t b = chain_lds_bytes(ch, sizeof(double), ch.hv64);
        if ((e = chain_set_lds<double, 0>(b)) != hipSuccess) return e;
        if ((e = chain_set_lds<double, 2>(b)) != hipSuccess) return e;
        if ((e = chain_set_lds<double, 3>(b)) != hipSuccess) return e;
        e = chain_set_lds<double, 4>(b);
    }
    return e;
}

template <typename T>
void launch_chain(hipStream_t st, const Ctx& c, int prec, int u, const T* kern, const T* elam, const T* hl,
                  const T* trapez, T* gC, T* gP, unsigned long long* smax) {
    const SpeciesHost& sh = c.species[u];
    const ChainPlanHost& ch = c.chain_plan[u];
    ChainArgs ca;
    memset(&ca, 0, sizeof(ca));
    ca.u = u; ca.nseg = sh.nseg; ca.S = sh.S; ca.state_off = sh.state_off;
    ca.GF = c.GF; ca.MB = c.MB; ca.NB = c.NB; ca.homog = ch.homog ? 1 : 0;
    ca.goff = sh.goff;
    for (int sw = 0; sw < 2; ++sw)
        for (int k = 0; k < SDFT_FAST_S * SDFT_FAST_S; ++k) ca.ck[sw][k] = ch.ck[sw][k];
    ca.npad_max = ch.npad_max; ca.be_max = ch.be_max; ca.xl_need = ch.xl_need;
    ca.shared_a = ch.shared_a ? 1 : 0;
    ca.rescale = sh.rescale; ca.nsp = c.nsp; ca.nmax = c.nmax_seg; ca.smax = smax;
    ca.info = ch.d_info;
    if (sizeof(T) == 4 ? ch.pp32 : ch.pp64) {
        const size_t lds = chain_pp_lds_bytes(ch, sizeof(T));
        if (ch.shared_a)
            hipLaunchKernelGGL((k_chain_pp<T, true>), dim3(c.P, 2), dim3(PP_THREADS), lds, st, c.d_sys, c.d_pts,
                               c.d_evalflag, prec, c.P, ch.pp_min, ca, c.d_band, kern, elam, hl, trapez, gC, gP);
        else
            hipLaunchKernelGGL((k_chain_pp<T, false>), dim3(c.P, 2), dim3(PP_THREADS), lds, st, c.d_sys, c.d_pts,
                               c.d_evalflag, prec, c.P, ch.pp_min, ca, c.d_band, kern, elam, hl, trapez, gC, gP);
        return;
    }
    const bool hvt = sizeof(T) == 4 ? ch.hv32 : ch.hv64;
    ca.hv = hvt ? 1 : 0;
    const size_t lds = chain_lds_bytes(ch, sizeof(T), hvt);
    const int xl = ch.npad_max + 2 * ch.be_max;
    const int epc = (xl + CHAIN_THREADS - 1) / CHAIN_THREADS;
    auto go = [&](auto kfn) {
        hipLaunchKernelGGL(kfn, dim3(c.P, 2), dim3(CHAIN_THREADS), lds, st, c.d_sys, c.d_pts, c.d_evalflag, prec, ca,
                           c.d_seg_states, c.d_band, kern, elam, hl, trapez, gC, gP);
    };
    const bool reread = sizeof(T) == 4 ? ch.reread32 : ch.reread64;
    if (reread || epc > 4) go(k_chain<T, 0>);
    else if (epc <= 2) go(k_chain<T, 2>);
    else if (epc <= 3) go(k_chain<T, 3>);
    else go(k_chain<T, 4>);
}

template void launch_chain<float>(hipStream_t, const Ctx&, int, int, const float*, const float*, const float*,
                                  const float*, float*, float*, unsigned long long*);
template void launch_chain<double>(hipStream_t, const Ctx&, int, int, const double*, const double*, const double*,
                                   const double*, double*, double*, unsigned long long*);


// --------------------------------------------------------------------------
// One banded propagator step for long grids (k_bandstep): the chain kernel's
// tile GEMM D[r][p] = sum_t c[r - t + Be] x_ext[16 p + t - Be] on the
// mirror-extended input, one launch per step, many workgroups per problem.
// A workgroup owns (point, sweep, output state a) and 1024 consecutive outputs
// (4 waves x 256); it stages the operator windows of the S input states b and
// their extended inputs x_ab = exp(-lambda) exp(lng/2) Q_b trapez over the
// 1024 + 2 Be elements the outputs need, then every wave runs its 256-output
// tile per b and writes g[outseg][b][a] = |A x_ab| exp(lng/2) / normative
// (child_propagators.jl:93-141 / parent_propagators.jl, per step), with the
// grafted chains' max-distance cut and the per-step scale maximum as in the
// per-step operator GEMM. Steps whose parent is the grafted segment are not
// convolutions (direct sums) and stay with k_prop_step.
// --------------------------------------------------------------------------
constexpr int BAND_OUT = 1024;  // outputs per workgroup (4 waves x 16 positions x 16 phases)

struct BandStepArgs {
    int32_t u, s, nseg, S, fixed_seg, state_off, maxd_off, GF, MB, NB;
    int64_t goff;
    int32_t skip_sweep, rescale, nsp, nmax;
    int32_t be_max, pad;
    unsigned long long* smax;
};

size_t bandstep_lds_bytes(int S, int be, size_t elem) {
    const int KC = 2 * be + 32, XL = BAND_OUT + 2 * be, XLP = XL + (XL >> 4) + 1;
    return (size_t)(S * KC + S * XLP + CHAIN_LDS_SLACK) * elem;
}

template <typename T>
__global__ void __launch_bounds__(256) k_bandstep(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                  const int32_t* __restrict__ evalflag, int prec, BandStepArgs A,
                                                  const int32_t* __restrict__ seg_states,
                                                  const double* __restrict__ maxd_tab,
                                                  const int32_t* __restrict__ band, const T* __restrict__ kern,
                                                  const T* __restrict__ elam, const T* __restrict__ hl,
                                                  const T* __restrict__ trapez, T* __restrict__ gC,
                                                  T* __restrict__ gP) {
    const int S = A.S, N = A.nseg;
    const int a = blockIdx.y % S;
    const int sw = (blockIdx.y / S) & 1;
    const int p = blockIdx.y / (2 * S);
    if (evalflag[p] != prec || sw == A.skip_sweep) return;
    const DevSystem& sys = *psys;
    const DevPoint& P = pts[p];
    const int NP = P.NP, L = P.L;
    const int n0 = blockIdx.x * BAND_OUT;
    if (n0 >= NP) return;
    const int outseg = sw == 0 ? N - 1 - A.s : A.s;
    const int inseg = sw == 0 ? outseg + 1 : outseg - 1;
    const int* st = seg_states + A.state_off;
    const int si = st[outseg * S + a];
    if (si < 0) return;
    const int Be = chain_half_width(band[P.kset]);
    const int KC = 2 * Be + 32, XL = BAND_OUT + 2 * Be, XLP = XL + (XL >> 4) + 1;
    const int KS = Be / 2 + 4;  // k-steps: K / 4 with K = 2 Be + 16
    extern __shared__ unsigned char band_lds[];
    T* cw = reinterpret_cast<T*>(band_lds);  // [S][KC] operator windows, reversed
    T* X = cw + S * KC;                      // [S][XLP] extended inputs (one pad word per 16)
    __shared__ int32_t colv[SDFT_FAST_S], colsj[SDFT_FAST_S], colbond[SDFT_FAST_S], colpi[SDFT_FAST_S], colpj[SDFT_FAST_S];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < S) {
        const int sj = st[inseg * S + tid];
        int bond = 0;
        if (sj >= 0) {
            const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
            for (int u2 = 0; u2 < A.MB; ++u2)
                if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
        }
        colv[tid] = sj >= 0 ? 1 : 0;
        colsj[tid] = sj >= 0 ? sj : 0;
        colbond[tid] = bond;
        colpi[tid] = si <= sj ? 0 : 1;  // output end of the bond
        colpj[tid] = si <= sj ? 1 : 0;  // input end
    }
    __syncthreads();
    // operator windows: cw[b][ri] = c[j], j = (KC - 1 - ri) - 15 - Be for |j| <= Be, else 0
    const T* cset = kern + (int64_t)P.kbase * sys.K;
    for (int e = tid; e < S * KC; e += 256) {
        const int b = e / KC, ri = e - b * KC;
        const int j = (KC - 1 - ri) - 15 - Be;
        T v = T(0);
        if (colv[b] && j >= -Be && j <= Be) {
            int jj = j % L;
            if (jj < 0) jj += L;
            v = cset[(int64_t)(3 * A.NB + colbond[b]) * L + jj];
        }
        cw[e] = v;
    }
    // extended inputs: element mi is grid point reflect(n0 - Be + mi)
    const double* maxd = maxd_tab + A.maxd_off;
    const int Kf = A.fixed_seg >= 0 ? P.fixed_coord[A.u] : -1;
    const bool stored = sw == 0 ? (inseg < N - 1) : (inseg > 0);
    const T qinv = (A.rescale && stored)
                       ? (T)(1.0 / scale_read(A.smax, scale_slot(p, A.u, sw, A.s - 1, A.nsp, A.nmax)))
                       : T(1);
    const T* gin = sw == 0 ? gC : gP;
    const int64_t gb = P.gbase + A.goff * P.gstride;
    const int64_t eb = P.base * (int64_t)A.NB, hb = P.base * (int64_t)(2 * A.MB);
    for (int e = tid; e < S * XL; e += 256) {
        const int b = e / XL, mi = e - b * XL;
        const int m = n0 - Be + mi;
        const int n = m < 0 ? -1 - m : (m < NP ? m : 2 * NP - 1 - m);
        T v = T(0);
        if (colv[b] && n >= 0 && n < NP) {
            T q;
            if (!stored) {
                q = base_value<T>(inseg, n, A.fixed_seg, Kf, P, maxd);
            } else {
                q = T(0);
                for (int b2 = 0; b2 < S; ++b2)
                    if (st[inseg * S + b2] >= 0) q += gin[gb + ((int64_t)(inseg * S + b2) * S + b) * P.gstride + n];
            }
            const int sj = colsj[b];
            v = elam[eb + (int64_t)sj * NP + n] * hl[hb + (int64_t)(2 * colbond[b] + colpj[b]) * NP + n] *
                (q * qinv) * trapez[eb + (int64_t)sj * NP + n];
        }
        X[b * XLP + chain_phys(mi)] = v;
    }
    __syncthreads();
    // ---- per wave: the 256 outputs n0 + 256 wave + 16 q' + r, for every input state b
    const int q = lane & 15, g = lane >> 4;
    const int p0 = 16 * wave;
    T* gout = sw == 0 ? gC : gP;
    const double norm = P.normative[A.u];
    double vmax = 0.0;
    if (n0 + 256 * wave < NP) {
        for (int b = 0; b < S; ++b) {
            if (!colv[b]) continue;
            T hv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int n = n0 + 16 * (p0 + q) + Mfma<T>::row(lane, v);
                hv[v] = n < NP ? hl[hb + (int64_t)(2 * colbond[b] + colpi[b]) * NP + n] : T(0);
            }
            const T* ap = cw + b * KC + (KC - 16 - 2 * Be) - q + g;
            const T* xp = X + b * XLP + 17 * (p0 + q) + g;
            typename Mfma<T>::acc_t acc;
            for (int v = 0; v < 4; ++v) acc[v] = T(0);
            T a0[4], b0[4], a1[4], b1[4];
            auto fetch = [&](int K, T (&av)[4], T (&bv)[4]) {
#pragma unroll
                for (int k4 = 0; k4 < 4; ++k4) {
                    av[k4] = ap[16 * K + 4 * k4];
                    bv[k4] = xp[17 * K + 4 * k4];
                }
            };
            const int KG = KS / 4;  // full 4-step groups, then KS - 4 KG single k-steps
            fetch(0, a0, b0);
            int K = 0;
            for (; K + 1 < KG; K += 2) {
                fetch(K + 1, a1, b1);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int k4 = 0; k4 < 4; ++k4) acc = Mfma<T>::mma(a0[k4], b0[k4], acc);
                __builtin_amdgcn_sched_barrier(0);
                fetch(K + 2, a0, b0);  // at most one group past the last: within the carve-out and slack, unused
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int k4 = 0; k4 < 4; ++k4) acc = Mfma<T>::mma(a1[k4], b1[k4], acc);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (K < KG) {  // odd group count: the last group, fetched above
#pragma unroll
                for (int k4 = 0; k4 < 4; ++k4) acc = Mfma<T>::mma(a0[k4], b0[k4], acc);
            }
            for (int ks = 4 * KG; ks < KS; ++ks) acc = Mfma<T>::mma(ap[4 * ks], xp[17 * (ks >> 2) + 4 * (ks & 3)], acc);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int n = n0 + 16 * (p0 + q) + Mfma<T>::row(lane, v);
                if (n >= NP) continue;
                const T y = acc[v];
                T val = (y < T(0) ? -y : y) * hv[v] / (T)norm;
                if (A.fixed_seg >= 0) {
                    const double d = (double)(n - Kf) * P.dz;
                    if (d * d > maxd[outseg] * maxd[outseg]) val = T(0);
                }
                gout[gb + ((int64_t)(outseg * S + b) * S + a) * P.gstride + n] = val;
                vmax = (double)val > vmax ? (double)val : vmax;
            }
        }
    }
    if (A.rescale) {
        for (int o = 32; o > 0; o >>= 1) {
            const double w = __shfl_xor(vmax, o, 64);
            vmax = w > vmax ? w : vmax;
        }
        if (lane == 0) scale_note(A.smax, scale_slot(p, A.u, sw, A.s, A.nsp, A.nmax), vmax);
    }
}

template <typename T>
void launch_bandstep(hipStream_t st, const Ctx& c, int prec, int u, int s, int skip_sweep, const T* kern,
                     const T* elam, const T* hl, const T* trapez, T* gC, T* gP, unsigned long long* smax) {
    const SpeciesHost& sh = c.species[u];
    BandStepArgs A;
    memset(&A, 0, sizeof(A));
    A.u = u; A.s = s; A.nseg = sh.nseg; A.S = sh.S; A.fixed_seg = sh.fixed_seg;
    A.state_off = sh.state_off; A.maxd_off = sh.maxd_off; A.GF = c.GF; A.MB = c.MB; A.NB = c.NB;
    A.goff = sh.goff; A.skip_sweep = skip_sweep; A.rescale = sh.rescale; A.nsp = c.nsp; A.nmax = c.nmax_seg;
    A.be_max = sh.band_be; A.smax = smax;
    const size_t lds = bandstep_lds_bytes(sh.S, sh.band_be, sizeof(T));
    hipLaunchKernelGGL(k_bandstep<T>, dim3((c.maxNP + BAND_OUT - 1) / BAND_OUT, c.P * 2 * sh.S), dim3(256), lds, st,
                       c.d_sys, c.d_pts, c.d_evalflag, prec, A, c.d_seg_states, c.d_maxd, c.d_band, kern, elam, hl,
                       trapez, gC, gP);
}

template void launch_bandstep<float>(hipStream_t, const Ctx&, int, int, int, int, const float*, const float*,
                                     const float*, const float*, float*, float*, unsigned long long*);
template void launch_bandstep<double>(hipStream_t, const Ctx&, int, int, int, int, const double*, const double*,
                                      const double*, const double*, double*, double*, unsigned long long*);

}  // namespace sdft
