// MFMA operator-GEMM for CDNA4 (gfx950).
//
// Every convolution of the SCF objective is y = A x with the Toeplitz +- Hankel
// operator A[n][n'] = s*c[n-n'] + c[n+n'+1] (indices mod L = 2 NP), c being the
// periodic IDFT of a sampled weight.  All state points of a kernel set (same NP,
// dz) share c, so one convolution stage over a batch is a GEMM
//     Y[NP x cols] = sum_terms A_term[NP x NP] * X_term[NP x cols]
// whose columns are (state point, field) pairs.  The A operand is generated on
// the fly from two LDS windows of c (no NP x NP matrix is ever stored); B is
// staged through LDS; products run on v_mfma_f32_16x16x4f32 /
// v_mfma_f64_16x16x4f64.  Lane layouts (verified on gfx950):
//   A: lane l holds A[l % 16][l / 16]      B: lane l holds B[l / 16][l % 16]
//   D f32: D[4 (l/16) + v][l % 16]         D f64: D[4 v + (l/16)][l % 16]
//
// Two input modes share the core:
//   MODE_CONV: X columns are plain fields (weighted densities / back projections)
//   MODE_PROP: X columns are propagator inputs exp(-lambda) exp(lng/2) Q trapez
//              built while staging; the epilogue writes |y| exp(lng/2)/normative
//              into the gC / gP families (child_propagators.jl:118-141).
#include <cstring>

#include "sdft_launch.h"
#include "sdft_mfma.h"

namespace sdft {

__device__ __forceinline__ int wrapLg(int d, int L) {
    d %= L;
    return d < 0 ? d + L : d;
}

// Workgroup tile: 4 waves stacked along rows; each wave WM x WN blocks of 16 x 16.
template <int WM, int WN>
struct TileShape {
    static constexpr int TM = 4 * WM * 16;  // rows per workgroup
    static constexpr int TN = WN * 16;      // columns per workgroup
    static constexpr int KC = 16;           // n' per staged chunk
    static constexpr int CW = TM + KC;      // c window length (TM + KC - 1 used)
};

struct PropStep {       // per-launch propagator parameters
    int32_t u, s, nseg, S, fixed_seg, state_off, maxd_off, GF, MB, NB;
    int64_t goff;
    int32_t skip_sweep; // sweep whose parent is grafted at this step (handled elsewhere), or -1
    int32_t banded;     // apply the group's band half-width
    int32_t fused;      // inputs of step s >= 2 come from fin; the epilogue writes fout for step s + 1
    int32_t rescale;    // per-step rescaling: inputs from step s - 1 divided by its recorded maximum,
                        // this step's maximum recorded (scale_slot)
    int32_t nsp, nmax;
    unsigned long long* smax;
};

template <typename T, int WM, int WN, bool PROP>
__global__ void __launch_bounds__(256) k_opgemm(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                const int32_t* __restrict__ evalflag, int prec,
                                                const OpTile* __restrict__ tiles, const OpGroup* __restrict__ groups,
                                                const void* __restrict__ colsv, const ConvJob* __restrict__ jobs,
                                                const T* __restrict__ kern,
                                                // MODE_CONV inputs / outputs
                                                const T* __restrict__ rhoT, int Frt, const T* __restrict__ dphi, int Fdp,
                                                T* __restrict__ nw, int Fnw, double* __restrict__ fx,
                                                double* __restrict__ wsum, int Fws,
                                                // MODE_PROP inputs / outputs
                                                PropStep ps, const int32_t* __restrict__ seg_states,
                                                const double* __restrict__ maxd_tab, const T* __restrict__ elam,
                                                const T* __restrict__ hl, const T* __restrict__ trapez,
                                                T* __restrict__ gC, T* __restrict__ gP,
                                                const T* __restrict__ fin, T* __restrict__ fout) {
    typedef TileShape<WM, WN> Sh;
    constexpr int TM = Sh::TM, TN = Sh::TN, KC = Sh::KC;
    const OpTile tile = tiles[blockIdx.x];
    const OpGroup G = groups[tile.group];
    const int NP = G.NP, L = G.L;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const DevSystem& sys = *psys;

    __shared__ T xs[KC][TN];
    __shared__ T cmw[Sh::CW];
    __shared__ T cpw[Sh::CW];
    __shared__ int32_t colp[TN];        // point of each column (-1 inactive)
    __shared__ int32_t colaux[TN][4];   // MODE_CONV: job; MODE_PROP: sweep, a, b, bond
    __shared__ int32_t colpp[TN][2];    // MODE_PROP: pi, pj
    __shared__ T colinv[TN];            // MODE_PROP: input scale of the column (1 / previous step's maximum)
    __shared__ int any_active;

    if (tid == 0) any_active = 0;
    __syncthreads();
    if (tid < TN) {
        int p = -1;
        if (tid < tile.ncols) {
            if (!PROP) {
                const OpCol c = ((const OpCol*)colsv)[G.col_off + tile.col0 + tid];
                p = c.p;
                colaux[tid][0] = c.job;
            } else {
                const PropCol c = ((const PropCol*)colsv)[G.col_off + tile.col0 + tid];
                p = c.p;
                const int N = ps.nseg, S = ps.S;
                const int outseg = c.sweep == 0 ? N - 1 - ps.s : ps.s;
                const int inseg = c.sweep == 0 ? outseg + 1 : outseg - 1;
                const int si = seg_states[ps.state_off + outseg * S + c.a];
                const int sj = seg_states[ps.state_off + inseg * S + c.b];
                if (si < 0 || sj < 0 || c.sweep == ps.skip_sweep) p = -1;
                int bond = 0;
                if (p >= 0) {
                    const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
                    for (int u2 = 0; u2 < ps.MB; ++u2)
                        if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
                }
                colaux[tid][0] = c.sweep;
                colaux[tid][1] = c.a;
                colaux[tid][2] = c.b;
                colaux[tid][3] = bond;
                colpp[tid][0] = si <= sj ? 0 : 1;
                colpp[tid][1] = si <= sj ? 1 : 0;
                // the input comes from step s - 1 when the input segment carries a stored propagator
                const bool stored = c.sweep == 0 ? (inseg < N - 1) : (inseg > 0);
                T inv = T(1);
                if (p >= 0 && ps.rescale && stored)
                    inv = (T)(1.0 / scale_read(ps.smax, scale_slot(p, ps.u, c.sweep, ps.s - 1, ps.nsp, ps.nmax)));
                colinv[tid] = inv;
            }
            if (p >= 0 && evalflag[p] != prec) p = -1;
        }
        colp[tid] = p;
        if (p >= 0) atomicOr(&any_active, 1);
    }
    __syncthreads();
    if (!any_active) return;

    typename Mfma<T>::acc_t acc[WM][WN];
    for (int i = 0; i < WM; ++i)
        for (int j = 0; j < WN; ++j)
            for (int v = 0; v < 4; ++v) acc[i][j][v] = T(0);

    const T* cset = kern + G.kbase * (int64_t)sys.K;
    const int r0 = tile.row0;
    int k_lo = 0, k_hi = NP;
    if (PROP && ps.banded && G.band >= 0) {
        k_lo = max(0, r0 - G.band);
        k_hi = min(NP, r0 + TM + G.band);
    }
    const int nterms = PROP ? 1 : G.nterms;
    // staging assignment: each thread loads KC*TN/256 consecutive-k elements of one column
    constexpr int PER = KC * TN / 256;
    const int scol = tid / (KC / PER);
    const int sk0 = (tid % (KC / PER)) * PER;

    for (int term = 0; term < nterms; ++term) {
        const T* c = cset + (int64_t)G.kern[term] * L;
        const T sgn = G.sign[term] < 0 ? T(-1) : T(1);
        // input base pointer of the staging column for this term (MODE_CONV)
        const T* inp = nullptr;
        int sp_ = -1;
        if (scol < TN) sp_ = colp[scol];
        if (!PROP && sp_ >= 0) {
            const ConvJob& J = jobs[colaux[scol][0]];
            const DevPoint& P = pts[sp_];
            const ConvTerm tm = J.t[term];
            inp = tm.infam == 0 ? rhoT + P.base * (int64_t)Frt + (int64_t)tm.infield * P.NP
                                : dphi + P.base * (int64_t)Fdp + (int64_t)tm.infield * P.NP;
        }
        for (int k0 = k_lo; k0 < k_hi; k0 += KC) {
            __syncthreads();
            // ---- stage X chunk (transposed into xs[k][col])
            if (scol < TN) {
                for (int q = 0; q < PER; ++q) {
                    const int nn = k0 + sk0 + q;
                    T v = T(0);
                    if (sp_ >= 0 && nn < NP) {
                        if (!PROP) {
                            v = inp[nn];
                        } else if (ps.fused && ps.s >= 2) {
                            const DevPoint& P = pts[sp_];
                            const int sweep = colaux[scol][0], a = colaux[scol][1], b = colaux[scol][2];
                            v = fin[P.base * (int64_t)(2 * ps.S * ps.S) + (((int64_t)sweep * ps.S + a) * ps.S + b) * P.NP + nn] *
                                colinv[scol];
                        } else {
                            const DevPoint& P = pts[sp_];
                            const int sweep = colaux[scol][0], b = colaux[scol][2], bond = colaux[scol][3];
                            const int N = ps.nseg, S = ps.S;
                            const int outseg = sweep == 0 ? N - 1 - ps.s : ps.s;
                            const int inseg = sweep == 0 ? outseg + 1 : outseg - 1;
                            const int sj = seg_states[ps.state_off + inseg * S + b];
                            const bool stored = sweep == 0 ? (inseg < N - 1) : (inseg > 0);
                            T q_in;
                            const T* gin = sweep == 0 ? gC : gP;
                            if (!stored) {
                                if (ps.fixed_seg < 0) {
                                    q_in = T(1);
                                } else {
                                    const int Kf = P.fixed_coord[ps.u];
                                    const double* maxd = maxd_tab + ps.maxd_off;
                                    if (inseg == ps.fixed_seg) q_in = nn == Kf ? T(1) : T(0);
                                    else {
                                        const double d = (double)(nn - Kf) * P.dz;
                                        q_in = (d * d > maxd[inseg] * maxd[inseg]) ? T(0) : T(1);
                                    }
                                }
                            } else {
                                q_in = T(0);
                                const int64_t gb = P.gbase + ps.goff * P.gstride;
                                for (int b2 = 0; b2 < S; ++b2)
                                    if (seg_states[ps.state_off + inseg * S + b2] >= 0)
                                        q_in += gin[gb + (((int64_t)inseg * S + b2) * S + b) * P.gstride + nn];
                            }
                            const int64_t eb = P.base * (int64_t)ps.NB;
                            v = elam[eb + (int64_t)sj * P.NP + nn] *
                                hl[P.base * (int64_t)(2 * ps.MB) + (int64_t)(2 * bond + colpp[scol][1]) * P.NP + nn] *
                                (q_in * colinv[scol]) * trapez[eb + (int64_t)sj * P.NP + nn];
                        }
                    }
                    xs[sk0 + q][scol] = v;
                }
            }
            // ---- stage the operator windows: cm[i] = s*c[r0 - k0 - (KC-1) + i], cp[i] = c[r0 + k0 + 1 + i]
            for (int i = tid; i < TM + KC - 1; i += 256) {
                cmw[i] = sgn * c[wrapLg(r0 - k0 - (KC - 1) + i, L)];
                cpw[i] = c[wrapLg(r0 + k0 + 1 + i, L)];
            }
            __syncthreads();
            // ---- MFMA over the chunk: KC/4 k-steps
#pragma unroll
            for (int ks = 0; ks < KC / 4; ++ks) {
                const int kl = ks * 4 + (lane >> 4);  // local k of this lane's A/B element
                T bv[WN];
#pragma unroll
                for (int j = 0; j < WN; ++j) bv[j] = xs[kl][j * 16 + (lane & 15)];
#pragma unroll
                for (int i = 0; i < WM; ++i) {
                    const int rl = wave * WM * 16 + i * 16 + (lane & 15);  // local row
                    const T av = cmw[rl - kl + KC - 1] + cpw[rl + kl];
#pragma unroll
                    for (int j = 0; j < WN; ++j) acc[i][j] = Mfma<T>::mma(av, bv[j], acc[i][j]);
                }
            }
        }
    }

    // ---- epilogue
#pragma unroll
    for (int i = 0; i < WM; ++i) {
#pragma unroll
        for (int j = 0; j < WN; ++j) {
            const int cl = j * 16 + (lane & 15);
            const int p = colp[cl];
            if (!PROP) {
                if (p < 0) continue;
                const DevPoint& P = pts[p];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int n = r0 + wave * WM * 16 + i * 16 + Mfma<T>::row(lane, v);
                    if (n >= NP) continue;
                    conv_store<T>(jobs[colaux[cl][0]], P, n, (double)acc[i][j][v], nw, Fnw, fx, wsum, Fws);
                }
            } else {
                // every lane takes part (the state sum below is a lane shuffle)
                const int sweep = colaux[cl][0], a = colaux[cl][1], b = colaux[cl][2], bond = colaux[cl][3];
                const int N = ps.nseg, S = ps.S;
                const int outseg = sweep == 0 ? N - 1 - ps.s : ps.s;
                const DevPoint& P = pts[p >= 0 ? p : 0];
                double vmax = 0.0;  // this lane's largest output (per-step rescaling)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int n = r0 + wave * WM * 16 + i * 16 + Mfma<T>::row(lane, v);
                    const bool ok = p >= 0 && n < NP;
                    const T y = acc[i][j][v];
                    T val = T(0);
                    if (ok) {
                        val = (y < T(0) ? -y : y) *
                              hl[P.base * (int64_t)(2 * ps.MB) + (int64_t)(2 * bond + colpp[cl][0]) * P.NP + n] /
                              (T)P.normative[ps.u];
                        if (ps.fixed_seg >= 0) {
                            const int Kf = P.fixed_coord[ps.u];
                            const double* maxd = maxd_tab + ps.maxd_off;
                            const double d = (double)(n - Kf) * P.dz;
                            if (d * d > maxd[outseg] * maxd[outseg]) val = T(0);
                        }
                        T* gout = sweep == 0 ? gC : gP;
                        gout[P.gbase + (ps.goff + ((int64_t)outseg * S + b) * S + a) * P.gstride + n] = val;
                        vmax = (double)val > vmax ? (double)val : vmax;
                    }
                    if (ps.fused) {
                        // Q(a) = sum_b g[outseg][b][a]: the S columns of one (point, sweep, a) are adjacent lanes
                        T q = val;
                        if (S == 2) q += __shfl_xor(val, 1, 64);
                        const int nextout = sweep == 0 ? outseg - 1 : outseg + 1;
                        if (ok && b == 0 && ps.s + 1 < N) {
                            const int sj = seg_states[ps.state_off + outseg * S + a];  // input bead of step s+1
                            const int64_t eb = P.base * (int64_t)ps.NB + (int64_t)sj * P.NP + n;
                            const T e = elam[eb] * q * trapez[eb];
                            for (int a2 = 0; a2 < S; ++a2) {
                                const int si2 = seg_states[ps.state_off + nextout * S + a2];
                                if (si2 < 0) continue;
                                const int lo = si2 <= sj ? si2 : sj, hi = si2 <= sj ? sj : si2;
                                int bond2 = 0;
                                for (int u2 = 0; u2 < ps.MB; ++u2)
                                    if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond2 = u2;
                                const int pj2 = si2 <= sj ? 1 : 0;
                                fout[P.base * (int64_t)(2 * S * S) + (((int64_t)sweep * S + a2) * S + a) * P.NP + n] =
                                    e * hl[P.base * (int64_t)(2 * ps.MB) + (int64_t)(2 * bond2 + pj2) * P.NP + n];
                            }
                        }
                    }
                }
                if (ps.rescale) {
                    // lanes l, l ^ 16, l ^ 32, l ^ 48 hold rows of the same column: one atomic per column
                    for (int o = 16; o < 64; o <<= 1) {
                        const double w = __shfl_xor(vmax, o, 64);
                        vmax = w > vmax ? w : vmax;
                    }
                    if (lane < 16 && p >= 0)
                        scale_note(ps.smax, scale_slot(p, ps.u, sweep, ps.s, ps.nsp, ps.nmax), vmax);
                }
            }
        }
    }
}

// --------------------------------------------------------------------------
// Active-column compaction of the convolution plans, per precision: points in the other precision
// or converged are dropped from every column group (order kept), so that in the late iterations of
// a batch, when few points are still active, the GEMM tiles hold active columns only. It runs when
// the set can grow (solve start, the host's polls, where precision switches happen) and the GEMMs
// still check each column's point, so points that converge between two polls just idle. A column's
// result does not depend on the other columns of its tile, so this leaves every value unchanged.
// --------------------------------------------------------------------------
constexpr int COMPACT_MAX_PLANS = 3 * CONV_NSHAPES;

struct CompactArgs {
    const OpGroup* groups[COMPACT_MAX_PLANS];
    const OpCol* cols[COMPACT_MAX_PLANS];
    OpCol* act[COMPACT_MAX_PLANS];
    int32_t* count[COMPACT_MAX_PLANS];
    int32_t ngroups[COMPACT_MAX_PLANS];
    int32_t nplans;
};

// one workgroup per group of the listed plans (in order): the group's columns whose point evaluates
// at `prec`, compacted in order, and their number
__global__ void __launch_bounds__(256) k_compact_cols(CompactArgs ca, const int32_t* __restrict__ evalflag, int prec) {
    int g = blockIdx.x, pl = 0;
    while (pl < ca.nplans - 1 && g >= ca.ngroups[pl]) g -= ca.ngroups[pl++];
    if (g >= ca.ngroups[pl]) return;
    const OpGroup& G = ca.groups[pl][g];
    const OpCol* in = ca.cols[pl] + G.col_off;
    OpCol* out = ca.act[pl] + G.col_off;
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int base = 0;
    for (int c0 = 0; c0 < G.ncols; c0 += 256) {
        const int ci = c0 + tid;
        OpCol c = {0, 0};
        bool act = false;
        if (ci < G.ncols) {
            c = in[ci];
            act = evalflag[c.p] == prec;
        }
        const unsigned long long m = __ballot(act);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        if (act) out[off + before] = c;
        base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    }
    if (tid == 0) ca.count[pl][g] = base;
}

void launch_compact_plans(hipStream_t st, const Ctx& c, int prec, const OpPlan* const* plans, int nplans) {
    CompactArgs ca;
    memset(&ca, 0, sizeof(ca));
    int tot = 0, k = 0;
    for (int i = 0; i < nplans && k < COMPACT_MAX_PLANS; ++i) {
        const OpPlan* pl = plans[i];
        if (pl->ntiles == 0) continue;
        ca.groups[k] = (const OpGroup*)pl->d_groups;
        ca.cols[k] = (const OpCol*)pl->d_cols;
        ca.act[k] = (OpCol*)pl->d_cols_act[prec == 64 ? 1 : 0];
        ca.count[k] = pl->d_count[prec == 64 ? 1 : 0];
        ca.ngroups[k] = pl->ngroups;
        tot += pl->ngroups;
        ++k;
    }
    ca.nplans = k;
    if (tot > 0) hipLaunchKernelGGL(k_compact_cols, dim3(tot), dim3(256), 0, st, ca, c.d_evalflag, prec);
}

void launch_compact_cols(hipStream_t st, const Ctx& c, int prec) {
    const OpPlan* pl[COMPACT_MAX_PLANS];
    int n = 0;
    for (const Ctx::ConvStage* stg : {&c.fwd_plan, &c.pre_plan, &c.back_plan})
        for (int sh = 0; sh < CONV_NSHAPES; ++sh) pl[n++] = &stg->sub[sh];
    launch_compact_plans(st, c, prec, pl, n);
}

// --------------------------------------------------------------------------
// Dense convolution GEMM (forward weighted densities, back projections).
// Tile 64 rows x 16*WN columns, 4 waves stacked along rows, K chunks of 64 with the
// next chunk (inputs and operator windows) prefetched into registers while the
// current one runs on the matrix cores; sums over a job's terms chain the chunks.
// Inputs are staged column-major in LDS ([column][k], rows padded by 2): each staging
// load reads 64 consecutive inputs of one column, and both the staging writes and the
// B reads are free of bank conflicts.
// --------------------------------------------------------------------------
// A plan tile of 64 * WM * RSPLIT rows is computed by RSPLIT workgroups of 64 * WM rows
// (blockIdx.x = tile + ntiles * part). The 128-row shape (CONV_WIDE128) runs FP32 as WM = 2 and
// FP64 as WM = 1, RSPLIT = 2; which shape a grid length takes is a plan property (heuristic, or
// measured by sdft_autotune).
// The FP64 64x32 tile (WM = 1, WN = 2) is held to 5 waves per SIMD (96 VGPRs instead of 98, a
// 12-byte spill): at full load the forward stage went 177 -> 169 us and the back stage 139 -> 134 us.
template <typename T, int WM, int WN, int RSPLIT>
__global__ void __launch_bounds__(256, (sizeof(T) == 8 && WM == 1 && WN == 2) ? 5 : 1) k_convgemm(const DevPoint* __restrict__ pts,
                                                  const int32_t* __restrict__ evalflag, int prec,
                                                  const OpTile* __restrict__ tiles, const OpGroup* __restrict__ groups,
                                                  const OpCol* __restrict__ cols, const int32_t* __restrict__ count,
                                                  const ConvJob* __restrict__ jobs,
                                                  const T* __restrict__ kern, int Ksys, const T* __restrict__ rhoT,
                                                  int Frt, const T* __restrict__ dphi, int Fdp, T* __restrict__ nw,
                                                  int Fnw, double* __restrict__ fx, double* __restrict__ wsum,
                                                  int Fws) {
    constexpr int TM = 64 * WM, TN = WN * 16, KC = 64;
    constexpr int XK = KC + 2;             // xs row (one column's KC inputs); the pad of 2 puts the B reads
                                           // of lanes (q, g) on banks 2q + g: conflict-free
    constexpr int PER = KC * TN / 256;     // staged inputs per thread and chunk
    constexpr int CWN = TM + KC - 1;       // operator window length
    static_assert(KC == 64, "one wavefront stages one column chunk");
    static_assert(CWN <= 256, "one window element per thread");
    const int ntiles = gridDim.x / RSPLIT;
    const int part = blockIdx.x / ntiles;
    const OpTile tile = tiles[blockIdx.x - part * ntiles];
    const OpGroup* __restrict__ Gp = groups + tile.group;  // read fields in place (dynamic term index)
    const int NP = Gp->NP, L = Gp->L;
    const int r0 = tile.row0 + part * TM;
    if (r0 >= NP) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    __shared__ T xs[TN * XK];              // [column][k]
    __shared__ T cmw[TM + KC];
    __shared__ T cpw[TM + KC];
    __shared__ int32_t colp[TN];
    __shared__ int32_t coljob[TN];
    __shared__ const T* colin[3][TN];      // input field of each column and term (nullptr: inactive)
    __shared__ int any_active;

    if (tid == 0) any_active = 0;
    __syncthreads();
    if (tid < TN) {
        int p = -1, job = 0;
        if (tid < tile.ncols && tile.col0 + tid < count[tile.group]) {
            const OpCol c = cols[Gp->col_off + tile.col0 + tid];
            p = c.p;
            job = c.job;
            if (evalflag[p] != prec) p = -1;
        }
        colp[tid] = p;
        coljob[tid] = job;
        if (p >= 0) atomicOr(&any_active, 1);
    }
    __syncthreads();
    if (!any_active) return;
    if (tid < TN) {
        const int sp = colp[tid];
        for (int term = 0; term < 3; ++term) {
            const T* in = nullptr;
            if (sp >= 0 && term < Gp->nterms) {
                const DevPoint& P = pts[sp];
                const ConvTerm* tm = jobs[coljob[tid]].t + term;
                in = tm->infam == 0 ? rhoT + P.base * (int64_t)Frt + (int64_t)tm->infield * P.NP
                                    : dphi + P.base * (int64_t)Fdp + (int64_t)tm->infield * P.NP;
            }
            colin[term][tid] = in;
        }
    }
    __syncthreads();

    const T* cset = kern + Gp->kbase * (int64_t)Ksys;
    const int nk = (NP + KC - 1) / KC;
    const int nchunks = Gp->nterms * nk;

    // staging: wavefront w loads columns w, w + 4, ... (PER of them), lane = k within the chunk,
    // so every load instruction reads 64 consecutive inputs of one column
    T xr[PER];
    T crm = T(0), crp = T(0);
    auto fetch = [&](int ch) {
        const int term = ch / nk, k0 = (ch - term * nk) * KC;
        const int nn = k0 + lane;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const T* in = colin[term][4 * q + wave];
            xr[q] = (in != nullptr && nn < NP) ? in[nn] : T(0);
        }
        if (tid < CWN) {
            const T* c = cset + (int64_t)Gp->kern[term] * L;
            const T sgn = Gp->sign[term] < 0 ? T(-1) : T(1);
            crm = sgn * c[wrapLg(r0 - k0 - (KC - 1) + tid, L)];
            crp = c[wrapLg(r0 + k0 + 1 + tid, L)];
        }
    };

    typename Mfma<T>::acc_t acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
            for (int v = 0; v < 4; ++v) acc[i][j][v] = T(0);

    fetch(0);
    const int rw = wave * 16 * WM + (lane & 15);  // row of this lane in row block 0
    for (int ch = 0; ch < nchunks; ++ch) {
        __syncthreads();  // the previous chunk's products are done with xs / windows
#pragma unroll
        for (int q = 0; q < PER; ++q) xs[(4 * q + wave) * XK + lane] = xr[q];
        if (tid < CWN) {
            cmw[tid] = crm;
            cpw[tid] = crp;
        }
        __syncthreads();
        if (ch + 1 < nchunks) fetch(ch + 1);  // in flight during the products below
        // operands of k-step pairs in ping-pong registers: the LDS reads of the next pair
        // are in flight while the current pair runs on the matrix core
        T cm0[2][WM], cp0[2][WM], bv0[2][WN], cm1[2][WM], cp1[2][WM], bv1[2][WN];
        auto lds_fetch = [&](int kp, T (&cm)[2][WM], T (&cp)[2][WM], T (&bv)[2][WN]) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int kl = (2 * kp + h) * 4 + (lane >> 4);
#pragma unroll
                for (int i = 0; i < WM; ++i) {
                    cm[h][i] = cmw[rw + 16 * i - kl + KC - 1];
                    cp[h][i] = cpw[rw + 16 * i + kl];
                }
#pragma unroll
                for (int j = 0; j < WN; ++j) bv[h][j] = xs[(j * 16 + (lane & 15)) * XK + kl];
            }
        };
        auto products = [&](T (&cm)[2][WM], T (&cp)[2][WM], T (&bv)[2][WN]) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int i = 0; i < WM; ++i) {
                    const T av = cm[h][i] + cp[h][i];
#pragma unroll
                    for (int j = 0; j < WN; ++j) acc[i][j] = Mfma<T>::mma(av, bv[h][j], acc[i][j]);
                }
            }
        };
        lds_fetch(0, cm0, cp0, bv0);
#pragma unroll
        for (int kp = 0; kp < KC / 8; kp += 2) {
            lds_fetch(kp + 1, cm1, cp1, bv1);
            __builtin_amdgcn_sched_barrier(0);
            products(cm0, cp0, bv0);
            __builtin_amdgcn_sched_barrier(0);
            if (kp + 2 < KC / 8) lds_fetch(kp + 2, cm0, cp0, bv0);
            __builtin_amdgcn_sched_barrier(0);
            products(cm1, cp1, bv1);
            __builtin_amdgcn_sched_barrier(0);
        }
    }

#pragma unroll
    for (int j = 0; j < WN; ++j) {
        const int cl = j * 16 + (lane & 15);
        const int p = colp[cl];
        if (p < 0) continue;
        const DevPoint& P = pts[p];
        const ConvJob& J = jobs[coljob[cl]];
#pragma unroll
        for (int i = 0; i < WM; ++i) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int n = r0 + wave * 16 * WM + 16 * i + Mfma<T>::row(lane, v);
                if (n >= NP) continue;
                conv_store<T>(J, P, n, (double)acc[i][j][v], nw, Fnw, fx, wsum, Fws);
            }
        }
    }
}

// --------------------------------------------------------------------------
// host launchers
// --------------------------------------------------------------------------
template <typename T>
void launch_opgemm_conv(hipStream_t st, const Ctx& c, int prec, const OpPlan& plan, const T* kern, const T* rhoT,
                        const T* dphi, T* nw, double* fx) {
    if (plan.shape == CONV_FFT) {
        launch_fft_conv<T>(st, c, prec, plan, rhoT, dphi, nw, fx);
        return;
    }
    if (plan.ntiles == 0) return;
    PropStep ps;
    memset(&ps, 0, sizeof(ps));
    const OpCol* act = (const OpCol*)plan.d_cols_act[prec == 64 ? 1 : 0];
    const int32_t* cnt = plan.d_count[prec == 64 ? 1 : 0];
    auto wide = [&](auto kfn, int rsplit) {
        hipLaunchKernelGGL(kfn, dim3(plan.ntiles * rsplit), dim3(256), 0, st, c.d_pts, c.d_evalflag, prec,
                           (const OpTile*)plan.d_tiles, (const OpGroup*)plan.d_groups, act, cnt, c.d_jobs, kern, c.K,
                           rhoT, c.Frt, dphi, c.Fdp, nw, c.Fnw, fx, c.d_wsum, c.Fws);
    };
    switch (plan.shape) {
        case CONV_WIDE64:
            wide(k_convgemm<T, 1, 4, 1>, 1);
            break;
        case CONV_MID:
            wide(k_convgemm<T, 1, 2, 1>, 1);
            break;
        case CONV_NARROW:
            wide(k_convgemm<T, 1, 1, 1>, 1);
            break;
        case CONV_WIDE128:
            // FP32 takes a 128-row tile in one workgroup; FP64 splits it over two 64-row workgroups
            // (one workgroup with WM = 2 measured 389 vs 356 us for the benchmark's forward + back
            // stages: its 266 registers allow one wave per SIMD)
            if (sizeof(T) == 4) wide(k_convgemm<T, 2, 4, 1>, 1);
            else wide(k_convgemm<T, 1, 4, 2>, 2);
            break;
        default:
            hipLaunchKernelGGL((k_opgemm<T, 2, 1, false>), dim3(plan.ntiles), dim3(256), 0, st, c.d_sys, c.d_pts,
                               c.d_evalflag, prec, (const OpTile*)plan.d_tiles, (const OpGroup*)plan.d_groups,
                               plan.d_cols, c.d_jobs, kern, rhoT, c.Frt, dphi, c.Fdp, nw, c.Fnw, fx, c.d_wsum, c.Fws,
                               ps, c.d_seg_states, c.d_maxd, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr,
                               (T*)nullptr, (T*)nullptr, (const T*)nullptr, (T*)nullptr);
            break;
    }
}

template <typename T>
void launch_opgemm_prop(hipStream_t st, const Ctx& c, int prec, const OpPlan& plan, int u, int s, int skip_sweep,
                        bool banded, const T* kern, const T* elam, const T* hl, const T* trapez, T* gC, T* gP,
                        T* fbuf, unsigned long long* smax) {
    if (plan.ntiles == 0) return;
    const SpeciesHost& sh = c.species[u];
    PropStep ps;
    memset(&ps, 0, sizeof(ps));
    ps.u = u; ps.s = s; ps.nseg = sh.nseg; ps.S = sh.S; ps.fixed_seg = sh.fixed_seg;
    ps.state_off = sh.state_off; ps.maxd_off = sh.maxd_off; ps.GF = c.GF; ps.MB = c.MB; ps.NB = c.NB;
    ps.goff = sh.goff; ps.skip_sweep = skip_sweep; ps.banded = banded ? 1 : 0;
    ps.rescale = sh.rescale; ps.nsp = c.nsp; ps.nmax = c.nmax_seg; ps.smax = smax;
    // fused input chaining: homogeneous, ungrafted chains whose every (sweep, a, b) column is present
    ps.fused = (plan.full_pairs && sh.fixed_seg < 0 && fbuf != nullptr) ? 1 : 0;
    const int64_t fstride = c.sumNP * 2 * sh.S * sh.S;
    const T* fin = fbuf ? fbuf + (int64_t)(s & 1) * fstride : nullptr;
    T* fout = fbuf ? fbuf + (int64_t)((s + 1) & 1) * fstride : nullptr;
    if (plan.wide)
        hipLaunchKernelGGL((k_opgemm<T, 1, 4, true>), dim3(plan.ntiles), dim3(256), 0, st, c.d_sys, c.d_pts,
                           c.d_evalflag, prec, (const OpTile*)plan.d_tiles, (const OpGroup*)plan.d_groups, plan.d_cols,
                           c.d_jobs, kern, (const T*)nullptr, c.Frt, (const T*)nullptr, c.Fdp, (T*)nullptr, c.Fnw,
                           (double*)nullptr, (double*)nullptr, 0, ps, c.d_seg_states, c.d_maxd, elam, hl, trapez, gC, gP, fin, fout);
    else
        hipLaunchKernelGGL((k_opgemm<T, 2, 1, true>), dim3(plan.ntiles), dim3(256), 0, st, c.d_sys, c.d_pts,
                           c.d_evalflag, prec, (const OpTile*)plan.d_tiles, (const OpGroup*)plan.d_groups, plan.d_cols,
                           c.d_jobs, kern, (const T*)nullptr, c.Frt, (const T*)nullptr, c.Fdp, (T*)nullptr, c.Fnw,
                           (double*)nullptr, (double*)nullptr, 0, ps, c.d_seg_states, c.d_maxd, elam, hl, trapez, gC, gP, fin, fout);
}

template void launch_opgemm_conv<float>(hipStream_t, const Ctx&, int, const OpPlan&, const float*, const float*,
                                        const float*, float*, double*);
template void launch_opgemm_conv<double>(hipStream_t, const Ctx&, int, const OpPlan&, const double*, const double*,
                                         const double*, double*, double*);
template void launch_opgemm_prop<float>(hipStream_t, const Ctx&, int, const OpPlan&, int, int, int, bool, const float*,
                                        const float*, const float*, const float*, float*, float*, float*,
                                        unsigned long long*);
template void launch_opgemm_prop<double>(hipStream_t, const Ctx&, int, const OpPlan&, int, int, int, bool, const double*,
                                         const double*, const double*, const double*, double*, double*, double*,
                                         unsigned long long*);

// row/column tile sizes of the two shapes (for the host planner)
int opgemm_tile_rows(bool wide) { return wide ? TileShape<1, 4>::TM : TileShape<2, 1>::TM; }
int opgemm_tile_cols(bool wide) { return wide ? TileShape<1, 4>::TN : TileShape<2, 1>::TN; }
int conv_tile_rows(int shape) {
    return shape == CONV_TALL ? TileShape<2, 1>::TM : shape == CONV_WIDE128 ? 128 : 64;
}
int conv_tile_cols(int shape) {
    return shape == CONV_TALL ? TileShape<2, 1>::TN : shape == CONV_MID ? 32 : shape == CONV_NARROW ? 16 : 64;
}

}  // namespace sdft
