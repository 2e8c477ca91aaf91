// Convolutions and propagator steps on two- and three-dimensional Cartesian grids (gfx950).
//
// The reference convolves on the calculation domain 2 NP_d per dimension: the field is reflected
// evenly in every dimension (apply_mirroring.jl, mirrored.jl), FFT'd, multiplied by a radial weight
// w(|k|) (or by i w_v(k) for the vector weights, k_v w / |k| odd in k_v), and transformed back,
// keeping the real part on the primary domain (calc_rho_hat.jl, mfmt.jl:65-103, 202-237,
// child_propagators.jl:120-141). For an evenly reflected field the DFT factorises into one
// cosine transform per dimension, C(m) = sum_j f(j) prod_d 2 cos(pi m_d (j_d + 1/2) / n_d), up to
// phases that cancel on the way back. The product with a kernel even in every k_d transforms back
// with cos sums, one odd in k_v with a sin sum along v. The m_d = n_d term of each dimension either
// vanishes or contributes only to the imaginary part the reference discards. So every convolution
// here is
//     forward:  one n_d x n_d real matrix per dimension (2 cos for even input, -2 sin for input
//               odd along that dimension, i.e. a vector back-projection field),
//     multiply: the real spectral weight w(k_m) at k_d = pi m_d / (n_d h_d), m_d < n_d,
//     inverse:  one n_d x n_d real matrix per dimension (c_m cos / (2 n_d), or -sin / n_d along the
//               axis a vector output is odd in),
// on the NP_1 x ... x NP_D grid itself: half the points per dimension of the mirrored FFT, all real.
// The per-dimension matrix products are what this file launches; each is memory- and LDS-bound at
// these sizes (a few to a few hundred points per dimension).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sdft_launch.h"

namespace sdft {

static constexpr int ND_LINES = 64;   // transform lines per workgroup (one wavefront wide)

namespace {
__device__ __forceinline__ int64_t nd_fidx(const DevPoint& P, int F, int f, int n) {
    return P.base * (int64_t)F + (int64_t)f * P.NP + n;
}
template <typename T> __device__ __forceinline__ T nd_abs(T v) { return v < T(0) ? -v : v; }
__device__ __forceinline__ double nd_warp_max(double v) {   // NaN-propagating
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = (w > v || w != w) ? w : v;
    }
    return v;
}
}  // namespace

static constexpr int ND_R = 4;        // outputs per thread and line (register blocking)

// LDS of one k_nd_axis workgroup: the input tile [n][ND_LINES + 1] and, for contiguous lines, the
// output tile of the same shape
size_t nd_axis_lds_bytes(int n, size_t elem) { return (size_t)2 * n * (ND_LINES + 1) * elem; }

// out[o] = sum_i A[o][i] in[i] along one axis of every field of every point evaluated at `prec`.
// The grid of one field is [outer][n][inner] (inner = product of the faster dimensions); a line is
// one (outer, inner) pair. AT holds A transposed: AT[i * n + o]. Fields whose parity entry names
// this axis use AT_odd, the others AT_even. A workgroup stages 64 lines in LDS; each lane owns one
// line and ND_R consecutive outputs of it, so one LDS read of the input feeds ND_R FMAs against
// matrix entries that are wave-uniform (scalar loads). Contiguous lines (inner == 1) go out through
// an LDS tile so that the stores run along the line.
template <typename T>
__global__ void __launch_bounds__(256) k_nd_axis(const T* __restrict__ in, T* __restrict__ out, int F, int64_t NPt,
                                                 int n, int inner, int nlines, const T* __restrict__ AT_even,
                                                 const T* __restrict__ AT_odd, const int8_t* __restrict__ fpar, int axis,
                                                 const int32_t* __restrict__ evalflag, int prec) {
    const int b = blockIdx.y;
    const int p = b / F, f = b - p * F;
    if (evalflag[p] != prec) return;
    const T* __restrict__ AT = fpar[f] == axis ? AT_odd : AT_even;
    const T* src = in + (int64_t)b * NPt;
    T* dst = out + (int64_t)b * NPt;
    extern __shared__ char nd_lds[];
    T* tile = reinterpret_cast<T*>(nd_lds);           // [n][ND_LINES + 1]
    T* otile = tile + n * (ND_LINES + 1);             // [n][ND_LINES + 1] (contiguous lines only)
    const int line0 = blockIdx.x * ND_LINES;
    const int total = n * ND_LINES;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        int j, l;
        if (inner == 1) { j = e % n; l = e / n; }        // contiguous lines: read along the line
        else { l = e % ND_LINES; j = e / ND_LINES; }     // strided lines: read across lines
        const int line = line0 + l;
        T v = T(0);
        if (line < nlines) {
            const int o = line / inner, i = line - o * inner;
            v = src[((int64_t)o * n + j) * inner + i];
        }
        tile[j * (ND_LINES + 1) + l] = v;
    }
    __syncthreads();
    const int l = threadIdx.x % ND_LINES;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / ND_LINES);
    const int nw = (int)(blockDim.x / ND_LINES);
    const int line = line0 + l;
    const int o = line / inner, i = line - o * inner;
    for (int m0 = w * ND_R; m0 < n; m0 += nw * ND_R) {
        T acc[ND_R];
#pragma unroll
        for (int r = 0; r < ND_R; ++r) acc[r] = T(0);
        if (m0 + ND_R <= n) {
            for (int j = 0; j < n; ++j) {
                const T a = tile[j * (ND_LINES + 1) + l];
                const T* arow = AT + (int64_t)j * n + m0;
#pragma unroll
                for (int r = 0; r < ND_R; ++r) acc[r] += arow[r] * a;
            }
        } else {   // last block of outputs: only the entries below n exist
            for (int j = 0; j < n; ++j) {
                const T a = tile[j * (ND_LINES + 1) + l];
                const T* arow = AT + (int64_t)j * n + m0;
#pragma unroll
                for (int r = 0; r < ND_R; ++r)
                    if (m0 + r < n) acc[r] += arow[r] * a;
            }
        }
        if (inner == 1) {
#pragma unroll
            for (int r = 0; r < ND_R; ++r)
                if (m0 + r < n) otile[(m0 + r) * (ND_LINES + 1) + l] = acc[r];
        } else if (line < nlines) {
#pragma unroll
            for (int r = 0; r < ND_R; ++r)
                if (m0 + r < n) dst[((int64_t)o * n + m0 + r) * inner + i] = acc[r];
        }
    }
    if (inner == 1) {
        __syncthreads();
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int m = e % n, ll = e / n;
            if (line0 + ll < nlines) dst[(int64_t)(line0 + ll) * n + m] = otile[m * (ND_LINES + 1) + ll];
        }
    }
}

// Axes of at most ND_SMAX points: the same transform with the matrix staged in LDS as well. The
// per-lane reads of wave-uniform matrix entries from global memory wait a cache round trip every
// few FMAs; from LDS they are broadcast reads. Each wave keeps all its output blocks in registers
// (ND_SMAX / (4 ND_R) of them), so contiguous lines reuse the input tile as the output tile. Every
// output sums over the input positions in the same order as k_nd_axis: the results are identical.
static constexpr int ND_SMAX = 64;
static constexpr int ND_SB = ND_SMAX / (4 * ND_R);   // output blocks per wave (4 waves)

static __host__ __device__ __forceinline__ int nd_round4(int v) { return (v + 3) & ~3; }

// input (and output) tile [n][ND_LINES + 1], then the matrix AT [n][round4(n)]
size_t nd_axis_s_lds_bytes(int n, size_t elem) {
    return (size_t)(nd_round4(n * (ND_LINES + 1)) + n * nd_round4(n)) * elem;
}

// The sums of one wave over its NBK output blocks (wave-uniform count, so the loop body has no
// branches and the compiler can overlap the next position's LDS reads with this one's FMAs).
template <typename T, int NBK>
__device__ __forceinline__ void nd_axis_sums(const T* tile, const T* Ms, int n, int n4, int l, int w,
                                             T (&acc)[ND_SB][ND_R]) {
#pragma unroll 2
    for (int j = 0; j < n; ++j) {
        const T a = tile[j * (ND_LINES + 1) + l];
        const T* arow = Ms + j * n4;
#pragma unroll
        for (int k = 0; k < NBK; ++k) {
            const int m0 = (w + 4 * k) * ND_R;   // wave w: output blocks w, w + 4, ... (as k_nd_axis)
#pragma unroll
            for (int r = 0; r < ND_R; ++r) acc[k][r] += arow[m0 + r] * a;
        }
    }
}

// e / d for 0 <= e < 4096 and 1 <= d <= ND_SMAX: multiply by ceil(2^20 / d) and shift (exact over
// that whole range, checked exhaustively), instead of a 32-bit integer division
__device__ __forceinline__ int nd_div(int e, uint32_t mul) { return (int)(((uint32_t)e * mul) >> 20); }
__host__ __device__ __forceinline__ uint32_t nd_divmul(int d) { return (1u << 20) / (uint32_t)d + 1u; }

static constexpr int ND_LB = 8;   // global loads a thread issues before it writes them to LDS

template <typename T>
__global__ void __launch_bounds__(256) k_nd_axis_s(const T* __restrict__ in, T* __restrict__ out, int F, int64_t NPt,
                                                   int n, int inner, int nlines, const T* __restrict__ AT_even,
                                                   const T* __restrict__ AT_odd, const int8_t* __restrict__ fpar,
                                                   int axis, const int32_t* __restrict__ evalflag, int prec) {
    const int b = blockIdx.y;
    const int p = b / F, f = b - p * F;
    const int ef = evalflag[p];
    const int par = fpar[f];
    if (ef != prec) return;
    const T* __restrict__ AT = par == axis ? AT_odd : AT_even;
    const T* src = in + (int64_t)b * NPt;
    T* dst = out + (int64_t)b * NPt;
    extern __shared__ __attribute__((aligned(16))) char nd_lds_s[];
    T* tile = reinterpret_cast<T*>(nd_lds_s);                      // [n][ND_LINES + 1]
    const int n4 = nd_round4(n);
    T* Ms = tile + nd_round4(n * (ND_LINES + 1));                   // [n][n4], zero past n
    const int line0 = blockIdx.x * ND_LINES;
    const int total = n * ND_LINES;
    // input tile and matrix: ND_LB loads in flight per thread (addresses clamped into the field,
    // values selected after), then the LDS writes
    const uint32_t mul_n = nd_divmul(n), mul_n4 = nd_divmul(n4);
    if (inner == 1) {   // contiguous lines: the tile is one run of memory
        const int64_t base = (int64_t)line0 * n, lim = (int64_t)nlines * n;
        for (int e0 = threadIdx.x; e0 < total; e0 += ND_LB * 256) {
            T v[ND_LB];
#pragma unroll
            for (int u = 0; u < ND_LB; ++u) {
                const int64_t g = base + e0 + u * 256;
                const T x = src[g < lim ? g : lim - 1];
                v[u] = g < lim ? x : T(0);
            }
#pragma unroll
            for (int u = 0; u < ND_LB; ++u) {
                const int e = e0 + u * 256;
                if (e < total) {
                    const int l = nd_div(e, mul_n), jj = e - l * n;
                    tile[jj * (ND_LINES + 1) + l] = v[u];
                }
            }
        }
    } else {   // strided lines: each thread keeps its line, positions 4 apart
        const int l = threadIdx.x % ND_LINES, j0 = threadIdx.x / ND_LINES;
        const int line = line0 + l;
        const int lc = line < nlines ? line : nlines - 1;
        const int o = lc / inner, i = lc - o * inner;
        const T* sl = src + (int64_t)o * n * inner + i;
        for (int jb = j0; jb < n; jb += ND_LB * 4) {
            T v[ND_LB];
#pragma unroll
            for (int u = 0; u < ND_LB; ++u) {
                const int jj = jb + 4 * u;
                const T x = sl[(int64_t)(jj < n ? jj : n - 1) * inner];
                v[u] = line < nlines ? x : T(0);
            }
#pragma unroll
            for (int u = 0; u < ND_LB; ++u) {
                const int jj = jb + 4 * u;
                if (jj < n) tile[jj * (ND_LINES + 1) + l] = v[u];
            }
        }
    }
    for (int e0 = threadIdx.x; e0 < n * n4; e0 += ND_LB * 256) {
        T v[ND_LB];
#pragma unroll
        for (int u = 0; u < ND_LB; ++u) {
            const int e = e0 + u * 256;
            const int ec = e < n * n4 ? e : n * n4 - 1;
            const int jj = nd_div(ec, mul_n4), m = ec - jj * n4;
            const T x = AT[jj * n + (m < n ? m : n - 1)];
            v[u] = m < n ? x : T(0);
        }
#pragma unroll
        for (int u = 0; u < ND_LB; ++u) {
            const int e = e0 + u * 256;
            if (e < n * n4) Ms[e] = v[u];
        }
    }
    __syncthreads();
    const int l = threadIdx.x % ND_LINES;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / ND_LINES);
    const int line = line0 + l;
    const int o = line / inner, i = line - o * inner;
    T acc[ND_SB][ND_R];
#pragma unroll
    for (int k = 0; k < ND_SB; ++k)
#pragma unroll
        for (int r = 0; r < ND_R; ++r) acc[k][r] = T(0);
    int nbk = 0;   // this wave's output blocks that start below n
#pragma unroll
    for (int k = 0; k < ND_SB; ++k) nbk = (w + 4 * k) * ND_R < n ? k + 1 : nbk;
    switch (nbk) {
        case 1: nd_axis_sums<T, 1>(tile, Ms, n, n4, l, w, acc); break;
        case 2: nd_axis_sums<T, 2>(tile, Ms, n, n4, l, w, acc); break;
        case 3: nd_axis_sums<T, 3>(tile, Ms, n, n4, l, w, acc); break;
        case 4: nd_axis_sums<T, 4>(tile, Ms, n, n4, l, w, acc); break;
        default: break;
    }
    if (inner == 1) {
        __syncthreads();   // every wave has read the input tile: it becomes the output tile
#pragma unroll
        for (int k = 0; k < ND_SB; ++k)
#pragma unroll
            for (int r = 0; r < ND_R; ++r) {
                const int m = (w + 4 * k) * ND_R + r;
                if (m < n) tile[m * (ND_LINES + 1) + l] = acc[k][r];
            }
        __syncthreads();
        const int64_t base = (int64_t)line0 * n, lim = (int64_t)nlines * n;
        for (int e0 = threadIdx.x; e0 < total; e0 += ND_LB * 256) {
            T v[ND_LB];
#pragma unroll
            for (int u = 0; u < ND_LB; ++u) {
                const int e = e0 + u * 256;
                const int ec = e < total ? e : total - 1;
                const int ll = nd_div(ec, mul_n), m = ec - ll * n;
                v[u] = tile[m * (ND_LINES + 1) + ll];
            }
#pragma unroll
            for (int u = 0; u < ND_LB; ++u) {
                const int64_t g = base + e0 + u * 256;
                if (e0 + u * 256 < total && g < lim) dst[g] = v[u];
            }
        }
    } else if (line < nlines) {
#pragma unroll
        for (int k = 0; k < ND_SB; ++k)
#pragma unroll
            for (int r = 0; r < ND_R; ++r) {
                const int m = (w + 4 * k) * ND_R + r;
                if (m < n) dst[((int64_t)o * n + m) * inner + i] = acc[k][r];
            }
    }
}

// out[p][job][m] = sum over the job's terms of kspec[kern][m] * S[p][infield][m]
template <typename T>
__global__ void __launch_bounds__(256) k_nd_combine(const NdJob* __restrict__ jobs, const T* __restrict__ S, int Fin,
                                                    const T* __restrict__ kspec, int64_t NPt, T* __restrict__ out,
                                                    int Fout, const int32_t* __restrict__ evalflag, int prec) {
    const int p = blockIdx.z;
    if (evalflag[p] != prec) return;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= NPt) return;
    const NdJob& J = jobs[blockIdx.y];
    T acc = T(0);
    for (int t = 0; t < J.nterms; ++t)
        acc += kspec[(int64_t)J.t[t].kern * NPt + m] * S[((int64_t)p * Fin + J.t[t].infield) * NPt + m];
    out[((int64_t)p * Fout + blockIdx.y) * NPt + m] = acc;
}

// stores transformed job outputs with the convolution-column semantics (abs, scale, destination)
template <typename T>
__global__ void __launch_bounds__(256) k_nd_store(const DevPoint* __restrict__ pts, const NdJob* __restrict__ jobs,
                                                  int j0, const T* __restrict__ R, int Fout, int64_t NPt, T* __restrict__ nw,
                                                  int Fnw, double* __restrict__ fx, double* __restrict__ wsum, int Fws,
                                                  const int32_t* __restrict__ evalflag, int prec) {
    const int p = blockIdx.z;
    if (evalflag[p] != prec) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= NPt) return;
    const int j = j0 + blockIdx.y;
    conv_store<T>(jobs[j].out, pts[p], n, (double)R[((int64_t)p * Fout + j) * NPt + n], nw, Fnw, fx, wsum, Fws);
}

// the (a, b) state pairs of one propagator step: a state of the output segment, b of the input one
struct NdPairs {
    int n;
    int a[SDFT_MAX_S * SDFT_MAX_S], b[SDFT_MAX_S * SDFT_MAX_S], bond[SDFT_MAX_S * SDFT_MAX_S];
    int pi[SDFT_MAX_S * SDFT_MAX_S], pj[SDFT_MAX_S * SDFT_MAX_S];
};

__device__ __forceinline__ void nd_pairs(const DevSystem& sys, const int* st, int S, int outseg, int inseg, NdPairs& q) {
    q.n = 0;
    for (int a = 0; a < S; ++a) {
        const int si = st[outseg * S + a];
        if (si < 0) continue;
        for (int b = 0; b < S; ++b) {
            const int sj = st[inseg * S + b];
            if (sj < 0) continue;
            const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
            int bond = 0;
            for (int u2 = 0; u2 < sys.MB; ++u2)
                if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;   // get_bond_info
            q.a[q.n] = a; q.b[q.n] = b; q.bond[q.n] = bond;
            q.pi[q.n] = si <= sj ? 0 : 1;
            q.pj[q.n] = si <= sj ? 1 : 0;
            ++q.n;
        }
    }
}

// The pairs of both sweeps of one step, built on the host from the same tables (nd_step_pairs) and
// passed by value: the elementwise step kernels then read them from the kernel arguments instead of
// searching the state and bond tables in global memory in every thread.
struct NdStep {
    NdPairs q[2];
    int32_t si[2][SDFT_MAX_S * SDFT_MAX_S], sj[2][SDFT_MAX_S * SDFT_MAX_S];   // bead types of pair k
    int32_t slot_bond[2][SDFT_MAX_S * SDFT_MAX_S];   // bond of slot a S + b, -1: no pair
    int32_t in_valid[2];                             // bit b2: input state b2 exists
};

static void nd_step_pairs(const Ctx& c, const SpeciesHost& sh, int s, NdStep& ns) {
    const int N = sh.nseg, S = sh.S;
    const int* st = c.h_seg_states.data() + sh.state_off;
    for (int sweep = 0; sweep < 2; ++sweep) {
        const int outseg = sweep == 0 ? N - 1 - s : s;
        const int inseg = sweep == 0 ? outseg + 1 : outseg - 1;
        NdPairs& q = ns.q[sweep];
        q.n = 0;
        ns.in_valid[sweep] = 0;
        for (int k = 0; k < SDFT_MAX_S * SDFT_MAX_S; ++k) ns.slot_bond[sweep][k] = -1;
        for (int b = 0; b < S; ++b)
            if (st[inseg * S + b] >= 0) ns.in_valid[sweep] |= 1 << b;
        for (int a = 0; a < S; ++a) {
            const int si = st[outseg * S + a];
            if (si < 0) continue;
            for (int b = 0; b < S; ++b) {
                const int sj = st[inseg * S + b];
                if (sj < 0) continue;
                const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
                int bond = 0;
                for (int u2 = 0; u2 < c.MB; ++u2)
                    if (c.h_bond_i[u2] == lo && c.h_bond_j[u2] == hi) bond = u2;   // get_bond_info
                q.a[q.n] = a; q.b[q.n] = b; q.bond[q.n] = bond;
                q.pi[q.n] = si <= sj ? 0 : 1;
                q.pj[q.n] = si <= sj ? 1 : 0;
                ns.si[sweep][q.n] = si;
                ns.sj[sweep][q.n] = sj;
                ns.slot_bond[sweep][a * S + b] = bond;
                ++q.n;
            }
        }
    }
}

// propagator step input f = e^{-lambda_j} e^{lng_j / 2} Q_in trapez_j for every (sweep, a, b) slot
// (slot = a * S + b; unused slots 0): calc_f_hat_child! / calc_f_hat_parent! before the FFT
template <typename T>
__global__ void __launch_bounds__(256) k_nd_prop_in(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                    const int32_t* __restrict__ evalflag, int prec, PropArgs A,
                                                    const int32_t* __restrict__ seg_states,
                                                    const double* __restrict__ maxd_tab, const T* __restrict__ elam,
                                                    const T* __restrict__ hl, const T* __restrict__ trapez,
                                                    const T* __restrict__ gC, const T* __restrict__ gP, T* __restrict__ W,
                                                    NdStep ns) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.z;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.NP) return;
    const int sweep = blockIdx.y;
    const int N = A.nseg, S = A.S, NB = sys.NB, MB = sys.MB;
    const int outseg = sweep == 0 ? N - 1 - A.s : A.s;
    const int inseg = sweep == 0 ? outseg + 1 : outseg - 1;
    const T* gin = sweep == 0 ? gC : gP;
    const bool in_stored = sweep == 0 ? (inseg < N - 1) : (inseg > 0);
    const T qinv = (A.rescale && in_stored)
                       ? (T)(1.0 / scale_read(A.smax, scale_slot(p, A.u, sweep, A.s - 1, A.nsp, A.nmax)))
                       : T(1);
    const NdPairs& q = ns.q[sweep];
    const int S2 = S * S;
    T* w = W + ((int64_t)p * 2 * S2 + (int64_t)sweep * S2) * P.NP + n;
    for (int k = 0; k < S2; ++k) w[(int64_t)k * P.NP] = T(0);
    for (int k = 0; k < q.n; ++k) {
        const int b = q.b[k];
        const int sj = ns.sj[sweep][k];
        T qv;
        if (!in_stored) {
            qv = base_value<T>(inseg, n, A.fixed_seg, A.fixed_seg >= 0 ? P.fixed_coord[A.u] : -1, P,
                               maxd_tab + A.maxd_off);
        } else {
            qv = T(0);
            for (int b2 = 0; b2 < S; ++b2)
                if ((ns.in_valid[sweep] >> b2) & 1)
                    qv += gin[P.gbase + (A.goff + ((int64_t)inseg * S + b2) * S + b) * P.gstride + n];
            qv *= qinv;
        }
        w[(int64_t)(q.a[k] * S + b) * P.NP] = elam[nd_fidx(P, NB, sj, n)] * hl[nd_fidx(P, 2 * MB, 2 * q.bond[k] + q.pj[k], n)] *
                                            qv * trapez[nd_fidx(P, NB, sj, n)];
    }
}

// spectral product with the filtered bond kernel of each slot's bond (apply_filter! x w_bond)
template <typename T>
__global__ void __launch_bounds__(256) k_nd_prop_mul(const DevSystem* __restrict__ psys, const int32_t* __restrict__ evalflag,
                                                     int prec, PropArgs A, const int32_t* __restrict__ seg_states,
                                                     const T* __restrict__ kspec, int kbond0, int64_t NPt, T* __restrict__ W,
                                                     NdStep ns) {
    const int p = blockIdx.z;
    if (evalflag[p] != prec) return;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= NPt) return;
    const int S = A.S, S2 = S * S;
    const int sweep = blockIdx.y / S2, slot = blockIdx.y - sweep * S2;
    T* w = W + ((int64_t)p * 2 * S2 + blockIdx.y) * NPt + m;
    const int bond = ns.slot_bond[sweep][slot];
    if (bond < 0) { *w = T(0); return; }
    *w *= kspec[(int64_t)(kbond0 + bond) * NPt + m];
}

// step output g = |conv| e^{lng_i / 2} / normative into gC (child sweep) or gP (parent sweep)
template <typename T>
__global__ void __launch_bounds__(256) k_nd_prop_out(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                     const int32_t* __restrict__ evalflag, int prec, PropArgs A,
                                                     const int32_t* __restrict__ seg_states,
                                                     const double* __restrict__ maxd_tab, const T* __restrict__ hl,
                                                     const T* __restrict__ R, T* __restrict__ gC, T* __restrict__ gP,
                                                     NdStep ns) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.z;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int sweep = blockIdx.y;
    const int N = A.nseg, S = A.S, MB = sys.MB;
    const int outseg = sweep == 0 ? N - 1 - A.s : A.s;
    const int inseg = sweep == 0 ? outseg + 1 : outseg - 1;
    T* gout = sweep == 0 ? gC : gP;
    double vmax = 0.0;
    if (n < P.NP) {
        const NdPairs& q = ns.q[sweep];
        const int S2 = S * S;
        const T* r = R + ((int64_t)p * 2 * S2 + (int64_t)sweep * S2) * P.NP + n;
        const double normative = P.normative[A.u];
        // beyond the segment's reach of the grafting point the propagator is 0 (child_propagators.jl:147-157)
        bool far = false;
        if (A.fixed_seg >= 0) {
            const double md = maxd_tab[A.maxd_off + outseg];
            far = graft_dist2(P, n, P.fixed_coord[A.u]) > md * md;
        }
        for (int k = 0; k < q.n; ++k) {
            const int a = q.a[k], b = q.b[k];
            T v = nd_abs(r[(int64_t)(a * S + b) * P.NP]) * hl[nd_fidx(P, 2 * MB, 2 * q.bond[k] + q.pi[k], n)] /
                  (T)normative;
            if (far) v = T(0);
            gout[P.gbase + (A.goff + ((int64_t)outseg * S + b) * S + a) * P.gstride + n] = v;
            vmax = (double)v > vmax ? (double)v : vmax;
        }
    }
    if (A.rescale) {
        vmax = nd_warp_max(vmax);
        if ((threadIdx.x & 63) == 0) scale_note(A.smax, scale_slot(p, A.u, sweep, A.s, A.nsp, A.nmax), vmax);
    }
}

// A step whose parent is the grafted segment (handle_fixed.jl / child_propagators.jl:74-117,
// parent_propagators.jl:80-100): direct sums over the ball |r - r_graft| <= D_ij instead of a
// convolution, with the 3-D bond renormalisation 4 pi D_ij^2 (get_renorm) and half weight on the ball's
// surface (|d^2 - D_ij^2| <= 1e-6, isapprox). Child sweep: only the grafting point gets a value,
//     gC(Kf) = sum_K1 |f(K1)| e^{lng_i(Kf)/2} dV trapez_j(K1) / renorm / normative;
// parent sweep: gP(n) = |f(Kf)| e^{lng_i(n)/2} / renorm / normative inside the ball, 0 outside.
template <typename T>
__global__ void __launch_bounds__(256) k_nd_prop_graft(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                       const int32_t* __restrict__ evalflag, int prec, PropArgs A, int sweep,
                                                       const int32_t* __restrict__ seg_states,
                                                       const double* __restrict__ maxd_tab, const T* __restrict__ elam,
                                                       const T* __restrict__ hl, const T* __restrict__ trapez,
                                                       T* __restrict__ gC, T* __restrict__ gP) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.z;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.NP) return;
    const int N = A.nseg, S = A.S, NB = sys.NB, MB = sys.MB;
    const int outseg = sweep == 0 ? N - 1 - A.s : A.s;
    const int inseg = sweep == 0 ? outseg + 1 : outseg - 1;
    const int* st = seg_states + A.state_off;
    const double* maxd = maxd_tab + A.maxd_off;
    const int Kf = P.fixed_coord[A.u];
    const T* gin = sweep == 0 ? gC : gP;
    T* gout = sweep == 0 ? gC : gP;
    const bool in_stored = sweep == 0 ? (inseg < N - 1) : (inseg > 0);
    const T qinv = (A.rescale && in_stored)
                       ? (T)(1.0 / scale_read(A.smax, scale_slot(p, A.u, sweep, A.s - 1, A.nsp, A.nmax)))
                       : T(1);
    auto qin = [&](int b, int nn) -> double {
        if (!in_stored) return (double)base_value<T>(inseg, nn, A.fixed_seg, Kf, P, maxd);
        T qv = T(0);
        for (int b2 = 0; b2 < S; ++b2)
            if (st[inseg * S + b2] >= 0) qv += gin[P.gbase + (A.goff + ((int64_t)inseg * S + b2) * S + b) * P.gstride + nn];
        return (double)(qv * qinv);
    };
    NdPairs q;
    nd_pairs(sys, st, S, outseg, inseg, q);
    const double normative = P.normative[A.u];
    const double md = maxd[outseg];
    const bool far = graft_dist2(P, n, Kf) > md * md;
    double vmax = 0.0;
    for (int k = 0; k < q.n; ++k) {
        const int a = q.a[k], b = q.b[k], bond = q.bond[k];
        const int si = st[outseg * S + a], sj = st[inseg * S + b];
        const double Dij = (sys.sigma[si] + sys.sigma[sj]) / 2.0;
        const double Dij2 = Dij * Dij, renorm = 4.0 * M_PI * Dij2;
        double val = 0.0;
        if (sweep == 0) {
            if (n == Kf) {
                // the ball's bounding box on the grid
                int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, c[3] = {0, 0, 0};
                int r = Kf;
                for (int v = 0; v < P.nd; ++v) {
                    c[v] = r % P.gn[v];
                    r /= P.gn[v];
                    const int w = (int)ceil(Dij / P.gh[v]) + 1;
                    lo[v] = max(0, c[v] - w);
                    hi[v] = min(P.gn[v] - 1, c[v] + w);
                }
                const double hlo = (double)hl[nd_fidx(P, 2 * MB, 2 * bond + q.pi[k], Kf)];
                for (int i2 = lo[2]; i2 <= hi[2]; ++i2)
                    for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
                        for (int i0 = lo[0]; i0 <= hi[0]; ++i0) {
                            const int K1 = i0 + P.gn[0] * (i1 + P.gn[1] * i2);
                            const double d2 = graft_dist2(P, K1, Kf);
                            if (d2 > Dij2) continue;
                            const double f = (double)elam[nd_fidx(P, NB, sj, K1)] *
                                             (double)hl[nd_fidx(P, 2 * MB, 2 * bond + q.pj[k], K1)] * qin(b, K1);
                            double t = fabs(f) * hlo * P.dz * (double)trapez[nd_fidx(P, NB, sj, K1)] / renorm;
                            if (fabs(d2 - Dij2) <= 1e-6) t /= 2.0;
                            val += t / normative;
                        }
            }
        } else {
            const double d2 = graft_dist2(P, n, Kf);
            if (d2 <= Dij2) {
                const double f = (double)elam[nd_fidx(P, NB, sj, Kf)] *
                                 (double)hl[nd_fidx(P, 2 * MB, 2 * bond + q.pj[k], Kf)] * qin(b, Kf);
                val = fabs(f) * (double)hl[nd_fidx(P, 2 * MB, 2 * bond + q.pi[k], n)] / renorm / normative;
                if (fabs(d2 - Dij2) <= 1e-6) val /= 2.0;
            }
        }
        if (far) val = 0.0;
        gout[P.gbase + (A.goff + ((int64_t)outseg * S + b) * S + a) * P.gstride + n] = (T)val;
        vmax = (double)(T)val > vmax ? (double)(T)val : vmax;
    }
    // per-lane atomic maximum (no wave-wide reduction: lanes past the grid returned above)
    if (A.rescale) scale_note(A.smax, scale_slot(p, A.u, sweep, A.s, A.nsp, A.nmax), vmax);
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------

hipError_t nd_prepare_attributes(const Ctx& c) {
    int nmax = 1;
    for (int d = 0; d < c.ndp.D; ++d) nmax = std::max(nmax, c.ndp.n[d]);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_nd_axis<double>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)nd_axis_lds_bytes(nmax, 8));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_nd_axis<float>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)nd_axis_lds_bytes(nmax, 4));
    const int smax = std::min(nmax, ND_SMAX);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_nd_axis_s<double>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)nd_axis_s_lds_bytes(smax, 8));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_nd_axis_s<float>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)nd_axis_s_lds_bytes(smax, 4));
    return e;
}

// transforms the F fields per point of `in` along every axis (forward or inverse), ping-ponging
// between the pipeline's two ND buffers; returns the buffer holding the result
template <typename T>
static T* nd_transform(hipStream_t st, const Ctx& c, int prec, const T* in, int F, const int8_t* fpar, bool inverse,
                       T* buf_a, T* buf_b) {
    const NdPlan& g = c.ndp;
    const T* src = in;
    T* dst = (in == buf_a) ? buf_b : buf_a;
    int inner = 1;
    for (int d = 0; d < g.D; ++d) {
        const int n = g.n[d];
        const int nlines = (int)(g.NPt / n);
        const T* const* mats = reinterpret_cast<const T* const*>(sizeof(T) == 8 ? (const void* const*)g.mat64[d]
                                                                               : (const void* const*)g.mat32[d]);
        const T* ev = mats[inverse ? 2 : 0];
        const T* od = mats[inverse ? 3 : 1];
        const dim3 grid((nlines + ND_LINES - 1) / ND_LINES, c.P * F);
        if (n <= ND_SMAX)
            hipLaunchKernelGGL(k_nd_axis_s<T>, grid, dim3(256), nd_axis_s_lds_bytes(n, sizeof(T)), st, src, dst, F,
                               g.NPt, n, inner, nlines, ev, od, fpar, d, c.d_evalflag, prec);
        else
            hipLaunchKernelGGL(k_nd_axis<T>, grid, dim3(256), nd_axis_lds_bytes(n, sizeof(T)), st, src, dst, F, g.NPt,
                               n, inner, nlines, ev, od, fpar, d, c.d_evalflag, prec);
        inner *= n;
        src = dst;
        dst = (dst == buf_a) ? buf_b : buf_a;
    }
    return const_cast<T*>(src);
}

template <typename T>
void Pipeline<T>::propagate_nd(hipStream_t st, const Ctx& c, int prec, int u) {
    const SpeciesHost& sh = c.species[u];
    if (sh.model == SDFT_MODEL_MONOMERBEAD || sh.nseg < 2) return;
    const NdPlan& g = c.ndp;
    PropArgs A;
    A.u = u; A.nseg = sh.nseg; A.S = sh.S; A.fixed_seg = sh.fixed_seg;
    A.state_off = sh.state_off; A.maxd_off = sh.maxd_off; A.goff = sh.goff; A.GF = c.GF;
    A.rescale = sh.rescale; A.nsp = c.nsp; A.nmax = c.nmax_seg; A.smax = smax; A.band = -1;
    const int S2 = sh.S * sh.S;
    const T* kspec = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)g.kspec64 : (const void*)g.kspec32);
    const unsigned gx = (unsigned)((g.NPt + 255) / 256);
    NdStep ns;
    for (int s = 1; s < sh.nseg; ++s) {
        A.s = s;
        nd_step_pairs(c, sh, s, ns);
        // sweeps whose parent is the grafted segment 0 take the direct sums; they are computed after
        // the spectral path, which also writes (and then has overwritten) their outputs
        const bool fix0 = sh.fixed_seg == 0 && s == sh.nseg - 1;   // child sweep, parent = segment 0
        const bool fix1 = sh.fixed_seg == 0 && s == 1;             // parent sweep, parent = segment 0
        hipLaunchKernelGGL(k_nd_prop_in<T>, dim3(gx, 2, c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag, prec, A,
                           c.d_seg_states, c.d_maxd, elam, hl, trapez, gC, gP, nd_a, ns);
        T* S = nd_transform<T>(st, c, prec, nd_a, 2 * S2, g.d_par_even, false, nd_a, nd_b);
        hipLaunchKernelGGL(k_nd_prop_mul<T>, dim3(gx, 2 * S2, c.P), dim3(256), 0, st, c.d_sys, c.d_evalflag, prec, A,
                           c.d_seg_states, kspec, g.kbond0, g.NPt, S, ns);
        T* R = nd_transform<T>(st, c, prec, S, 2 * S2, g.d_par_even, true, nd_a, nd_b);
        hipLaunchKernelGGL(k_nd_prop_out<T>, dim3(gx, 2, c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag, prec,
                           A, c.d_seg_states, c.d_maxd, hl, R, gC, gP, ns);
        if (fix0)
            hipLaunchKernelGGL(k_nd_prop_graft<T>, dim3(gx, 1, c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag,
                               prec, A, 0, c.d_seg_states, c.d_maxd, elam, hl, trapez, gC, gP);
        if (fix1)
            hipLaunchKernelGGL(k_nd_prop_graft<T>, dim3(gx, 1, c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag,
                               prec, A, 1, c.d_seg_states, c.d_maxd, elam, hl, trapez, gC, gP);
    }
    if (sh.rescale) launch_scale_prefix(st, c, prec, u, smax, lsc);
    if (sh.fixed_seg >= 0) launch_fixed_lambda<T>(st, c, prec, A, elam, trapez, gC, gP);
}

template <typename T>
void Pipeline<T>::conv_nd(hipStream_t st, const Ctx& c, int prec, bool back, double* fx) {
    const NdPlan& g = c.ndp;
    const T* kspec = reinterpret_cast<const T*>(sizeof(T) == 8 ? (const void*)g.kspec64 : (const void*)g.kspec32);
    const unsigned gx = (unsigned)((g.NPt + 255) / 256);
    const T* in = back ? dphi : rhoT;
    const int Fin = back ? c.Fdp : c.Frt;
    const int nj = back ? g.npre + g.nback : g.nfwd;
    const NdJob* jobs = back ? g.d_back : g.d_fwd;
    if (nj == 0) return;
    T* S = nd_transform<T>(st, c, prec, in, Fin, back ? g.d_par_dphi : g.d_par_even, false, nd_a, nd_b);
    T* O = (S == nd_a) ? nd_b : nd_a;
    hipLaunchKernelGGL(k_nd_combine<T>, dim3(gx, nj, c.P), dim3(256), 0, st, jobs, S, Fin, kspec, g.NPt, O, nj,
                       c.d_evalflag, prec);
    T* R = nd_transform<T>(st, c, prec, O, nj, back ? g.d_par_even : g.d_par_fwdout, true, nd_a, nd_b);
    if (!back) {
        hipLaunchKernelGGL(k_nd_store<T>, dim3(gx, nj, c.P), dim3(256), 0, st, c.d_pts, jobs, 0, R, nj, g.NPt, nw, c.Fnw,
                           fx, c.d_wsum, c.Fws, c.d_evalflag, prec);
        return;
    }
    // shared partials (and the lng / Psi columns) first, then the projections that add the partials
    if (g.npre > 0)
        hipLaunchKernelGGL(k_nd_store<T>, dim3(gx, g.npre, c.P), dim3(256), 0, st, c.d_pts, jobs, 0, R, nj, g.NPt, nw,
                           c.Fnw, fx, c.d_wsum, c.Fws, c.d_evalflag, prec);
    if (g.nback > 0)
        hipLaunchKernelGGL(k_nd_store<T>, dim3(gx, g.nback, c.P), dim3(256), 0, st, c.d_pts, jobs, g.npre, R, nj, g.NPt,
                           nw, c.Fnw, fx, c.d_wsum, c.Fws, c.d_evalflag, prec);
}

template void Pipeline<float>::propagate_nd(hipStream_t, const Ctx&, int, int);
template void Pipeline<double>::propagate_nd(hipStream_t, const Ctx&, int, int);
template void Pipeline<float>::conv_nd(hipStream_t, const Ctx&, int, bool, double*);
template void Pipeline<double>::conv_nd(hipStream_t, const Ctx&, int, bool, double*);

// ---------------------------------------------------------------------------------------------
// setup: spectral weights, per-axis matrices and job tables
// ---------------------------------------------------------------------------------------------

// weights at |k| (delta_func.jl, step_func.jl) and the vector weight's radial factor
static double nd_w_delta(double k, double R) { return k == 0.0 ? 4.0 * M_PI * R * R : 4.0 * M_PI * R * sin(k * R) / k; }
static double nd_w_step(double k, double R) {
    return k == 0.0 ? 4.0 * M_PI * R * R * R / 3.0 : 4.0 * M_PI * (sin(k * R) - k * R * cos(k * R)) / (k * k * k);
}

// maps a one-dimensional kernel index (3 j + r, 3 NB + bond, 3 NB + MB) to the ND index; r = 2
// (vector) takes the axis
static int nd_kern(int k1, int NB, int D, int axis) {
    if (k1 < 3 * NB) {
        const int j = k1 / 3, r = k1 % 3;
        return (2 + D) * j + (r < 2 ? r : 2 + axis);
    }
    return (2 + D) * NB + (k1 - 3 * NB);
}
// maps a one-dimensional nw / dphi field index to the ND layout (vector block of NB per axis)
static int nd_field(int f1, int NB, int D, int axis) {
    if (f1 < 2 * NB) return f1;
    if (f1 < 3 * NB) return (2 + axis) * NB + (f1 - 2 * NB);
    return f1 + (D - 1) * NB;
}

int nd_setup(Ctx& c, const DevSystem& sys, const std::vector<ConvJob>& fwd1, const std::vector<ConvJob>& pre1,
             const std::vector<ConvJob>& back1, std::vector<void*>& allocs, std::string& err) {
    NdPlan& g = c.ndp;
    const int D = g.D, NB = sys.NB, MB = sys.MB;
    const int64_t NPt = g.NPt;
    g.K = (2 + D) * NB + MB + 1;
    g.kbond0 = (2 + D) * NB;
    auto dalloc_raw = [&](void** p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes) != hipSuccess) { err = "hipMalloc failed (ND setup)"; return false; }
        allocs.push_back(*p);
        return true;
    };
    auto upload_raw = [&](void** p, const void* h, size_t bytes) -> bool {
        if (!dalloc_raw(p, bytes)) return false;
        if (hipMemcpy(*p, h, bytes, hipMemcpyHostToDevice) != hipSuccess) { err = "hipMemcpy failed (ND setup)"; return false; }
        return true;
    };
    // spectral weights on the NP_1 x ... x NP_D mode grid, first dimension fastest
    std::vector<double> ks((size_t)g.K * NPt, 0.0);
    double hmax = 0.0;
    for (int d = 0; d < D; ++d) hmax = std::max(hmax, g.h[d]);
    const double kcut = 0.5 * (M_PI / hmax);   // apply_filter!: frac_cut 0.5 of the smallest Nyquist
    for (int64_t m = 0; m < NPt; ++m) {
        double kv[3] = {0, 0, 0}, k2 = 0.0, lap = 0.0;
        int64_t r = m;
        for (int d = 0; d < D; ++d) {
            const int md = (int)(r % g.n[d]);
            r /= g.n[d];
            kv[d] = M_PI * md / (g.n[d] * g.h[d]);
            k2 += kv[d] * kv[d];
            lap += (1.0 - cos(kv[d] * g.h[d])) / (g.h[d] * g.h[d]);
        }
        const double k = sqrt(k2);
        for (int j = 0; j < NB; ++j) {
            const double R = sys.sigma[j] / 2.0;
            ks[(size_t)((2 + D) * j + 0) * NPt + m] = nd_w_delta(k, R);
            ks[(size_t)((2 + D) * j + 1) * NPt + m] = nd_w_step(k, R);
            for (int d = 0; d < D; ++d)   // vectored_delta_func_fft!: -4 pi (sin kR - kR cos kR) / k^3 k_v
                ks[(size_t)((2 + D) * j + 2 + d) * NPt + m] =
                    k == 0.0 ? 0.0 : -4.0 * M_PI * (sin(k * R) - k * R * cos(k * R)) / (k * k * k) * kv[d];
        }
        const double filt = exp(-pow(k / kcut, 4.0));
        for (int u = 0; u < MB; ++u) {
            const double sij = (sys.sigma[sys.bond_i[u]] + sys.sigma[sys.bond_j[u]]) / 2.0;
            ks[(size_t)(g.kbond0 + u) * NPt + m] = nd_w_delta(k, sij) / (4.0 * M_PI * sij * sij) * filt;
        }
        ks[(size_t)(g.kbond0 + MB) * NPt + m] = m == 0 ? 0.0 : 2.0 * M_PI * sys.lB / lap;   // coul_func_fft!
    }
    if (!upload_raw(reinterpret_cast<void**>(&g.kspec64), ks.data(), ks.size() * sizeof(double))) return SDFT_E_HIP;
    {
        std::vector<float> kf(ks.begin(), ks.end());
        if (!upload_raw(reinterpret_cast<void**>(&g.kspec32), kf.data(), kf.size() * sizeof(float))) return SDFT_E_HIP;
    }
    // per-axis matrices, transposed (AT[i][o] = A[o][i]): 0 forward even, 1 forward odd,
    // 2 inverse even, 3 inverse odd
    for (int d = 0; d < D; ++d) {
        const int n = g.n[d];
        std::vector<double> M((size_t)4 * n * n, 0.0);
        for (int j = 0; j < n; ++j)
            for (int m = 0; m < n; ++m) {
                const double th = M_PI * m * (j + 0.5) / n;
                M[(size_t)0 * n * n + (size_t)j * n + m] = 2.0 * cos(th);                        // A[m][j]
                M[(size_t)1 * n * n + (size_t)j * n + m] = m == 0 ? 0.0 : -2.0 * sin(th);
                M[(size_t)2 * n * n + (size_t)m * n + j] = (m == 0 ? 1.0 : 2.0) * cos(th) / (2.0 * n);   // A[j][m]
                M[(size_t)3 * n * n + (size_t)m * n + j] = m == 0 ? 0.0 : -sin(th) / n;
            }
        for (int t = 0; t < 4; ++t) {
            if (!upload_raw(reinterpret_cast<void**>(&g.mat64[d][t]), M.data() + (size_t)t * n * n,
                            (size_t)n * n * sizeof(double)))
                return SDFT_E_HIP;
            std::vector<float> Mf(M.begin() + (size_t)t * n * n, M.begin() + (size_t)(t + 1) * n * n);
            if (!upload_raw(reinterpret_cast<void**>(&g.mat32[d][t]), Mf.data(), Mf.size() * sizeof(float)))
                return SDFT_E_HIP;
        }
    }
    // job tables from the one-dimensional ones: every vector term becomes one term per axis
    std::vector<NdJob> fj, bj;
    std::vector<int8_t> fpar;
    for (const ConvJob& J1 : fwd1) {
        const bool vec = J1.nterms == 1 && J1.t[0].odd == 1;
        for (int ax = 0; ax < (vec ? D : 1); ++ax) {
            NdJob J;
            memset(&J, 0, sizeof(J));
            J.out = J1;
            J.nterms = J1.nterms;
            for (int t = 0; t < J1.nterms; ++t) {
                J.t[t].kern = nd_kern(J1.t[t].kern, NB, D, ax);
                J.t[t].infield = J1.t[t].infield;   // rhoT fields keep their layout
            }
            if (J1.outkind == 0) J.out.outfield = nd_field(J1.outfield, NB, D, ax);
            J.odd_axis = vec ? ax : -1;
            fj.push_back(J);
            fpar.push_back((int8_t)J.odd_axis);
        }
    }
    auto map_back = [&](const ConvJob& J1) {
        NdJob J;
        memset(&J, 0, sizeof(J));
        J.out = J1;
        J.odd_axis = -1;
        for (int t = 0; t < J1.nterms; ++t) {
            if (J1.t[t].odd == 2) {   // transposed vector term: one per axis on that axis' dphi field
                for (int ax = 0; ax < D; ++ax) {
                    J.t[J.nterms].kern = nd_kern(J1.t[t].kern, NB, D, ax);
                    J.t[J.nterms].infield = nd_field(J1.t[t].infield, NB, D, ax);
                    ++J.nterms;
                }
            } else {
                J.t[J.nterms].kern = nd_kern(J1.t[t].kern, NB, D, 0);
                J.t[J.nterms].infield = nd_field(J1.t[t].infield, NB, D, 0);
                ++J.nterms;
            }
        }
        return J;
    };
    for (const ConvJob& J1 : pre1) bj.push_back(map_back(J1));
    for (const ConvJob& J1 : back1) bj.push_back(map_back(J1));
    g.nfwd = (int)fj.size();
    g.npre = (int)pre1.size();
    g.nback = (int)back1.size();
    if (!fj.empty() && !upload_raw(reinterpret_cast<void**>(&g.d_fwd), fj.data(), fj.size() * sizeof(NdJob))) return SDFT_E_HIP;
    if (!bj.empty() && !upload_raw(reinterpret_cast<void**>(&g.d_back), bj.data(), bj.size() * sizeof(NdJob))) return SDFT_E_HIP;
    // parities: forward job outputs, dphi fields (vector block of axis v odd along v), all even
    int Smax = 1;
    for (const auto& sh : c.species) Smax = std::max(Smax, sh.S);
    g.maxF = std::max({c.Frt, c.Fdp, g.nfwd, g.npre + g.nback, 2 * Smax * Smax});
    std::vector<int8_t> even((size_t)g.maxF, -1), dpar((size_t)c.Fdp, -1);
    for (int ax = 0; ax < D; ++ax)
        for (int j = 0; j < NB; ++j) dpar[(size_t)(2 + ax) * NB + j] = (int8_t)ax;
    if (fpar.empty()) fpar.push_back(-1);
    if (!upload_raw(reinterpret_cast<void**>(&g.d_par_fwdout), fpar.data(), fpar.size())) return SDFT_E_HIP;
    if (!upload_raw(reinterpret_cast<void**>(&g.d_par_dphi), dpar.data(), dpar.size())) return SDFT_E_HIP;
    if (!upload_raw(reinterpret_cast<void**>(&g.d_par_even), even.data(), even.size())) return SDFT_E_HIP;
    return SDFT_OK;
}

}  // namespace sdft
