// Device kernels of libsoftdft_hip for CDNA4 (gfx950).
//
// Convolution semantics.  The reference evaluates every convolution as
//   Re IFFT( FFT(mirror(f)) * w(k) )  on the even extension of length L = 2 NP
// (src/iDFT/coordinate_system/Cartesian/fourier_transforms/calc_rho_hat.jl:17-97,
//  utilities/apply_mirroring.jl:1-23).  With c = periodic IDFT of the sampled
// weight this is exactly the real NP x NP operator
//   y[n] = sum_{n'} f[n'] * ( s * c[n - n'] + c[n + n' + 1] )     (indices mod L)
// with s = +1 for even weights (and the forward odd vector weight) and s = -1
// for the adjoint of the odd vector weight (the -i w back-projection).  The
// kernels below apply that operator directly (Toeplitz +- Hankel), exact to
// round-off, so no arbitrary-length FFT is needed; the c vectors are built
// once per distinct (NP, dz) by k_build_cvec.
#include "sdft_launch.h"

#include <math.h>

namespace sdft {

__device__ __forceinline__ int64_t fidx(const DevPoint& P, int F, int f, int n) {
    return P.base * (int64_t)F + (int64_t)f * P.NP + n;
}

__device__ __forceinline__ int wrapL(int d, int L) {
    d %= L;
    return d < 0 ? d + L : d;
}

template <typename T> __device__ __forceinline__ T absT(T v) { return v < T(0) ? -v : v; }

__device__ __forceinline__ double warp_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// NaN-propagating max (fmax would drop a NaN residual)
__device__ __forceinline__ double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ __forceinline__ double warp_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
    return v;
}

// --------------------------------------------------------------------------
// setup: sampled weights w_k(m) and their periodic IDFT c_k[d]
// --------------------------------------------------------------------------

__device__ double weight_value(const DevSystem& sys, int k, int m, int L, double dz) {
    // signed wavenumber as in delta_func.jl:12-20 (m <= L/2 positive)
    const int nn = (m <= L / 2) ? m : m - L;
    const double kv = 2.0 * M_PI * nn / (L * dz);
    const double a = fabs(kv);
    const int NB = sys.NB, MB = sys.MB;
    if (k < 3 * NB) {
        const int j = k / 3, kind = k % 3;
        const double b = sys.sigma[j] / 2.0;
        if (kind == 0) {  // delta_func_fft!
            return (m == 0) ? 4.0 * M_PI * b * b : 4.0 * M_PI * b * sin(b * a) / a;
        } else if (kind == 1) {  // step_func_fft!
            return (m == 0) ? (4.0 / 3.0) * M_PI * b * b * b
                            : 4.0 * M_PI * (sin(b * a) - b * a * cos(b * a)) / (a * a * a);
        } else {  // vectored_delta_func_fft! (1-D: k_v = k)
            return (m == 0) ? 0.0 : -4.0 * M_PI * (sin(b * a) - b * a * cos(b * a)) / (a * a * a) * kv;
        }
    } else if (k < 3 * NB + MB) {  // bond: process_propagators.jl:13-34, then apply_filter.jl
        const int u = k - 3 * NB;
        const double sij = (sys.sigma[sys.bond_i[u]] + sys.sigma[sys.bond_j[u]]) / 2.0;
        double w = (m == 0) ? 4.0 * M_PI * sij * sij : 4.0 * M_PI * sij * sin(sij * a) / a;
        w /= 4.0 * M_PI * sij * sij;
        const double kc = 0.5 * (M_PI / dz);
        const double r = a / kc;
        return w * exp(-(r * r) * (r * r));
    } else {  // coul_func_fft!: 2 pi lB / ((1 - cos(k dz)) / dz^2), k = 0 -> 0
        if (m == 0) return 0.0;
        return 2.0 * M_PI * sys.lB / ((1.0 - cos(kv * dz)) / (dz * dz));
    }
}

__global__ void k_build_weights(const DevSystem* __restrict__ psys, const int32_t* __restrict__ set_L,
                                const int64_t* __restrict__ set_kbase, const double* __restrict__ set_dz,
                                double* __restrict__ wbuf) {
    const DevSystem& sys = *psys;
    const int s = blockIdx.z, k = blockIdx.y;
    const int L = set_L[s];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    wbuf[set_kbase[s] * sys.K + (int64_t)k * L + m] = weight_value(sys, k, m, L, set_dz[s]);
}

// c[d] = (1/L) sum_m w(m) cos(2 pi m d / L)      (even kernels)
// c[d] = -(1/L) sum_m w(m) sin(2 pi m d / L)     (odd vector kernels: Re IFFT(i w))
__global__ void __launch_bounds__(256) k_build_cvec(const DevSystem* __restrict__ psys, const int32_t* __restrict__ set_L,
                                                    const int64_t* __restrict__ set_kbase, const double* __restrict__ wbuf,
                                                    double* __restrict__ c64, float* __restrict__ c32) {
    const DevSystem& sys = *psys;
    const int s = blockIdx.z, k = blockIdx.y;
    const int L = set_L[s];
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const bool odd = (k < 3 * sys.NB) && (k % 3 == 2);
    const double* w = wbuf + set_kbase[s] * sys.K + (int64_t)k * L;
    __shared__ double ws[256];
    double acc = 0.0;
    for (int m0 = 0; m0 < L; m0 += 256) {
        __syncthreads();
        ws[threadIdx.x] = (m0 + threadIdx.x < L) ? w[m0 + threadIdx.x] : 0.0;
        __syncthreads();
        if (d < L) {
            const int mmax = min(256, L - m0);
            for (int i = 0; i < mmax; ++i) {
                const int m = m0 + i;
                const long long r = ((long long)m * d) % L;  // exact argument reduction
                double sn, cs;
                sincospi(2.0 * (double)r / (double)L, &sn, &cs);
                acc += ws[i] * (odd ? sn : cs);
            }
        }
    }
    if (d < L) {
        const double v = (odd ? -acc : acc) / (double)L;
        const int64_t o = set_kbase[s] * sys.K + (int64_t)k * L + d;
        c64[o] = v;
        c32[o] = (float)v;
    }
}

// --------------------------------------------------------------------------
// objective stage 1: one-body potential (get_onebody_potential!, eval_spatial_density.jl:138-149)
//   elam_j = exp(-lambda_j), lambda_j = mu_ex_j + (Psi + PsiC) z_j + Ext_j + delta_muH_j
//   hl_{u,e} = exp(lng_{u,e} / 2)
// --------------------------------------------------------------------------

template <typename T>
__global__ void k_prepare(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                          const int32_t* __restrict__ evalflag, int prec, const double* __restrict__ x,
                          const double* __restrict__ Ext, T* __restrict__ elam, T* __restrict__ hl) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.y;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.NP) return;
    const int NB = sys.NB, MB = sys.MB, NP = P.NP;
    const double* xp = x + P.xoff;
    const double psi = xp[(int64_t)(NB + 2 * MB) * NP + n] + xp[P.n - 1];
    for (int j = 0; j < NB; ++j) {
        const double lam = xp[(int64_t)j * NP + n] + psi * sys.z[j] + Ext[fidx(P, NB, j, n)] + P.delta_muH[j];
        elam[fidx(P, NB, j, n)] = (T)exp(-lam);
    }
    for (int f = 0; f < 2 * MB; ++f) hl[fidx(P, 2 * MB, f, n)] = (T)exp(xp[(int64_t)(NB + f) * NP + n] / 2.0);
}

// --------------------------------------------------------------------------
// chain propagator step (child_propagators.jl:53-159, parent_propagators.jl:54-139)
//
// sweep 0 (child, leaves -> root): output segment q = N-1-s, input segment c = q+1
//   gC[q][b][a](n) = | conv_bond( elam_sj * hl_{bond,pj} * Qc(c,b) * trapez_sj ) | * hl_{bond,pi}(n) / normative
//   Qc(c,b) = sum_b' gC[c][b'][b]  (1 for the leaf)
// sweep 1 (parent, root -> leaves): output segment c = s, input segment q = s-1
//   gP[c][b][a](n) = | conv_bond( elam_sj * hl_{bond,pj} * Gp(q,b) * trapez_sj ) | * hl_{bond,pi}(n) / normative
//   Gp(q,b) = sum_b' gP[q][b'][b]  (1 for the root)
// a = state of the output segment (bead si), b = state of the input segment (bead sj),
// (bond, pi, pj) = get_bond_info(si, sj) (child_propagators.jl:204-215).
// Grafted segment 0 (fixed_seg == 0) follows the direct real-space sums of
// child_propagators.jl:74-117 / parent_propagators.jl:76-95 and the
// max_distance zeroing of child_propagators.jl:147-157.
// --------------------------------------------------------------------------

struct PropArgs {
    int u;            // species
    int s;            // step 1..N-1
    int nseg, S;
    int fixed_seg;
    int state_off, maxd_off;
    int64_t goff;     // field offset of this species in the gC / gP families
    int GF;           // fields per point of the gC / gP families
    int band;         // >= 0: banded Toeplitz window (FP32 phase), -1: dense
};

template <typename T>
__device__ __forceinline__ T base_value(int seg, int n, int fixed_seg, int Kf, double dz, const double* maxd) {
    if (fixed_seg < 0) return T(1);
    if (seg == fixed_seg) return n == Kf ? T(1) : T(0);
    const double d = (double)(n - Kf) * dz;
    return (d * d > maxd[seg] * maxd[seg]) ? T(0) : T(1);
}

template <typename T, int TB>
__global__ void __launch_bounds__(TB) k_prop_step(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                  const int32_t* __restrict__ evalflag, int prec, PropArgs A,
                                                  const int32_t* __restrict__ seg_states, const double* __restrict__ maxd_tab,
                                                  const T* __restrict__ kern, const T* __restrict__ elam,
                                                  const T* __restrict__ hl, const T* __restrict__ trapez,
                                                  T* __restrict__ gC, T* __restrict__ gP, const int32_t* __restrict__ band_of_set,
                                                  int sweep_base) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.y;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int NP = P.NP, L = P.L, NB = sys.NB, MB = sys.MB, K = sys.K;
    const int n0 = blockIdx.x * TB;
    if (n0 >= NP) return;
    const int sweep = blockIdx.z + sweep_base;
    const int N = A.nseg, S = A.S;
    const int outseg = sweep == 0 ? N - 1 - A.s : A.s;
    const int inseg = sweep == 0 ? outseg + 1 : outseg - 1;
    const int* st = seg_states + A.state_off;
    const double* maxd = maxd_tab + A.maxd_off;
    const int Kf = A.fixed_seg >= 0 ? P.fixed_coord[A.u] : -1;
    const double dz = P.dz;
    const double normative = P.normative[A.u];
    const int t = threadIdx.x;
    const int n = n0 + t;
    T* gout = sweep == 0 ? gC : gP;
    const T* gin = sweep == 0 ? gC : gP;
    // does the input segment carry a stored propagator (else base value)?
    const bool in_stored = sweep == 0 ? (inseg < N - 1) : (inseg > 0);

    auto gidx = [&](int seg, int b, int a, int nn) -> int64_t {
        return P.base * (int64_t)A.GF + (A.goff + ((int64_t)seg * S + b) * S + a) * NP + nn;
    };
    // Q(n') = sum_b' g_in[inseg][b'][b](n')  or base
    auto qin = [&](int b, int nn) -> T {
        if (!in_stored) return base_value<T>(inseg, nn, A.fixed_seg, Kf, dz, maxd);
        T q = 0;
        for (int b2 = 0; b2 < S; ++b2)
            if (st[inseg * S + b2] >= 0) q += gin[gidx(inseg, b2, b, nn)];
        return q;
    };

    int npair = 0;
    int pa[SDFT_MAX_S * SDFT_MAX_S], pb[SDFT_MAX_S * SDFT_MAX_S], pbond[SDFT_MAX_S * SDFT_MAX_S];
    int ppi[SDFT_MAX_S * SDFT_MAX_S], ppj[SDFT_MAX_S * SDFT_MAX_S];
    for (int a = 0; a < S; ++a) {
        const int si = st[outseg * S + a];
        if (si < 0) continue;
        for (int b = 0; b < S; ++b) {
            const int sj = st[inseg * S + b];
            if (sj < 0) continue;
            const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
            int bond = 0;
            for (int u2 = 0; u2 < MB; ++u2)
                if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
            pa[npair] = a; pb[npair] = b; pbond[npair] = bond;
            ppi[npair] = si <= sj ? 0 : 1;
            ppj[npair] = si <= sj ? 1 : 0;
            ++npair;
        }
    }

    // grafted parent: direct real-space sums (no convolution)
    const int parent_seg = sweep == 0 ? outseg : inseg;
    if (A.fixed_seg >= 0 && parent_seg == A.fixed_seg) {
        if (n >= NP) return;
        for (int q = 0; q < npair; ++q) {
            const int a = pa[q], b = pb[q], bond = pbond[q];
            const int si = st[outseg * S + a], sj = st[inseg * S + b];
            const double Dij = (sys.sigma[si] + sys.sigma[sj]) / 2.0;
            const double Dij2 = Dij * Dij, renorm = 2.0 * Dij;
            double val = 0.0;
            if (sweep == 0) {
                // gC[Kf] = sum_{K1 within Dij} |f(K1)| hl_pi(Kf) dz trapez(K1) / renorm (/2 at the edge) / normative
                if (n == Kf) {
                    const int r = (int)ceil(Dij / dz) + 1;
                    for (int K1 = max(0, Kf - r); K1 <= min(NP - 1, Kf + r); ++K1) {
                        const double d2 = ((double)(K1 - Kf) * dz) * ((double)(K1 - Kf) * dz);
                        if (d2 <= Dij2) {
                            const double f = (double)elam[fidx(P, NB, sj, K1)] *
                                             (double)hl[fidx(P, 2 * MB, 2 * bond + ppj[q], K1)] * (double)qin(b, K1);
                            double tmp = fabs(f) * (double)hl[fidx(P, 2 * MB, 2 * bond + ppi[q], Kf)] * dz *
                                         (double)trapez[fidx(P, NB, sj, K1)] / renorm;
                            if (fabs(d2 - Dij2) <= 1e-6) tmp /= 2.0;
                            val += tmp / normative;
                        }
                    }
                }
            } else {
                // gP(n) = |f(Kf)| hl_pi(n) / renorm / normative within Dij of Kf (/2 at the edge)
                const double d2 = ((double)(n - Kf) * dz) * ((double)(n - Kf) * dz);
                if (d2 <= Dij2) {
                    const double f = (double)elam[fidx(P, NB, sj, Kf)] *
                                     (double)hl[fidx(P, 2 * MB, 2 * bond + ppj[q], Kf)] * (double)qin(b, Kf);
                    val = fabs(f) * (double)hl[fidx(P, 2 * MB, 2 * bond + ppi[q], n)] / renorm / normative;
                    if (fabs(d2 - Dij2) <= 1e-6) val /= 2.0;
                }
            }
            if (A.fixed_seg >= 0) {
                const double d = (double)(n - Kf) * dz;
                if (d * d > maxd[outseg] * maxd[outseg]) val = 0.0;
            }
            gout[gidx(outseg, b, a, n)] = (T)val;
        }
        return;
    }

    // convolution path
    __shared__ T fs[SDFT_MAX_S * SDFT_MAX_S][TB];
    __shared__ T cm[SDFT_MAX_S * SDFT_MAX_S][2 * TB];
    __shared__ T cp[SDFT_MAX_S * SDFT_MAX_S][2 * TB];
    T acc[SDFT_MAX_S * SDFT_MAX_S];
    for (int q = 0; q < SDFT_MAX_S * SDFT_MAX_S; ++q) acc[q] = T(0);
    const T* cset = kern + P.kbase * (int64_t)K;
    const int band = A.band >= 0 ? band_of_set[P.kset] : -1;
    int k_lo = 0, k_hi = NP;
    if (band >= 0) {
        k_lo = max(0, n0 - band);
        k_hi = min(NP, n0 + TB - 1 + band + 1);
        k_lo = (k_lo / TB) * TB;
    }
    for (int k0 = k_lo; k0 < k_hi; k0 += TB) {
        __syncthreads();
        {
            const int nn = k0 + t;
            for (int q = 0; q < npair; ++q) {
                T v = T(0);
                if (nn < NP) {
                    const int b = pb[q], bond = pbond[q];
                    const int sj = st[inseg * S + b];
                    v = elam[fidx(P, NB, sj, nn)] * hl[fidx(P, 2 * MB, 2 * bond + ppj[q], nn)] * qin(b, nn) *
                        trapez[fidx(P, NB, sj, nn)];
                }
                fs[q][t] = v;
            }
        }
        for (int q = 0; q < npair; ++q) {
            const T* c = cset + (int64_t)(3 * NB + pbond[q]) * L;
            for (int i = t; i < 2 * TB - 1; i += TB) {
                cm[q][i] = c[wrapL(n0 - k0 - (TB - 1) + i, L)];
                cp[q][i] = c[wrapL(n0 + k0 + 1 + i, L)];
            }
        }
        __syncthreads();
        const int jmax = min(TB, NP - k0);
        for (int q = 0; q < npair; ++q) {
            T a = acc[q];
            for (int j = 0; j < jmax; ++j) a += fs[q][j] * (cm[q][t - j + TB - 1] + cp[q][t + j]);
            acc[q] = a;
        }
    }
    if (n >= NP) return;
    for (int q = 0; q < npair; ++q) {
        const int a = pa[q], b = pb[q], bond = pbond[q];
        T v = absT(acc[q]) * hl[fidx(P, 2 * MB, 2 * bond + ppi[q], n)] / (T)normative;
        if (A.fixed_seg >= 0) {
            const double d = (double)(n - Kf) * dz;
            if (d * d > maxd[outseg] * maxd[outseg]) v = T(0);
        }
        gout[gidx(outseg, b, a, n)] = v;
    }
}

// --------------------------------------------------------------------------
// grafted chain normalisation (handle_fixed.jl:41-119): lambda = sum_K sum_a exp(-lambda_K) Gp Qc trapez dz
// at the first non-fixed segment; norm = sigma_graft / lambda.
// --------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_fixed_lambda(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                     const int32_t* __restrict__ evalflag, int prec, PropArgs A,
                                                     const int32_t* __restrict__ seg_states, const double* __restrict__ maxd_tab,
                                                     const T* __restrict__ elam, const T* __restrict__ trapez,
                                                     const T* __restrict__ gC, const T* __restrict__ gP,
                                                     double* __restrict__ fixed_lambda) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.x;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int NP = P.NP, NB = sys.NB, S = A.S, N = A.nseg;
    const int* st = seg_states + A.state_off;
    const double* maxd = maxd_tab + A.maxd_off;
    const int Kf = P.fixed_coord[A.u];
    const int q = A.fixed_seg == 0 ? 1 : 0;  // first non-fixed segment
    auto gidx = [&](int seg, int b, int a, int nn) -> int64_t {
        return P.base * (int64_t)A.GF + (A.goff + ((int64_t)seg * S + b) * S + a) * NP + nn;
    };
    double acc = 0.0;
    for (int n = threadIdx.x; n < NP; n += blockDim.x) {
        for (int a = 0; a < S; ++a) {
            const int si = st[q * S + a];
            if (si < 0) continue;
            double Gp = 0.0;
            if (q == 0) Gp = (double)base_value<T>(0, n, A.fixed_seg, Kf, P.dz, maxd);
            else
                for (int b = 0; b < S; ++b)
                    if (st[(q - 1) * S + b] >= 0) Gp += (double)gP[gidx(q, b, a, n)];
            double Qc = 0.0;
            if (q == N - 1) Qc = (a == 0) ? (double)base_value<T>(q, n, A.fixed_seg, Kf, P.dz, maxd) : 0.0;
            else
                for (int b = 0; b < S; ++b)
                    if (st[(q + 1) * S + b] >= 0) Qc += (double)gC[gidx(q, b, a, n)];
            acc += (double)elam[fidx(P, NB, si, n)] * Gp * Qc * (double)trapez[fidx(P, NB, si, n)] * P.dz;
        }
    }
    __shared__ double red[4];
    acc = warp_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
        fixed_lambda[(int64_t)p * SDFT_MAX_SPECIES + A.u] = s;
    }
}

// --------------------------------------------------------------------------
// densities (accumulate_densities!, accumulate_densities.jl:95-184; update_spatial_structures!)
// rho_seg(q,a) = norm e^{mu} exp(-lambda_si) Gp(q,a) Qc(q,a)
// bond end 0 (parent end):  += norm e^{mu} exp(-lambda_si) Gp(q,a) gC[q][b][a]
// bond end 1 (child end):   += norm e^{mu} exp(-lambda_si) Qc(q,a) gP[q][b][a]
// Also per-species totals for the neutrality shift (neutrality_shift.jl:20-47):
//   acc[u][0] = sum_seg,state num / N_u,  acc[u][1] = sum num * z
// --------------------------------------------------------------------------
template <typename T, int TB>
__global__ void __launch_bounds__(TB) k_accumulate(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                   const int32_t* __restrict__ evalflag, int prec,
                                                   const DevSpecies* __restrict__ spec, const int64_t* __restrict__ goffs, int GF,
                                                   const int32_t* __restrict__ seg_states, const double* __restrict__ maxd_tab,
                                                   const T* __restrict__ elam, const T* __restrict__ trapez,
                                                   const T* __restrict__ gC, const T* __restrict__ gP,
                                                   const double* __restrict__ fixed_lambda,
                                                   T* __restrict__ rho_beads, T* __restrict__ rho_bonds, T* __restrict__ rhoT,
                                                   double* __restrict__ acc_part, int32_t* __restrict__ errflag) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.y;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int NP = P.NP, NB = sys.NB, MB = sys.MB;
    const int n = blockIdx.x * TB + threadIdx.x;
    const bool valid = n < NP;
    const int nc = valid ? n : NP - 1;
    double rb[SDFT_MAX_NB];
    double rbo[2 * SDFT_MAX_MB];
    for (int j = 0; j < NB; ++j) rb[j] = 0.0;
    for (int f = 0; f < 2 * MB; ++f) rbo[f] = 0.0;
    __shared__ double red[2 * SDFT_MAX_SPECIES][TB / 64];
    for (int u = 0; u < sys.n_species; ++u) {
        const DevSpecies sp = spec[u];
        const int N = sp.nseg, S = sp.S;
        const int* st = seg_states + sp.state_off;
        const double* maxd = maxd_tab + sp.maxd_off;
        const int Kf = sp.fixed_seg >= 0 ? P.fixed_coord[u] : -1;
        double norm;
        if (sp.fixed_seg >= 0) {
            const double lamb = fixed_lambda[(int64_t)p * SDFT_MAX_SPECIES + u];
            if (!(isfinite(lamb) && lamb > 0.0)) {
                if (threadIdx.x == 0 && blockIdx.x == 0) errflag[p] = SDFT_E_NORMALIZATION;
                norm = 0.0;
            } else {
                norm = sp.fixed_density / lamb;
            }
        } else {
            norm = P.chain_norm[u];
        }
        const double pref = norm * exp(P.mu_species[u]);
        auto gidx = [&](int seg, int b, int a, int nn) -> int64_t {
            return P.base * (int64_t)GF + (goffs[u] + ((int64_t)seg * S + b) * S + a) * NP + nn;
        };
        double num = 0.0, chg = 0.0;
        for (int q = 0; q < N; ++q) {
            for (int a = 0; a < S; ++a) {
                const int si = st[q * S + a];
                if (si < 0) continue;
                double Gp = 0.0, Qc = 0.0;
                if (q == 0) Gp = (double)base_value<T>(0, nc, sp.fixed_seg, Kf, P.dz, maxd);
                else
                    for (int b = 0; b < S; ++b)
                        if (st[(q - 1) * S + b] >= 0) Gp += (double)gP[gidx(q, b, a, nc)];
                // leaf: sum gC[K, :, 1, idx_i, leaf] is the base case 1 for every state
                if (q == N - 1) Qc = (double)base_value<T>(q, nc, sp.fixed_seg, Kf, P.dz, maxd);
                else
                    for (int b = 0; b < S; ++b)
                        if (st[(q + 1) * S + b] >= 0) Qc += (double)gC[gidx(q, b, a, nc)];
                const double el = (double)elam[fidx(P, NB, si, nc)];
                const double contr = pref * el * Gp;
                const double rs = contr * Qc;
                const double tz = (double)trapez[fidx(P, NB, si, nc)];
                if (valid) {
                    if (q == sp.fixed_seg) {
                        if (nc == Kf) {
                            rb[si] += rs / P.dz / tz;
                            num += rs;
                            chg += rs * sys.z[si];
                        }
                    } else {
                        rb[si] += rs;
                        num += rs * tz * P.dz;
                        chg += rs * tz * P.dz * sys.z[si];
                    }
                    if (q < N - 1) {
                        for (int b = 0; b < S; ++b) {
                            const int sj = st[(q + 1) * S + b];
                            if (sj < 0) continue;
                            const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
                            int bond = 0;
                            for (int u2 = 0; u2 < MB; ++u2)
                                if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
                            rbo[2 * bond + 0] += contr * (double)gC[gidx(q, b, a, nc)];
                        }
                    }
                    if (q > 0) {
                        const double c1 = pref * el * Qc;
                        for (int b = 0; b < S; ++b) {
                            const int sj = st[(q - 1) * S + b];
                            if (sj < 0) continue;
                            const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
                            int bond = 0;
                            for (int u2 = 0; u2 < MB; ++u2)
                                if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
                            rbo[2 * bond + 1] += c1 * (double)gP[gidx(q, b, a, nc)];
                        }
                    }
                }
            }
        }
        // Num_u = sum num / N_u (chains); charge = sum num * z (not divided): neutrality_shift.jl:34-45
        num = warp_sum(num / N);
        chg = warp_sum(chg);
        if ((threadIdx.x & 63) == 0) {
            red[2 * u][threadIdx.x >> 6] = num;
            red[2 * u + 1][threadIdx.x >> 6] = chg;
        }
    }
    __syncthreads();
    // per-block partial sums, reduced in a fixed order by k_neutrality (deterministic, batch-invariant)
    if (threadIdx.x < 2 * sys.n_species && blockIdx.x < SDFT_MAX_TILES) {
        double s = 0.0;
        for (int w = 0; w < TB / 64; ++w) s += red[threadIdx.x][w];
        acc_part[((int64_t)p * SDFT_MAX_TILES + blockIdx.x) * 2 * SDFT_MAX_SPECIES + threadIdx.x] = s;
    }
    if (!valid) return;
    for (int j = 0; j < NB; ++j) {
        rho_beads[fidx(P, NB, j, n)] = (T)rb[j];
        rhoT[fidx(P, NB + 2 * MB, j, n)] = (T)(rb[j] * (double)trapez[fidx(P, NB, j, n)]);
    }
    for (int u = 0; u < MB; ++u) {
        for (int e = 0; e < 2; ++e) {
            const int bt = e == 0 ? sys.bond_i[u] : sys.bond_j[u];
            rho_bonds[fidx(P, 2 * MB, 2 * u + e, n)] = (T)rbo[2 * u + e];
            rhoT[fidx(P, NB + 2 * MB, NB + 2 * u + e, n)] = (T)(rbo[2 * u + e] * (double)trapez[fidx(P, NB, bt, n)]);
        }
    }
}

// --------------------------------------------------------------------------
// generic dense convolution job: y = scale * sum_terms A_term * in_term (Toeplitz +- Hankel)
// --------------------------------------------------------------------------
template <typename T, int TB>
__global__ void __launch_bounds__(TB) k_conv(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                             const int32_t* __restrict__ evalflag, int prec,
                                             const ConvJob* __restrict__ jobs, const T* __restrict__ kern,
                                             const T* __restrict__ rhoT, int Frt, const T* __restrict__ dphi, int Fdp,
                                             T* __restrict__ nw, int Fnw, double* __restrict__ fx) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.z;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int NP = P.NP, L = P.L;
    const int n0 = blockIdx.x * TB;
    if (n0 >= NP) return;
    const ConvJob J = jobs[blockIdx.y];
    const int t = threadIdx.x;
    const int n = n0 + t;
    __shared__ T xs[TB];
    __shared__ T cm[2 * TB];
    __shared__ T cp[2 * TB];
    T acc = T(0);
    const T* cset = kern + P.kbase * (int64_t)sys.K;
    for (int term = 0; term < J.nterms; ++term) {
        const ConvTerm tm = J.t[term];
        const T* c = cset + (int64_t)tm.kern * L;
        const T* in = tm.infam == 0 ? rhoT + fidx(P, Frt, tm.infield, 0) : dphi + fidx(P, Fdp, tm.infield, 0);
        const T sm = tm.odd == 2 ? T(-1) : T(1);
        for (int k0 = 0; k0 < NP; k0 += TB) {
            __syncthreads();
            xs[t] = (k0 + t < NP) ? in[k0 + t] : T(0);
            for (int i = t; i < 2 * TB - 1; i += TB) {
                cm[i] = sm * c[wrapL(n0 - k0 - (TB - 1) + i, L)];
                cp[i] = c[wrapL(n0 + k0 + 1 + i, L)];
            }
            __syncthreads();
            const int jmax = min(TB, NP - k0);
            for (int j = 0; j < jmax; ++j) acc += xs[j] * (cm[t - j + TB - 1] + cp[t + j]);
        }
    }
    if (n >= NP) return;
    double y = (double)acc * J.scale;
    if (J.absval) y = fabs(y);
    if (J.outkind == 0) nw[fidx(P, Fnw, J.outfield, n)] = (T)y;
    else fx[P.xoff + (int64_t)J.outfield * NP + n] = y;
}

// --------------------------------------------------------------------------
// pointwise functional derivatives, FP64 arithmetic
//   FMT (mfmt.jl:106-199 / TPT1mfmt.jl:135-241 with the iLST kernels mfmt.jl:210-244,
//   TPT1mfmt.jl:311-374), MSA (msa.jl / TPT1msa.jl with Gamma_msa msa.jl:181-230)
// outputs the back-projection inputs:
//   dphi[j]          = dPhi1/(pi s^2) + dPhi2/(2 pi s) + dPhi3  (+ MSA mu_j / (pi s^2))   -> w2_j
//   dphi[NB + j]     = dPhi4                                                              -> w3_j
//   dphi[2NB + j]    = dPhi5/(2 pi s) + dPhi8                                             -> wV_j (adjoint)
//   dphi[3NB + 2u+e] = (lng_fmt_u + lng_msa_u) / (pi s_e^2)                               -> w2_e
//   dphi[3NB + 2MB]  = surf + sum_j z_j rho_j trapez_j                                     -> Coulomb
// --------------------------------------------------------------------------

struct GammaOut { double G, eta, H; };

__device__ GammaOut gamma_msa(const double* rho, const DevSystem& sys) {
    const int NB = sys.NB;
    const double tol = 1e-10;
    double s_rz = 0.0, s_r = 0.0, s_rs = 0.0, s_rs3 = 0.0;
    for (int j = 0; j < NB; ++j) {
        s_rz += rho[j] * sys.z[j] * sys.z[j];
        s_r += rho[j];
        s_rs += rho[j] * sys.sigma[j];
        s_rs3 += rho[j] * sys.sigma[j] * sys.sigma[j] * sys.sigma[j];
    }
    const double kap = sqrt(4.0 * M_PI * sys.lB * s_rz);
    GammaOut o;
    if (kap < tol) {
        o.G = kap / 2.0; o.eta = 0.0; o.H = 2.0 / M_PI;
        return o;
    }
    const double seff = s_rs / s_r;
    const double x = seff * sqrt(4.0 * M_PI * sys.lB * s_rz);
    double G = (sqrt(1.0 + 2.0 * x) - 1.0) / (2.0 * seff);
    const double term2 = 2.0 / M_PI * (1.0 - M_PI / 6.0 * s_rs3);
    double H = 0.0, eta = 0.0, G1 = 0.0;
    int it = 0;
    while (true) {
        double h = 0.0, e = 0.0;
        for (int j = 0; j < NB; ++j) {
            const double s = sys.sigma[j];
            h += rho[j] * s * s * s / (1.0 + G * s);
            e += rho[j] * sys.z[j] * s / (1.0 + G * s);
        }
        H = h + term2;
        eta = e / H;
        double g1 = 0.0;
        for (int j = 0; j < NB; ++j) {
            const double s = sys.sigma[j];
            const double q = sys.z[j] - eta * s * s;
            g1 += rho[j] * q * q / (1.0 + G * s) / (1.0 + G * s);
        }
        G1 = sqrt(M_PI * sys.lB * g1);
        const double Gn = 0.9 * G + 0.1 * G1;
        ++it;
        if (fabs(G1 - G) < tol || it >= 1000) {
            G = Gn;
            break;
        }
        G = Gn;
    }
    o.G = G; o.eta = eta; o.H = H;
    return o;
}

// compute_msa_deriv! (iLST msa.jl:160-170)
__device__ void msa_deriv(const double* rho, const DevSystem& sys, const GammaOut& g, double* dphi) {
    const int NB = sys.NB;
    double ustar = 0.0;
    for (int j = 0; j < NB; ++j) {
        const double s = sys.sigma[j];
        ustar += rho[j] * s * s * ((sys.z[j] - g.eta * s * s) / (1.0 + g.G * s) + sys.z[j]);
    }
    ustar *= -M_PI / 6.0;
    for (int j = 0; j < NB; ++j) {
        const double s = sys.sigma[j], z = sys.z[j];
        dphi[j] = -sys.lB * (g.G * z * z / (1.0 + g.G * s) +
                             g.eta * s * ((2.0 * z - g.eta * s * s) / (1.0 + g.G * s) + g.eta * s * s / 3.0) + z * ustar);
    }
}

// compute_TPT1msa_deriv! (iLST TPT1msa.jl:210-282) -- adds the chain term to dphi, fills lng
__device__ void tpt1msa_extra(const double* rho, const double* rbond /*[2MB]*/, const DevSystem& sys, const GammaOut& g,
                              double* dphi, double* lng) {
    const int NB = sys.NB, MB = sys.MB;
    const double lB = sys.lB, G = g.G, eta = g.eta, H = g.H;
    for (int u = 0; u < MB; ++u) {
        const int i = sys.bond_i[u], j = sys.bond_j[u];
        const double sij = (sys.sigma[i] + sys.sigma[j]) / 2.0;
        if (G == 0.0) {
            lng[u] = -lB / sij * sys.z[i] * sys.z[j];
        } else {
            const double Xi = (sys.z[i] - eta * sys.sigma[i] * sys.sigma[i]) / (1.0 + G * sys.sigma[i]);
            const double Xj = (sys.z[j] - eta * sys.sigma[j] * sys.sigma[j]) / (1.0 + G * sys.sigma[j]);
            lng[u] = -lB / sij * Xi * Xj;
        }
    }
    if (G == 0.0) return;
    double dfG = 0.0, dfe = 0.0, deG = 0.0, dHG = 0.0;
    for (int w = 0; w < NB; ++w) {
        const double s = sys.sigma[w], z = sys.z[w];
        const double opg = 1.0 + G * s;
        const double q = z - eta * s * s;
        dfG += rho[w] * s * q * q / opg / opg / opg;
        dfe += rho[w] * s * s * q / opg / opg;
        deG += rho[w] * z * s * s / opg / opg;
        dHG += rho[w] * s * s * s * s / opg / opg;
    }
    dfG *= -2.0 * M_PI * lB;
    dfe *= -2.0 * M_PI * lB;
    deG *= -1.0 / H;
    dHG = -dHG;
    const double deH = -eta / H;
    const double detaG = deG + deH * dHG;
    const double denom = 1.0 / (2.0 * G - dfe * detaG - dfG);
    for (int w = 0; w < NB; ++w) {
        const double s = sys.sigma[w], z = sys.z[w];
        const double ig = 1.0 / (1.0 + G * s);
        const double Xw = (z - eta * s * s) * ig;
        const double dfr = M_PI * lB * Xw * Xw;
        const double der = s / H * z * ig;
        const double dHr = s * s * s * (ig - 1.0 / 3.0);
        const double dGr = (dfr + dfe * (der + dfe * deH * dHr)) * denom;
        const double detr = detaG * dGr;
        double add = 0.0;
        for (int u = 0; u < MB; ++u) {
            const int i = sys.bond_i[u], j = sys.bond_j[u];
            const double sij = (sys.sigma[i] + sys.sigma[j]) / 2.0;
            const double igi = 1.0 / (1.0 + G * sys.sigma[i]);
            const double Xi = (sys.z[i] - eta * sys.sigma[i] * sys.sigma[i]) * igi;
            const double dXi = -(sys.sigma[i] * sys.sigma[i] * igi * detr + sys.sigma[i] * Xi * igi * dGr);
            const double igj = 1.0 / (1.0 + G * sys.sigma[j]);
            const double Xj = (sys.z[j] - eta * sys.sigma[j] * sys.sigma[j]) * igj;
            const double dXj = -(sys.sigma[j] * sys.sigma[j] * igj * detr + sys.sigma[j] * Xj * igj * dGr);
            add += (rbond[2 * u] + rbond[2 * u + 1]) / 2.0 * lB / sij * (dXi * Xj + Xi * dXj);
        }
        dphi[w] += add;
    }
}

struct FmtIn {
    double n0, n1, n2, n3, nV1, nV2;
    double s3;   // sum of capped n3_i (TPT1) or capped n3
};

__device__ FmtIn fmt_collect(const double* n2j, const double* n3j, const double* nVj, const DevSystem& sys) {
    FmtIn f{0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < sys.NB; ++j) {
        const double s = sys.sigma[j];
        f.n0 += n2j[j] / (M_PI * s * s);
        f.n1 += n2j[j] / (2.0 * M_PI * s);
        f.n2 += n2j[j];
        f.n3 += n3j[j];
        f.nV1 += nVj[j] / (2.0 * M_PI * s);
        f.nV2 += nVj[j];
    }
    if (f.n3 > 0.95) {  // cap (mfmt.jl:126-131 / TPT1mfmt.jl:154-159: uniform scaling)
        const double fac = 0.95 / f.n3;
        f.n0 *= fac; f.n1 *= fac; f.n2 *= fac; f.n3 *= fac; f.nV1 *= fac; f.nV2 *= fac;
    }
    f.s3 = f.n3;
    return f;
}

// compute_mfmt_deriv!: rows 1..4, 5 (nV1_z), 8 (nV2_z)
__device__ void fmt_deriv(const FmtIn& f, double* d) {
    const double om = 1.0 - f.n3;
    const double lg = log(om);
    const double i1 = 1.0 / om;
    const double i3 = 1.0 / f.n3;
    const double c = 1.0 / (36.0 * M_PI);
    const double dot12 = f.nV1 * f.nV2, dot22 = f.nV2 * f.nV2;
    d[0] = -lg;
    d[1] = f.n2 * i1;
    d[2] = f.n1 * i1 + (f.n2 * f.n2 - dot22) * 3.0 * c * i3 * (lg * i3 + i1 * i1);
    d[3] = f.n0 * i1 + (f.n1 * f.n2 - dot12) * i1 * i1 -
           (lg * i3 * i3 * i3 * 2.0 * c + i3 * i3 * i1 * c + (1.0 - 3.0 * f.n3) * i3 * i3 * c * i1 * i1 * i1) *
               (f.n2 * f.n2 * f.n2 - 3.0 * f.n2 * dot22);
    d[4] = -f.nV2 * i1;
    d[5] = -f.nV1 * i1 - (lg * i3 + i1 * i1) * f.n2 * f.nV2 * i3 * 6.0 * c;
}

// free_energy_mfmt (inhomogeneous); cubic term takes its n3 -> 0 limit (0) at n3 == 0
__device__ double fmt_phi(const FmtIn& f) {
    const double om = 1.0 - f.n3;
    const double lg = log(om);
    const double dot12 = f.nV1 * f.nV2, dot22 = f.nV2 * f.nV2;
    double cubic = 0.0;
    if (f.n3 != 0.0)
        cubic = (f.n2 * f.n2 * f.n2 - 3.0 * f.n2 * dot22) * (f.n3 + om * om * lg) / (36.0 * M_PI * f.n3 * f.n3 * om * om);
    return -f.n0 * lg + (f.n1 * f.n2 - dot12) / om + cubic;
}

template <typename T>
__device__ void load_point_inputs(const DevSystem& sys, const DevPoint& P, int n, const T* nw, int Fnw,
                                  double* n2j, double* n3j, double* nVj, double* rbond) {
    const int NB = sys.NB, MB = sys.MB;
    for (int j = 0; j < NB; ++j) {
        n2j[j] = (double)nw[fidx(P, Fnw, j, n)];
        n3j[j] = (double)nw[fidx(P, Fnw, NB + j, n)];
        nVj[j] = (double)nw[fidx(P, Fnw, 2 * NB + j, n)];
    }
    for (int f = 0; f < 2 * MB; ++f) rbond[f] = (double)nw[fidx(P, Fnw, 3 * NB + f, n)];
}

template <typename T>
__global__ void __launch_bounds__(256) k_pointwise(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                  const int32_t* __restrict__ evalflag, int prec,
                                                  const T* __restrict__ nw, int Fnw, const T* __restrict__ rhoT, int Frt,
                                                  const double* __restrict__ surf, T* __restrict__ dphi, int Fdp) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.y;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P.NP) return;
    const int NB = sys.NB, MB = sys.MB;
    double n2j[SDFT_MAX_NB], n3j[SDFT_MAX_NB], nVj[SDFT_MAX_NB], rbond[2 * SDFT_MAX_MB];
    load_point_inputs(sys, P, n, nw, Fnw, n2j, n3j, nVj, rbond);
    double ch0[SDFT_MAX_NB], ch1[SDFT_MAX_NB], chv[SDFT_MAX_NB], lngs[SDFT_MAX_MB];
    for (int j = 0; j < NB; ++j) ch0[j] = ch1[j] = chv[j] = 0.0;
    for (int u = 0; u < MB; ++u) lngs[u] = 0.0;
    const int F = sys.functionals;
    if (F & (SDFT_F_MFMT | SDFT_F_TPT1MFMT)) {
        const FmtIn f = fmt_collect(n2j, n3j, nVj, sys);
        if (f.s3 > 1e-10) {
            double d[6];
            fmt_deriv(f, d);
            if (F & SDFT_F_TPT1MFMT) {
                // TPT1 chain term (TPT1mfmt.jl:387-445) with the bond weighted densities
                const double Xi = fmax(1.0 - f.nV2 * f.nV2 / (f.n2 * f.n2), 0.0);
                const double i1 = 1.0 / (1.0 - f.n3);
                const double t2 = f.n2 * Xi / 4.0 * i1 * i1;
                const double t3 = f.n2 * f.n2 * Xi / 72.0 * i1 * i1 * i1;
                for (int u = 0; u < MB; ++u) {
                    const double si = sys.sigma[sys.bond_i[u]], sj = sys.sigma[sys.bond_j[u]];
                    const double sij = (si + sj) / 2.0, tau = si * sj / sij;
                    const double y = i1 + t2 * tau + t3 * tau * tau;
                    const double dn2 = (tau / 4.0 * i1 * i1 + f.n2 * tau * tau * i1 * i1 * i1 / 36.0) / y;
                    const double dn3 = (i1 * i1 + f.n2 * Xi / 2.0 * i1 * i1 * i1 * tau +
                                        f.n2 * f.n2 * Xi / 24.0 * i1 * i1 * i1 * i1 * tau * tau) / y;
                    const double dv = -f.nV2 * (tau / (2.0 * f.n2) * i1 * i1 + tau * tau / 36.0 * i1 * i1 * i1) / y;
                    const double rb = rbond[2 * u] + rbond[2 * u + 1];
                    d[2] -= 0.5 * rb * dn2;
                    d[3] -= 0.5 * rb * dn3;
                    d[5] -= 0.5 * rb * dv;
                    lngs[u] += log(y);
                }
            }
            for (int j = 0; j < NB; ++j) {
                const double s = sys.sigma[j];
                ch0[j] = d[0] / (M_PI * s * s) + d[1] / (2.0 * M_PI * s) + d[2];
                ch1[j] = d[3];
                chv[j] = d[4] / (2.0 * M_PI * s) + d[5];
            }
        }
    }
    if (F & (SDFT_F_MSA | SDFT_F_TPT1MSA)) {
        double rho[SDFT_MAX_NB], dm[SDFT_MAX_NB], lm[SDFT_MAX_MB];
        for (int j = 0; j < NB; ++j) rho[j] = n2j[j] / (M_PI * sys.sigma[j] * sys.sigma[j]);
        const GammaOut g = gamma_msa(rho, sys);
        msa_deriv(rho, sys, g, dm);
        if (F & SDFT_F_TPT1MSA) {
            tpt1msa_extra(rho, rbond, sys, g, dm, lm);
            for (int u = 0; u < MB; ++u) lngs[u] += lm[u];
        }
        for (int j = 0; j < NB; ++j) ch0[j] += dm[j] / (M_PI * sys.sigma[j] * sys.sigma[j]);
    }
    for (int j = 0; j < NB; ++j) {
        dphi[fidx(P, Fdp, j, n)] = (T)ch0[j];
        dphi[fidx(P, Fdp, NB + j, n)] = (T)ch1[j];
        dphi[fidx(P, Fdp, 2 * NB + j, n)] = (T)chv[j];
    }
    for (int u = 0; u < MB; ++u) {
        for (int e = 0; e < 2; ++e) {
            const int bt = e == 0 ? sys.bond_i[u] : sys.bond_j[u];
            dphi[fidx(P, Fdp, 3 * NB + 2 * u + e, n)] = (T)(lngs[u] / (M_PI * sys.sigma[bt] * sys.sigma[bt]));
        }
    }
    double q = surf[P.base + n];
    for (int j = 0; j < NB; ++j) q += sys.z[j] * (double)rhoT[fidx(P, Frt, j, n)];
    dphi[fidx(P, Fdp, 3 * NB + 2 * MB, n)] = (T)q;
}

// --------------------------------------------------------------------------
// neutrality shift (compute_neutrality_shift, neutrality_shift.jl:1-99): one thread per point
// --------------------------------------------------------------------------
__global__ void k_neutrality(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                             const DevSpecies* __restrict__ spec, const int32_t* __restrict__ evalflag, int prec,
                             const double* __restrict__ acc_part, double* __restrict__ acc_species, const double* __restrict__ x,
                             double* __restrict__ fx, int P_count) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_count || evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int nsp = sys.n_species;
    double Num[SDFT_MAX_SPECIES], qs[SDFT_MAX_SPECIES];
    double Qtot = P.total_charge;
    const int ntiles = (P.NP + 255) / 256;
    for (int k = 0; k < 2 * nsp; ++k) {
        double t = 0.0;
        for (int b = 0; b < ntiles; ++b) t += acc_part[((int64_t)p * SDFT_MAX_TILES + b) * 2 * SDFT_MAX_SPECIES + k];
        acc_species[(int64_t)p * 2 * SDFT_MAX_SPECIES + k] = t;
    }
    for (int u = 0; u < nsp; ++u) {
        Num[u] = acc_species[(int64_t)p * 2 * SDFT_MAX_SPECIES + 2 * u];
        const double ch = acc_species[(int64_t)p * 2 * SDFT_MAX_SPECIES + 2 * u + 1];
        qs[u] = Num[u] <= 0.0 ? 0.0 : ch / Num[u];
        if (spec[u].multistate || spec[u].fixed_seg >= 0) {
            Qtot += Num[u] * qs[u];
            qs[u] = 0.0;
        }
    }
    double s = 0.0;
    for (int u = 0; u < nsp; ++u) s += Num[u] * qs[u];
    double err = fabs(Qtot + s);
    double psi1 = 0.0, cst = 1.0;
    int it = 1;
    while (err > 1e-10) {
        double f = Qtot, f1 = 0.0;
        for (int i = 0; i < nsp; ++i) {
            const double t = pow(cst, qs[i]) * Num[i] * qs[i];
            f += t;
            f1 -= t * qs[i];
        }
        err = fabs(f);
        if (fabs(f1) < 1e-10) break;
        const double del = fmax(1.0, fabs(f / f1));
        psi1 -= f / (f1 * del);
        cst = exp(-psi1);
        ++it;
        if (it > 1000) break;
    }
    fx[P.xoff + P.n - 1] = x[P.xoff + P.n - 1] + psi1;
}

// --------------------------------------------------------------------------
// Anderson mixing on the device (AndersonMixing.jl:56-169)
// --------------------------------------------------------------------------
__global__ void k_and_residual(const DevPoint* __restrict__ pts, int32_t* __restrict__ evalflag_unused,
                               const DevSolveState* __restrict__ stt, const double* __restrict__ x,
                               const double* __restrict__ fx, double* __restrict__ Xh, double* __restrict__ Gh,
                               unsigned long long* __restrict__ errbits, double* __restrict__ gram, int m,
                               int32_t* __restrict__ n_active) {
    const int p = blockIdx.y;
    if (blockIdx.x == 0 && p == 0 && threadIdx.x == 0) *n_active = 0;
    const DevSolveState& S = stt[p];
    if (S.converged) return;
    const DevPoint& P = pts[p];
    const int col = S.iter % m;
    const int64_t hoff = P.xoff * m;
    double mx = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += (int64_t)gridDim.x * blockDim.x) {
        const double xi = x[P.xoff + i];
        const double g = fx[P.xoff + i] - xi;
        Xh[hoff + (int64_t)col * P.n + i] = xi;
        Gh[hoff + (int64_t)col * P.n + i] = g;
        const double a = fabs(g);
        mx = (a > mx || a != a) ? a : mx;  // NaN propagates
    }
    __shared__ double red[4];
    mx = warp_max(mx);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v = nanmax(v, red[w]);
        if (v != v) v = INFINITY;
        atomicMax(&errbits[p], (unsigned long long)__double_as_longlong(v));
    }
}

__device__ __forceinline__ int hist_len(const DevSolveState& S, int m) {
    const int it = S.iter + 1;          // 1-based index of this iteration
    const int local = it - S.hist_base; // iterations in the current history
    return local < m ? local : m;
}

__global__ void k_and_gram(const DevPoint* __restrict__ pts, const DevSolveState* __restrict__ stt,
                           const double* __restrict__ Xh, const double* __restrict__ Gh, double* __restrict__ gram, int m) {
    const int p = blockIdx.y;
    const DevSolveState& S = stt[p];
    if (S.converged) return;
    const int hl = hist_len(S, m);
    const int nc = hl - 1;
    if (nc < 1) return;
    const DevPoint& P = pts[p];
    const int col = S.iter % m;
    const int64_t hoff = P.xoff * m;
    const double* g = Gh + hoff + (int64_t)col * P.n;
    double acc[SDFT_MAX_M * (SDFT_MAX_M + 1) / 2 + SDFT_MAX_M];
    const int ntri = nc * (nc + 1) / 2;
    for (int k = 0; k < ntri + nc; ++k) acc[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += (int64_t)gridDim.x * blockDim.x) {
        double dg[SDFT_MAX_M];
        const double gi = g[i];
        for (int a = 0; a < nc; ++a) {
            const int ca = ((col - 1 - a) % m + m) % m;
            dg[a] = Gh[hoff + (int64_t)ca * P.n + i] - gi;
        }
        int k = 0;
        for (int a = 0; a < nc; ++a)
            for (int b = a; b < nc; ++b) acc[k++] += dg[a] * dg[b];
        for (int a = 0; a < nc; ++a) acc[ntri + a] += dg[a] * gi;
    }
    __shared__ double red[SDFT_MAX_M * (SDFT_MAX_M + 1) / 2 + SDFT_MAX_M][4];
    for (int k = 0; k < ntri + nc; ++k) {
        const double v = warp_sum(acc[k]);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < ntri + nc) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[threadIdx.x][w];
        gram[((int64_t)p * SDFT_GRAM_BLOCKS + blockIdx.x) * SDFT_GRAM_LEN + threadIdx.x] = s;
    }
}

// minimum-norm least squares through the normal equations with a Jacobi
// eigen-decomposition and a relative cutoff (the reference calls lsqr!)
__device__ void small_lstsq(int nc, const double* tri, const double* rhs, double* lam) {
    double A[SDFT_MAX_M][SDFT_MAX_M], V[SDFT_MAX_M][SDFT_MAX_M];
    int k = 0;
    for (int a = 0; a < nc; ++a)
        for (int b = a; b < nc; ++b) { A[a][b] = tri[k]; A[b][a] = tri[k]; ++k; }
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) V[a][b] = a == b ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0;
        for (int a = 0; a < nc; ++a)
            for (int b = a + 1; b < nc; ++b) off += A[a][b] * A[a][b];
        if (off < 1e-300) break;
        for (int a = 0; a < nc; ++a) {
            for (int b = a + 1; b < nc; ++b) {
                if (fabs(A[a][b]) < 1e-300) continue;
                const double th = (A[b][b] - A[a][a]) / (2.0 * A[a][b]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < nc; ++r) {
                    const double ar = A[r][a], br = A[r][b];
                    A[r][a] = c * ar - s * br;
                    A[r][b] = s * ar + c * br;
                }
                for (int r = 0; r < nc; ++r) {
                    const double ar = A[a][r], br = A[b][r];
                    A[a][r] = c * ar - s * br;
                    A[b][r] = s * ar + c * br;
                }
                for (int r = 0; r < nc; ++r) {
                    const double vr = V[r][a], wr = V[r][b];
                    V[r][a] = c * vr - s * wr;
                    V[r][b] = s * vr + c * wr;
                }
            }
        }
    }
    double emax = 0.0;
    for (int a = 0; a < nc; ++a) emax = fmax(emax, fabs(A[a][a]));
    for (int a = 0; a < nc; ++a) lam[a] = 0.0;
    for (int e = 0; e < nc; ++e) {
        const double ev = A[e][e];
        if (!(ev > 1e-14 * emax)) continue;
        double proj = 0.0;
        for (int r = 0; r < nc; ++r) proj += V[r][e] * rhs[r];
        proj /= ev;
        for (int r = 0; r < nc; ++r) lam[r] -= V[r][e] * proj;  // solves M lam = -rhs
    }
}

__global__ void k_and_solve(DevSolveState* __restrict__ stt, int32_t* __restrict__ evalflag,
                            unsigned long long* __restrict__ errbits, const double* __restrict__ gram,
                            SolveParams sp, int P_count, int32_t* __restrict__ n_active, const int32_t* __restrict__ errflag) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_count) return;
    DevSolveState S = stt[p];
    if (S.converged) {
        S.mode = -1;
        evalflag[p] = 0;
        stt[p] = S;
        return;
    }
    const int m = sp.m;
    double err = __longlong_as_double((long long)errbits[p]);
    errbits[p] = 0ull;
    if (errflag[p] != 0) err = INFINITY;
    err = fmax(err, 1e-20);
    const double amix = fmin(sp.mixing_max, sp.damping / err);
    const int it = S.iter + 1;
    const int local = it - S.hist_base;
    S.err = err;
    S.amix = amix;
    if (!(err < INFINITY)) {  // non-finite residual: stop this point
        S.failed = 2;
        S.converged = 1;
        S.mode = -1;
        S.iter = it;
        evalflag[p] = 0;
        stt[p] = S;
        return;
    }
    bool picard = (local < m) || (err > sp.switch_tole && S.use_picard);
    if (S.restarted && local >= 2 && !(err > sp.switch_tole && S.use_picard)) picard = false;
    if (picard) {
        S.mode = 0;
        S.ncols = 0;
    } else {
        S.use_picard = 0;
        const int nc = (local < m ? local : m) - 1;
        const int ntri = nc * (nc + 1) / 2;
        double gp[SDFT_GRAM_LEN];
        for (int k = 0; k < ntri + nc; ++k) {
            double t = 0.0;
            for (int b = 0; b < SDFT_GRAM_BLOCKS; ++b) t += gram[((int64_t)p * SDFT_GRAM_BLOCKS + b) * SDFT_GRAM_LEN + k];
            gp[k] = t;
        }
        double rhs[SDFT_MAX_M];
        for (int a = 0; a < nc; ++a) rhs[a] = gp[ntri + a];
        small_lstsq(nc, gp, rhs, S.lam);
        S.mode = 1;
        S.ncols = nc;
    }
    S.iter = it;
    if (err < sp.tole) {
        S.converged = 1;
    } else if (it >= sp.max_iters) {
        S.converged = 1;
        S.failed = 1;
    }
    if (!S.converged && sp.mixed && S.prec == 32 && err < sp.fp32_switch) {
        S.prec = 64;
        S.switch_iter = it;
        S.hist_base = it;
        S.restarted = 1;
    }
    evalflag[p] = S.converged ? 0 : S.prec;
    if (!S.converged) atomicAdd(n_active, 1);
    stt[p] = S;
}

__global__ void k_and_update(const DevPoint* __restrict__ pts, const DevSolveState* __restrict__ stt,
                             double* __restrict__ x, const double* __restrict__ fx, const double* __restrict__ Xh,
                             const double* __restrict__ Gh, int m) {
    const int p = blockIdx.y;
    const DevSolveState& S = stt[p];
    if (S.mode < 0) return;
    const DevPoint& P = pts[p];
    const int64_t hoff = P.xoff * m;
    const int col = (S.iter - 1) % m;  // column written by this iteration
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += (int64_t)gridDim.x * blockDim.x) {
        const double xi = Xh[hoff + (int64_t)col * P.n + i];
        const double gi = Gh[hoff + (int64_t)col * P.n + i];
        double nx;
        if (S.mode == 0) {
            nx = (1.0 - S.amix) * xi + S.amix * fx[P.xoff + i];
        } else {
            double dx = 0.0, dg = 0.0;
            for (int a = 0; a < S.ncols; ++a) {
                const int ca = ((col - 1 - a) % m + m) % m;
                dx += (Xh[hoff + (int64_t)ca * P.n + i] - xi) * S.lam[a];
                dg += (Gh[hoff + (int64_t)ca * P.n + i] - gi) * S.lam[a];
            }
            nx = xi + dx + S.amix * (gi + dg);
        }
        x[P.xoff + i] = nx;
    }
}

// --------------------------------------------------------------------------
// observables
// --------------------------------------------------------------------------
// grand potential integrand on the primary points (grand_potential.jl:4-28):
//   f_ex(K) - sum_j rho_j mu_ex_j trapez_j + 1/2 sum rho_bond lng
template <typename T>
__global__ void __launch_bounds__(256) k_omega(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                              const DevSolveState* __restrict__ stt,
                                              const T* __restrict__ nw, int Fnw, const T* __restrict__ rhoT, int Frt,
                                              const T* __restrict__ rho_bonds, const double* __restrict__ x,
                                              double* __restrict__ omega_acc) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.y;
    if ((stt[p].prec == 64 ? 64 : 32) != (sizeof(T) == 8 ? 64 : 32)) return;
    const DevPoint& P = pts[p];
    const int NB = sys.NB, MB = sys.MB, NP = P.NP;
    double acc = 0.0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < NP; n += gridDim.x * blockDim.x) {
        double n2j[SDFT_MAX_NB], n3j[SDFT_MAX_NB], nVj[SDFT_MAX_NB], rbond[2 * SDFT_MAX_MB];
        load_point_inputs(sys, P, n, nw, Fnw, n2j, n3j, nVj, rbond);
        const int F = sys.functionals;
        double f = 0.0;
        if (F & (SDFT_F_MFMT | SDFT_F_TPT1MFMT)) {
            const FmtIn fi = fmt_collect(n2j, n3j, nVj, sys);
            f += fmt_phi(fi);
            if (F & SDFT_F_TPT1MFMT) {
                const double Xi = 1.0 - fi.nV2 * fi.nV2 / (fi.n2 * fi.n2);  // no max(., 0) in the free energy
                const double i1 = 1.0 / (1.0 - fi.n3);
                const double t2 = fi.n2 * Xi / 4.0 * i1 * i1;
                const double t3 = fi.n2 * fi.n2 * Xi / 72.0 * i1 * i1 * i1;
                for (int u = 0; u < MB; ++u) {
                    const double si = sys.sigma[sys.bond_i[u]], sj = sys.sigma[sys.bond_j[u]];
                    const double sij = (si + sj) / 2.0, tau = si * sj / sij;
                    const double y = i1 + t2 * tau + t3 * tau * tau;
                    f -= 0.5 * (rbond[2 * u] + rbond[2 * u + 1]) * log(y);
                }
            }
        }
        if (F & (SDFT_F_MSA | SDFT_F_TPT1MSA)) {
            double rho[SDFT_MAX_NB];
            for (int j = 0; j < NB; ++j) rho[j] = n2j[j] / (M_PI * sys.sigma[j] * sys.sigma[j]);
            const GammaOut g = gamma_msa(rho, sys);
            double fe = g.G * g.G * g.G / (3.0 * M_PI);
            for (int j = 0; j < NB; ++j) {
                const double s = sys.sigma[j], z = sys.z[j];
                fe -= sys.lB * rho[j] * z * (z * g.G + g.eta * s) / (1.0 + g.G * s);
            }
            if (F & SDFT_F_TPT1MSA) {
                for (int u = 0; u < MB; ++u) {
                    const int i = sys.bond_i[u], j = sys.bond_j[u];
                    const double sij = (sys.sigma[i] + sys.sigma[j]) / 2.0;
                    const double Xi = (sys.z[i] - g.eta * sys.sigma[i] * sys.sigma[i]) / (1.0 + g.G * sys.sigma[i]);
                    const double Xj = (sys.z[j] - g.eta * sys.sigma[j] * sys.sigma[j]) / (1.0 + g.G * sys.sigma[j]);
                    fe -= 0.5 * (rbond[2 * u] + rbond[2 * u + 1]) * (-sys.lB / sij * Xi * Xj);
                }
            }
            f += fe;
        }
        const double* xp = x + P.xoff;
        for (int j = 0; j < NB; ++j) f -= (double)rhoT[fidx(P, Frt, j, n)] * xp[(int64_t)j * NP + n];
        for (int k = 0; k < 2 * MB; ++k) f += (double)rho_bonds[fidx(P, 2 * MB, k, n)] * xp[(int64_t)(NB + k) * NP + n] / 2.0;
        acc += f;
    }
    __shared__ double red[4];
    acc = warp_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
        omega_acc[(int64_t)p * gridDim.x + blockIdx.x] = s;
    }
}

// per species / segment / state raw sums of rho_seg (degree_of_dissociation.jl:33-63)
template <typename T>
__global__ void __launch_bounds__(256) k_segment_sums(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                     const DevSolveState* __restrict__ stt,
                                                     const DevSpecies* __restrict__ spec, const int64_t* __restrict__ goffs, int GF,
                                                     const int32_t* __restrict__ seg_states, const double* __restrict__ maxd_tab,
                                                     const T* __restrict__ elam, const T* __restrict__ gC, const T* __restrict__ gP,
                                                     const double* __restrict__ fixed_lambda, int u, double* __restrict__ out,
                                                     int64_t out_stride) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.y;
    if ((stt[p].prec == 64 ? 64 : 32) != (sizeof(T) == 8 ? 64 : 32)) return;
    const DevPoint& P = pts[p];
    const DevSpecies sp = spec[u];
    const int q = blockIdx.x;  // segment
    const int S = sp.S, N = sp.nseg, NP = P.NP, NB = sys.NB;
    const int* st = seg_states + sp.state_off;
    const double* maxd = maxd_tab + sp.maxd_off;
    const int Kf = sp.fixed_seg >= 0 ? P.fixed_coord[u] : -1;
    double norm = sp.fixed_seg >= 0 ? sp.fixed_density / fixed_lambda[(int64_t)p * SDFT_MAX_SPECIES + u] : P.chain_norm[u];
    const double pref = norm * exp(P.mu_species[u]);
    auto gidx = [&](int seg, int b, int a, int nn) -> int64_t {
        return P.base * (int64_t)GF + (goffs[u] + ((int64_t)seg * S + b) * S + a) * NP + nn;
    };
    __shared__ double red[4];
    for (int a = 0; a < S; ++a) {
        const int si = st[q * S + a];
        double acc = 0.0;
        if (si >= 0) {
            for (int n = threadIdx.x; n < NP; n += blockDim.x) {
                double Gp = 0.0, Qc = 0.0;
                if (q == 0) Gp = (double)base_value<T>(0, n, sp.fixed_seg, Kf, P.dz, maxd);
                else
                    for (int b = 0; b < S; ++b)
                        if (st[(q - 1) * S + b] >= 0) Gp += (double)gP[gidx(q, b, a, n)];
                if (q == N - 1) Qc = (double)base_value<T>(q, n, sp.fixed_seg, Kf, P.dz, maxd);
                else
                    for (int b = 0; b < S; ++b)
                        if (st[(q + 1) * S + b] >= 0) Qc += (double)gC[gidx(q, b, a, n)];
                acc += pref * (double)elam[fidx(P, NB, si, n)] * Gp * Qc;
            }
        }
        acc = warp_sum(acc);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
            out[(int64_t)p * out_stride + (int64_t)q * S + a] = si >= 0 ? s : 0.0;
        }
    }
}

__global__ void k_set_evalflag(int32_t* evalflag, int value, int P_count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P_count) evalflag[p] = value;
}

__global__ void k_zero_i32(int32_t* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}
__global__ void k_zero_f64(double* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}
__global__ void k_convert_f32(const double* a, float* b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// ==========================================================================
// host launchers
// ==========================================================================
static inline dim3 grid2(int64_t n, int tb, int y, int z = 1) {
    return dim3((unsigned)((n + tb - 1) / tb), (unsigned)y, (unsigned)z);
}

void launch_build_kernels(hipStream_t st, const DevSystem* sys, int K, int nsets, int maxL, const int32_t* set_L,
                          const int64_t* set_kbase, const double* set_dz, double* wbuf, double* c64, float* c32) {
    hipLaunchKernelGGL(k_build_weights, grid2(maxL, 256, K, nsets), dim3(256), 0, st, sys, set_L, set_kbase, set_dz, wbuf);
    hipLaunchKernelGGL(k_build_cvec, grid2(maxL, 256, K, nsets), dim3(256), 0, st, sys, set_L, set_kbase, wbuf, c64, c32);
}

template <typename T>
void Pipeline<T>::objective(hipStream_t st, const Ctx& c, int prec, const double* x, double* fx) {
    const int P = c.P;
    const int maxNP = c.maxNP;
    hipLaunchKernelGGL(k_prepare<T>, grid2(maxNP, 256, P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag, prec, x,
                       c.d_Ext, elam, hl);
    for (int u = 0; u < c.nsp; ++u) {
        const SpeciesHost& sh = c.species[u];
        PropArgs A;
        A.u = u; A.nseg = sh.nseg; A.S = sh.S; A.fixed_seg = sh.fixed_seg;
        A.state_off = sh.state_off; A.maxd_off = sh.maxd_off; A.goff = sh.goff; A.GF = c.GF;
        const bool banded = (sizeof(T) == 4) ? c.band_fp32 != 0 : c.band_fp64 != 0;
        A.band = banded ? 1 : -1;
        const bool gemm = c.use_gemm && c.prop_plan[u].ntiles > 0;
        for (int s = 1; s < sh.nseg; ++s) {
            A.s = s;
            // sweeps whose parent is the grafted segment use the direct real-space sums
            const bool fix0 = sh.fixed_seg == 0 && s == sh.nseg - 1;  // child sweep, parent = segment 0
            const bool fix1 = sh.fixed_seg == 0 && s == 1;            // parent sweep, parent = segment 0
            if (!gemm) {
                hipLaunchKernelGGL((k_prop_step<T, 256>), grid2(maxNP, 256, P, 2), dim3(256), 0, st, c.d_sys,
                                   c.d_pts, c.d_evalflag, prec, A, c.d_seg_states, c.d_maxd, kern, elam, hl, trapez,
                                   gC, gP, c.d_band, 0);
                continue;
            }
            const int skip = (fix0 && fix1) ? 2 : fix0 ? 0 : fix1 ? 1 : -1;
            if (skip != 2)
                launch_opgemm_prop<T>(st, c, prec, c.prop_plan[u], u, s, skip, banded, kern, elam, hl, trapez, gC, gP,
                                      fbuf);
            if (skip >= 0)
                hipLaunchKernelGGL((k_prop_step<T, 256>), grid2(maxNP, 256, P, skip == 2 ? 2 : 1), dim3(256), 0, st,
                                   c.d_sys, c.d_pts, c.d_evalflag, prec, A, c.d_seg_states, c.d_maxd, kern, elam, hl,
                                   trapez, gC, gP, c.d_band, skip == 2 ? 0 : skip);
        }
        if (sh.fixed_seg >= 0) {
            hipLaunchKernelGGL(k_fixed_lambda<T>, dim3(P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag, prec, A,
                               c.d_seg_states, c.d_maxd, elam, trapez, gC, gP, c.d_fixed_lambda);
        }
    }
    hipLaunchKernelGGL((k_accumulate<T, 256>), grid2(maxNP, 256, P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag,
                       prec, c.d_spec, c.d_goffs, c.GF, c.d_seg_states, c.d_maxd, elam, trapez, gC, gP,
                       c.d_fixed_lambda, rho_beads, rho_bonds, rhoT, c.d_acc_part, c.d_errflag);
    if (c.use_gemm) launch_opgemm_conv<T>(st, c, prec, c.fwd_plan, kern, rhoT, dphi, nw, fx);
    else if (c.n_fwd_jobs > 0)
        hipLaunchKernelGGL((k_conv<T, 256>), grid2(maxNP, 256, c.n_fwd_jobs, P), dim3(256), 0, st, c.d_sys, c.d_pts,
                           c.d_evalflag, prec, c.d_jobs, kern, rhoT, c.Frt, dphi, c.Fdp, nw, c.Fnw, fx);
    hipLaunchKernelGGL(k_pointwise<T>, grid2(maxNP, 256, P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag, prec,
                       nw, c.Fnw, rhoT, c.Frt, c.d_surf, dphi, c.Fdp);
    if (c.use_gemm) launch_opgemm_conv<T>(st, c, prec, c.back_plan, kern, rhoT, dphi, nw, fx);
    else
        hipLaunchKernelGGL((k_conv<T, 256>), grid2(maxNP, 256, c.n_back_jobs, P), dim3(256), 0, st, c.d_sys, c.d_pts,
                           c.d_evalflag, prec, c.d_jobs + c.n_fwd_jobs, kern, rhoT, c.Frt, dphi, c.Fdp, nw, c.Fnw, fx);
    hipLaunchKernelGGL(k_neutrality, dim3((P + 63) / 64), dim3(64), 0, st, c.d_sys, c.d_pts, c.d_spec, c.d_evalflag,
                       prec, c.d_acc_part, c.d_acc_species, x, fx, P);
}

template <typename T>
void Pipeline<T>::launch_prop_step(hipStream_t st, const Ctx& c, int prec, int u, int s, bool banded) {
    const SpeciesHost& sh = c.species[u];
    PropArgs A;
    A.u = u; A.nseg = sh.nseg; A.S = sh.S; A.fixed_seg = sh.fixed_seg;
    A.state_off = sh.state_off; A.maxd_off = sh.maxd_off; A.goff = sh.goff; A.GF = c.GF;
    A.band = banded ? 1 : -1;
    A.s = s;
    if (c.use_gemm && c.prop_plan[u].ntiles > 0 && sh.fixed_seg < 0)
        launch_opgemm_prop<T>(st, c, prec, c.prop_plan[u], u, s, -1, banded, kern, elam, hl, trapez, gC, gP,
                              fbuf);
    else
        hipLaunchKernelGGL((k_prop_step<T, 256>), grid2(c.maxNP, 256, c.P, 2), dim3(256), 0, st, c.d_sys, c.d_pts,
                           c.d_evalflag, prec, A, c.d_seg_states, c.d_maxd, kern, elam, hl, trapez, gC, gP, c.d_band, 0);
}

template <typename T>
void Pipeline<T>::launch_conv(hipStream_t st, const Ctx& c, int prec, bool back, double* fx) {
    const int nj = back ? c.n_back_jobs : c.n_fwd_jobs;
    if (nj == 0) return;
    if (c.use_gemm) {
        launch_opgemm_conv<T>(st, c, prec, back ? c.back_plan : c.fwd_plan, kern, rhoT, dphi, nw, fx);
        return;
    }
    hipLaunchKernelGGL((k_conv<T, 256>), grid2(c.maxNP, 256, nj, c.P), dim3(256), 0, st, c.d_sys, c.d_pts,
                       c.d_evalflag, prec, c.d_jobs + (back ? c.n_fwd_jobs : 0), kern, rhoT, c.Frt, dphi, c.Fdp, nw,
                       c.Fnw, fx);
}

template <typename T>
void Pipeline<T>::launch_pointwise(hipStream_t st, const Ctx& c, int prec) {
    hipLaunchKernelGGL(k_pointwise<T>, grid2(c.maxNP, 256, c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag,
                       prec, nw, c.Fnw, rhoT, c.Frt, c.d_surf, dphi, c.Fdp);
}

template <typename T>
void Pipeline<T>::omega(hipStream_t st, const Ctx& c, const double* x, double* omega_acc) {
    hipLaunchKernelGGL(k_omega<T>, dim3(SDFT_OMEGA_BLOCKS, c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_state, nw, c.Fnw, rhoT, c.Frt,
                       rho_bonds, x, omega_acc);
}

template <typename T>
void Pipeline<T>::segment_sums(hipStream_t st, const Ctx& c, int u, double* out, int64_t stride) {
    const SpeciesHost& sh = c.species[u];
    hipLaunchKernelGGL(k_segment_sums<T>, dim3(sh.nseg, c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_state, c.d_spec,
                       c.d_goffs, c.GF, c.d_seg_states, c.d_maxd, elam, gC, gP, c.d_fixed_lambda, u, out, stride);
}

template struct Pipeline<float>;
template struct Pipeline<double>;

void launch_anderson(hipStream_t st, const Ctx& c, const SolveParams& sp) {
    const int P = c.P;
    const int tiles = SDFT_GRAM_BLOCKS;  // fixed per point: the reductions do not depend on the batch
    hipLaunchKernelGGL(k_and_residual, dim3(tiles, P), dim3(256), 0, st, c.d_pts, c.d_evalflag, c.d_state, c.d_x, c.d_fx,
                       c.d_Xh, c.d_Gh, c.d_errbits, c.d_gram, sp.m, c.d_n_active);
    hipLaunchKernelGGL(k_and_gram, dim3(tiles, P), dim3(256), 0, st, c.d_pts, c.d_state, c.d_Xh, c.d_Gh, c.d_gram, sp.m);
    hipLaunchKernelGGL(k_and_solve, dim3((P + 63) / 64), dim3(64), 0, st, c.d_state, c.d_evalflag, c.d_errbits, c.d_gram,
                       sp, P, c.d_n_active, c.d_errflag);
    hipLaunchKernelGGL(k_and_update, dim3(tiles, P), dim3(256), 0, st, c.d_pts, c.d_state, c.d_x, c.d_fx, c.d_Xh, c.d_Gh,
                       sp.m);
}

void launch_set_evalflag(hipStream_t st, int32_t* evalflag, int value, int P) {
    hipLaunchKernelGGL(k_set_evalflag, dim3((P + 255) / 256), dim3(256), 0, st, evalflag, value, P);
}
void launch_zero_i32(hipStream_t st, int32_t* a, int64_t n) {
    if (n > 0) hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n);
}
void launch_zero_f64(hipStream_t st, double* a, int64_t n) {
    if (n > 0) hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n);
}
void launch_convert_f32(hipStream_t st, const double* a, float* b, int64_t n) {
    if (n > 0) hipLaunchKernelGGL(k_convert_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, n);
}

}  // namespace sdft
