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
#include <type_traits>
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
    } else if (k == 3 * NB + MB) {  // coul_func_fft!: 2 pi lB / ((1 - cos(k dz)) / dz^2), k = 0 -> 0
        if (m == 0) return 0.0;
        return 2.0 * M_PI * sys.lB / ((1.0 - cos(kv * dz)) / (dz * dz));
    } else {  // square well (square_well.jl:37-50): eps (step(lambda sigma_ij) - step(sigma_ij))
        const int u = (k - sys.sw_kbase) >> 1;
        const double sij = (sys.sigma[sys.sw_i[u]] + sys.sigma[sys.sw_j[u]]) / 2.0;
        auto step = [&](double b) {
            return (m == 0) ? (4.0 / 3.0) * M_PI * b * b * b
                            : 4.0 * M_PI * (sin(b * a) - b * a * cos(b * a)) / (a * a * a);
        };
        return sys.sw_eps[u] * (step(sys.sw_lam[u] * sij) - step(sij));
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
// square-well mu kernels: (c[d] + c[d-1]) / 2, the real part of the k-space product on the
// half-sample mirrored grid (x_ext[-n] = x_ext[n-1])
__global__ void __launch_bounds__(256) k_build_cvec(const DevSystem* __restrict__ psys, const int32_t* __restrict__ set_L,
                                                    const int64_t* __restrict__ set_kbase, const double* __restrict__ wbuf,
                                                    double* __restrict__ c64, float* __restrict__ c32) {
    const DevSystem& sys = *psys;
    const int s = blockIdx.z, k = blockIdx.y;
    const int L = set_L[s];
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const bool odd = (k < 3 * sys.NB) && (k % 3 == 2);
    const bool halfshift = sys.n_sw > 0 && k >= sys.sw_kbase && ((k - sys.sw_kbase) & 1) == 0;
    const int dm = d > 0 ? d - 1 : L - 1;
    const double* w = wbuf + set_kbase[s] * sys.K + (int64_t)k * L;
    __shared__ double ws[256];
    double acc = 0.0, acc2 = 0.0;
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
                if (halfshift) {
                    const long long r2 = ((long long)m * dm) % L;
                    acc2 += ws[i] * cospi(2.0 * (double)r2 / (double)L);
                }
            }
        }
    }
    if (d < L) {
        if (halfshift) acc = 0.5 * (acc + acc2);
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



template <typename T, int TB, int SM>
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
        return P.gbase + (A.goff + ((int64_t)seg * S + b) * S + a) * P.gstride + nn;
    };
    // input scale: 1 / the previous step's largest output (per-step rescaling)
    const T qinv = (A.rescale && in_stored)
                       ? (T)(1.0 / scale_read(A.smax, scale_slot(p, A.u, sweep, A.s - 1, A.nsp, A.nmax)))
                       : T(1);
    // Q(n') = sum_b' g_in[inseg][b'][b](n')  or base
    auto qin = [&](int b, int nn) -> T {
        if (!in_stored) return base_value<T>(inseg, nn, A.fixed_seg, Kf, P, maxd);
        T q = 0;
        for (int b2 = 0; b2 < S; ++b2)
            if (st[inseg * S + b2] >= 0) q += gin[gidx(inseg, b2, b, nn)];
        return q * qinv;
    };
    const int64_t oslot = scale_slot(p, A.u, sweep, A.s, A.nsp, A.nmax);

    int npair = 0;
    int pa[SM * SM], pb[SM * SM], pbond[SM * SM];
    int ppi[SM * SM], ppj[SM * SM];
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
            if (A.rescale) scale_note(A.smax, oslot, (double)(T)val);
        }
        return;
    }

    // convolution path
    __shared__ T fs[SM * SM][TB];
    __shared__ T cm[SM * SM][2 * TB];
    __shared__ T cp[SM * SM][2 * TB];
    T acc[SM * SM];
    for (int q = 0; q < SM * SM; ++q) acc[q] = T(0);
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
    double vmax = 0.0;
    if (n < NP) {
        for (int q = 0; q < npair; ++q) {
            const int a = pa[q], b = pb[q], bond = pbond[q];
            T v = absT(acc[q]) * hl[fidx(P, 2 * MB, 2 * bond + ppi[q], n)] / (T)normative;
            if (A.fixed_seg >= 0) {
                const double d = (double)(n - Kf) * dz;
                if (d * d > maxd[outseg] * maxd[outseg]) v = T(0);
            }
            gout[gidx(outseg, b, a, n)] = v;
            vmax = (double)v > vmax ? (double)v : vmax;
        }
    }
    if (A.rescale) {
        vmax = warp_max(vmax);
        if ((t & 63) == 0) scale_note(A.smax, oslot, vmax);
    }
}

// --------------------------------------------------------------------------
// per-step rescaling: E(q) = LC(q) + LP(q), the summed logarithms of the input factors divided
// out on the way to gC[q] (child sweep step N-1-q) and gP[q] (parent sweep step q); a step's
// stored output carries the factors of all earlier steps of its sweep. Slot nmax: E(1), the
// grafted chain's normalisation segment.
// --------------------------------------------------------------------------
__global__ void k_scale_prefix(const int32_t* __restrict__ evalflag, int prec, int P, int u, int N, int fixed_seg,
                               const unsigned long long* __restrict__ smax, double* __restrict__ lsc, int nsp,
                               int nmax) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || evalflag[p] != prec) return;
    double* E = lsc + ((int64_t)p * nsp + u) * (nmax + 1);
    double acc = 0.0;
    for (int q = 0; q < N; ++q) {  // parent sweep: gP[q] from step q, inputs divided by steps 1..q-1
        if (q >= 2) acc += log(scale_read(smax, scale_slot(p, u, 1, q - 1, nsp, nmax)));
        E[q] = acc;
    }
    acc = 0.0;
    for (int q = N - 1; q >= 0; --q) {  // child sweep: gC[q] from step N-1-q
        const int s = N - 1 - q;
        if (s >= 2) acc += log(scale_read(smax, scale_slot(p, u, 0, s - 1, nsp, nmax)));
        E[q] += acc;
    }
    E[nmax] = (fixed_seg == 0 && N > 1) ? E[1] : 0.0;
}

// log of the factor restoring the unscaled product Gp(q) Qc(q) relative to the species' norm: for a
// grafted chain norm = sigma / lambda with lambda from the stored (scaled) propagators at segment 1,
// so the factor is exp(E(q) - E(1)); a free chain's norm normative^(N-1) is folded in as
// (N-1) log normative + E(q) (the power alone can leave the double range)
__device__ __forceinline__ double seg_log_scale(const DevSpecies& sp, const DevPoint& P, int p, int u, int q,
                                                const double* __restrict__ lsc, int nsp, int nmax) {
    const double* E = scale_row(lsc, p, u, nsp, nmax);
    if (sp.fixed_seg >= 0) return E[q] - E[nmax];
    return (double)(sp.nseg - 1) * log(P.normative[u]) + E[q];
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
        return P.gbase + (A.goff + ((int64_t)seg * S + b) * S + a) * P.gstride + nn;
    };
    double acc = 0.0;
    for (int n = threadIdx.x; n < NP; n += blockDim.x) {
        for (int a = 0; a < S; ++a) {
            const int si = st[q * S + a];
            if (si < 0) continue;
            double Gp = 0.0;
            if (q == 0) Gp = (double)base_value<T>(0, n, A.fixed_seg, Kf, P, maxd);
            else
                for (int b = 0; b < S; ++b)
                    if (st[(q - 1) * S + b] >= 0) Gp += (double)gP[gidx(q, b, a, n)];
            double Qc = 0.0;
            if (q == N - 1) Qc = (a == 0) ? (double)base_value<T>(q, n, A.fixed_seg, Kf, P, maxd) : 0.0;
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

// grafted monomerbead normalisation (monomerbead/handle_fixed.jl:25-48): lambda = sum over the fixed
// segment's states of exp(-mu_ex - Ext - delta_muH) at the grafting point (no electrostatic term),
// from the solution vector in FP64 at either precision; norm = sigma_graft / lambda
__global__ void k_fixed_lambda_mb(const DevPoint* __restrict__ pts, const int32_t* __restrict__ evalflag, int prec,
                                  int P_count, int u, const int32_t* __restrict__ states, int S,
                                  const double* __restrict__ x, const double* __restrict__ Ext, int NB,
                                  double* __restrict__ fixed_lambda) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_count || evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int K = P.fixed_coord[u];
    double lam = 0.0;
    for (int a = 0; a < S; ++a) {
        const int si = states[a];   // the grafted segment is segment 0
        if (si < 0) continue;
        lam += exp(-x[P.xoff + (int64_t)si * P.NP + K] - Ext[P.base * NB + (int64_t)si * P.NP + K] - P.delta_muH[si]);
    }
    fixed_lambda[(int64_t)p * SDFT_MAX_SPECIES + u] = lam;
}

constexpr int ACC_POINTS = SDFT_ACC_POINTS;  // grid points per k_accumulate block (one wavefront wide)
constexpr int ACC_GROUPS = 4;   // wavefronts per block, each taking every 4th segment
static_assert(ACC_POINTS * ACC_GROUPS == 256, "k_accumulate runs 256-thread blocks");

// --------------------------------------------------------------------------
// densities (accumulate_densities!, accumulate_densities.jl:95-184; update_spatial_structures!)
// rho_seg(q,a) = norm e^{mu} exp(-lambda_si) Gp(q,a) Qc(q,a)
// bond end 0 (parent end):  += norm e^{mu} exp(-lambda_si) Gp(q,a) gC[q][b][a]
// bond end 1 (child end):   += norm e^{mu} exp(-lambda_si) Qc(q,a) gP[q][b][a]
// Also per-species totals for the neutrality shift (neutrality_shift.jl:20-47):
//   acc[u][0] = sum_seg,state num / N_u,  acc[u][1] = sum num * z
// --------------------------------------------------------------------------
// The FP64 variant is held to 4 waves per SIMD (128 VGPRs, a 12-byte spill): the kernel is bound by
// load latency, and the extra wave made the full-load launch 156 -> 133 us. The FP32 variant reaches
// 4 waves without a bound (held to 5 it spills 84 bytes per lane).
template <typename T, int TB, int SM>
__global__ void __launch_bounds__(TB, sizeof(T) == 4 ? 1 : 4) k_accumulate(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                   const int32_t* __restrict__ evalflag, int prec,
                                                   const DevSpecies* __restrict__ spec, const int64_t* __restrict__ goffs, int GF,
                                                   const int32_t* __restrict__ seg_states, const double* __restrict__ maxd_tab,
                                                   const T* __restrict__ elam, const T* __restrict__ trapez,
                                                   const T* __restrict__ gC, const T* __restrict__ gP,
                                                   const double* __restrict__ fixed_lambda,
                                                   T* __restrict__ rho_beads, T* __restrict__ rho_bonds, T* __restrict__ rhoT,
                                                   int Frt, double* __restrict__ acc_part, int32_t* __restrict__ errflag,
                                                   const double* __restrict__ lsc, int nmax, int nsp) {
    // A block covers ACC_POINTS grid points; its ACC_GROUPS wavefronts split the segments of every
    // species, dealt round-robin continuing from one species to the next (so single-bead species do
    // not all fall to group 0), which keeps more independent loads in flight per point; the groups'
    // partial densities are added in group order at the end.
    // The point's, the system's and the species' descriptors (nsp of them) are copied to LDS first:
    // read through the scalar cache, their many fields (a different point in every block) cost a
    // chain of dependent global round trips per species and segment. The words are read in one
    // pass, issued together with the point's evaluation flag, and written to LDS after its test.
    const int p = blockIdx.y;
    const int ef = evalflag[p];
    __shared__ DevPoint Ps;
    __shared__ DevSystem Ss;
    __shared__ DevSpecies Us[SDFT_MAX_SPECIES];
    {
        static_assert(sizeof(DevPoint) % 4 == 0 && sizeof(DevSystem) % 4 == 0 && sizeof(DevSpecies) % 4 == 0,
                      "descriptor copies move 4-byte words");
        constexpr int WP = (int)(sizeof(DevPoint) / 4), WS = (int)(sizeof(DevSystem) / 4);
        constexpr int WU = (int)(sizeof(DevSpecies) / 4), WR = 3;
        static_assert(WP + WS + SDFT_MAX_SPECIES * WU <= WR * TB, "descriptor words per thread");
        const int wtot = WP + WS + nsp * WU;
        const int32_t* sp_src = reinterpret_cast<const int32_t*>(pts + p);
        const int32_t* ss_src = reinterpret_cast<const int32_t*>(psys);
        const int32_t* su_src = reinterpret_cast<const int32_t*>(spec);
        int32_t wv[WR];
#pragma unroll
        for (int r = 0; r < WR; ++r) {
            const int i = threadIdx.x + r * TB;
            int32_t v = 0;
            if (i < WP) v = sp_src[i];
            else if (i < WP + WS) v = ss_src[i - WP];
            else if (i < wtot) v = su_src[i - WP - WS];
            wv[r] = v;
        }
        if (ef != prec) return;
        int32_t* sp_dst = reinterpret_cast<int32_t*>(&Ps);
        int32_t* ss_dst = reinterpret_cast<int32_t*>(&Ss);
        int32_t* su_dst = reinterpret_cast<int32_t*>(Us);
#pragma unroll
        for (int r = 0; r < WR; ++r) {
            const int i = threadIdx.x + r * TB;
            if (i < WP) sp_dst[i] = wv[r];
            else if (i < WP + WS) ss_dst[i - WP] = wv[r];
            else if (i < wtot) su_dst[i - WP - WS] = wv[r];
        }
        __syncthreads();
    }
    const DevSystem& sys = Ss;
    const DevPoint& P = Ps;
    const int NP = P.NP, NB = sys.NB, MB = sys.MB;
    const int nl = threadIdx.x % ACC_POINTS;
    const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x / ACC_POINTS);  // wave-uniform: scalar segment loop
    const int n = blockIdx.x * ACC_POINTS + nl;
    const bool valid = n < NP;
    const int nc = valid ? n : NP - 1;
    // per-thread accumulators in LDS, one column per thread: rb[j * TB], rbo[f * TB]
    extern __shared__ double acc_lds[];
    double* rb = acc_lds + threadIdx.x;
    double* rbo = acc_lds + NB * TB + threadIdx.x;
    for (int j = 0; j < NB; ++j) rb[j * TB] = 0.0;
    for (int f = 0; f < 2 * MB; ++f) rbo[f * TB] = 0.0;
    __shared__ double red[2 * SDFT_MAX_SPECIES][TB / 64];
    int dealt = 0;  // segments dealt to the groups by the previous species (round-robin across species)
    for (int u = 0; u < sys.n_species; ++u) {
        const DevSpecies sp = Us[u];
        const int N = sp.nseg, S = sp.S;
        const int q0 = (grp - dealt % ACC_GROUPS + ACC_GROUPS) % ACC_GROUPS;  // this group's first segment
        dealt += N;
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
                norm = P.graft_density[u] / lamb;
            }
        } else {
            norm = sp.rescale ? 1.0 : P.chain_norm[u];  // rescaled: normative^(N-1) enters per segment
        }
        double num = 0.0, chg = 0.0;
        // a group with no segment of this species adds nothing (its accumulators stay +0)
        if (q0 < N) {
            const double pref0 = norm * exp(P.mu_species[u]);
            auto gidx = [&](int seg, int b, int a, int nn) -> int64_t {
                return P.gbase + (goffs[u] + ((int64_t)seg * S + b) * S + a) * P.gstride + nn;
            };
            // homogeneous species (every segment carries the same state beads): the bead of state a and
            // the bond of each (a, b) link are the same at every segment, so the contributions are summed
            // in registers and added to the LDS accumulators once per species
            const bool hom = sp.homog != 0;
            // states of segment q (bead index, -1 absent); a homogeneous species reads segment 0's once
            int st0[SM];
#pragma unroll
            for (int a = 0; a < SM; ++a) st0[a] = a < S ? st[a] : -1;
            auto stq = [&](int qq, int a) -> int { return hom ? st0[a] : st[qq * S + a]; };
            // monomerbead (monomerbead/accumulate_densities.jl:46-85): every segment independent,
            // rho = norm exp(mu_u - lambda), no propagators and no bonds
            const bool mb = sp.model == SDFT_MODEL_MONOMERBEAD;
            double racc[SM], rc0[SM][SM], rc1[SM][SM];
#pragma unroll
            for (int a = 0; a < SM; ++a) {
                racc[a] = 0.0;
#pragma unroll
                for (int b = 0; b < SM; ++b) rc0[a][b] = rc1[a][b] = 0.0;
            }
            for (int q = q0; q < N; q += ACC_GROUPS) {
                const double pref = sp.rescale ? pref0 * exp(seg_log_scale(sp, P, p, u, q, lsc, sys.n_species, nmax))
                                               : pref0;
                // every propagator value of this segment is loaded up front (independent loads in flight),
                // gcv[b][a] = gC[q][b][a] (child side, b: state of q+1), gpv[b][a] = gP[q][b][a] (parent side)
                // The loads are unconditional, from clamped (always in-bounds) indices, and masked after
                // the fact, so that they issue back to back instead of one conditional load at a time.
                T gcv[SM][SM], gpv[SM][SM], elv[SM], tzv[SM];
#pragma unroll
                for (int a = 0; a < SM; ++a) {
                    const int ac = a < S ? a : 0;
                    const int si = a < S ? stq(q, a) : -1;
                    const int sic = si >= 0 ? si : 0;
                    const T e = elam[fidx(P, NB, sic, nc)], t = trapez[fidx(P, NB, sic, nc)];
                    elv[a] = si >= 0 ? e : T(0);
                    tzv[a] = si >= 0 ? t : T(0);
#pragma unroll
                    for (int b = 0; b < SM; ++b) {
                        const int bc = b < S ? b : 0;
                        const bool okc = si >= 0 && b < S && q < N - 1 && stq(q + 1, b) >= 0;
                        const bool okp = si >= 0 && b < S && q > 0 && stq(q - 1, b) >= 0;
                        // a single-bead species has no propagators: no loads (N is wave-uniform)
                        const T vc = (N > 1 && !mb) ? gC[gidx(q, bc, ac, nc)] : T(0);
                        const T vp = (N > 1 && !mb) ? gP[gidx(q, bc, ac, nc)] : T(0);
                        gcv[b][a] = okc ? vc : T(0);
                        gpv[b][a] = okp ? vp : T(0);
                    }
                }
#pragma unroll
                for (int a = 0; a < SM; ++a) {
                    const int si = a < S ? stq(q, a) : -1;
                    if (si < 0) continue;
                    double Gp = 0.0, Qc = 0.0;
                    if (mb) {
                        Gp = Qc = 1.0;
                    } else {
                        if (q == 0) Gp = (double)base_value<T>(0, nc, sp.fixed_seg, Kf, P, maxd);
                        else
#pragma unroll
                            for (int b = 0; b < SM; ++b)
                                if (b < S && stq(q - 1, b) >= 0) Gp += (double)gpv[b][a];
                        // leaf: sum gC[K, :, 1, idx_i, leaf] is the base case 1 for every state
                        if (q == N - 1) Qc = (double)base_value<T>(q, nc, sp.fixed_seg, Kf, P, maxd);
                        else
#pragma unroll
                            for (int b = 0; b < SM; ++b)
                                if (b < S && stq(q + 1, b) >= 0) Qc += (double)gcv[b][a];
                    }
                    const double el = (double)elv[a];
                    const double contr = pref * el * Gp;
                    const double rs = contr * Qc;
                    const double tz = (double)tzv[a];
                    if (valid) {
                        if (q == sp.fixed_seg) {
                            if (nc == Kf) {
                                rb[si * TB] += rs / P.dz / tz;
                                num += rs;
                                chg += rs * sys.z[si];
                            }
                        } else {
                            if (hom) racc[a] += rs;
                            else rb[si * TB] += rs;
                            num += rs * tz * P.dz;
                            chg += rs * tz * P.dz * sys.z[si];
                        }
                        if (q < N - 1 && !mb) {
#pragma unroll
                            for (int b = 0; b < SM; ++b) {
                                const int sj = b < S ? stq(q + 1, b) : -1;
                                if (sj < 0) continue;
                                if (hom) {
                                    rc0[a][b] += contr * (double)gcv[b][a];
                                    continue;
                                }
                                const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
                                int bond = 0;
                                for (int u2 = 0; u2 < MB; ++u2)
                                    if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
                                rbo[(2 * bond + 0) * TB] += contr * (double)gcv[b][a];
                            }
                        }
                        if (q > 0 && !mb) {
                            const double c1 = pref * el * Qc;
#pragma unroll
                            for (int b = 0; b < SM; ++b) {
                                const int sj = b < S ? stq(q - 1, b) : -1;
                                if (sj < 0) continue;
                                if (hom) {
                                    rc1[a][b] += c1 * (double)gpv[b][a];
                                    continue;
                                }
                                const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
                                int bond = 0;
                                for (int u2 = 0; u2 < MB; ++u2)
                                    if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
                                rbo[(2 * bond + 1) * TB] += c1 * (double)gpv[b][a];
                            }
                        }
                    }
                }
            }
            if (hom && valid) {
                // segment 0's states and their links stand for every segment (homogeneous species)
#pragma unroll
                for (int a = 0; a < SM; ++a) {
                    const int si = a < S ? st[a] : -1;
                    if (si < 0) continue;
                    rb[si * TB] += racc[a];
#pragma unroll
                    for (int b = 0; b < SM; ++b) {
                        const int sj = b < S ? st[b] : -1;
                        if (sj < 0 || N < 2 || mb) continue;
                        const int lo = si <= sj ? si : sj, hi = si <= sj ? sj : si;
                        int bond = 0;
                        for (int u2 = 0; u2 < MB; ++u2)
                            if (sys.bond_i[u2] == lo && sys.bond_j[u2] == hi) bond = u2;
                        rbo[(2 * bond + 0) * TB] += rc0[a][b];
                        rbo[(2 * bond + 1) * TB] += rc1[a][b];
                    }
                }
            }
        }
        // Num_u = sum num / N_u (chains); charge = sum num * z (not divided): neutrality_shift.jl:34-45
        // a wave whose lanes all hold +0 (no items of this species, or no charge) skips the shuffle
        // rounds: the butterfly's sum of zeros is +0 as well
        num = __ballot(num != 0.0) ? warp_sum(num / N) : 0.0;
        chg = __ballot(chg != 0.0) ? warp_sum(chg) : 0.0;
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
    // densities of this grid point: the segment groups' partials in group order; the four groups
    // share the output columns (group g writes the beads, bond ends and classes congruent to g)
    auto total = [&](int f) {
        double v = 0.0;
        for (int g = 0; g < ACC_GROUPS; ++g) v += acc_lds[f * TB + g * ACC_POINTS + nl];
        return v;
    };
    for (int j = grp; j < NB; j += ACC_GROUPS) {
        const double v = total(j);
        rho_beads[fidx(P, NB, j, n)] = (T)v;
        rhoT[fidx(P, Frt, j, n)] = (T)(v * (double)trapez[fidx(P, NB, j, n)]);
    }
    for (int f = grp; f < 2 * MB; f += ACC_GROUPS) {
        const int u = f >> 1, e = f & 1;
        const int bt = e == 0 ? sys.bond_i[u] : sys.bond_j[u];
        const double v = total(NB + f);
        rho_bonds[fidx(P, 2 * MB, f, n)] = (T)v;
        rhoT[fidx(P, Frt, NB + f, n)] = (T)(v * (double)trapez[fidx(P, NB, bt, n)]);
    }
    // per diameter class: sum of rho_j trapez_j over its beads (input of the class-shared vector weight)
    for (int c = grp; c < sys.ncls; c += ACC_GROUPS) {
        double t = 0.0;
        for (int j = 0; j < NB; ++j)
            if (sys.cls[j] == c) t += total(j) * (double)trapez[fidx(P, NB, j, n)];
        rhoT[fidx(P, Frt, NB + 2 * MB + c, n)] = (T)t;
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
                                             T* __restrict__ nw, int Fnw, double* __restrict__ fx,
                                             double* __restrict__ wsum, int Fws) {
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
    conv_store<T>(J, P, n, (double)acc, nw, Fnw, fx, wsum, Fws);
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

// weighted densities of the mixture at one point; D vector components (one per grid dimension)
template <int D>
struct FmtIn {
    double n0, n1, n2, n3, nV1[D], nV2[D];
    double s3;   // sum of capped n3_i (TPT1) or capped n3
};

template <int D>
__device__ __forceinline__ double dotv(const double* a, const double* b) {
    double r = a[0] * b[0];
#pragma unroll
    for (int v = 1; v < D; ++v) r += a[v] * b[v];
    return r;
}

// weighted densities of the mixture at point n, read from the nw family: n2_j at j, n3_j at NB + j,
// the vector component v of diameter class c at (2 + v) NB + c
template <int D, typename T>
__device__ FmtIn<D> fmt_load(const DevSystem& sys, const DevPoint& P, int n, const T* __restrict__ nw, int Fnw) {
    FmtIn<D> f;
    f.n0 = f.n1 = f.n2 = f.n3 = f.s3 = 0.0;
#pragma unroll
    for (int v = 0; v < D; ++v) f.nV1[v] = f.nV2[v] = 0.0;
    const int NB = sys.NB;
    for (int j = 0; j < NB; ++j) {
        const double s = sys.sigma[j];
        const double n2 = (double)nw[fidx(P, Fnw, j, n)];
        f.n0 += n2 / (M_PI * s * s);
        f.n1 += n2 / (2.0 * M_PI * s);
        f.n2 += n2;
        f.n3 += (double)nw[fidx(P, Fnw, NB + j, n)];
    }
    // the vector weighted densities are linear in rho: one per diameter class (beads share w_V)
    for (int c = 0; c < sys.ncls; ++c) {
        const double s = sys.cls_sigma[c];
#pragma unroll
        for (int v = 0; v < D; ++v) {
            const double nv = (double)nw[fidx(P, Fnw, (2 + v) * NB + c, n)];
            f.nV1[v] += nv / (2.0 * M_PI * s);
            f.nV2[v] += nv;
        }
    }
    if (f.n3 > 0.95) {  // cap (mfmt.jl:126-131 / TPT1mfmt.jl:154-159: uniform scaling)
        const double fac = 0.95 / f.n3;
        f.n0 *= fac; f.n1 *= fac; f.n2 *= fac; f.n3 *= fac;
#pragma unroll
        for (int v = 0; v < D; ++v) { f.nV1[v] *= fac; f.nV2[v] *= fac; }
    }
    f.s3 = f.n3;
    return f;
}

// compute_mfmt_deriv! (mfmt.jl:210-244): d[0..3] the scalar rows, dv1[v] / dv2[v] the vector rows
// 4 + v and 7 + v
template <int D>
__device__ void fmt_deriv(const FmtIn<D>& f, double* d, double* dv1, double* dv2) {
    const double om = 1.0 - f.n3;
    const double lg = log(om);
    const double i1 = 1.0 / om;
    const double i3 = 1.0 / f.n3;
    const double c = 1.0 / (36.0 * M_PI);
    const double dot12 = dotv<D>(f.nV1, f.nV2), dot22 = dotv<D>(f.nV2, f.nV2);
    d[0] = -lg;
    d[1] = f.n2 * i1;
    d[2] = f.n1 * i1 + (f.n2 * f.n2 - dot22) * 3.0 * c * i3 * (lg * i3 + i1 * i1);
    d[3] = f.n0 * i1 + (f.n1 * f.n2 - dot12) * i1 * i1 -
           (lg * i3 * i3 * i3 * 2.0 * c + i3 * i3 * i1 * c + (1.0 - 3.0 * f.n3) * i3 * i3 * c * i1 * i1 * i1) *
               (f.n2 * f.n2 * f.n2 - 3.0 * f.n2 * dot22);
#pragma unroll
    for (int v = 0; v < D; ++v) {
        dv1[v] = -f.nV2[v] * i1;
        dv2[v] = -f.nV1[v] * i1 - (lg * i3 + i1 * i1) * f.n2 * f.nV2[v] * i3 * 6.0 * c;
    }
}

// free_energy_mfmt (inhomogeneous); cubic term takes its n3 -> 0 limit (0) at n3 == 0
template <int D>
__device__ double fmt_phi(const FmtIn<D>& f) {
    const double om = 1.0 - f.n3;
    const double lg = log(om);
    const double dot12 = dotv<D>(f.nV1, f.nV2), dot22 = dotv<D>(f.nV2, f.nV2);
    double cubic = 0.0;
    if (f.n3 != 0.0)
        cubic = (f.n2 * f.n2 * f.n2 - 3.0 * f.n2 * dot22) * (f.n3 + om * om * lg) / (36.0 * M_PI * f.n3 * f.n3 * om * om);
    return -f.n0 * lg + (f.n1 * f.n2 - dot12) / om + cubic;
}

// per-class sums of the MSA densities over the beads of one diameter:
// C = sum rho, B = sum rho z, A = sum rho z^2. Every MSA sum over beads
// (msa.jl:181-230, TPT1msa.jl:210-282) depends on a bead only through
// (sigma, rho, rho z, rho z^2), so it runs over the NCL classes instead.
template <int NCL>
struct MsaState {
    double C[NCL], B[NCL], A[NCL];
    double ig[NCL];           // 1 / (1 + G sigma_c) at the final G
    double G, eta, H;
};

// Gamma_msa (msa.jl:181-230): damped fixed point G <- 0.9 G + 0.1 G1, stop at |G1 - G| < 1e-10
template <int NCL>
__device__ __forceinline__ void gamma_msa_cls(MsaState<NCL>& m, const DevSystem& sys) {
    const double tol = 1e-10, lB = sys.lB;
    double s_rz = 0.0, s_r = 0.0, s_rs = 0.0, s_rs3 = 0.0;
#pragma unroll
    for (int c = 0; c < NCL; ++c) {
        const double s = sys.cls_sigma[c];
        s_rz += m.A[c];
        s_r += m.C[c];
        s_rs += m.C[c] * s;
        s_rs3 += m.C[c] * s * s * s;
    }
    const double kap = sqrt(4.0 * M_PI * lB * s_rz);
    if (kap < tol) {
        m.G = kap / 2.0; m.eta = 0.0; m.H = 2.0 / M_PI;
#pragma unroll
        for (int c = 0; c < NCL; ++c) m.ig[c] = 1.0 / (1.0 + m.G * sys.cls_sigma[c]);
        return;
    }
    const double seff = s_rs / s_r;
    const double x = seff * kap;
    double G = (sqrt(1.0 + 2.0 * x) - 1.0) / (2.0 * seff);
    const double term2 = 2.0 / M_PI * (1.0 - M_PI / 6.0 * s_rs3);
    double c3[NCL], bs[NCL];
#pragma unroll
    for (int c = 0; c < NCL; ++c) {
        const double s = sys.cls_sigma[c];
        c3[c] = m.C[c] * s * s * s;
        bs[c] = m.B[c] * s;
    }
    double H = 0.0, eta = 0.0;
    for (int it = 1;; ++it) {
        double ig[NCL], h = 0.0, e = 0.0;
#pragma unroll
        for (int c = 0; c < NCL; ++c) {
            ig[c] = 1.0 / (1.0 + G * sys.cls_sigma[c]);
            h += c3[c] * ig[c];
            e += bs[c] * ig[c];
        }
        H = h + term2;
        eta = e / H;
        double g1 = 0.0;
#pragma unroll
        for (int c = 0; c < NCL; ++c) {
            const double s2 = sys.cls_sigma[c] * sys.cls_sigma[c];
            // sum over the class of rho (z - eta s^2)^2
            const double q2 = m.A[c] - 2.0 * eta * s2 * m.B[c] + eta * eta * s2 * s2 * m.C[c];
            g1 += q2 * ig[c] * ig[c];
        }
        const double G1 = sqrt(M_PI * lB * g1);
        const double Gn = 0.9 * G + 0.1 * G1;
        const bool done = fabs(G1 - G) < tol || it >= 1000;
        G = Gn;
        if (done) break;
    }
    m.G = G; m.eta = eta; m.H = H;
#pragma unroll
    for (int c = 0; c < NCL; ++c) m.ig[c] = 1.0 / (1.0 + G * sys.cls_sigma[c]);
}

// class sums of rho = n2 / (pi sigma^2) at point n
template <typename T, int NCL>
__device__ __forceinline__ void msa_load(MsaState<NCL>& m, const DevSystem& sys, const DevPoint& P, int n, const T* __restrict__ nw,
                         int Fnw) {
#pragma unroll
    for (int c = 0; c < NCL; ++c) m.C[c] = m.B[c] = m.A[c] = 0.0;
    for (int j = 0; j < sys.NB; ++j) {
        const double s = sys.sigma[j], z = sys.z[j];
        const double rho = (double)nw[fidx(P, Fnw, j, n)] / (M_PI * s * s);
        const int cj = sys.cls[j];
#pragma unroll
        for (int c = 0; c < NCL; ++c) {
            if (c == cj) {
                m.C[c] += rho;
                m.B[c] += rho * z;
                m.A[c] += rho * z * z;
            }
        }
    }
}

// calls f(std::integral_constant<int, NCL>) with the number of diameter classes
// rounded up to a power of two (padding classes hold zero density and add exact zeros)
template <class F>
static void cls_dispatch(int ncls, F&& f) {
    if (ncls <= 1) f(std::integral_constant<int, 1>{});
    else if (ncls <= 2) f(std::integral_constant<int, 2>{});
    else if (ncls <= 4) f(std::integral_constant<int, 4>{});
    else if (ncls <= 8) f(std::integral_constant<int, 8>{});
    else f(std::integral_constant<int, SDFT_MAX_NB>{});
}

// calls f(std::integral_constant<int, D>) with the grid dimension (1, 2 or 3)
template <class F>
static void nd_dispatch(int nd, F&& f) {
    if (nd == 3) f(std::integral_constant<int, 3>{});
    else if (nd == 2) f(std::integral_constant<int, 2>{});
    else f(std::integral_constant<int, 1>{});
}

// pointwise functional derivatives, FP64 arithmetic, no per-bead arrays:
// the weighted densities are re-read from nw where a second pass needs them
template <typename T, int NCL, int D>
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
    const int F = sys.functionals;
    const bool has_fmt = (F & (SDFT_F_MFMT | SDFT_F_TPT1MFMT)) != 0;
    const bool has_msa = (F & (SDFT_F_MSA | SDFT_F_TPT1MSA)) != 0;
    const bool tpt1msa = (F & SDFT_F_TPT1MSA) != 0;

    // ---- FMT (mfmt.jl:106-199, TPT1mfmt.jl:135-241)
    constexpr int VB = 2 + D;   // bond fields start at VB * NB in nw and dphi
    FmtIn<D> f;
    f.n0 = f.n1 = f.n2 = f.n3 = f.s3 = 0.0;
    bool fmt_on = false;
    double d[4] = {0, 0, 0, 0}, dv1[D], dv2[D];
#pragma unroll
    for (int v = 0; v < D; ++v) { f.nV1[v] = f.nV2[v] = 0.0; dv1[v] = dv2[v] = 0.0; }
    if (has_fmt) {
        f = fmt_load<D>(sys, P, n, nw, Fnw);
        fmt_on = f.s3 > 1e-10;
        if (fmt_on) fmt_deriv<D>(f, d, dv1, dv2);
    }

    // ---- MSA (msa.jl:160-230, TPT1msa.jl:210-282)
    MsaState<NCL> m;
    double ustar = 0.0, dfe = 0.0, deH = 0.0, detaG = 0.0, denom = 0.0;
    if (has_msa) {
        msa_load<T, NCL>(m, sys, P, n, nw, Fnw);
        gamma_msa_cls<NCL>(m, sys);
#pragma unroll
        for (int c = 0; c < NCL; ++c) {
            const double s = sys.cls_sigma[c];
            ustar += s * s * (m.ig[c] * (m.B[c] - m.eta * s * s * m.C[c]) + m.B[c]);
        }
        ustar *= -M_PI / 6.0;
        if (tpt1msa && m.G != 0.0) {
            double dfG = 0.0, deG = 0.0, dHG = 0.0;
#pragma unroll
            for (int c = 0; c < NCL; ++c) {
                const double s = sys.cls_sigma[c], s2 = s * s, ig = m.ig[c];
                const double q2 = m.A[c] - 2.0 * m.eta * s2 * m.B[c] + m.eta * m.eta * s2 * s2 * m.C[c];
                dfG += s * q2 * ig * ig * ig;
                dfe += s2 * (m.B[c] - m.eta * s2 * m.C[c]) * ig * ig;
                deG += s2 * m.B[c] * ig * ig;
                dHG += s2 * s2 * m.C[c] * ig * ig;
            }
            dfG *= -2.0 * M_PI * sys.lB;
            dfe *= -2.0 * M_PI * sys.lB;
            deG *= -1.0 / m.H;
            dHG = -dHG;
            deH = -m.eta / m.H;
            detaG = deG + deH * dHG;
            denom = 1.0 / (2.0 * m.G - dfe * detaG - dfG);
        }
    }

    // ---- bonds: TPT1 chain terms and the bond outputs (TPT1mfmt.jl:387-445, TPT1msa.jl:232-282)
    double Sa = 0.0, Sb = 0.0;
    const bool tpt1fmt_on = fmt_on && (F & SDFT_F_TPT1MFMT);
    double Xif = 0.0, i1 = 0.0, t2 = 0.0, t3 = 0.0;
    if (tpt1fmt_on) {
        Xif = fmax(1.0 - dotv<D>(f.nV2, f.nV2) / (f.n2 * f.n2), 0.0);
        i1 = 1.0 / (1.0 - f.n3);
        t2 = f.n2 * Xif / 4.0 * i1 * i1;
        t3 = f.n2 * f.n2 * Xif / 72.0 * i1 * i1 * i1;
    }
    for (int u = 0; u < MB; ++u) {
        const int bi = sys.bond_i[u], bj = sys.bond_j[u];
        const double si = sys.sigma[bi], sj = sys.sigma[bj];
        const double sij = (si + sj) / 2.0;
        double lng = 0.0;
        if (tpt1fmt_on || tpt1msa) {
            const double rb = (double)nw[fidx(P, Fnw, VB * NB + 2 * u, n)] + (double)nw[fidx(P, Fnw, VB * NB + 2 * u + 1, n)];
            if (tpt1fmt_on) {
                const double tau = si * sj / sij;
                const double y = i1 + t2 * tau + t3 * tau * tau;
                const double dn2 = (tau / 4.0 * i1 * i1 + f.n2 * tau * tau * i1 * i1 * i1 / 36.0) / y;
                const double dn3 = (i1 * i1 + f.n2 * Xif / 2.0 * i1 * i1 * i1 * tau +
                                    f.n2 * f.n2 * Xif / 24.0 * i1 * i1 * i1 * i1 * tau * tau) / y;
                d[2] -= 0.5 * rb * dn2;
                d[3] -= 0.5 * rb * dn3;
#pragma unroll
                for (int v = 0; v < D; ++v) {
                    const double dv = -f.nV2[v] * (tau / (2.0 * f.n2) * i1 * i1 + tau * tau / 36.0 * i1 * i1 * i1) / y;
                    dv2[v] -= 0.5 * rb * dv;
                }
                lng += log(y);
            }
            if (tpt1msa) {
                const double zi = sys.z[bi], zj = sys.z[bj];
                if (m.G == 0.0) {
                    lng += -sys.lB / sij * zi * zj;
                } else {
                    const double igi = 1.0 / (1.0 + m.G * si), igj = 1.0 / (1.0 + m.G * sj);
                    const double Xi = (zi - m.eta * si * si) * igi, Xj = (zj - m.eta * sj * sj) * igj;
                    lng += -sys.lB / sij * Xi * Xj;
                    // d(X_i X_j)/d rho_w = detr_w (a_i X_j + X_i a_j) + dGr_w (b_i X_j + X_i b_j)
                    const double cu = rb / 2.0 * sys.lB / sij;
                    const double ai = -si * si * igi, aj = -sj * sj * igj;
                    const double b_i = -si * Xi * igi, b_j = -sj * Xj * igj;
                    Sa += cu * (ai * Xj + Xi * aj);
                    Sb += cu * (b_i * Xj + Xi * b_j);
                }
            }
        }
        dphi[fidx(P, Fdp, VB * NB + 2 * u, n)] = (T)(lng / (M_PI * si * si));
        dphi[fidx(P, Fdp, VB * NB + 2 * u + 1, n)] = (T)(lng / (M_PI * sj * sj));
    }

    // ---- bead outputs
    double q = surf[P.base + n];
    for (int j = 0; j < NB; ++j) {
        const double s = sys.sigma[j], z = sys.z[j];
        double ch0 = 0.0, ch1 = 0.0, chv[D];
#pragma unroll
        for (int v = 0; v < D; ++v) chv[v] = 0.0;
        if (fmt_on) {
            ch0 = d[0] / (M_PI * s * s) + d[1] / (2.0 * M_PI * s) + d[2];
            ch1 = d[3];
#pragma unroll
            for (int v = 0; v < D; ++v) chv[v] = dv1[v] / (2.0 * M_PI * s) + dv2[v];
        }
        if (has_msa) {
            const double ig = 1.0 / (1.0 + m.G * s);
            const double G = m.G, eta = m.eta;
            double dm = -sys.lB * (G * z * z * ig + eta * s * ((2.0 * z - eta * s * s) * ig + eta * s * s / 3.0) + z * ustar);
            if (tpt1msa && G != 0.0) {
                const double Xw = (z - eta * s * s) * ig;
                const double dfr = M_PI * sys.lB * Xw * Xw;
                const double der = s / m.H * z * ig;
                const double dHr = s * s * s * (ig - 1.0 / 3.0);
                const double dGr = (dfr + dfe * (der + dfe * deH * dHr)) * denom;
                const double detr = detaG * dGr;
                dm += detr * Sa + dGr * Sb;
            }
            ch0 += dm / (M_PI * s * s);
        }
        dphi[fidx(P, Fdp, j, n)] = (T)ch0;
        dphi[fidx(P, Fdp, NB + j, n)] = (T)ch1;
#pragma unroll
        for (int v = 0; v < D; ++v) dphi[fidx(P, Fdp, (2 + v) * NB + j, n)] = (T)chv[v];
        q += z * (double)rhoT[fidx(P, Frt, j, n)];
    }
    dphi[fidx(P, Fdp, VB * NB + 2 * MB, n)] = (T)q;
}

// --------------------------------------------------------------------------
// neutrality shift (compute_neutrality_shift, neutrality_shift.jl:1-99): one thread per point
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_neutrality(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                                   const DevSpecies* __restrict__ spec,
                                                   const int32_t* __restrict__ evalflag, int prec,
                                                   const double* __restrict__ acc_part,
                                                   double* __restrict__ acc_species, const double* __restrict__ x,
                                                   double* __restrict__ fx) {
    // one wavefront per point: lane k < 2 n_species adds the per-tile partials of quantity k in
    // tile order; lane u < n_species then carries species u through the Newton iteration, whose
    // sums over species are butterfly reductions (identical in every lane)
    const DevSystem& sys = *psys;
    const int p = blockIdx.x, lane = threadIdx.x;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int nsp = sys.n_species;
    const int ntiles = (P.NP + ACC_POINTS - 1) / ACC_POINTS;
    double t = 0.0;
    if (lane < 2 * nsp) {
        for (int b = 0; b < ntiles; ++b) t += acc_part[((int64_t)p * SDFT_MAX_TILES + b) * 2 * SDFT_MAX_SPECIES + lane];
        acc_species[(int64_t)p * 2 * SDFT_MAX_SPECIES + lane] = t;
    }
    const double num_u = __shfl(t, (2 * lane) & 63, 64);
    const double ch_u = __shfl(t, (2 * lane + 1) & 63, 64);
    double Num = 0.0, q = 0.0, fold = 0.0;
    if (lane < nsp) {
        Num = num_u;
        q = Num <= 0.0 ? 0.0 : ch_u / Num;
        if (spec[lane].multistate || spec[lane].fixed_seg >= 0) {  // folded into the fixed charge
            fold = Num * q;
            q = 0.0;
        }
    }
    const double Qtot = P.total_charge + warp_sum(fold);
    double err = fabs(Qtot + warp_sum(Num * q));
    double psi1 = 0.0, cst = 1.0;
    int it = 1;
    while (err > 1e-10) {
        const double tq = lane < nsp ? pow(cst, q) * Num * q : 0.0;
        const double f = Qtot + warp_sum(tq);
        const double f1 = -warp_sum(tq * q);
        err = fabs(f);
        if (fabs(f1) < 1e-10) break;
        const double del = fmax(1.0, fabs(f / f1));
        psi1 -= f / (f1 * del);
        cst = exp(-psi1);
        ++it;
        if (it > 1000) break;
    }
    if (lane == 0) fx[P.xoff + P.n - 1] = x[P.xoff + P.n - 1] + psi1;
}

// --------------------------------------------------------------------------
// Anderson mixing on the device (AndersonMixing.jl:56-169)
// --------------------------------------------------------------------------
// The history depth M is a template parameter so that the Gram accumulators,
// column offsets and the small least-squares solve live in registers; a point
// whose history holds fewer than M columns masks the missing ones to zero,
// which leaves every sum and every Jacobi rotation unchanged.
//
// Per-point block count of the reductions: a function of the point's own
// length only, so the summation order (and the result) never depends on the
// rest of the batch.
__host__ __device__ __forceinline__ int gram_blocks(int64_t n) {
    const int64_t b = (n + SDFT_GRAM_CHUNK - 1) / SDFT_GRAM_CHUNK;
    return (int)(b < 1 ? 1 : (b > SDFT_GRAM_BLOCKS ? SDFT_GRAM_BLOCKS : b));
}

__device__ __forceinline__ int hist_len(const DevSolveState& S, int m) {
    const int it = S.iter + 1;          // 1-based index of this iteration
    const int local = it - S.hist_base; // iterations in the current history
    return local < m ? local : m;
}

// packed layout of the normal equations for NC columns: the upper triangle
// (a <= b) row by row, then the NC right-hand sides
template <int NC> struct GramLayout {
    static constexpr int NTRI = NC * (NC + 1) / 2;
    static constexpr int LEN = NTRI + NC;
    static_assert(LEN <= SDFT_GRAM_LEN, "Gram layout exceeds its slot");
    __host__ __device__ static constexpr int tri(int a, int b) { return a * NC - a * (a - 1) / 2 + (b - a); }
};

// residual g = F(x) - x into the history, its max norm, and the per-block
// partial Gram sums  <dG_a, dG_b>, <dG_a, g>  with dG_a = G_{col-1-a} - g
template <int M>
__global__ void __launch_bounds__(256) k_and_resgram(const DevPoint* __restrict__ pts,
                                                    const DevSolveState* __restrict__ stt, const double* __restrict__ x,
                                                    const double* __restrict__ fx, double* __restrict__ Xh,
                                                    double* __restrict__ Gh, unsigned long long* __restrict__ errbits,
                                                    double* __restrict__ gram, int32_t* __restrict__ n_active) {
    constexpr int NC = M - 1;
    using L = GramLayout<NC>;
    const int p = blockIdx.y;
    if (blockIdx.x == 0 && p == 0 && threadIdx.x < 3) n_active[threadIdx.x] = 0;
    const DevSolveState& S = stt[p];
    if (S.converged) return;
    const DevPoint& P = pts[p];
    const int nb = gram_blocks(P.n);
    if ((int)blockIdx.x >= nb) return;
    const int col = S.iter % M;
    const int nc = hist_len(S, M) - 1;
    const int64_t hoff = P.xoff * M, n = P.n;
    const int64_t cur = hoff + (int64_t)col * n;
    int64_t coff[NC > 0 ? NC : 1];
#pragma unroll
    for (int a = 0; a < NC; ++a) coff[a] = hoff + (int64_t)((col - 1 - a + M) % M) * n;
    double acc[L::LEN];
#pragma unroll
    for (int k = 0; k < L::LEN; ++k) acc[k] = 0.0;
    double mx = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)nb * blockDim.x) {
        const double xi = x[P.xoff + i];
        const double g = fx[P.xoff + i] - xi;
        Xh[cur + i] = xi;
        Gh[cur + i] = g;
        const double ag = fabs(g);
        mx = (ag > mx || ag != ag) ? ag : mx;  // NaN propagates
        double dg[NC > 0 ? NC : 1];
#pragma unroll
        for (int a = 0; a < NC; ++a) dg[a] = a < nc ? Gh[coff[a] + i] - g : 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
#pragma unroll
            for (int b = a; b < NC; ++b) acc[L::tri(a, b)] += dg[a] * dg[b];
            acc[L::NTRI + a] += dg[a] * g;
        }
    }
    __shared__ double red[L::LEN + 1][4];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < L::LEN; ++k) {
        const double v = warp_sum(acc[k]);
        if ((threadIdx.x & 63) == 0) red[k][w] = v;
    }
    mx = warp_max(mx);
    if ((threadIdx.x & 63) == 0) red[L::LEN][w] = mx;
    __syncthreads();
    if ((int)threadIdx.x < L::LEN) {
        const double s = ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
        gram[((int64_t)p * SDFT_GRAM_BLOCKS + blockIdx.x) * SDFT_GRAM_LEN + threadIdx.x] = s;
    } else if ((int)threadIdx.x == L::LEN) {
        double v = nanmax(nanmax(red[L::LEN][0], red[L::LEN][1]), nanmax(red[L::LEN][2], red[L::LEN][3]));
        if (v != v) v = INFINITY;
        atomicMax(&errbits[p], (unsigned long long)__double_as_longlong(v));
    }
}

// minimum-norm least squares through the normal equations with a Jacobi
// eigen-decomposition and a relative eigenvalue cutoff (the reference calls
// lsqr!). Columns a >= nc are zero-padded: their rotations are skipped and
// their zero eigenvalues fall under the cutoff.
template <int NC>
__device__ void small_lstsq(int nc, const double* __restrict__ gp, double* __restrict__ lam) {
    using L = GramLayout<NC>;
    double A[NC][NC], V[NC][NC], rhs[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        rhs[a] = a < nc ? gp[L::NTRI + a] : 0.0;
#pragma unroll
        for (int b = 0; b < NC; ++b) {
            const int lo = a < b ? a : b, hi = a < b ? b : a;
            A[a][b] = (a < nc && b < nc) ? gp[L::tri(lo, hi)] : 0.0;
            V[a][b] = a == b ? 1.0 : 0.0;
        }
    }
    // cyclic Jacobi until the off-diagonal mass is at round-off level of the matrix
    double fro = 0.0;
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
        for (int b = 0; b < NC; ++b) fro += A[a][b] * A[a][b];
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0;
#pragma unroll
        for (int a = 0; a < NC; ++a)
#pragma unroll
            for (int b = a + 1; b < NC; ++b) off += A[a][b] * A[a][b];
        if (!(off > 1e-32 * fro)) break;
#pragma unroll
        for (int a = 0; a < NC; ++a) {
#pragma unroll
            for (int b = a + 1; b < NC; ++b) {
                if (fabs(A[a][b]) < 1e-300) continue;
                const double th = (A[b][b] - A[a][a]) / (2.0 * A[a][b]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                for (int r = 0; r < NC; ++r) {
                    const double ar = A[r][a], br = A[r][b];
                    A[r][a] = c * ar - s * br;
                    A[r][b] = s * ar + c * br;
                }
#pragma unroll
                for (int r = 0; r < NC; ++r) {
                    const double ar = A[a][r], br = A[b][r];
                    A[a][r] = c * ar - s * br;
                    A[b][r] = s * ar + c * br;
                }
#pragma unroll
                for (int r = 0; r < NC; ++r) {
                    const double vr = V[r][a], wr = V[r][b];
                    V[r][a] = c * vr - s * wr;
                    V[r][b] = s * vr + c * wr;
                }
            }
        }
    }
    double emax = 0.0;
#pragma unroll
    for (int a = 0; a < NC; ++a) emax = fmax(emax, fabs(A[a][a]));
    double out[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) out[a] = 0.0;
#pragma unroll
    for (int e = 0; e < NC; ++e) {
        const double ev = A[e][e];
        if (!(ev > 1e-14 * emax)) continue;
        double proj = 0.0;
#pragma unroll
        for (int r = 0; r < NC; ++r) proj += V[r][e] * rhs[r];
        proj /= ev;
#pragma unroll
        for (int r = 0; r < NC; ++r) out[r] -= V[r][e] * proj;  // solves M lam = -rhs
    }
#pragma unroll
    for (int a = 0; a < NC; ++a) lam[a] = out[a];
}

// one wavefront per point: lanes add the block partials in a fixed order,
// lane 0 takes the Picard/Anderson decision (AndersonMixing.jl:95-150)
template <int M>
__global__ void __launch_bounds__(64) k_and_solve(const DevPoint* __restrict__ pts, DevSolveState* __restrict__ stt,
                                                  int32_t* __restrict__ evalflag, unsigned long long* __restrict__ errbits,
                                                  const double* __restrict__ gram, SolveParams sp,
                                                  int32_t* __restrict__ n_active, const int32_t* __restrict__ errflag) {
    constexpr int NC = M - 1;
    using L = GramLayout<NC>;
    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    __shared__ double gp[L::LEN];
    if (stt[p].converged) {
        if (lane == 0) {
            stt[p].mode = -1;
            evalflag[p] = 0;
        }
        return;
    }
    if (lane < L::LEN) {
        const int nb = gram_blocks(pts[p].n);
        const double* src = gram + (int64_t)p * SDFT_GRAM_BLOCKS * SDFT_GRAM_LEN + lane;
        double t = 0.0;
        for (int b = 0; b < nb; ++b) t += src[(int64_t)b * SDFT_GRAM_LEN];
        gp[lane] = t;
    }
    __syncthreads();
    if (lane != 0) return;
    DevSolveState S = stt[p];
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
    bool picard = (local < M) || (err > sp.switch_tole && S.use_picard);
    if (S.restarted && local >= 2 && !(err > sp.switch_tole && S.use_picard)) picard = false;
    if (picard) {
        S.mode = 0;
        S.ncols = 0;
    } else {
        S.use_picard = 0;
        const int nc = (local < M ? local : M) - 1;
        small_lstsq<NC>(nc, gp, S.lam);
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
    // the FP32 -> FP64 switch itself is applied by k_apply_switch at the host's next poll, so that
    // the host always knows which precisions have active points
    if (!S.converged && sp.mixed && S.prec == 32 && err < sp.fp32_switch) S.want64 = 1;
    evalflag[p] = S.converged ? 0 : S.prec;
    if (!S.converged) {
        atomicAdd(&n_active[0], 1);
        if (S.prec == 32) atomicAdd(&n_active[1], 1);
        if (S.want64) atomicAdd(&n_active[2], 1);
    }
    stt[p] = S;
}

// mixed precision: move the points that asked for it to FP64 and restart their history
// (the FP32 history is not reused by the FP64 iterations)
__global__ void k_apply_switch(DevSolveState* __restrict__ stt, int32_t* __restrict__ evalflag, int P_count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_count) return;
    DevSolveState S = stt[p];
    if (!S.want64) return;
    S.want64 = 0;
    if (!S.converged) {
        S.prec = 64;
        S.switch_iter = S.iter;
        S.hist_base = S.iter;
        S.restarted = 1;
        evalflag[p] = 64;
    }
    stt[p] = S;
}

// x <- Picard mix or the Anderson extrapolation (AndersonMixing.jl:120-150)
template <int M>
__global__ void __launch_bounds__(256) k_and_update(const DevPoint* __restrict__ pts,
                                                   const DevSolveState* __restrict__ stt, double* __restrict__ x,
                                                   const double* __restrict__ fx, const double* __restrict__ Xh,
                                                   const double* __restrict__ Gh) {
    constexpr int NC = M - 1;
    const int p = blockIdx.y;
    const DevSolveState& S = stt[p];
    const int mode = S.mode;
    if (mode < 0) return;
    const DevPoint& P = pts[p];
    const int nb = gram_blocks(P.n);
    if ((int)blockIdx.x >= nb) return;
    const int64_t hoff = P.xoff * M, n = P.n;
    const int col = (S.iter - 1) % M;  // column written by this iteration
    const int64_t cur = hoff + (int64_t)col * n;
    const double amix = S.amix;
    const int nc = mode == 1 ? S.ncols : 0;
    int64_t coff[NC > 0 ? NC : 1];
    double lam[NC > 0 ? NC : 1];
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        coff[a] = hoff + (int64_t)((col - 1 - a + M) % M) * n;
        lam[a] = S.lam[a];
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)nb * blockDim.x) {
        const double xi = Xh[cur + i];
        double nx;
        if (mode == 0) {
            nx = (1.0 - amix) * xi + amix * fx[P.xoff + i];
        } else {
            const double gi = Gh[cur + i];
            double dx = 0.0, dg = 0.0;
#pragma unroll
            for (int a = 0; a < NC; ++a) {
                if (a < nc) {
                    dx += (Xh[coff[a] + i] - xi) * lam[a];
                    dg += (Gh[coff[a] + i] - gi) * lam[a];
                }
            }
            nx = xi + dx + amix * (gi + dg);
        }
        x[P.xoff + i] = nx;
    }
}

// --------------------------------------------------------------------------
// observables
// --------------------------------------------------------------------------
// grand potential integrand on the primary points (grand_potential.jl:4-28):
//   f_ex(K) - sum_j rho_j mu_ex_j trapez_j + 1/2 sum rho_bond lng
template <typename T, int NCL, int D>
__global__ void __launch_bounds__(256) k_omega(const DevSystem* __restrict__ psys, const DevPoint* __restrict__ pts,
                                              const DevSolveState* __restrict__ stt,
                                              const T* __restrict__ nw, int Fnw, const T* __restrict__ rhoT, int Frt,
                                              const T* __restrict__ rho_bonds, const T* __restrict__ rho_beads,
                                              const double* __restrict__ x, double* __restrict__ omega_acc) {
    const DevSystem& sys = *psys;
    const int p = blockIdx.y;
    if ((stt[p].prec == 64 ? 64 : 32) != (sizeof(T) == 8 ? 64 : 32)) return;
    const DevPoint& P = pts[p];
    const int NB = sys.NB, MB = sys.MB, NP = P.NP;
    const int F = sys.functionals;
    double acc = 0.0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < NP; n += gridDim.x * blockDim.x) {
        double f = 0.0;
        if (F & SDFT_F_SW) {
            // eval_fe_functional (square_well.jl:64-95): -1/2 rho_i sum_j eps_ij (w_ij * rho_j trapez_j)
            for (int s = 0; s < sys.n_sw_beads; ++s)
                f -= 0.5 * (double)rho_beads[fidx(P, NB, sys.sw_bead[s], n)] *
                     (double)nw[fidx(P, Fnw, sys.nw_sw_base + s, n)];
        }
        if (F & (SDFT_F_MFMT | SDFT_F_TPT1MFMT)) {
            const FmtIn<D> fi = fmt_load<D>(sys, P, n, nw, Fnw);
            f += fmt_phi<D>(fi);
            if (F & SDFT_F_TPT1MFMT) {
                const double Xi = 1.0 - dotv<D>(fi.nV2, fi.nV2) / (fi.n2 * fi.n2);  // no max(., 0) in the free energy
                const double i1 = 1.0 / (1.0 - fi.n3);
                const double t2 = fi.n2 * Xi / 4.0 * i1 * i1;
                const double t3 = fi.n2 * fi.n2 * Xi / 72.0 * i1 * i1 * i1;
                for (int u = 0; u < MB; ++u) {
                    const double si = sys.sigma[sys.bond_i[u]], sj = sys.sigma[sys.bond_j[u]];
                    const double sij = (si + sj) / 2.0, tau = si * sj / sij;
                    const double y = i1 + t2 * tau + t3 * tau * tau;
                    const double rb = (double)nw[fidx(P, Fnw, (2 + D) * NB + 2 * u, n)] +
                                      (double)nw[fidx(P, Fnw, (2 + D) * NB + 2 * u + 1, n)];
                    f -= 0.5 * rb * log(y);
                }
            }
        }
        if (F & (SDFT_F_MSA | SDFT_F_TPT1MSA)) {
            // free_energy_msa (msa.jl:232-250): G^3/(3 pi) - lB sum rho z (z G + eta s)/(1 + G s)
            MsaState<NCL> m;
            msa_load<T, NCL>(m, sys, P, n, nw, Fnw);
            gamma_msa_cls<NCL>(m, sys);
            double fe = m.G * m.G * m.G / (3.0 * M_PI);
#pragma unroll
            for (int c = 0; c < NCL; ++c)
                fe -= sys.lB * m.ig[c] * (m.A[c] * m.G + m.eta * sys.cls_sigma[c] * m.B[c]);
            if (F & SDFT_F_TPT1MSA) {
                for (int u = 0; u < MB; ++u) {
                    const int i = sys.bond_i[u], j = sys.bond_j[u];
                    const double si = sys.sigma[i], sj = sys.sigma[j];
                    const double sij = (si + sj) / 2.0;
                    const double Xi = (sys.z[i] - m.eta * si * si) / (1.0 + m.G * si);
                    const double Xj = (sys.z[j] - m.eta * sj * sj) / (1.0 + m.G * sj);
                    const double rb = (double)nw[fidx(P, Fnw, (2 + D) * NB + 2 * u, n)] +
                                      (double)nw[fidx(P, Fnw, (2 + D) * NB + 2 * u + 1, n)];
                    fe -= 0.5 * rb * (-sys.lB / sij * Xi * Xj);
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
                                                     int64_t out_stride, const double* __restrict__ lsc, int nmax) {
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
    double norm = sp.fixed_seg >= 0 ? P.graft_density[u] / fixed_lambda[(int64_t)p * SDFT_MAX_SPECIES + u]
                                    : (sp.rescale ? 1.0 : P.chain_norm[u]);
    double pref = norm * exp(P.mu_species[u]);
    if (sp.rescale) pref *= exp(seg_log_scale(sp, P, p, u, q, lsc, sys.n_species, nmax));
    auto gidx = [&](int seg, int b, int a, int nn) -> int64_t {
        return P.gbase + (goffs[u] + ((int64_t)seg * S + b) * S + a) * P.gstride + nn;
    };
    __shared__ double red[4];
    for (int a = 0; a < S; ++a) {
        const int si = st[q * S + a];
        double acc = 0.0;
        if (si >= 0) {
            for (int n = threadIdx.x; n < NP; n += blockDim.x) {
                double Gp = 0.0, Qc = 0.0;
                if (sp.model == SDFT_MODEL_MONOMERBEAD) Gp = Qc = 1.0;  // independent beads, full profiles
                else {
                    if (q == 0) Gp = (double)base_value<T>(0, n, sp.fixed_seg, Kf, P, maxd);
                    else
                        for (int b = 0; b < S; ++b)
                            if (st[(q - 1) * S + b] >= 0) Gp += (double)gP[gidx(q, b, a, n)];
                    if (q == N - 1) Qc = (double)base_value<T>(q, n, sp.fixed_seg, Kf, P, maxd);
                    else
                        for (int b = 0; b < S; ++b)
                            if (st[(q + 1) * S + b] >= 0) Qc += (double)gC[gidx(q, b, a, n)];
                }
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

// segment density profiles of species u at point p (rho_K.segments, accumulate_densities.jl:95-184):
// out[n + NP * (a + S * q)] = norm e^{mu} e^{-lambda} Gp Qc, 0 for absent states
template <typename T>
__global__ void __launch_bounds__(256) k_segment_profile(const DevSystem* __restrict__ psys,
                                                        const DevPoint* __restrict__ pts,
                                                        const DevSpecies* __restrict__ spec,
                                                        const int64_t* __restrict__ goffs, int GF,
                                                        const int32_t* __restrict__ seg_states,
                                                        const double* __restrict__ maxd_tab,
                                                        const T* __restrict__ elam, const T* __restrict__ gC,
                                                        const T* __restrict__ gP,
                                                        const double* __restrict__ fixed_lambda, int u, int p,
                                                        double* __restrict__ out, const double* __restrict__ lsc,
                                                        int nmax) {
    const DevSystem& sys = *psys;
    const DevPoint& P = pts[p];
    const DevSpecies sp = spec[u];
    const int q = blockIdx.y;
    const int S = sp.S, N = sp.nseg, NP = P.NP, NB = sys.NB;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= NP) return;
    const int* st = seg_states + sp.state_off;
    const double* maxd = maxd_tab + sp.maxd_off;
    const int Kf = sp.fixed_seg >= 0 ? P.fixed_coord[u] : -1;
    const double norm = sp.fixed_seg >= 0 ? P.graft_density[u] / fixed_lambda[(int64_t)p * SDFT_MAX_SPECIES + u]
                                          : (sp.rescale ? 1.0 : P.chain_norm[u]);
    double pref = norm * exp(P.mu_species[u]);
    if (sp.rescale) pref *= exp(seg_log_scale(sp, P, p, u, q, lsc, sys.n_species, nmax));
    auto gidx = [&](int seg, int b, int a) -> int64_t {
        return P.gbase + (goffs[u] + ((int64_t)seg * S + b) * S + a) * P.gstride + n;
    };
    for (int a = 0; a < S; ++a) {
        const int si = st[q * S + a];
        double v = 0.0;
        if (si >= 0) {
            double Gp = 0.0, Qc = 0.0;
            if (sp.model == SDFT_MODEL_MONOMERBEAD) Gp = Qc = 1.0;  // independent beads, full profiles
            else {
                if (q == 0) Gp = (double)base_value<T>(0, n, sp.fixed_seg, Kf, P, maxd);
                else
                    for (int b = 0; b < S; ++b)
                        if (st[(q - 1) * S + b] >= 0) Gp += (double)gP[gidx(q, b, a)];
                if (q == N - 1) Qc = (double)base_value<T>(q, n, sp.fixed_seg, Kf, P, maxd);
                else
                    for (int b = 0; b < S; ++b)
                        if (st[(q + 1) * S + b] >= 0) Qc += (double)gC[gidx(q, b, a)];
            }
            v = pref * (double)elam[fidx(P, NB, si, n)] * Gp * Qc;
        }
        out[n + (int64_t)NP * (a + (int64_t)S * q)] = v;
    }
}

// Contact-value theorem (contact_value_theorem.jl:1-20): per point and wall side the sum of the
// table's entries, rho_beads at a (bead field, grid index) offset times a factor, or a constant term
// (index < 0: charged walls). Densities come from the pipeline the point was last evaluated in.
__global__ void k_cvt_gather(const DevPoint* __restrict__ pts, const DevSolveState* __restrict__ stt,
                             const int32_t* __restrict__ off, const int32_t* __restrict__ idx,
                             const int32_t* __restrict__ side, const double* __restrict__ fac,
                             const float* __restrict__ rb32, const double* __restrict__ rb64, int NB, int P_count,
                             double* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_count) return;
    const int64_t base = pts[p].base * (int64_t)NB;
    const bool f32 = stt[p].prec == 32;
    double s[2] = {0.0, 0.0};
    for (int e = off[p]; e < off[p + 1]; ++e) {
        const int32_t i = idx[e];
        const double v = i < 0 ? 1.0 : (f32 ? (double)rb32[base + i] : rb64[base + i]);
        s[side[e]] += v * fac[e];
    }
    out[2 * p] = s[0];
    out[2 * p + 1] = s[1];
}

// Observables of every point from the partials of k_omega, k_neutrality and k_segment_sums
// (grand_potential.jl:4-28 with num_species.jl:1-38; degree_of_dissociation.jl:1-78): Omega =
// dz * (sum of the point's block partials) - sum_u N_u; per multistate species, the state fractions
// of each segment with more than one state, averaged over those segments (NaN where undefined).
// One thread per point; the sums run in the same order as a host loop over blocks, species,
// segments and states.
__global__ void k_obs_finalize(const DevPoint* __restrict__ pts, const DevSpecies* __restrict__ spec,
                               const int32_t* __restrict__ seg_states, ObsLayout lay, int nsp, int Smax,
                               int P_count, const double* __restrict__ om_blocks,
                               const double* __restrict__ acc_species, const double* __restrict__ segsums,
                               double* __restrict__ omega, double* __restrict__ alpha) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_count) return;
    const double qnan = __longlong_as_double(0x7ff8000000000000LL);
    double om = 0.0;
    for (int b = 0; b < SDFT_OMEGA_BLOCKS; ++b) om += om_blocks[(int64_t)p * SDFT_OMEGA_BLOCKS + b];
    double Nsum = 0.0;
    for (int u = 0; u < nsp; ++u) Nsum += acc_species[(int64_t)p * 2 * SDFT_MAX_SPECIES + 2 * u];
    omega[p] = om * pts[p].dz - Nsum;
    for (int u = 0; u < nsp; ++u) {
        double* al = alpha + ((int64_t)p * nsp + u) * Smax;
        for (int a = 0; a < Smax; ++a) al[a] = qnan;
        const DevSpecies sp = spec[u];
        if (!sp.multistate) continue;
        const int N = sp.nseg, S = sp.S;
        const double* ss = segsums + lay.ss_base[u] + (int64_t)p * N * S;
        const int32_t* st = seg_states + sp.state_off;
        double acc[SDFT_MAX_S], cnt[SDFT_MAX_S];
        for (int a = 0; a < S; ++a) acc[a] = cnt[a] = 0.0;
        for (int q = 0; q < N; ++q) {
            int ns = 0;
            for (int a = 0; a < S; ++a) if (st[q * S + a] >= 0) ++ns;
            if (ns == 1) continue;
            double tot = 0.0;
            for (int a = 0; a < ns; ++a) tot += ss[q * S + a];
            for (int a = 0; a < ns; ++a) {
                acc[a] += tot > 0 ? ss[q * S + a] / tot : 0.0;
                cnt[a] += 1.0;
            }
        }
        for (int a = 0; a < S; ++a) al[a] = cnt[a] > 0 ? acc[a] / cnt[a] : qnan;
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

// LDS of k_accumulate: one FP64 accumulator column per bead and per bond end
static size_t accumulate_lds_bytes(const Ctx& c) { return (size_t)(c.NB + 2 * c.MB) * 256 * sizeof(double); }

hipError_t prepare_kernel_attributes(const Ctx& c) {
    const size_t lds = accumulate_lds_bytes(c);
    if (lds <= 65536) return hipSuccess;
    const void* fns[4] = {reinterpret_cast<const void*>(&k_accumulate<float, 256, SDFT_FAST_S>),
                          reinterpret_cast<const void*>(&k_accumulate<double, 256, SDFT_FAST_S>),
                          reinterpret_cast<const void*>(&k_accumulate<float, 256, SDFT_MAX_S>),
                          reinterpret_cast<const void*>(&k_accumulate<double, 256, SDFT_MAX_S>)};
    for (const void* f : fns) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Merged back stage: the back projections were stored without their shared partials (computed in
// the same launch); add them here, in the order of conv_store (projection + partial + second partial).
__global__ void __launch_bounds__(256) k_add_partials(const DevPoint* __restrict__ pts,
                                                      const int32_t* __restrict__ evalflag, int prec,
                                                      const ConvJob* __restrict__ jobs, int njobs,
                                                      double* __restrict__ fx, const double* __restrict__ wsum,
                                                      int Fws) {
    const int p = blockIdx.y;
    if (evalflag[p] != prec) return;
    const DevPoint& P = pts[p];
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= P.NP) return;
    for (int j = 0; j < njobs; ++j) {
        const ConvJob& J = jobs[j];
        if (J.outkind != 1 || (J.addfield < 0 && J.addfield2 < 0)) continue;
        double* f = fx + P.xoff + (int64_t)J.outfield * P.NP + n;
        double y = *f;
        if (J.addfield >= 0) y += wsum[P.base * (int64_t)Fws + (int64_t)J.addfield * P.NP + n];
        if (J.addfield2 >= 0) y += wsum[P.base * (int64_t)Fws + (int64_t)J.addfield2 * P.NP + n];
        *f = y;
        if (J.outfield2 >= 0) fx[P.xoff + (int64_t)J.outfield2 * P.NP + n] = y;
    }
}

// the direct propagator step kernel: 256-point blocks for up to two states per segment, 64-point
// blocks (LDS for 16 state pairs) beyond
template <typename T>
static void launch_prop_step_direct(hipStream_t st, const Ctx& c, int prec, const PropArgs& A, int nsweeps,
                                    int sweep_base, const T* kern, const T* elam, const T* hl, const T* trapez, T* gC,
                                    T* gP) {
    if (A.S <= SDFT_FAST_S)
        hipLaunchKernelGGL((k_prop_step<T, 256, SDFT_FAST_S>), grid2(c.maxNP, 256, c.P, nsweeps), dim3(256), 0, st,
                           c.d_sys, c.d_pts, c.d_evalflag, prec, A, c.d_seg_states, c.d_maxd, kern, elam, hl, trapez,
                           gC, gP, c.d_band, sweep_base);
    else
        hipLaunchKernelGGL((k_prop_step<T, 64, SDFT_MAX_S>), grid2(c.maxNP, 64, c.P, nsweeps), dim3(64), 0, st,
                           c.d_sys, c.d_pts, c.d_evalflag, prec, A, c.d_seg_states, c.d_maxd, kern, elam, hl, trapez,
                           gC, gP, c.d_band, sweep_base);
}

// propagators of species u (child_propagators.jl / parent_propagators.jl / handle_fixed.jl): the
// persistent chain kernel where the plan allows it (and allow_chain), else one launch per step
template <typename T>
void Pipeline<T>::propagate(hipStream_t st, const Ctx& c, int prec, int u, bool allow_chain) {
    const int P = c.P;
    const int maxNP = c.maxNP;
    const SpeciesHost& sh = c.species[u];
    PropArgs A;
    A.u = u; A.nseg = sh.nseg; A.S = sh.S; A.fixed_seg = sh.fixed_seg;
    A.state_off = sh.state_off; A.maxd_off = sh.maxd_off; A.goff = sh.goff; A.GF = c.GF;
    A.rescale = sh.rescale; A.nsp = c.nsp; A.nmax = c.nmax_seg; A.smax = smax;
    if (sh.model == SDFT_MODEL_MONOMERBEAD) {
        // no propagators; a grafted monomerbead species only needs its normalisation
        if (sh.fixed_seg >= 0)
            hipLaunchKernelGGL(k_fixed_lambda_mb, dim3((P + 63) / 64), dim3(64), 0, st, c.d_pts, c.d_evalflag, prec, P,
                               u, c.d_seg_states + sh.state_off, sh.S, c.d_x, c.d_Ext, c.NB, c.d_fixed_lambda);
        return;
    }
    const bool banded = (sizeof(T) == 4) ? c.band_fp32 != 0 : c.band_fp64 != 0;
    A.band = banded ? 1 : -1;
    const bool gemm = c.use_gemm && c.prop_plan[u].ntiles > 0;
    if (allow_chain && banded && (sizeof(T) == 4 ? c.chain_plan[u].ok32 : c.chain_plan[u].ok64)) {
        launch_chain<T>(st, c, prec, u, kern, elam, hl, trapez, gC, gP, smax);
        if (sh.rescale)
            hipLaunchKernelGGL(k_scale_prefix, dim3((P + 63) / 64), dim3(64), 0, st, c.d_evalflag, prec, P, u,
                               sh.nseg, sh.fixed_seg, smax, lsc, c.nsp, c.nmax_seg);
        return;
    }
    const bool bandstep = banded && sh.band_step != 0;
    for (int s = 1; s < sh.nseg; ++s) {
        A.s = s;
        // sweeps whose parent is the grafted segment use the direct real-space sums
        const bool fix0 = sh.fixed_seg == 0 && s == sh.nseg - 1;  // child sweep, parent = segment 0
        const bool fix1 = sh.fixed_seg == 0 && s == 1;            // parent sweep, parent = segment 0
        if (bandstep) {
            const int skip = (fix0 && fix1) ? 2 : fix0 ? 0 : fix1 ? 1 : -1;
            if (skip != 2) launch_bandstep<T>(st, c, prec, u, s, skip, kern, elam, hl, trapez, gC, gP, smax);
            if (skip >= 0)
                launch_prop_step_direct<T>(st, c, prec, A, skip == 2 ? 2 : 1, skip == 2 ? 0 : skip, kern, elam, hl, trapez, gC, gP);
            continue;
        }
        if (!gemm) {
            launch_prop_step_direct<T>(st, c, prec, A, 2, 0, kern, elam, hl, trapez, gC, gP);
            continue;
        }
        const int skip = (fix0 && fix1) ? 2 : fix0 ? 0 : fix1 ? 1 : -1;
        if (skip != 2)
            launch_opgemm_prop<T>(st, c, prec, c.prop_plan[u], u, s, skip, banded, kern, elam, hl, trapez, gC, gP,
                                  fbuf, smax);
        if (skip >= 0)
            launch_prop_step_direct<T>(st, c, prec, A, skip == 2 ? 2 : 1, skip == 2 ? 0 : skip, kern, elam, hl, trapez, gC, gP);
    }
    if (sh.rescale)
        hipLaunchKernelGGL(k_scale_prefix, dim3((P + 63) / 64), dim3(64), 0, st, c.d_evalflag, prec, P, u, sh.nseg,
                           sh.fixed_seg, smax, lsc, c.nsp, c.nmax_seg);
    if (sh.fixed_seg >= 0) {
        hipLaunchKernelGGL(k_fixed_lambda<T>, dim3(P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag, prec, A,
                           c.d_seg_states, c.d_maxd, elam, trapez, gC, gP, c.d_fixed_lambda);
    }
}

template <typename T>
void launch_fixed_lambda(hipStream_t st, const Ctx& c, int prec, const PropArgs& A, const T* elam, const T* trapez,
                         const T* gC, const T* gP) {
    hipLaunchKernelGGL(k_fixed_lambda<T>, dim3(c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag, prec, A,
                       c.d_seg_states, c.d_maxd, elam, trapez, gC, gP, c.d_fixed_lambda);
}
template void launch_fixed_lambda<float>(hipStream_t, const Ctx&, int, const PropArgs&, const float*, const float*,
                                         const float*, const float*);
template void launch_fixed_lambda<double>(hipStream_t, const Ctx&, int, const PropArgs&, const double*, const double*,
                                          const double*, const double*);

void launch_scale_prefix(hipStream_t st, const Ctx& c, int prec, int u, unsigned long long* smax, double* lsc) {
    const SpeciesHost& sh = c.species[u];
    hipLaunchKernelGGL(k_scale_prefix, dim3((c.P + 63) / 64), dim3(64), 0, st, c.d_evalflag, prec, c.P, u, sh.nseg,
                       sh.fixed_seg, smax, lsc, c.nsp, c.nmax_seg);
}

template <typename T>
void Pipeline<T>::objective(hipStream_t st, const Ctx& c, int prec, const double* x, double* fx) {
    const int P = c.P;
    const int maxNP = c.maxNP;
    hipLaunchKernelGGL(k_prepare<T>, grid2(maxNP, 256, P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_evalflag, prec, x,
                       c.d_Ext, elam, hl);
    if (c.any_rescale) {  // per-step maxima start at 0 every evaluation (a fill kernel: capture-safe)
        const int64_t ns = (int64_t)P * c.nsp * 2 * c.nmax_seg;
        hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st,
                           reinterpret_cast<double*>(smax), ns);
    }
    const bool nd = c.nd > 1;
    for (int u = 0; u < c.nsp; ++u) {
        if (nd) propagate_nd(st, c, prec, u);
        else propagate(st, c, prec, u, true);
    }
    launch_accumulate(st, c, prec);
    if (nd) conv_nd(st, c, prec, false, fx);
    else launch_conv(st, c, prec, false, fx);
    launch_pointwise(st, c, prec);
    if (nd) conv_nd(st, c, prec, true, fx);
    else launch_conv(st, c, prec, true, fx);
    hipLaunchKernelGGL(k_neutrality, dim3(P), dim3(64), 0, st, c.d_sys, c.d_pts, c.d_spec, c.d_evalflag, prec,
                       c.d_acc_part, c.d_acc_species, x, fx);
}

template <typename T>
void Pipeline<T>::launch_accumulate(hipStream_t st, const Ctx& c, int prec) {
    // species of up to two states per segment (every benchmark configuration) take the two-state
    // instantiation, whose per-state arrays stay in registers
    if (c.states_max <= SDFT_FAST_S)
        hipLaunchKernelGGL((k_accumulate<T, 256, SDFT_FAST_S>), grid2(c.maxNP, ACC_POINTS, c.P), dim3(256),
                           accumulate_lds_bytes(c), st, c.d_sys, c.d_pts, c.d_evalflag, prec, c.d_spec, c.d_goffs, c.GF,
                           c.d_seg_states, c.d_maxd, elam, trapez, gC, gP, c.d_fixed_lambda, rho_beads, rho_bonds, rhoT,
                           c.Frt, c.d_acc_part, c.d_errflag, lsc, c.nmax_seg, c.nsp);
    else
        hipLaunchKernelGGL((k_accumulate<T, 256, SDFT_MAX_S>), grid2(c.maxNP, ACC_POINTS, c.P), dim3(256),
                           accumulate_lds_bytes(c), st, c.d_sys, c.d_pts, c.d_evalflag, prec, c.d_spec, c.d_goffs, c.GF,
                           c.d_seg_states, c.d_maxd, elam, trapez, gC, gP, c.d_fixed_lambda, rho_beads, rho_bonds, rhoT,
                           c.Frt, c.d_acc_part, c.d_errflag, lsc, c.nmax_seg, c.nsp);
}

template <typename T>
void Pipeline<T>::launch_prop_step(hipStream_t st, const Ctx& c, int prec, int u, int s, bool banded) {
    const SpeciesHost& sh = c.species[u];
    PropArgs A;
    A.u = u; A.nseg = sh.nseg; A.S = sh.S; A.fixed_seg = sh.fixed_seg;
    A.state_off = sh.state_off; A.maxd_off = sh.maxd_off; A.goff = sh.goff; A.GF = c.GF;
    A.band = banded ? 1 : -1;
    A.s = s;
    A.rescale = sh.rescale; A.nsp = c.nsp; A.nmax = c.nmax_seg; A.smax = smax;
    if (c.use_gemm && c.prop_plan[u].ntiles > 0 && sh.fixed_seg < 0)
        launch_opgemm_prop<T>(st, c, prec, c.prop_plan[u], u, s, -1, banded, kern, elam, hl, trapez, gC, gP,
                              fbuf, smax);
    else
        launch_prop_step_direct<T>(st, c, prec, A, 2, 0, kern, elam, hl, trapez, gC, gP);
}

// forward: weighted densities; back: shared partials first, then the projections that add them
template <typename T>
void Pipeline<T>::launch_conv(hipStream_t st, const Ctx& c, int prec, bool back, double* fx) {
    auto run = [&](const Ctx::ConvStage& stage, int job0, int nj) {
        if (nj == 0) return;
        if (c.use_gemm) {
            for (int sh = 0; sh < CONV_NSHAPES; ++sh)
                launch_opgemm_conv<T>(st, c, prec, stage.sub[sh], kern, rhoT, dphi, nw, fx);
            return;
        }
        hipLaunchKernelGGL((k_conv<T, 256>), grid2(c.maxNP, 256, nj, c.P), dim3(256), 0, st, c.d_sys, c.d_pts,
                           c.d_evalflag, prec, c.d_jobs + job0, kern, rhoT, c.Frt, dphi, c.Fdp, nw, c.Fnw, fx,
                           c.d_wsum, c.Fws);
    };
    if (!back) {
        run(c.fwd_plan, 0, c.n_fwd_jobs);
    } else if (c.use_gemm && c.merge_back) {
        run(c.back_plan, c.n_jobs + c.n_fwd_jobs, c.n_pre_jobs + c.n_back_jobs);
        if (c.n_back_jobs > 0)
            hipLaunchKernelGGL(k_add_partials, grid2(c.maxNP, 256, c.P), dim3(256), 0, st, c.d_pts, c.d_evalflag, prec,
                               c.d_jobs + c.n_fwd_jobs + c.n_pre_jobs, c.n_back_jobs, fx, c.d_wsum, c.Fws);
    } else {
        run(c.pre_plan, c.n_fwd_jobs, c.n_pre_jobs);
        run(c.back_plan, c.n_fwd_jobs + c.n_pre_jobs, c.n_back_jobs);
    }
}

template <typename T>
void Pipeline<T>::launch_pointwise(hipStream_t st, const Ctx& c, int prec) {
    const T *nw_ = nw, *rhoT_ = rhoT;
    T* dphi_ = dphi;
    cls_dispatch(c.ncls, [&](auto k) {
        constexpr int NCL = decltype(k)::value;
        nd_dispatch(c.nd, [&](auto dd) {
            constexpr int D = decltype(dd)::value;
            hipLaunchKernelGGL((k_pointwise<T, NCL, D>), grid2(c.maxNP, 256, c.P), dim3(256), 0, st, c.d_sys, c.d_pts,
                               c.d_evalflag, prec, nw_, c.Fnw, rhoT_, c.Frt, c.d_surf, dphi_, c.Fdp);
        });
    });
}

template <typename T>
void Pipeline<T>::omega(hipStream_t st, const Ctx& c, const double* x, double* omega_acc) {
    const T *nw_ = nw, *rhoT_ = rhoT, *rb_ = rho_bonds, *rbe_ = rho_beads;
    cls_dispatch(c.ncls, [&](auto k) {
        constexpr int NCL = decltype(k)::value;
        nd_dispatch(c.nd, [&](auto dd) {
            constexpr int D = decltype(dd)::value;
            hipLaunchKernelGGL((k_omega<T, NCL, D>), dim3(SDFT_OMEGA_BLOCKS, c.P), dim3(256), 0, st, c.d_sys, c.d_pts,
                               c.d_state, nw_, c.Fnw, rhoT_, c.Frt, rb_, rbe_, x, omega_acc);
        });
    });
}

template <typename T>
void Pipeline<T>::segment_sums(hipStream_t st, const Ctx& c, int u, double* out, int64_t stride) {
    const SpeciesHost& sh = c.species[u];
    hipLaunchKernelGGL(k_segment_sums<T>, dim3(sh.nseg, c.P), dim3(256), 0, st, c.d_sys, c.d_pts, c.d_state, c.d_spec,
                       c.d_goffs, c.GF, c.d_seg_states, c.d_maxd, elam, gC, gP, c.d_fixed_lambda, u, out, stride, lsc,
                       c.nmax_seg);
}

template <typename T>
void Pipeline<T>::segment_profile(hipStream_t st, const Ctx& c, int u, int p, double* out) {
    const SpeciesHost& sh = c.species[u];
    hipLaunchKernelGGL(k_segment_profile<T>, dim3((c.maxNP + 255) / 256, sh.nseg), dim3(256), 0, st, c.d_sys,
                       c.d_pts, c.d_spec, c.d_goffs, c.GF, c.d_seg_states, c.d_maxd, elam, gC, gP,
                       c.d_fixed_lambda, u, p, out, lsc, c.nmax_seg);
}

template struct Pipeline<float>;
template struct Pipeline<double>;

template <int M>
static void launch_anderson_m(hipStream_t st, const Ctx& c, const SolveParams& sp) {
    const int P = c.P;
    const int tiles = gram_blocks(c.max_n);  // each point uses gram_blocks(its n) of these
    hipLaunchKernelGGL(k_and_resgram<M>, dim3(tiles, P), dim3(256), 0, st, c.d_pts, c.d_state, c.d_x, c.d_fx, c.d_Xh,
                       c.d_Gh, c.d_errbits, c.d_gram, c.d_n_active);
    hipLaunchKernelGGL(k_and_solve<M>, dim3(P), dim3(64), 0, st, c.d_pts, c.d_state, c.d_evalflag, c.d_errbits, c.d_gram,
                       sp, c.d_n_active, c.d_errflag);
    hipLaunchKernelGGL(k_and_update<M>, dim3(tiles, P), dim3(256), 0, st, c.d_pts, c.d_state, c.d_x, c.d_fx, c.d_Xh,
                       c.d_Gh);
}

void launch_anderson(hipStream_t st, const Ctx& c, const SolveParams& sp) {
    switch (sp.m) {
        case 2: launch_anderson_m<2>(st, c, sp); break;
        case 3: launch_anderson_m<3>(st, c, sp); break;
        case 4: launch_anderson_m<4>(st, c, sp); break;
        case 5: launch_anderson_m<5>(st, c, sp); break;
        case 6: launch_anderson_m<6>(st, c, sp); break;
        case 7: launch_anderson_m<7>(st, c, sp); break;
        default: launch_anderson_m<8>(st, c, sp); break;
    }
}

void launch_apply_switch(hipStream_t st, const Ctx& c) {
    hipLaunchKernelGGL(k_apply_switch, dim3((c.P + 63) / 64), dim3(64), 0, st, c.d_state, c.d_evalflag, c.P);
}

void launch_cvt_gather(hipStream_t st, const Ctx& c, const int32_t* off, const int32_t* idx, const int32_t* side,
                       const double* fac, const float* rb32, const double* rb64, double* out) {
    hipLaunchKernelGGL(k_cvt_gather, dim3((c.P + 63) / 64), dim3(64), 0, st, c.d_pts, c.d_state, off, idx, side, fac,
                       rb32, rb64, c.NB, c.P, out);
}

void launch_obs_finalize(hipStream_t st, const Ctx& c, const ObsLayout& lay, int Smax, const double* om_blocks,
                         const double* segsums, double* omega, double* alpha) {
    hipLaunchKernelGGL(k_obs_finalize, dim3((c.P + 63) / 64), dim3(64), 0, st, c.d_pts, c.d_spec, c.d_seg_states,
                       lay, c.nsp, Smax, c.P, om_blocks, c.d_acc_species, segsums, omega, alpha);
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
