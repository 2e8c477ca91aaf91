// Bulk (homogeneous) Ising state of every state point of a batch on the device: the input of
// sdft_create (mu_species, bulk mu_ex and lng, bead and bond densities, pressure).
//
// Reference: the iLST bulk solve, src/iLST/evaluation/objective/solver_bulkstate.jl:31-103 with
// residual_bulkstate.jl:26-42 (densities from the bulk chain propagators, charge balance, bulk
// chemical potentials), functions/balance_charge.jl:21-79, the bulk propagators
// src/iLST/species_model/freely_jointed/*.jl and monomerbead, the excess free energies
// src/iLST/free_energy/{excluded_volume,electrostaticcorrelations,short_range_interactions}, and
// the pressure src/iLST/evaluation/analysis/eval_pressure.jl:10-25.
//
// The same fixed point is reached by Newton's method on the self-consistency of (mu_ex, lng)
// (the reference iterates Anderson on log densities; only the converged state is shared). One
// 64-lane wavefront owns one state point: lane k < n evaluates the residual at y + h_k e_k, lane n at
// y, so a forward-difference Jacobian costs one residual evaluation in time; lane 0 solves the
// n x n system (n = NB + MB <= 32) in LDS. The chain recursion runs with a rescaled propagator per
// segment (its logarithm carried along), equal to the direct products wherever those are finite and
// free of their overflow for long chains at high charge.
#include <cmath>

#include "sdft_launch.h"

namespace sdft {

constexpr int BULK_LANES = 64;

__device__ __forceinline__ int bulk_bond(const BulkSysDev& s, int i, int j) {
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    for (int u = 0; u < s.MB; ++u)
        if (s.bi[u] == lo && s.bj[u] == hi) return u;
    return -1;
}

// Gamma_msa (src/iLST/free_energy/electrostaticcorrelations/msa.jl:181-230): the damped fixed point
// Gamma <- 0.9 Gamma + 0.1 Gamma_1 from (sqrt(1 + 2 x) - 1) / (2 sigma_eff), tolerance 1e-10, at most
// 1000 iterations; kappa < 1e-10 gives Gamma = kappa / 2, eta = 0, H = 2 / pi
__device__ void bulk_gamma_msa(const BulkSysDev& s, const double* rho, double& G, double& eta, double& H) {
    const int NB = s.NB;
    double rz2 = 0.0, rs = 0.0, rsum = 0.0, rs3 = 0.0;
    for (int j = 0; j < NB; ++j) {
        rz2 += rho[j] * s.z[j] * s.z[j];
        rs += rho[j] * s.sig[j];
        rsum += rho[j];
        rs3 += rho[j] * s.sig[j] * s.sig[j] * s.sig[j];
    }
    const double kap = sqrt(4.0 * M_PI * s.lB * rz2);
    if (kap < 1e-10) {
        G = kap / 2.0;
        eta = 0.0;
        H = 2.0 / M_PI;
        return;
    }
    const double seff = rs / rsum;
    const double x = seff * kap;
    G = (sqrt(1.0 + 2.0 * x) - 1.0) / (2.0 * seff);
    const double term2 = 2.0 / M_PI * (1.0 - M_PI / 6.0 * rs3);
    for (int it = 1;; ++it) {
        double h = 0.0, e = 0.0;
        for (int j = 0; j < NB; ++j) {
            const double og = 1.0 + G * s.sig[j];
            h += rho[j] * s.sig[j] * s.sig[j] * s.sig[j] / og;
            e += rho[j] * s.z[j] * s.sig[j] / og;
        }
        H = h + term2;
        eta = e / H;
        double g2 = 0.0;
        for (int j = 0; j < NB; ++j) {
            const double og = 1.0 + G * s.sig[j];
            const double t = s.z[j] - eta * s.sig[j] * s.sig[j];
            g2 += rho[j] * t * t / (og * og);
        }
        const double g1 = sqrt(M_PI * s.lB * g2);
        const double gn = 0.9 * G + 0.1 * g1;
        const bool done = fabs(g1 - G) < 1e-10 || it >= 1000;
        G = gn;
        if (done) break;
    }
}

// Excess chemical potentials, ln y per bond and the excess free-energy density of a bulk state
// (bulk_chemical_potential.jl:8-24 over the model list)
__device__ void bulk_fe(const BulkSysDev& s, const double* rho, const double* rb0, const double* rb1, double* mu,
                        double* lng, double* f_out, bool want_f) {
    const int NB = s.NB, MB = s.MB, F = s.F;
    for (int j = 0; j < NB; ++j) mu[j] = 0.0;
    for (int u = 0; u < MB; ++u) lng[u] = 0.0;
    double f = 0.0;
    if (F & (SDFT_F_MFMT | SDFT_F_TPT1MFMT)) {
        // compute_mfmt_deriv! / free_energy_mfmt (src/iLST/free_energy/excluded_volume/mfmt.jl:143-300)
        double n0 = 0.0, n1 = 0.0, n2 = 0.0, n3 = 0.0;
        for (int j = 0; j < NB; ++j) {
            const double sg = s.sig[j];
            n0 += rho[j];
            n1 += rho[j] * sg / 2.0;
            n2 += rho[j] * M_PI * sg * sg;
            n3 += rho[j] * M_PI * sg * sg * sg / 6.0;
        }
        const double om = 1.0 - n3, lg = log(om), i1 = 1.0 / om, i3 = 1.0 / n3, c = 1.0 / (36.0 * M_PI);
        double d0 = -lg;
        double d1 = n2 * i1;
        double d2 = n1 * i1 + n2 * n2 * 3.0 * c * i3 * (lg * i3 + i1 * i1);
        double d3 = n0 * i1 + n1 * n2 * i1 * i1 -
                    (lg * i3 * i3 * i3 * 2.0 * c + i3 * i3 * i1 * c + (1.0 - 3.0 * n3) * i3 * i3 * c * i1 * i1 * i1) *
                        n2 * n2 * n2;
        if (want_f) {
            const double inv_den = 1.0 / (36.0 * M_PI * n3 * n3 * om * om);
            f += -n0 * lg + n1 * n2 / om + n2 * n2 * n2 * (n3 + om * om * lg) * inv_den;
        }
        if (F & SDFT_F_TPT1MFMT) {
            // TPT1 chain term in the bulk (Xi = 1): src/iLST/free_energy/excluded_volume/TPT1mfmt.jl:398-445
            const double t2 = n2 / 4.0 * i1 * i1;
            const double t3 = n2 * n2 / 72.0 * i1 * i1 * i1;
            for (int u = 0; u < MB; ++u) {
                const double si = s.sig[s.bi[u]], sj = s.sig[s.bj[u]];
                const double sij = (si + sj) / 2.0, tau = si * sj / sij;
                const double y = i1 + t2 * tau + t3 * tau * tau;
                const double dn2 = (tau / 4.0 * i1 * i1 + n2 * tau * tau * i1 * i1 * i1 / 36.0) / y;
                const double dn3 =
                    (i1 * i1 + n2 / 2.0 * i1 * i1 * i1 * tau + n2 * n2 / 24.0 * i1 * i1 * i1 * i1 * tau * tau) / y;
                const double rbs = rb0[u] + rb1[u];
                d2 -= 0.5 * rbs * dn2;
                d3 -= 0.5 * rbs * dn3;
                lng[u] += log(y);
                if (want_f) f -= 0.5 * rbs * log(y);
            }
        }
        for (int j = 0; j < NB; ++j) {
            const double sg = s.sig[j];
            mu[j] += d0 + d1 * sg / 2.0 + d2 * M_PI * sg * sg + d3 * M_PI * sg * sg * sg / 6.0;
        }
    }
    if (F & (SDFT_F_MSA | SDFT_F_TPT1MSA)) {
        // compute_msa_deriv! / free_energy_msa (msa.jl:115-230) and the TPT1 electrostatic chain term
        // (TPT1msa.jl:138-282)
        double G, eta, H;
        bulk_gamma_msa(s, rho, G, eta, H);
        double ustar = 0.0;
        for (int j = 0; j < NB; ++j) {
            const double sg = s.sig[j];
            ustar += rho[j] * sg * sg * ((s.z[j] - eta * sg * sg) / (1.0 + G * sg) + s.z[j]);
        }
        ustar *= -M_PI / 6.0;
        for (int j = 0; j < NB; ++j) {
            const double sg = s.sig[j], z = s.z[j], og = 1.0 + G * sg;
            mu[j] += -s.lB * (G * z * z / og + eta * sg * ((2.0 * z - eta * sg * sg) / og + eta * sg * sg / 3.0) +
                              z * ustar);
        }
        if (want_f) {
            double fm = G * G * G / (3.0 * M_PI);
            for (int j = 0; j < NB; ++j)
                fm -= s.lB * rho[j] * s.z[j] * (s.z[j] * G + eta * s.sig[j]) / (1.0 + G * s.sig[j]);
            f += fm;
        }
        if (F & SDFT_F_TPT1MSA) {
            for (int u = 0; u < MB; ++u) {
                const int i = s.bi[u], j = s.bj[u];
                const double sij = (s.sig[i] + s.sig[j]) / 2.0;
                const double Xi_ = (s.z[i] - eta * s.sig[i] * s.sig[i]) / (1.0 + G * s.sig[i]);
                const double Xj_ = (s.z[j] - eta * s.sig[j] * s.sig[j]) / (1.0 + G * s.sig[j]);
                lng[u] += -s.lB / sij * Xi_ * Xj_;
                if (want_f) f -= 0.5 * (rb0[u] + rb1[u]) * (-s.lB / sij * Xi_ * Xj_);
            }
            if (G != 0.0) {
                double dfG = 0.0, dfe = 0.0, deGs = 0.0, dHG = 0.0;
                for (int w = 0; w < NB; ++w) {
                    const double sg = s.sig[w], og = 1.0 + G * sg, t = s.z[w] - eta * sg * sg;
                    dfG += rho[w] * sg * t * t / (og * og * og);
                    dfe += rho[w] * sg * sg * t / (og * og);
                    deGs += rho[w] * s.z[w] * sg * sg / (og * og);
                    dHG += rho[w] * sg * sg * sg * sg / (og * og);
                }
                dfG *= -2.0 * M_PI * s.lB;
                dfe *= -2.0 * M_PI * s.lB;
                const double deG = -1.0 / H * deGs, deH = -eta / H;
                dHG = -dHG;
                const double detaG = deG + deH * dHG;
                const double denom = 1.0 / (2.0 * G - dfe * detaG - dfG);
                for (int w = 0; w < NB; ++w) {
                    const double ig = 1.0 / (1.0 + G * s.sig[w]);
                    const double Xw = (s.z[w] - eta * s.sig[w] * s.sig[w]) * ig;
                    const double dfr = M_PI * s.lB * Xw * Xw;
                    const double der = s.sig[w] / H * s.z[w] * ig;
                    const double dHr = s.sig[w] * s.sig[w] * s.sig[w] * (ig - 1.0 / 3.0);
                    const double dGr = (dfr + dfe * (der + dfe * deH * dHr)) * denom;
                    const double detr = detaG * dGr;
                    for (int u = 0; u < MB; ++u) {
                        const int i = s.bi[u], j = s.bj[u];
                        const double sij = (s.sig[i] + s.sig[j]) / 2.0;
                        const double igi = 1.0 / (1.0 + G * s.sig[i]);
                        const double Xi_ = (s.z[i] - eta * s.sig[i] * s.sig[i]) * igi;
                        const double dXi = -(s.sig[i] * s.sig[i] * igi * detr + s.sig[i] * Xi_ * igi * dGr);
                        const double igj = 1.0 / (1.0 + G * s.sig[j]);
                        const double Xj_ = (s.z[j] - eta * s.sig[j] * s.sig[j]) * igj;
                        const double dXj = -(s.sig[j] * s.sig[j] * igj * detr + s.sig[j] * Xj_ * igj * dGr);
                        mu[w] += (rb0[u] + rb1[u]) / 2.0 * s.lB / sij * (dXi * Xj_ + Xi_ * dXj);
                    }
                }
            }
        }
    }
    if (F & SDFT_F_SW) {
        // mean-field square well (src/iLST/free_energy/short_range_interactions/square_well.jl:64-150):
        // the double loop over the participating bead types enters an unlike pair twice
        for (int k = 0; k < s.n_sw; ++k) {
            const int i = s.swi[k], j = s.swj[k];
            const double ev = s.sweps[k] * s.swV[k];
            mu[i] -= rho[j] * ev;
            if (i != j) mu[j] -= rho[i] * ev;
            if (want_f) f -= (i != j ? 1.0 : 0.5) * rho[i] * rho[j] * ev;
        }
    }
    if (f_out) *f_out = f;
}

// Bulk state of one point at y = (mu_ex, lng): species densities (with the pSpecies activities and
// the charge balance), bead and bond densities; the recursion's workspace ws holds
// max_u (S_u + 1) N_u doubles of this lane. Returns false when no compensating species can absorb
// the net charge (balance_charge.jl).
struct BulkState {
    double rho[SDFT_MAX_SPECIES];
    double logXi[SDFT_MAX_SPECIES];
    double beads[SDFT_MAX_NB];
    double rb[SDFT_MAX_MB];
};

__device__ bool bulk_state(const BulkSysDev& s, const int32_t* __restrict__ seg_states, const double* dmu,
                           const double* base, const double* logact, const double* qch, const double* y,
                           double* ws, BulkState& st) {
    const int NB = s.NB, MB = s.MB, nsp = s.nsp;
    const double* mu = y;
    const double* lng = y + NB;
    double bead_u[SDFT_MAX_SPECIES][SDFT_MAX_NB];
    double bond_u[SDFT_MAX_SPECIES][SDFT_MAX_MB];
    for (int u = 0; u < nsp; ++u) {
        for (int j = 0; j < NB; ++j) bead_u[u][j] = 0.0;
        for (int k = 0; k < MB; ++k) bond_u[u][k] = 0.0;
        const int N = s.nseg[u], S = s.S[u];
        const int* sq = seg_states + s.soff[u];
        if (s.model[u] == SDFT_MODEL_MONOMERBEAD) {
            // monomerbead (src/iLST/species_model/monomerbead): Xi from the first segment's states
            double Xi = 0.0;
            for (int a = 0; a < S; ++a)
                if (sq[a] >= 0) Xi += exp(-mu[sq[a]] - dmu[sq[a]]);
            st.logXi[u] = log(Xi);
            for (int i = 0; i < N; ++i)
                for (int a = 0; a < S; ++a) {
                    const int si = sq[i * S + a];
                    if (si >= 0) bead_u[u][si] += exp(-mu[si] - dmu[si]) / Xi;
                }
            continue;
        }
        // child sweep from the free end (child_propagators.jl): Qc[a][p] = sum_b w(p,a; p+1,b) Qc[b][p+1],
        // w = exp(-dmuH_sj - mu_ex_sj + lng_bond); stored normalised by its largest state, the
        // logarithms of the factors summed in lq[p]
        double* Qs = ws;                 // [N][S]
        double* lq = ws + (int64_t)N * S;  // [N]
        for (int a = 0; a < S; ++a) Qs[(N - 1) * S + a] = 1.0;
        lq[N - 1] = 0.0;
        for (int p = N - 2; p >= 0; --p) {
            double mx = 0.0;
            for (int a = 0; a < S; ++a) {
                const int si = sq[p * S + a];
                double v = 0.0;
                if (si >= 0)
                    for (int b = 0; b < S; ++b) {
                        const int sj = sq[(p + 1) * S + b];
                        if (sj < 0) continue;
                        v += exp(-dmu[sj] - mu[sj] + lng[bulk_bond(s, si, sj)]) * Qs[(p + 1) * S + b];
                    }
                Qs[p * S + a] = v;
                mx = v > mx ? v : mx;
            }
            const double m = mx > 0.0 ? mx : 1.0;
            for (int a = 0; a < S; ++a) Qs[p * S + a] /= m;
            lq[p] = lq[p + 1] + log(m);
        }
        // partition function Xi = sum_a exp(-(mu_ex + dmuH)) Qc[a][0] (partition_function.jl)
        double xs = 0.0;
        for (int a = 0; a < S; ++a) {
            const int si = sq[a];
            if (si >= 0) xs += exp(-(mu[si] + dmu[si])) * Qs[a];
        }
        const double lXi = log(xs) + lq[0];
        st.logXi[u] = lXi;
        // parent sweep from the root, densities on the way (accumulate_densities.jl:32-96):
        // seg[a][i] = exp(-(mu_ex + dmuH)) Gp[a][i] Qc[a][i] / Xi; bond (i, i+1): seg times the share
        // gC[b][a][i] / Qc[a][i]
        double Gp[SDFT_MAX_S], Gn[SDFT_MAX_S];
        for (int a = 0; a < S; ++a) Gp[a] = 1.0;
        double lg = 0.0;
        for (int i = 0; i < N; ++i) {
            for (int a = 0; a < S; ++a) {
                const int si = sq[i * S + a];
                if (si < 0) continue;
                const double lval = -(mu[si] + dmu[si]) + log(Gp[a]) + lg + log(Qs[i * S + a]) + lq[i] - lXi;
                const double val = Gp[a] > 0.0 && Qs[i * S + a] > 0.0 ? exp(lval) : 0.0;
                bead_u[u][si] += val;
                if (i < N - 1 && val > 0.0) {
                    double tot = 0.0;
                    double gc[SDFT_MAX_S];
                    for (int b = 0; b < S; ++b) {
                        const int sj = sq[(i + 1) * S + b];
                        gc[b] = sj < 0 ? 0.0 : exp(-dmu[sj] - mu[sj] + lng[bulk_bond(s, si, sj)]) * Qs[(i + 1) * S + b];
                        tot += gc[b];
                    }
                    for (int b = 0; b < S; ++b) {
                        const int sj = sq[(i + 1) * S + b];
                        if (sj >= 0) bond_u[u][bulk_bond(s, si, sj)] += val * gc[b] / tot;
                    }
                }
            }
            if (i == N - 1) break;
            // Gp[b][i+1] = sum_a exp(-dmuH_si - mu_ex_si + lng_bond) Gp[a][i] (parent_propagators.jl)
            double mx = 0.0;
            for (int b = 0; b < S; ++b) {
                const int sj = sq[(i + 1) * S + b];
                double v = 0.0;
                if (sj >= 0)
                    for (int a = 0; a < S; ++a) {
                        const int si = sq[i * S + a];
                        if (si < 0) continue;
                        v += exp(-dmu[si] - mu[si] + lng[bulk_bond(s, si, sj)]) * Gp[a];
                    }
                Gn[b] = v;
                mx = v > mx ? v : mx;
            }
            const double m = mx > 0.0 ? mx : 1.0;
            for (int b = 0; b < S; ++b) Gp[b] = Gn[b] / m;
            lg += log(m);
        }
    }
    // species densities: input densities, or pSpecies activities rho = exp(ln c - pSpecies ln 10 + ln Xi)
    for (int u = 0; u < nsp; ++u) st.rho[u] = isnan(logact[u]) ? base[u] : exp(logact[u] + st.logXi[u]);
    // charge balance (balance_charge.jl:21-79): opposite-sign compensators absorb the net charge
    double R = 0.0;
    for (int u = 0; u < nsp; ++u) R += qch[u] * st.rho[u];
    if (R != 0.0) {
        const double sgn = R > 0.0 ? 1.0 : -1.0;
        int nadj = 0;
        for (int u = 0; u < nsp; ++u)
            if (s.adjustable[u] && ((qch[u] > 0.0) - (qch[u] < 0.0)) == -sgn) ++nadj;
        if (nadj == 0) return false;
        for (int u = 0; u < nsp; ++u)
            if (s.adjustable[u] && ((qch[u] > 0.0) - (qch[u] < 0.0)) == -sgn) {
                const double add = -R / nadj / qch[u];
                st.rho[u] = base[u] + (add > 0.0 ? add : 0.0);
            }
    }
    for (int j = 0; j < NB; ++j) st.beads[j] = 0.0;
    for (int k = 0; k < MB; ++k) st.rb[k] = 0.0;
    for (int u = 0; u < nsp; ++u) {
        for (int j = 0; j < NB; ++j) st.beads[j] += bead_u[u][j] * st.rho[u];
        for (int k = 0; k < MB; ++k) st.rb[k] += bond_u[u][k] * st.rho[u];
    }
    return true;
}

__global__ void __launch_bounds__(BULK_LANES) k_bulk_solve(BulkSysDev s, const int32_t* __restrict__ seg_states,
                                                          const double* __restrict__ dmu_all,
                                                          const double* __restrict__ base_all,
                                                          const double* __restrict__ logact_all,
                                                          const double* __restrict__ qch_all, double tol,
                                                          int max_newton, double* __restrict__ ws_all, int64_t ws_len,
                                                          BulkOutDev out) {
    const int p = blockIdx.x, lane = threadIdx.x;
    const int NB = s.NB, MB = s.MB, nsp = s.nsp, n = NB + MB;
    const double* dmu = dmu_all + (int64_t)p * NB;
    const double* base = base_all + (int64_t)p * nsp;
    const double* logact = logact_all + (int64_t)p * nsp;
    const double* qch = qch_all + (int64_t)p * nsp;
    double* ws = ws_all + ((int64_t)p * BULK_LANES + lane) * ws_len;
    __shared__ double y[SDFT_MAX_NB + SDFT_MAX_MB];
    __shared__ double gv[SDFT_MAX_NB + SDFT_MAX_MB + 1][SDFT_MAX_NB + SDFT_MAX_MB];
    __shared__ double J[SDFT_MAX_NB + SDFT_MAX_MB][SDFT_MAX_NB + SDFT_MAX_MB + 1];
    __shared__ int flag;   // 0 iterate, 1 converged, -1 charge-balance failure
    if (lane < n) y[lane] = 0.0;
    if (lane == 0) flag = 0;
    __syncthreads();
    int it = 0;
    for (; it <= max_newton; ++it) {
        if (lane <= n) {
            // lane k < n: residual at y + h_k e_k; lane n: at y (forward differences,
            // h_k = 1e-7 max(1, |y_k|))
            double yl[SDFT_MAX_NB + SDFT_MAX_MB];
            for (int k = 0; k < n; ++k) yl[k] = y[k];
            if (lane < n) yl[lane] += 1e-7 * fmax(1.0, fabs(y[lane]));
            BulkState st;
            if (!bulk_state(s, seg_states, dmu, base, logact, qch, yl, ws, st)) flag = -1;
            double mu[SDFT_MAX_NB], lng[SDFT_MAX_MB];
            double rb[SDFT_MAX_MB];
            for (int k = 0; k < MB; ++k) rb[k] = st.rb[k];
            bulk_fe(s, st.beads, rb, rb, mu, lng, nullptr, false);
            for (int k = 0; k < NB; ++k) gv[lane][k] = mu[k] - yl[k];
            for (int k = 0; k < MB; ++k) gv[lane][NB + k] = lng[k] - yl[NB + k];
        }
        __syncthreads();
        if (lane == 0 && flag == 0) {
            double gmax = 0.0;
            for (int k = 0; k < n; ++k) gmax = fmax(gmax, fabs(gv[n][k]));
            if (gmax < tol) {
                flag = 1;
            } else if (it < max_newton) {
                // Newton step: J dy = -g, Gaussian elimination with partial pivoting; a singular J
                // falls back to dy = -g
                for (int k = 0; k < n; ++k) {
                    const double h = 1e-7 * fmax(1.0, fabs(y[k]));
                    for (int r = 0; r < n; ++r) J[r][k] = (gv[k][r] - gv[n][r]) / h;
                }
                for (int r = 0; r < n; ++r) J[r][n] = -gv[n][r];
                bool singular = false;
                for (int c = 0; c < n && !singular; ++c) {
                    int piv = c;
                    for (int r = c + 1; r < n; ++r)
                        if (fabs(J[r][c]) > fabs(J[piv][c])) piv = r;
                    if (!(fabs(J[piv][c]) > 0.0)) {
                        singular = true;
                        break;
                    }
                    if (piv != c)
                        for (int k = c; k <= n; ++k) {
                            const double t = J[c][k];
                            J[c][k] = J[piv][k];
                            J[piv][k] = t;
                        }
                    for (int r = c + 1; r < n; ++r) {
                        const double fct = J[r][c] / J[c][c];
                        for (int k = c; k <= n; ++k) J[r][k] -= fct * J[c][k];
                    }
                }
                if (singular) {
                    for (int k = 0; k < n; ++k) y[k] += -gv[n][k];
                } else {
                    double dy[SDFT_MAX_NB + SDFT_MAX_MB];
                    for (int r = n - 1; r >= 0; --r) {
                        double v = J[r][n];
                        for (int k = r + 1; k < n; ++k) v -= J[r][k] * dy[k];
                        dy[r] = v / J[r][r];
                    }
                    for (int k = 0; k < n; ++k) y[k] += dy[k];
                }
            }
        }
        __syncthreads();
        if (flag != 0) break;
    }
    if (lane != 0) return;
    if (flag != 1) {
        out.iterations[p] = flag == -1 ? -2 : -1;
        return;
    }
    // converged: the state at y, its excess terms and the pressure (eval_pressure.jl:10-25)
    double yl[SDFT_MAX_NB + SDFT_MAX_MB];
    for (int k = 0; k < n; ++k) yl[k] = y[k];
    BulkState st;
    bulk_state(s, seg_states, dmu, base, logact, qch, yl, ws, st);
    double mu[SDFT_MAX_NB], lng[SDFT_MAX_MB], rb[SDFT_MAX_MB], f = 0.0;
    for (int k = 0; k < MB; ++k) rb[k] = st.rb[k];
    bulk_fe(s, st.beads, rb, rb, mu, lng, &f, true);
    double P = -f;
    for (int u = 0; u < nsp; ++u) P += st.rho[u];
    for (int j = 0; j < NB; ++j) P += st.beads[j] * mu[j];
    for (int k = 0; k < MB; ++k) P -= rb[k] * lng[k] / 2.0 + rb[k] * lng[k] / 2.0;
    for (int u = 0; u < nsp; ++u) {
        const double ideal = st.rho[u] > 1e-100 ? log(st.rho[u]) : -1e20;
        out.mu_species[(int64_t)p * nsp + u] = ideal - st.logXi[u];
        out.rho_species[(int64_t)p * nsp + u] = st.rho[u];
    }
    for (int j = 0; j < NB; ++j) {
        out.mu_ex[(int64_t)p * NB + j] = mu[j];
        out.rho_beads[(int64_t)p * NB + j] = st.beads[j];
    }
    for (int k = 0; k < MB; ++k) {
        out.lng[(int64_t)p * MB + k] = lng[k];
        out.rho_bonds[((int64_t)p * 2 + 0) * MB + k] = rb[k];
        out.rho_bonds[((int64_t)p * 2 + 1) * MB + k] = rb[k];
    }
    out.pressure[p] = P;
    out.iterations[p] = it;
}

void launch_bulk_solve(hipStream_t st, int P, const BulkSysDev& s, const int32_t* seg_states, const double* dmu,
                       const double* base, const double* logact, const double* qch, double tol, int max_newton,
                       double* ws, int64_t ws_len, const BulkOutDev& out) {
    hipLaunchKernelGGL(k_bulk_solve, dim3(P), dim3(BULK_LANES), 0, st, s, seg_states, dmu, base, logact, qch, tol,
                       max_newton, ws, ws_len, out);
}

}  // namespace sdft
