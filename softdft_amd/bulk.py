"""Bulk (homogeneous) Ising liquid-state state point that seeds every SCF solve.

Host-side restatement of the reference's iLST bulk solve.  Outputs the
quantities the SCF path consumes: per-species chemical potentials mu_species,
per-bead excess chemical potentials mu_ex, per-bond ln y (lng), bulk bead and
bond densities, and the bulk pressure.

Reference: src/iLST/setup/setup_steps/initialize_densities.jl:31-121 (initial
densities, ideal species charge), src/iLST/evaluation/objective/
{solver_bulkstate.jl:31-104, residual_bulkstate.jl:26-42,
functions/balance_charge.jl:21-79, functions/eval_residuals.jl:26-60},
src/iLST/species_model/freely_jointed/*.jl (bulk propagators),
src/iLST/evaluation/analysis/eval_pressure.jl:10-25.

The reference drives the residuals to 1e-10 with Anderson mixing on
(log rho, Q); here the same equations are solved to ~1e-14 by eliminating
the pSpecies densities and the charge variable in closed form and running
Newton on the remaining (mu_ex, lng) self-consistency.  Only the converged
state is shared with the reference, not the iteration path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

from .molsys import LOG10, MolSys, SetupError

PI = math.pi


# ------------------------------------------------------------ pointwise FE
# src/iLST/free_energy/excluded_volume/mfmt.jl:143-300

def free_energy_mfmt(n0, n1, n2, n3, dot12=0.0, dot22=0.0):
    om = 1.0 - n3
    lg = math.log(om)
    inv_den = 1.0 / (36.0 * PI * n3 ** 2 * om ** 2)
    return -n0 * lg + (n1 * n2 - dot12) / om + (n2 ** 3 - 3.0 * n2 * dot22) * (n3 + om ** 2 * lg) * inv_den


def mfmt_deriv(n0, n1, n2, n3):
    om = 1.0 - n3
    lg = math.log(om)
    i1 = 1.0 / om
    i3 = 1.0 / n3
    c = 1.0 / (36.0 * PI)
    d = np.zeros(4)
    d[0] = -lg
    d[1] = n2 * i1
    d[2] = n1 * i1 + n2 ** 2 * 3.0 * c * i3 * (lg * i3 + i1 ** 2)
    d[3] = (n0 * i1 + n1 * n2 * i1 ** 2
            - (lg * i3 ** 3 * 2.0 * c + i3 ** 2 * i1 * c + (1.0 - 3.0 * n3) * i3 ** 2 * c * i1 ** 3) * n2 ** 3)
    return d


def gamma_msa(rho, sig, Z, lB):
    """Gamma_msa (src/iLST/free_energy/electrostaticcorrelations/msa.jl:181-230), same damped iteration."""
    tol = 1e-10
    g1 = 0.0
    kap = math.sqrt(4.0 * PI * lB * float(np.sum(rho * Z * Z)))
    if kap < tol:
        return kap / 2.0, 0.0, 2.0 / PI
    seff = float(np.sum(rho * sig)) / float(np.sum(rho))
    x = seff * math.sqrt(4.0 * PI * lB * float(np.sum(rho * Z * Z)))
    G = (math.sqrt(1.0 + 2.0 * x) - 1.0) / (2.0 * seff)
    term2 = 2.0 / PI * (1.0 - PI / 6.0 * float(np.sum(rho * sig ** 3)))
    it = 0
    while True:
        H = float(np.sum(rho * sig ** 3 / (1.0 + G * sig))) + term2
        eta = float(np.sum(rho * Z * sig / (1.0 + G * sig))) / H
        g1 = math.sqrt(PI * lB * float(np.sum(rho * (Z - eta * sig ** 2) ** 2 / (1.0 + G * sig) ** 2)))
        gn = 0.9 * G + 0.1 * g1
        it += 1
        if abs(g1 - G) < tol or it >= 1000:
            G = gn
            break
        G = gn
    if it == 1000 and math.isnan(abs(g1 - G)):
        raise SetupError("Gamma is NaN")
    return G, eta, H


def msa_deriv(rho, sig, Z, lB):
    G, eta, _ = gamma_msa(rho, sig, Z, lB)
    ustar = -PI / 6.0 * float(np.sum(rho * sig * sig * ((Z - eta * sig * sig) / (1.0 + G * sig) + Z)))
    return -lB * (G * Z ** 2 / (1.0 + G * sig) + eta * sig * ((2.0 * Z - eta * sig ** 2) / (1.0 + G * sig)
                  + eta * sig ** 2 / 3.0) + Z * ustar)


def free_energy_msa(rho, sig, Z, lB):
    G, eta, _ = gamma_msa(rho, sig, Z, lB)
    return G ** 3 / (3.0 * PI) - lB * float(np.sum(rho * Z * (Z * G + eta * sig) / (1.0 + G * sig)))


def tpt1msa_deriv(rho_bonds, bond_types, rho, sig, Z, lB):
    """compute_TPT1msa_deriv! (src/iLST/free_energy/electrostaticcorrelations/TPT1msa.jl:138-282)."""
    dphi = msa_deriv(rho, sig, Z, lB)
    G, eta, H = gamma_msa(rho, sig, Z, lB)
    lng = np.zeros(len(bond_types))
    if G == 0.0:
        for u, (i, j) in enumerate(bond_types):
            lng[u] = -lB / ((sig[i] + sig[j]) / 2.0) * Z[i] * Z[j]
        return dphi, lng
    for u, (i, j) in enumerate(bond_types):
        sij = (sig[i] + sig[j]) / 2.0
        Xi_ = (Z[i] - eta * sig[i] ** 2) / (1.0 + G * sig[i])
        Xj_ = (Z[j] - eta * sig[j] ** 2) / (1.0 + G * sig[j])
        lng[u] = -lB / sij * Xi_ * Xj_
    opg = 1.0 + G * sig
    dfG = -2.0 * PI * lB * float(np.sum(rho * sig * (Z - eta * sig * sig) ** 2 / opg ** 3))
    dfe = -2.0 * PI * lB * float(np.sum(rho * sig * sig * (Z - eta * sig * sig) / opg ** 2))
    deG = -1.0 / H * float(np.sum(rho * Z * sig * sig / opg ** 2))
    deH = -eta / H
    dHG = -float(np.sum(rho * sig ** 4 / opg ** 2))
    detaG = deG + deH * dHG
    denom = 1.0 / (2.0 * G - dfe * detaG - dfG)
    for w in range(len(rho)):
        ig = 1.0 / (1.0 + G * sig[w])
        Xw = (Z[w] - eta * sig[w] ** 2) * ig
        dfr = PI * lB * Xw ** 2
        der = sig[w] / H * Z[w] * ig
        dHr = sig[w] ** 3 * (ig - 1.0 / 3.0)
        dGr = (dfr + dfe * (der + dfe * deH * dHr)) * denom
        detr = detaG * dGr
        for u, (i, j) in enumerate(bond_types):
            sij = (sig[i] + sig[j]) / 2.0
            igi = 1.0 / (1.0 + G * sig[i])
            Xi_ = (Z[i] - eta * sig[i] ** 2) * igi
            dXi = -(sig[i] ** 2 * igi * detr + sig[i] * Xi_ * igi * dGr)
            igj = 1.0 / (1.0 + G * sig[j])
            Xj_ = (Z[j] - eta * sig[j] ** 2) * igj
            dXj = -(sig[j] ** 2 * igj * detr + sig[j] * Xj_ * igj * dGr)
            dphi[w] += (rho_bonds[0, u] + rho_bonds[1, u]) / 2.0 * lB / sij * (dXi * Xj_ + Xi_ * dXj)
    return dphi, lng


def free_energy_tpt1msa(rho_bonds, bond_types, rho, sig, Z, lB):
    f = free_energy_msa(rho, sig, Z, lB)
    G, eta, _ = gamma_msa(rho, sig, Z, lB)
    for u, (i, j) in enumerate(bond_types):
        sij = (sig[i] + sig[j]) / 2.0
        Xi_ = (Z[i] - eta * sig[i] ** 2) / (1.0 + G * sig[i])
        Xj_ = (Z[j] - eta * sig[j] ** 2) / (1.0 + G * sig[j])
        f -= 0.5 * (rho_bonds[0, u] + rho_bonds[1, u]) * (-lB / sij * Xi_ * Xj_)
    return f


def sw_volumes(ms: MolSys) -> np.ndarray:
    """V_ij = 4 pi / 3 sigma_ij^3 (lambda_ij^3 - 1) per square-well pair
    (src/iLST/free_energy/short_range_interactions/square_well.jl:73-74)."""
    sw, sig = ms.sw, ms.diameters
    return np.array([4.0 * PI / 3.0 * ((sig[i] + sig[j]) / 2.0) ** 3 * (lam ** 3 - 1.0)
                     for (i, j), lam in zip(sw.pairs, sw.lambdas)])


def free_energy_sw(ms: MolSys, rho) -> float:
    """Mean-field square-well free energy density (square_well.jl:64-81): the double loop runs over
    the participating bead types, so each unlike pair enters twice with the factor 1/2."""
    sw, V = ms.sw, sw_volumes(ms)
    f = 0.0
    for i in sw.species:
        for j in sw.species:
            k = sw.index(i, j)
            if k is not None:
                f -= 0.5 * rho[i] * rho[j] * sw.energys[k] * V[k]
    return f


def sw_deriv(ms: MolSys, rho) -> np.ndarray:
    """d f_sw / d rho_i (compute_sw_deriv!, square_well.jl:129-150)."""
    sw, V = ms.sw, sw_volumes(ms)
    d = np.zeros(ms.NB)
    for i in sw.species:
        for j in sw.species:
            k = sw.index(i, j)
            if k is not None:
                d[i] -= rho[j] * sw.energys[k] * V[k]
    return d


def _fmt_scalars(rho, sig):
    return rho, rho * sig / 2, rho * PI * sig ** 2, rho * PI * sig ** 3 / 6


def bulk_fe_terms(ms: MolSys, rho_beads, rho_bonds):
    """(mu_ex[NB], lng[MB], f_ex) summed over the excess models (bulk_chemical_potential.jl:8-24)."""
    sig, Z, lB, bt = ms.diameters, ms.valences, ms.bjerrum, ms.bond_types
    mu = np.zeros(ms.NB)
    lng = np.zeros(ms.MB)
    f = 0.0
    for m in ms.fe_model:
        if m == "ideal":
            continue
        if m in ("mfmt", "TPT1mfmt"):
            n0i, n1i, n2i, n3i = _fmt_scalars(rho_beads, sig)
            n0, n1, n2, n3 = n0i.sum(), n1i.sum(), n2i.sum(), n3i.sum()
            d = mfmt_deriv(n0, n1, n2, n3)
            dd = np.tile(d[:, None], (1, ms.NB))
            f += free_energy_mfmt(n0, n1, n2, n3)
            if m == "TPT1mfmt":
                # src/iLST/free_energy/excluded_volume/TPT1mfmt.jl:398-445 (bulk: Xi = 1)
                i1 = 1.0 / (1.0 - n3)
                t2 = n2 / 4.0 * i1 ** 2
                t3 = n2 ** 2 / 72.0 * i1 ** 3
                for u, (i, j) in enumerate(bt):
                    sij = (sig[i] + sig[j]) / 2
                    tau = sig[i] * sig[j] / sij
                    y = i1 + t2 * tau + t3 * tau ** 2
                    dn2 = (tau / 4.0 * i1 ** 2 + n2 * tau ** 2 * i1 ** 3 / 36.0) / y
                    dn3 = (i1 ** 2 + n2 / 2.0 * i1 ** 3 * tau + n2 ** 2 / 24.0 * i1 ** 4 * tau ** 2) / y
                    rb = rho_bonds[0, u] + rho_bonds[1, u]
                    dd[2, :] -= 0.5 * rb * dn2
                    dd[3, :] -= 0.5 * rb * dn3
                    lng[u] += math.log(y)
                    f -= 0.5 * rb * math.log(y)
            mu += dd[0] + dd[1] * sig / 2 + dd[2] * PI * sig ** 2 + dd[3] * PI * sig ** 3 / 6
        elif m == "msa":
            mu += msa_deriv(rho_beads, sig, Z, lB)
            f += free_energy_msa(rho_beads, sig, Z, lB)
        elif m == "TPT1msa":
            d, l = tpt1msa_deriv(rho_bonds, bt, rho_beads, sig, Z, lB)
            mu += d
            lng += l
            f += free_energy_tpt1msa(rho_bonds, bt, rho_beads, sig, Z, lB)
        elif m == "sw":
            mu += sw_deriv(ms, rho_beads)
            f += free_energy_sw(ms, rho_beads)
    return mu, lng, f


# ------------------------------------------------------------ bulk chains

def _lse(v) -> float:
    v = np.asarray(v, dtype=np.float64)
    m = float(np.max(v)) if v.size else -math.inf
    if m == -math.inf:
        return -math.inf
    return m + math.log(float(np.sum(np.exp(v - m))))


def _chain_bulk_log(ms: MolSys, u: int, mu_ex, lng):
    """The recursion of _chain_bulk carried out on logarithms of the propagators, for chains whose
    propagator products overflow a double (500-mers at high pH): (log Xi, rho_seg, bonds). The
    reference evaluates the same products directly and returns Inf/NaN there."""
    sp = ms.species[u]
    fam, ch, pa, lv = sp.state_family, sp.topology.children, sp.topology.parents, sp.topology.levels
    dmu = ms.delta_muH
    N = sp.n_seg
    S = max(len(f) for f in fam)
    C = max(max((len(c) for c in ch), default=0), 1)
    lC = np.full((S, C, S, N), -math.inf)
    for level in reversed(lv):
        for p in level:
            for a, si in enumerate(fam[p]):
                lC[0, 0, a, p] = 0.0
                for c_idx, c in enumerate(ch[p]):
                    for b, sj in enumerate(fam[c]):
                        bond = ms.bond_types.index((min(si, sj), max(si, sj)))
                        v = -dmu[sj] - mu_ex[sj] + lng[bond]
                        for gci, _ in enumerate(ch[c]):
                            v += _lse(lC[:, gci, b, c])
                        lC[b, c_idx, a, p] = v
    lP = np.full((S, S, N), -math.inf)
    for level in lv:
        for c in level:
            for b, sj in enumerate(fam[c]):
                lP[0, b, c] = 0.0
                for p in pa[c]:
                    for a, si in enumerate(fam[p]):
                        bond = ms.bond_types.index((min(si, sj), max(si, sj)))
                        v = -dmu[si] - mu_ex[si] + lng[bond] + _lse(lP[:, a, p])
                        for oci, oc in enumerate(ch[p]):
                            if oc != c:
                                v += _lse(lC[:, oci, a, p])
                        lP[a, b, c] = v
    terms = []
    for a, si in enumerate(fam[0]):
        t = -(mu_ex[si] + dmu[si]) + _lse(lP[:, a, 0])
        for ci, _ in enumerate(ch[0]):
            t += _lse(lC[:, ci, a, 0])
        terms.append(t)
    logXi = _lse(terms)
    seg = np.zeros((S, N))
    bonds = np.zeros(ms.MB)
    for i in range(N):
        for a, si in enumerate(fam[i]):
            lpref = -(mu_ex[si] + dmu[si]) + _lse(lP[:, a, i]) - logXi
            if ch[i]:
                c0 = ch[i][0]
                cc = []
                for b, _ in enumerate(fam[c0]):
                    t = lC[b, 0, a, i]
                    for oci, oc in enumerate(ch[i]):
                        if oc != c0:
                            t += _lse(lC[:, oci, a, i])
                    cc.append(t)
                lcc = _lse(cc)
            else:
                lcc = 0.0
            lval = lpref + lcc
            seg[a, i] = math.exp(lval)
            for ci, c in enumerate(ch[i]):
                lQ = _lse(lC[:, ci, a, i])
                for b, sj in enumerate(fam[c]):
                    bonds[ms.bond_types.index((min(si, sj), max(si, sj)))] += math.exp(lval + lC[b, ci, a, i] - lQ)
    return logXi, seg, bonds


def _chain_bulk(ms: MolSys, u: int, mu_ex, lng):
    """(Xi, rho_seg[S,N], bond contributions [MB], log Xi) per unit species density;
    src/iLST/species_model/freely_jointed/{child,parent}_propagators.jl,
    partition_function.jl, accumulate_densities.jl:32-96. Products that overflow are redone on
    logarithms (_chain_bulk_log), which agrees with the direct form wherever that is finite."""
    with np.errstate(over="ignore", invalid="ignore"):
        Xi, seg, bonds = _chain_bulk_direct(ms, u, mu_ex, lng)
    if math.isfinite(Xi) and Xi > 0 and np.all(np.isfinite(seg)) and np.all(np.isfinite(bonds)):
        return Xi, seg, bonds, math.log(Xi)
    logXi, seg, bonds = _chain_bulk_log(ms, u, mu_ex, lng)
    return math.exp(logXi) if logXi < 709.0 else math.inf, seg, bonds, logXi


def _chain_bulk_direct(ms: MolSys, u: int, mu_ex, lng):
    sp = ms.species[u]
    fam, ch, pa, lv = sp.state_family, sp.topology.children, sp.topology.parents, sp.topology.levels
    dmu = ms.delta_muH
    N = sp.n_seg
    S = max(len(f) for f in fam)
    if sp.model == "monomerbead":
        Xi = sum(math.exp(-mu_ex[s] - dmu[s]) for s in fam[0])
        seg = np.zeros((S, N))
        for i in range(N):
            for a, s in enumerate(fam[i]):
                seg[a, i] += math.exp(-mu_ex[s] - dmu[s]) / Xi
        return Xi, seg, np.zeros(ms.MB)
    C = max(max((len(c) for c in ch), default=0), 1)
    gC = np.zeros((S, C, S, N))
    for level in reversed(lv):
        for p in level:
            for a, si in enumerate(fam[p]):
                gC[0, 0, a, p] = 1.0
                for c_idx, c in enumerate(ch[p]):
                    for b, sj in enumerate(fam[c]):
                        bond = ms.bond_types.index((min(si, sj), max(si, sj)))
                        v = math.exp(-dmu[sj] - mu_ex[sj] + lng[bond])
                        for gci, _ in enumerate(ch[c]):
                            v *= gC[:, gci, b, c].sum()
                        gC[b, c_idx, a, p] = v
    gP = np.zeros((S, S, N))
    for level in lv:
        for c in level:
            for b, sj in enumerate(fam[c]):
                gP[0, b, c] = 1.0
                for p in pa[c]:
                    for a, si in enumerate(fam[p]):
                        bond = ms.bond_types.index((min(si, sj), max(si, sj)))
                        v = math.exp(-dmu[si] - mu_ex[si] + lng[bond]) * gP[:, a, p].sum()
                        for oci, oc in enumerate(ch[p]):
                            if oc != c:
                                v *= gC[:, oci, a, p].sum()
                        gP[a, b, c] = v
    Xi = 0.0
    for a, si in enumerate(fam[0]):
        t = math.exp(-(mu_ex[si] + dmu[si])) * gP[:, a, 0].sum()
        for ci, _ in enumerate(ch[0]):
            t *= gC[:, ci, a, 0].sum()
        Xi += t
    seg = np.zeros((S, N))
    bonds = np.zeros(ms.MB)
    for i in range(N):
        for a, si in enumerate(fam[i]):
            pref = math.exp(-(mu_ex[si] + dmu[si])) * gP[:, a, i].sum() / Xi
            if ch[i]:
                c0 = ch[i][0]
                cc = 0.0
                for b, _ in enumerate(fam[c0]):
                    t = gC[b, 0, a, i]
                    for oci, oc in enumerate(ch[i]):
                        if oc != c0:
                            t *= gC[:, oci, a, i].sum()
                    cc += t
            else:
                cc = 1.0
            val = pref * cc
            seg[a, i] = val
            for ci, c in enumerate(ch[i]):
                Qc = gC[:, ci, a, i].sum()
                for b, sj in enumerate(fam[c]):
                    fval = val * gC[b, ci, a, i] / Qc
                    bonds[ms.bond_types.index((min(si, sj), max(si, sj)))] += fval
    return Xi, seg, bonds


@dataclass
class BulkState:
    rho_species: np.ndarray   # per species (chain) density
    rho_segments: List[np.ndarray]
    rho_beads: np.ndarray
    rho_bonds: np.ndarray     # [2, MB]
    mu_species: np.ndarray
    mu_ex: np.ndarray
    lng: np.ndarray
    Xi: np.ndarray
    pressure: float
    species_charge: np.ndarray


def _ideal_species_charge(ms: MolSys) -> np.ndarray:
    """Per-chain charge from delta_muH only (initialize_densities.jl:49,84)."""
    out = np.zeros(len(ms.species))
    for u, sp in enumerate(ms.species):
        for fam in sp.state_family:
            Q = sum(math.exp(-ms.delta_muH[s]) for s in fam)
            for s in fam:
                out[u] += math.exp(-ms.delta_muH[s]) / Q * ms.valences[s]
    return out


def solve_bulk(ms: MolSys, tol: float = 1e-14, max_newton: int = 100) -> BulkState:
    nsp = len(ms.species)
    unit_conv = ms.constants.concentration_to_density(1.0)
    qch = _ideal_species_charge(ms)
    ms.species_charge = qch
    base = np.zeros(nsp)
    for u, sp in enumerate(ms.species):
        if not math.isnan(sp.pSpecies):
            base[u] = ms.constants.concentration_to_density(10.0 ** (-sp.pSpecies)) / sp.n_seg
        else:
            base[u] = sp.input_density / sp.n_seg

    def species_densities(mu_ex, lng):
        rho = base.copy()
        Xi = np.zeros(nsp)
        segs, bonds = [], []
        logXi = np.zeros(nsp)
        for u in range(nsp):
            X, seg, b, lX = _chain_bulk(ms, u, mu_ex, lng)
            Xi[u] = X
            logXi[u] = lX
            segs.append(seg)
            bonds.append(b)
        for u, sp in enumerate(ms.species):
            if not math.isnan(sp.pSpecies):
                # mu_species = ln rho - ln Xi = ln(unit_conv) - pSpecies ln10
                rho[u] = math.exp(math.log(unit_conv) - sp.pSpecies * LOG10 + logXi[u])
        # charge balance (balance_charge.jl:21-79): Q_solver equals the net charge with
        # compensators at their input densities; opposite-sign compensators absorb it.
        comp = [sp.adjustable for sp in ms.species]
        R = float(np.sum(qch * rho))
        if R != 0.0:
            sgn = 1.0 if R > 0 else -1.0
            adj = [u for u in range(nsp) if comp[u] and np.sign(qch[u]) == -sgn]
            if not adj:
                raise SetupError(f"No compensating species available to absorb net charge {R}")
            for u in adj:
                rho[u] = base[u] + max(-R / len(adj) / qch[u], 0.0)
        return rho, Xi, segs, bonds, logXi

    def state(y):
        mu_ex, lng = y[:ms.NB], y[ms.NB:]
        rho, Xi, segs, bonds, logXi = species_densities(mu_ex, lng)
        beads = np.zeros(ms.NB)
        rb = np.zeros((2, ms.MB))
        rseg = []
        for u, sp in enumerate(ms.species):
            s = segs[u] * rho[u]
            rseg.append(s)
            for i in range(sp.n_seg):
                for a, st in enumerate(sp.state_family[i]):
                    beads[st] += s[a, i]
            rb[0] += bonds[u] * rho[u]
            rb[1] += bonds[u] * rho[u]
        return rho, Xi, rseg, beads, rb, logXi

    def G(y):
        _, _, _, beads, rb, _ = state(y)
        mu, lng, _ = bulk_fe_terms(ms, beads, rb)
        return np.concatenate([mu, lng]) - y

    n = ms.NB + ms.MB
    y = np.zeros(n)
    for _ in range(max_newton):
        g = G(y)
        if np.max(np.abs(g)) < tol:
            break
        J = np.zeros((n, n))
        for k in range(n):
            h = 1e-7 * max(1.0, abs(y[k]))
            yp = y.copy()
            yp[k] += h
            J[:, k] = (G(yp) - g) / h
        try:
            dy = np.linalg.solve(J, -g)
        except np.linalg.LinAlgError:
            dy = -g
        y = y + dy
    else:
        raise SetupError("bulk state did not converge")
    rho, Xi, rseg, beads, rb, logXi = state(y)
    mu_ex, lng, f_ex = bulk_fe_terms(ms, beads, rb)
    mu_sp = np.zeros(nsp)
    for u in range(nsp):
        ideal = math.log(rho[u]) if rho[u] > 1e-100 else -1e20
        mu_sp[u] = ideal - logXi[u]
    P = float(np.sum(rho) - f_ex + np.sum(beads * mu_ex) - np.sum(rb[0] * lng) / 2 - np.sum(rb[1] * lng) / 2)
    return BulkState(rho_species=rho, rho_segments=rseg, rho_beads=beads, rho_bonds=rb,
                     mu_species=mu_sp, mu_ex=mu_ex, lng=lng, Xi=Xi, pressure=P, species_charge=qch)
