"""Batched bulk states through libsoftdft_hip's sdft_bulk_solve.

The device restatement of the iLST bulk solve (src/iLST/evaluation/objective/solver_bulkstate.jl:31-103,
eval_pressure.jl:10-25) for every state point of a batch in one launch, the batched counterpart of
`bulk.solve_bulk`.  The host still builds each point's molecular system (delta_muH from pH / pK,
state families, topology); the per-point inputs the device needs follow from it exactly as in
`bulk.solve_bulk`: the ideal (delta_muH-only) charge per chain (initialize_densities.jl:49,84), the
chain densities before the charge balance, and ln(c 10^-pSpecies) for species given by pSpecies.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List

import numpy as np

from . import abi
from .bulk import BulkState, _ideal_species_charge, sw_volumes
from .molsys import LOG10, MolSys, SetupError

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class SdftBulkDesc(C.Structure):
    _fields_ = [("abi_version", C.c_int32), ("n_points", C.c_int32), ("NB", C.c_int32), ("MB", C.c_int32),
                ("n_species", C.c_int32), ("functionals", C.c_int32), ("bjerrum", C.c_double),
                ("diameters", _f64p), ("valences", _f64p), ("bond_types", _i32p),
                ("species_nseg", _i32p), ("species_model", _i32p), ("species_nstates", _i32p),
                ("seg_states", _i32p), ("adjustable", _i32p),
                ("delta_muH", _f64p), ("base_density", _f64p), ("log_activity", _f64p), ("species_charge", _f64p),
                ("n_sw_pairs", C.c_int32), ("sw_pairs", _i32p), ("sw_energy", _f64p), ("sw_volume", _f64p)]


class SdftBulkOut(C.Structure):
    _fields_ = [("mu_species", _f64p), ("rho_species", _f64p), ("mu_ex", _f64p), ("rho_beads", _f64p),
                ("lng", _f64p), ("rho_bonds", _f64p), ("pressure", _f64p), ("iterations", _i32p)]


def _lib():
    lib = abi.load_library()
    if not getattr(lib, "_bulk_bound", False):
        lib.sdft_bulk_solve.restype = C.c_int32
        lib.sdft_bulk_solve.argtypes = [C.POINTER(SdftBulkDesc), C.c_double, C.c_int32, C.POINTER(SdftBulkOut)]
        lib._bulk_bound = True
    return lib


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def solve_bulk_batch(mss: List[MolSys], tol: float = 1e-13, max_newton: int = 100) -> List[BulkState]:
    """Bulk state of every molecular system (all sharing one layout) in one device launch. Raises
    abi.SdftError on failure (no HIP device, no convergence, no compensating species). The returned
    states carry what the SCF setup and the analysis read (mu_species, mu_ex, lng, densities,
    pressure); the per-segment bulk densities and Xi, which nothing downstream reads, are not
    returned (rho_segments empty, Xi NaN)."""
    if not mss:
        return []
    m0 = mss[0]
    for ms in mss[1:]:
        if (ms.monomers != m0.monomers or ms.bond_types != m0.bond_types or len(ms.species) != len(m0.species)
                or ms.fe_model != m0.fe_model or not np.array_equal(ms.diameters, m0.diameters)
                or not np.array_equal(ms.valences, m0.valences) or ms.bjerrum != m0.bjerrum or ms.sw != m0.sw
                or [s.sequence for s in ms.species] != [s.sequence for s in m0.species]
                or [s.state_family for s in ms.species] != [s.state_family for s in m0.species]):
            raise ValueError("molecular systems of one bulk batch must share their layout")
    lib = _lib()
    P, NB, MB, nsp = len(mss), m0.NB, m0.MB, len(m0.species)
    keep = {}
    keep["diam"] = np.ascontiguousarray(m0.diameters, dtype=np.float64)
    keep["val"] = np.ascontiguousarray(m0.valences, dtype=np.float64)
    keep["bonds"] = np.ascontiguousarray(np.array(m0.bond_types, dtype=np.int32).reshape(-1))
    nseg, model, nst, states = [], [], [], []
    for u, s in enumerate(m0.species):
        for q in range(s.n_seg):
            if len(s.topology.children[q]) > 1:
                raise abi.SdftError(-5, f"species {u}: branched chains are outside this build")
        S = max(len(f) for f in s.state_family)
        nseg.append(s.n_seg)
        model.append(0 if s.model == "freely_jointed" else 1)
        nst.append(S)
        for f in s.state_family:
            states.extend(list(f) + [-1] * (S - len(f)))
    keep["nseg"] = np.array(nseg, dtype=np.int32)
    keep["model"] = np.array(model, dtype=np.int32)
    keep["nst"] = np.array(nst, dtype=np.int32)
    keep["states"] = np.array(states, dtype=np.int32)
    keep["adj"] = np.array([1 if s.adjustable else 0 for s in m0.species], dtype=np.int32)
    dmu = np.zeros((P, NB))
    base = np.zeros((P, nsp))
    lact = np.full((P, nsp), np.nan)
    qch = np.zeros((P, nsp))
    for k, ms in enumerate(mss):
        dmu[k] = ms.delta_muH
        q = _ideal_species_charge(ms)
        ms.species_charge = q
        qch[k] = q
        unit = ms.constants.concentration_to_density(1.0)
        for u, s in enumerate(ms.species):
            if not math.isnan(s.pSpecies):
                base[k, u] = ms.constants.concentration_to_density(10.0 ** (-s.pSpecies)) / s.n_seg
                lact[k, u] = math.log(unit) - s.pSpecies * LOG10
            else:
                base[k, u] = s.input_density / s.n_seg
    keep["dmu"], keep["base"], keep["lact"], keep["qch"] = dmu, base, lact, qch
    d = SdftBulkDesc()
    d.abi_version = abi.SDFT_ABI_VERSION
    d.n_points, d.NB, d.MB, d.n_species = P, NB, MB, nsp
    d.functionals = abi.functional_flags([m for m in m0.fe_model if m in ("mfmt", "TPT1mfmt", "msa", "TPT1msa", "sw")])
    d.bjerrum = m0.bjerrum
    d.diameters, d.valences, d.bond_types = _p(keep["diam"], C.c_double), _p(keep["val"], C.c_double), \
        _p(keep["bonds"], C.c_int32)
    d.species_nseg, d.species_model = _p(keep["nseg"], C.c_int32), _p(keep["model"], C.c_int32)
    d.species_nstates, d.seg_states = _p(keep["nst"], C.c_int32), _p(keep["states"], C.c_int32)
    d.adjustable = _p(keep["adj"], C.c_int32)
    d.delta_muH, d.base_density = _p(dmu, C.c_double), _p(base, C.c_double)
    d.log_activity, d.species_charge = _p(lact, C.c_double), _p(qch, C.c_double)
    if m0.sw is not None and "sw" in m0.fe_model:
        # first occurrence of each pair, as sw.index finds it in the bulk double loop
        seen, pairs, eps, vol = set(), [], [], []
        V = sw_volumes(m0)
        for k, pr in enumerate(m0.sw.pairs):
            key = (min(pr), max(pr))
            if key in seen:
                continue
            seen.add(key)
            pairs.extend(key)
            eps.append(m0.sw.energys[k])
            vol.append(V[k])
        keep["swp"] = np.array(pairs, dtype=np.int32)
        keep["swe"] = np.array(eps, dtype=np.float64)
        keep["swv"] = np.array(vol, dtype=np.float64)
        d.n_sw_pairs = len(eps)
        d.sw_pairs, d.sw_energy, d.sw_volume = _p(keep["swp"], C.c_int32), _p(keep["swe"], C.c_double), \
            _p(keep["swv"], C.c_double)
    out = {name: np.zeros(shape) for name, shape in (("mu_species", (P, nsp)), ("rho_species", (P, nsp)),
                                                     ("mu_ex", (P, NB)), ("rho_beads", (P, NB)), ("lng", (P, MB)),
                                                     ("rho_bonds", (P, 2, MB)), ("pressure", (P,)))}
    its = np.zeros(P, dtype=np.int32)
    o = SdftBulkOut(*[_p(out[n], C.c_double) for n in ("mu_species", "rho_species", "mu_ex", "rho_beads", "lng",
                                                      "rho_bonds", "pressure")], _p(its, C.c_int32))
    rc = lib.sdft_bulk_solve(C.byref(d), tol, max_newton, C.byref(o))
    if rc != 0:
        raise abi.SdftError(rc, lib.sdft_last_error(None).decode())
    res = []
    for k in range(P):
        res.append(BulkState(rho_species=out["rho_species"][k].copy(), rho_segments=[],
                             rho_beads=out["rho_beads"][k].copy(), rho_bonds=out["rho_bonds"][k].copy(),
                             mu_species=out["mu_species"][k].copy(), mu_ex=out["mu_ex"][k].copy(),
                             lng=out["lng"][k].copy(), Xi=np.full(nsp, np.nan), pressure=float(out["pressure"][k]),
                             species_charge=qch[k].copy()))
    return res


__all__ = ["solve_bulk_batch", "SetupError"]
