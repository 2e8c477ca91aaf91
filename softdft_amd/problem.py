"""One SCF state point: grid, walls, grafting, bulk state and initial guess.

Restates the reference's inhomogeneous setup for 1-D Cartesian slits:

* NP = cld(H, dz) + 1, evaluated in exact rational arithmetic
  (src/geometry/setup/coordinate_system/Cartesian/create_cartesian.jl:66);
  the calculation domain is the even (mirrored) extension of length 2*NP with
  zero offset (process_cartesian.jl:20-31, iDFT/.../utilities/mirrored.jl:58-62);
* hard walls: Ext = 1e20 outside [lo, hi], trapez = 0.5 at lo and hi, with
  lo = round(d/2/dz) + 1 and hi = round((H - d/2)/dz) + 1, Julia round =
  ties-to-even (src/iDFT/coordinate_system/Cartesian/external_fields/hard_walls.jl:6-46);
* charged walls: q/dz placed at round(x_wall/dz) + 1, mirrored
  (charged_walls.jl:7-57); total wall charge enters the neutrality shift;
* grafted segments at wall index +- round(sigma/dz/2)
  (src/iDFT/coordinate_system/Cartesian/fixed_species/determine_fixed.jl:1-29,88-92);
* initial guess x0 = (bulk mu_ex, bulk lng, 0, 0)
  (src/iDFT/setup/setup_steps/process_guess.jl:5-54).

Indices stored here are 0-based.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from fractions import Fraction
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .bulk import BulkState, solve_bulk
from .molsys import MolSys, SetupError, build_molsys

FUNCTIONAL_KEYS = ("mfmt", "TPT1mfmt", "msa", "TPT1msa", "sw")


def julia_cld_plus1(H: float, dz: float) -> int:
    q = Fraction(H) / Fraction(dz)
    return int(math.ceil(q)) + 1


def julia_round(x: float) -> int:
    return int(round(x))  # Python rounds half to even, like Julia's default


@dataclass
class FixedSpec:
    segments: np.ndarray          # bool [N]
    coords: np.ndarray            # int [N], 0-based grid index (only meaningful where fixed)
    density: float = 0.0
    lam: float = 1.0

    @property
    def any(self) -> bool:
        return bool(self.segments.any())


@dataclass
class StatePoint:
    ms: MolSys
    bulk: BulkState
    H: float
    dz: float
    NP: int
    Ext: np.ndarray               # [NP, NB]
    trapez: np.ndarray            # [NP, NB]
    surf: np.ndarray              # [2NP] real-space wall charge on the mirrored domain
    total_charge: float
    hard_walls: bool
    wall_lo: np.ndarray           # [NB] 0-based contact index (hard walls)
    wall_hi: np.ndarray
    wall_charges: Optional[Tuple[float, float]]
    fixed: List[FixedSpec]
    mu_species: np.ndarray        # bulk mu_species with grafted species zeroed (handle_fixed.jl:43)
    functionals: List[str]
    numerics: dict
    label: str = ""
    # square-well walls (sw_walls.jl): per bead type the 0-based grid index of the well's edge at each
    # wall (-1: no well for that bead) and the contact-theorem factor exp(energy) - 1
    sw_wall_idx: Optional[np.ndarray] = None   # [NB, 2] int
    sw_wall_fac: Optional[np.ndarray] = None   # [NB, 2]
    # multi-dimensional grids (geometry_nd): points and bin width per dimension, first fastest; then
    # NP = prod(grid), dz = prod(bins) (the volume element), Ext / trapez / surf flattened over the grid
    grid: Optional[tuple] = None
    bins: Optional[tuple] = None

    @property
    def ndim(self) -> int:
        return 1 if self.grid is None else len(self.grid)

    @property
    def NB(self) -> int:
        return self.ms.NB

    @property
    def MB(self) -> int:
        return self.ms.MB

    @property
    def n(self) -> int:
        return self.NP * (self.NB + 2 * self.MB + 1) + 1

    def x0(self) -> np.ndarray:
        mu = np.tile(self.bulk.mu_ex[None, :], (self.NP, 1))
        lng = np.tile(self.bulk.lng[None, None, :], (self.NP, 2, 1))
        return flatten(mu, lng, np.zeros(self.NP), 0.0)


def flatten(mu_ex_K, lng_K, Psi, PsiC) -> np.ndarray:
    """Julia column-major flatten of (mu_ex_K, lng_K, Psi, PsiC)
    (src/NonlinearSolvers/core/helpers/flat_solution.jl:25-54)."""
    return np.concatenate([np.asarray(mu_ex_K).ravel(order="F"), np.asarray(lng_K).ravel(order="F"),
                           np.asarray(Psi).ravel(), np.atleast_1d(float(PsiC))])


def unflatten(x: np.ndarray, NP: int, NB: int, MB: int):
    a = NP * NB
    b = a + NP * 2 * MB
    mu = x[:a].reshape((NP, NB), order="F")
    lng = x[a:b].reshape((NP, 2, MB), order="F")
    Psi = x[b:b + NP]
    return mu, lng, Psi, float(x[b + NP])


def build_state_point(dataset: dict, label: str = "", bulk: Optional[BulkState] = None,
                      ms: Optional[MolSys] = None) -> StatePoint:
    """Host setup of one state point. `bulk` and `ms` (the molecular system it was solved for, built
    from the same dataset) skip the host bulk solve, e.g. with states from bulk_device.solve_bulk_batch."""
    ds = copy.deepcopy(dataset)
    if ms is None:
        ms = build_molsys(ds)
    if bulk is None:
        bulk = solve_bulk(ms)
    geom = ds["geometry"]
    if geom.get("coordinate_system", "Cartesian") != "Cartesian" or len(geom["dimensions"]) != 1:
        raise SetupError("only 1-D Cartesian slits are supported")
    H = float(geom["dimensions"][0])
    dz = float(geom["bin_width"][0])
    NP = julia_cld_plus1(H, dz)
    NB = ms.NB
    Ext = np.zeros((NP, NB))
    trapez = np.ones((NP, NB))
    surf = np.zeros(2 * NP)
    total_charge = 0.0
    ext = geom.get("external_field") or {}
    sw_idx = sw_fac = None
    hard = "hard_walls" in ext
    lo = np.zeros(NB, dtype=np.int64)
    hi = np.zeros(NB, dtype=np.int64)
    wall_q = None
    for name, desc in ext.items():
        if name == "hard_walls":
            for j, d in enumerate(ms.diameters):
                lo1 = julia_round((0.0 + d / 2.0) / dz) + 1
                hi1 = julia_round((H - d / 2.0) / dz) + 1
                lo[j], hi[j] = lo1 - 1, hi1 - 1
                for K in range(NP):
                    idx = K + 1
                    if idx < lo1:
                        Ext[K, j] += 1.0e20
                        trapez[K, j] = 1.0
                    elif idx == lo1:
                        trapez[K, j] = 0.5
                    if idx > hi1:
                        Ext[K, j] += 1.0e20
                        trapez[K, j] = 1.0
                    elif idx == hi1:
                        trapez[K, j] = 0.5
        elif name in ("charged_walls", "pointcharge_walls"):
            # charged_walls.jl:6-58 / pointcharge_walls.jl:6-48: in one dimension both put q / dz on the
            # wall's grid point; only charged walls add a contact-theorem term (contact_value_theorem
            # has no method for point-charge walls and contributes zero)
            q = desc["charge"]
            if isinstance(q[0], (list, tuple)):
                q = q[0]
            qL, qR = float(q[0]), float(q[1])
            if name == "charged_walls":
                wall_q = (qL, qR)
            total_charge += qL + qR
            for qq, pos in ((qL, 0.0), (qR, H)):
                g = julia_round(pos / dz) + 1
                if 1 <= g <= NP:
                    surf[g - 1] += qq / dz
            for K in range(NP):           # apply_mirroring!
                surf[2 * NP - 1 - K] = surf[K]
        elif name == "sw_walls":
            # sw_walls.jl:13-58 (geometry parsing: geometry/.../sw_walls.jl:73-131): a well of depth
            # energy[w] up to grid index round((wall +- lambda[w] d / 2) / dz) + 1 at each wall
            mons = (desc or {}).get("monomers")
            if not mons:
                raise SetupError("Missing required section: 'monomers' under 'sw_walls'.")
            sw_idx = np.full((NB, 2), -1, dtype=np.int64)
            sw_fac = np.zeros((NB, 2))
            for j, mname in enumerate(ms.monomers):
                if mname not in mons:
                    continue
                e, lam = mons[mname].get("energy"), mons[mname].get("lambda")
                if e is None or lam is None:
                    raise SetupError("Each monomer entry under 'sw_walls.monomers' must define both 'energy' and 'lambda'.")
                e, lam = [float(v) for v in e[0]], [float(v) for v in lam[0]]
                d = float(ms.diameters[j])
                lo1 = julia_round((0.0 + lam[0] * d / 2.0) / dz) + 1
                hi1 = julia_round((H - lam[1] * d / 2.0) / dz) + 1
                for K in range(NP):
                    idx = K + 1
                    if idx <= lo1:
                        Ext[K, j] += e[0]
                    elif idx >= hi1:
                        Ext[K, j] += e[1]
                sw_idx[j] = (lo1 - 1, hi1 - 1)
                sw_fac[j] = (math.exp(e[0]) - 1.0, math.exp(e[1]) - 1.0)
            if np.any(sw_idx < -1) or np.any(sw_idx >= NP):
                raise SetupError("sw_walls: well edge outside the grid")
        elif name == "input_external":
            # input_external.jl:6-28: rows "i, Ext_1, ..., Ext_NB" (0-based grid index) overwrite Ext
            path = (desc or {}).get("file", "external_potential.csv")
            data = np.loadtxt(path, delimiter=",", ndmin=2)
            if data.shape[1] != 1 + NB:
                raise SetupError(f"Expected {1 + NB} columns: dims + {NB} species.")
            for row in data:
                K = int(round(row[0]))
                if not 0 <= K < NP:
                    raise SetupError(f"input_external: grid index {K} outside 0..{NP - 1}")
                Ext[K, :] = row[1:]
        else:
            raise SetupError(f"external field '{name}' is outside this build's scope")

    fixed = []
    for u, sp in enumerate(ms.species):
        fs = FixedSpec(segments=np.zeros(sp.n_seg, dtype=bool), coords=np.zeros(sp.n_seg, dtype=np.int64))
        fd = sp.fixed
        if fd:
            if sp.input_density > 0.0:
                raise SetupError(f"Species {u} defines a fixed segment but also has nonzero input density")
            if "density" not in fd:
                raise SetupError(f"Species {u} contains 'fixed' but does not specify a density.")
            fs.density = float(fd["density"])
            segs, poss = fd["segment"], fd["position"]
            if len(segs) != len(poss):
                raise SetupError("Mismatch in number of fixed segments and positions")
            idx_LW = julia_round(0.0 / dz) + 1
            idx_RW = julia_round(H / dz) + 1
            for L, pos in zip(segs, poss):
                seg = int(L) - 1
                fs.segments[seg] = True
                d = julia_round(ms.diameters[sp.sequence[seg]] / dz / 2.0)
                if pos == "left":
                    fs.coords[seg] = idx_LW + d - 1
                elif pos == "right":
                    fs.coords[seg] = idx_RW - d - 1
                else:
                    raise SetupError(f"Invalid grafting position tag '{pos}'")
        fixed.append(fs)
    mu_sp = bulk.mu_species.copy()
    for u, fs in enumerate(fixed):
        if fs.any:
            mu_sp[u] = 0.0
    funcs = [m for m in ms.fe_model if m in FUNCTIONAL_KEYS]
    num = {k: float(v) for k, v in (ds.get("numerical") or {}).items()}
    sp_out = StatePoint(ms=ms, bulk=bulk, H=H, dz=dz, NP=NP, Ext=Ext, trapez=trapez, surf=surf,
                      total_charge=total_charge, hard_walls=hard, wall_lo=lo, wall_hi=hi,
                      wall_charges=wall_q, fixed=fixed, mu_species=mu_sp, functionals=funcs,
                      numerics=num, label=label, sw_wall_idx=sw_idx, sw_wall_fac=sw_fac)
    return sp_out


def build_state_points(datasets: Sequence[dict], workers: int = 0, device_bulk: bool = False) -> List[StatePoint]:
    """build_state_point over many datasets; with workers > 1 the host setups (bulk solves) run in
    that many spawned processes (no GPU state is shared with them), results in input order. With
    device_bulk the bulk states of all points are solved in one sdft_bulk_solve launch (datasets
    sharing one molecular-system layout) and only the rest of the setup runs on the host."""
    if device_bulk:
        from .bulk_device import solve_bulk_batch
        mss = [build_molsys(copy.deepcopy(d)) for d in datasets]
        bulks = solve_bulk_batch(mss)
        return [build_state_point(d, bulk=b, ms=m) for d, b, m in zip(datasets, bulks, mss)]
    if workers <= 1 or len(datasets) < 2:
        return [build_state_point(d) for d in datasets]
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(min(workers, len(datasets)), mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(build_state_point, datasets, chunksize=max(1, len(datasets) // (4 * workers))))
