"""Host setup of two- and three-dimensional Cartesian state points.

The reference builds a D-dimensional Cartesian grid from `geometry.dimensions` and `geometry.bin_width`
(src/geometry/setup/coordinate_system/Cartesian/create_cartesian.jl:25-68):
* NP_d = cld(L_d, h_d) + 1 points per dimension;
* every dimension is mirrored (process_cartesian.jl:14-45), so each convolution runs on the
  2 NP_d-periodic even extension in every dimension (compute_full_domain, mirrored.jl:58-62).
Grid arrays are flattened with the first dimension fastest, as in Julia, so the solution vector keeps
the one-dimensional flatten order with NP = prod NP_d (flat_solution.jl:25-54).

External fields:
* hard walls at 0 and L_v along each listed dimension v (hard_walls.jl:6-46). For every point and
  every bead j the rule is applied dimension by dimension in order: Ext += 1e20 with trapez = 1
  outside, trapez = 0.5 on the contact plane; a later dimension overwrites trapez.
* charged walls at 0 and L_v along each listed dimension (charged_walls.jl:6-57):
  - the wall charge q goes onto the wall's grid plane times invbin_area, which the reference
    forms as (1/h_v) * prod_{v1 != v} h_v1;
  - the neutrality constraint uses the sum of all wall charges (neutrality_shift.jl:17).

The contact-value theorem (contact_value_theorem.jl:1-20, hard_walls.jl:49-79, charged_walls.jl:60-81)
returns one (left, right) pair per dimension. The reference reads the hard-wall contact density of
dimension v at the grid point (NP_1, ..., NP_{D-1}, contact index along v), whatever v is, and that
is kept here.

Grafted (fixed) segments are placed as the reference's 3-D method does (determine_fixed.jl:31-86): one
chain per wall, grafted at the centre of the wall, with the grafting density reset to 1 / (L_x L_y).
Two-dimensional grids have no such method and are rejected.
Chains of more than one segment need three dimensions: the reference's bond renormalisation has no
two-dimensional form (get_renorm, child_propagators.jl:218-226).
"""
from __future__ import annotations

import copy
import math
from typing import List, Optional, Sequence

import numpy as np

from .bulk import BulkState, solve_bulk
from .molsys import MolSys, SetupError, build_molsys
from .problem import FUNCTIONAL_KEYS, FixedSpec, StatePoint, julia_cld_plus1, julia_round


def grid_index_tuples(grid: Sequence[int]) -> np.ndarray:
    """0-based index of every point along each dimension, flat order with the first dimension fastest:
    shape [D, prod(grid)]."""
    idx = np.indices(tuple(grid[::-1])).reshape(len(grid), -1)[::-1]
    return np.ascontiguousarray(idx)


def build_state_point_nd(dataset: dict, label: str = "", bulk: Optional[BulkState] = None,
                         ms: Optional[MolSys] = None) -> StatePoint:
    ds = copy.deepcopy(dataset)
    if ms is None:
        ms = build_molsys(ds)
    if bulk is None:
        bulk = solve_bulk(ms)
    geom = ds["geometry"]
    if geom.get("coordinate_system", "Cartesian") != "Cartesian":
        raise SetupError("only Cartesian geometries are supported")
    dims = [float(v) for v in geom["dimensions"]]
    bins = [float(v) for v in geom["bin_width"]]
    D = len(dims)
    if len(bins) != D:
        raise SetupError("Error: Mismatch in number of `dimensions` and `bin_width` entries.")
    if D not in (2, 3):
        raise SetupError("multi-dimensional setup needs 2 or 3 dimensions")
    grid = tuple(julia_cld_plus1(L, h) for L, h in zip(dims, bins))
    NPt = int(np.prod(grid))
    NB = ms.NB
    for u, s in enumerate(ms.species):
        if s.fixed and D != 3:
            raise SetupError(f"species {u}: grafted segments need a 3-D grid (determine_fixed has no 2-D method)")
        if D == 2 and s.n_seg > 1 and s.model != "monomerbead":
            raise SetupError("Not added yet.")   # get_renorm has no 2-D form (child_propagators.jl:218-226)
    idx = grid_index_tuples(grid) + 1        # 1-based, as in the reference
    Ext = np.zeros((NPt, NB))
    trapez = np.ones((NPt, NB))
    surf = np.zeros(NPt)
    total_charge = 0.0
    ext = geom.get("external_field") or {}
    hard_dims: List[int] = []
    wall_q = None
    for name, desc in ext.items():
        if name == "hard_walls":
            hard_dims = sorted(int(v) - 1 for v in desc["dims"])
            for v in hard_dims:
                if not 0 <= v < D:
                    raise SetupError(f"Invalid dimension index: {v + 1} (expected between 1 and {D})")
        elif name == "charged_walls":
            wdims = [int(v) - 1 for v in desc["dims"]]
            q = desc["charge"]
            if not isinstance(q[0], (list, tuple)):
                q = [q]
            if len(q) != len(wdims):
                raise SetupError("Mismatch: number of charge entries does not match dims.")
            wall_q = {}
            for v, (qL, qR) in zip(wdims, q):
                qL, qR = float(qL), float(qR)
                wall_q[v] = (qL, qR)
                total_charge += qL + qR
                inv_area = 1.0 / bins[v]
                for v1 in range(D):
                    if v1 != v:
                        inv_area *= bins[v1]   # as written (charged_walls.jl:28-32)
                for qq, pos in ((qL, 0.0), (qR, dims[v])):
                    g = julia_round(pos / bins[v]) + 1
                    surf[idx[v] == g] += qq * inv_area
        else:
            raise SetupError(f"external field '{name}' is outside this build's multi-dimensional scope")
    # hard walls: dimension by dimension in order, every bead (hard_walls.jl:6-46)
    for v in range(D):
        if v not in hard_dims:
            continue
        for j, d in enumerate(ms.diameters):
            lo1 = julia_round((0.0 + d / 2.0) / bins[v]) + 1
            hi1 = julia_round((dims[v] - d / 2.0) / bins[v]) + 1
            iv = idx[v]
            out_lo, on_lo = iv < lo1, iv == lo1
            out_hi, on_hi = iv > hi1, iv == hi1
            Ext[out_lo, j] += 1.0e20
            trapez[out_lo, j] = 1.0
            trapez[on_lo, j] = 0.5
            Ext[out_hi, j] += 1.0e20
            trapez[out_hi, j] = 1.0
            trapez[on_hi, j] = 0.5
    fixed = determine_fixed_xyz(ms, grid, bins, dims) if any(s.fixed for s in ms.species) else \
        [FixedSpec(segments=np.zeros(s.n_seg, dtype=bool), coords=np.zeros(s.n_seg, dtype=np.int64))
         for s in ms.species]
    mu_sp = bulk.mu_species.copy()
    for u, fs in enumerate(fixed):
        if fs.any:
            mu_sp[u] = 0.0   # handle_fixed.jl:43
    funcs = [m for m in ms.fe_model if m in FUNCTIONAL_KEYS]
    if "sw" in funcs:
        raise SetupError("the square-well functional is outside this build's multi-dimensional scope")
    num = {k: float(v) for k, v in (ds.get("numerical") or {}).items()}
    sp = StatePoint(ms=ms, bulk=bulk, H=float(dims[-1]), dz=float(np.prod(bins)), NP=NPt, Ext=Ext,
                    trapez=trapez, surf=surf, total_charge=total_charge, hard_walls=False,
                    wall_lo=np.zeros(NB, dtype=np.int64), wall_hi=np.zeros(NB, dtype=np.int64),
                    wall_charges=None, fixed=fixed, mu_species=mu_sp, functionals=funcs,
                    numerics=num, label=label, grid=grid, bins=tuple(bins))
    sp.nd_hard_dims = hard_dims
    sp.nd_wall_q = wall_q
    sp.nd_dims = tuple(dims)
    return sp


def determine_fixed_xyz(ms: MolSys, grid, bins, dims) -> List[FixedSpec]:
    """determine_fixed for CartesianXYZ (determine_fixed.jl:31-86): one grafted segment per chain, at
    (round(NP_x / 2) + 1, round(NP_y / 2) + 1) in x and y and round(d / h_z / 2) grid steps from the
    wall along z (1-based), the grafting density set to 1 / (L_x L_y); x and y must have the same
    length and each wall carries at most one brush. Coordinates are returned as flat 0-based indices."""
    out = []
    left = right = 0
    for u, s in enumerate(ms.species):
        fs = FixedSpec(segments=np.zeros(s.n_seg, dtype=bool), coords=np.zeros(s.n_seg, dtype=np.int64))
        fd = s.fixed
        if fd:
            if s.input_density > 0.0:
                raise SetupError(f"Species {u} defines a fixed segment but also has nonzero input density")
            if "density" not in fd:
                raise SetupError(f"Species {u} contains 'fixed' but does not specify a density.")
            segs, poss = fd["segment"], fd["position"]
            if len(segs) != len(poss):
                raise SetupError("Mismatch in number of fixed segments and positions")
            if len(segs) > 1:
                raise SetupError(f"Only one fixed segment per chain is allowed in 3D (species {u}).")
            if dims[0] != dims[1]:
                raise SetupError("For 3D brushes, wall dimensions in x and y must match.")
            fs.density = 1.0 / (dims[0] * dims[1])
            idx_lw = julia_round(0.0 / bins[-1]) + 1
            idx_rw = julia_round(dims[-1] / bins[-1]) + 1
            for L, pos in zip(segs, poss):
                seg = int(L) - 1
                fs.segments[seg] = True
                dzi = julia_round(ms.diameters[s.sequence[seg]] / bins[-1] / 2.0)
                xc = julia_round(grid[0] / 2) + 1
                yc = julia_round(grid[1] / 2) + 1
                if pos == "left":
                    right += 1          # the reference's counters, as written
                    zc = idx_lw + dzi
                elif pos == "right":
                    left += 1
                    zc = idx_rw - dzi
                else:
                    raise SetupError(f"Invalid grafting position tag for species {u}, segment {L}: '{pos}'")
                fs.coords[seg] = (xc - 1) + grid[0] * ((yc - 1) + grid[1] * (zc - 1))
        out.append(fs)
    if left > 1 or right > 1:
        raise SetupError("Only one brush may be attached to each wall in 3D.")
    return out


def build_state_points_nd(datasets: Sequence[dict], device_bulk: bool = False) -> List[StatePoint]:
    if device_bulk:
        from .bulk_device import solve_bulk_batch
        mss = [build_molsys(copy.deepcopy(d)) for d in datasets]
        bulks = solve_bulk_batch(mss)
        return [build_state_point_nd(d, bulk=b, ms=m) for d, b, m in zip(datasets, bulks, mss)]
    return [build_state_point_nd(d) for d in datasets]


def contact_value_theorem(sp: StatePoint, rho_beads: np.ndarray) -> np.ndarray:
    """contr[2, D] of contact_value_theorem (contact_value_theorem.jl:5-20) from the bead densities
    (NPt, NB) of a multi-dimensional point: hard walls (hard_walls.jl:49-79, contact point
    (NP_1, ..., NP_{D-1}, index) for every walled dimension), charged walls (charged_walls.jl:60-81)."""
    grid, bins, dims = sp.grid, sp.bins, sp.nd_dims
    D = len(grid)
    contr = np.zeros((2, D))
    rb = rho_beads.reshape(tuple(grid) + (sp.NB,), order="F")
    for v in sp.nd_hard_dims:
        for j, d in enumerate(sp.ms.diameters):
            lo1 = julia_round((0.0 + d / 2.0) / bins[v]) + 1
            hi1 = julia_round((dims[v] - d / 2.0) / bins[v]) + 1
            head = tuple(n - 1 for n in grid[:-1])
            contr[0, v] += rb[head + (lo1 - 1, j)]
            contr[1, v] += rb[head + (hi1 - 1, j)]
    if sp.nd_wall_q:
        lB = sp.ms.bjerrum
        for v, (qL, qR) in sp.nd_wall_q.items():
            contr[0, v] += -2.0 * math.pi * lB * qL * qL
            contr[1, v] += -2.0 * math.pi * lB * qR * qR
    return contr
