"""Analysis outputs of a converged state point in the reference's text formats.

call_analysis mirrors src/iDFT/evaluation/analysis/analysis.jl:9-16: it writes, into one directory,
  cvt.txt                          contact_value_theorem.jl:5-20
  grand_potential.txt              print_array.jl:1-9 (write_scalar_value) of grand_potential.jl:4-28
  brush_height.txt                 brush_height.jl:1-64 (species with a grafted segment only)
  MEP.txt, rho_beads_K.txt         output_arrays.jl:2-4 via print_array.jl:11-36 (write_field_array)
  rho_segments_<species>.txt       print_array.jl:37-84 (species with more than one segment)
  state_fraction_<species>.txt     print_array.jl:87-151 (species with a multistate segment)
  total_state_fractions_<species>.txt  degree_of_dissociation.jl:1-71

Numbers are printed as Julia's print/println prints them (julia_repr): the shortest decimal that
reads back to the same Float64, in fixed notation when its decimal exponent lies in -4 < pt <= 6
and in Julia's e-notation otherwise; grid indices are 1-based as in Julia's CartesianIndices.
"""
from __future__ import annotations

import math
import os
from decimal import Decimal
from typing import List, Optional, Sequence

import numpy as np

from . import abi
from .problem import StatePoint, unflatten


def julia_repr(x) -> str:
    """print(x::Float64) of Julia (Base.Ryu shortest form); integers print as integers."""
    if isinstance(x, (int, np.integer)) and not isinstance(x, bool):
        return str(int(x))
    if isinstance(x, str):
        return x
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + "0.0"
    # shortest round-trip digits (Python's repr picks the same digits as Ryu: shortest, then
    # closest to the exact binary value)
    t = Decimal(repr(abs(x))).as_tuple()
    digits = "".join(str(d) for d in t.digits).rstrip("0") or "0"
    ndrop = len(t.digits) - len(digits)
    e = t.exponent + ndrop               # value = int(digits) * 10**e
    olength = len(digits)
    pt = e + olength                     # value = 0.digits * 10**pt
    if -4 < pt <= 6:
        if pt <= 0:
            return sign + "0." + "0" * (-pt) + digits
        if pt < olength:
            return sign + digits[:pt] + "." + digits[pt:]
        return sign + digits + "0" * (pt - olength) + ".0"
    mant = digits[0] + "." + (digits[1:] or "0")
    return f"{sign}{mant}e{pt - 1}"


def _join(vals) -> str:
    return " ".join(julia_repr(v) for v in vals)


def write_scalar_value(value: float, filename: str, label: Optional[str] = None) -> None:
    """print_array.jl:1-9."""
    with open(filename, "w") as f:
        if label is None:
            f.write(julia_repr(value) + "\n")
        else:
            f.write(f"{label}: {julia_repr(value)}\n")


def write_field_array(arr: np.ndarray, filename: str, NP: Sequence[int],
                      names: Optional[Sequence[str]] = None) -> None:
    """print_array.jl:11-36: one line per grid point, the 1-based index then the values at it."""
    NP = tuple(int(n) for n in NP)
    arr = np.asarray(arr, dtype=np.float64)
    scalar = arr.shape == NP
    nd = len(NP)
    vals = arr.reshape(NP + (1,), order="F") if scalar else arr.reshape(NP + (arr.shape[nd],), order="F")
    lines = []
    if names is not None:
        lines.append(" ".join(f"Index_{i + 1}" for i in range(nd)) + "  " + " ".join(names))
    for idx in np.ndindex(*NP[::-1]):
        coord = idx[::-1]                # Julia's CartesianIndices: first index fastest
        lines.append(" ".join(str(c + 1) for c in coord) + "  " + _join(vals[coord]))
    with open(filename, "w") as f:
        f.write("\n".join(lines) + "\n")


def _segment_headers(species, monomers: Sequence[str]) -> List[str]:
    return [f"seg{q + 1}_{monomers[g]}" for q, fam in enumerate(species.state_family) for g in fam]


def write_segment_array(rho_segments: Sequence[np.ndarray], ms, NP: Sequence[int], directory: str) -> None:
    """print_array.jl:37-84: rho_segments_<species>.txt for each species with more than one segment;
    rho_segments[u] has shape (NP, S_u, N_u)."""
    NP = tuple(int(n) for n in NP)
    for rho, sp in zip(rho_segments, ms.species):
        nseg = rho.shape[-1]
        if nseg <= 1:
            continue
        lines = [" ".join(f"Index_{d + 1}" for d in range(len(NP))) + "  "
                 + " ".join(_segment_headers(sp, ms.monomers))]
        for k in range(NP[0]):
            vals = [rho[k, a, q] for q, fam in enumerate(sp.state_family) for a in range(len(fam))]
            lines.append(f"{k + 1}  {_join(vals)}")
        with open(os.path.join(directory, f"rho_segments_{sp.name}.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")


def write_state_fraction_profiles(rho_segments: Sequence[np.ndarray], ms, NP: Sequence[int], directory: str,
                                  out_prefix: str = "state_fraction_") -> None:
    """print_array.jl:87-151: per grid point, each segment's state densities over their sum (zeros
    where the sum is not positive), for species with a multistate segment."""
    NP = tuple(int(n) for n in NP)
    for rho, sp in zip(rho_segments, ms.species):
        if not any(len(f) > 1 for f in sp.state_family):
            continue
        lines = [" ".join(f"Index_{d + 1}" for d in range(len(NP))) + "  "
                 + " ".join(_segment_headers(sp, ms.monomers))]
        for k in range(NP[0]):
            vals = []
            for q, fam in enumerate(sp.state_family):
                loc = [float(rho[k, a, q]) for a in range(len(fam))]
                den = 0.0
                for v in loc:
                    den += v
                vals.extend([v / den for v in loc] if den > 0 else [0.0] * len(loc))
            lines.append(f"{k + 1}  {_join(vals)}")
        with open(os.path.join(directory, f"{out_prefix}{sp.name}.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")


def write_total_state_fractions(rho_segments: Sequence[np.ndarray], ms, NP: Sequence[int], directory: str,
                                out_prefix: str = "total_state_fractions_") -> None:
    """degree_of_dissociation.jl:1-71: per multistate segment the grid-summed state densities over
    their total, then the per-state average over those segments."""
    for rho, sp in zip(rho_segments, ms.species):
        fams = sp.state_family
        if not any(len(f) > 1 for f in fams):
            continue
        max_states = max(len(f) for f in fams)
        sums = [0.0] * max_states
        counts = [0] * max_states
        lines = ["segment " + " ".join(f"state{n + 1}" for n in range(max_states))]
        for q, fam in enumerate(fams):
            ns = len(fam)
            if ns == 1:
                continue
            loc = [float(np.sum(rho[:, a, q])) for a in range(ns)]
            total = 0.0
            for v in loc:
                total += v
            normed = [v / total for v in loc] if total > 0 else [0.0] * ns
            for j in range(ns):
                sums[j] += normed[j]
                counts[j] += 1
            lines.append(f"{q + 1} " + " ".join([julia_repr(v) for v in normed] + [""] * (max_states - ns)))
        lines.append("")
        lines.append("# Average state fractions across all dissociating segments:")
        avg = [julia_repr(sums[j] / counts[j]) if counts[j] > 0 else "" for j in range(max_states)]
        lines.append("avg " + " ".join(avg))
        with open(os.path.join(directory, f"{out_prefix}{sp.name}.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")


def brush_heights(rho_segments: Sequence[np.ndarray], sp: StatePoint) -> List[tuple]:
    """brush_height.jl:1-64: 2 <z> over the segment density of each species with a grafted segment;
    grafted segments weigh rho / dz, free ones rho * trapez of their state, z_k = (k - 1) dz."""
    out = []
    ms = sp.ms
    z = np.arange(sp.NP) * sp.dz
    for u, spec in enumerate(ms.species):
        fx = sp.fixed[u]
        if not fx.any:
            continue
        num = den = 0.0
        rho = rho_segments[u]
        for q, fam in enumerate(spec.state_family):
            for a, g in enumerate(fam):
                if fx.segments[q]:
                    val = rho[:, a, q] / sp.dz
                else:
                    val = rho[:, a, q] * sp.trapez[:sp.NP, g]
                for k in range(sp.NP):   # the reference's running sums, in its order
                    num += float(val[k]) * float(z[k])
                    den += float(val[k])
        out.append((spec.name, 2.0 * num / den if den > 0 else 0.0))
    return out


def write_brush_height(heights: Sequence[tuple], filename: str) -> None:
    with open(filename, "w") as f:
        f.write("# Brush heights per species (computed as 2 * <z>)\n")
        f.write("# Format: species brush_height\n")
        for name, h in heights:
            f.write(f"{name} {julia_repr(h)}\n")


def write_cvt(contr: Sequence[float], filename: str) -> None:
    """contact_value_theorem.jl:5-20 for the planar slit (one dimension, two walls)."""
    with open(filename, "w") as f:
        f.write("# Contact Value Theorem results\n")
        f.write(f"dim1  {julia_repr(contr[0])}  {julia_repr(contr[1])}\n")


def call_analysis(batch: "abi.Batch", x: np.ndarray, point: int, directory: str) -> dict:
    """Write the analysis files of batch point `point` (analysis.jl:9-16) after a solve or objective
    call whose solution vector is `x` (the whole batch's). Returns the values written."""
    os.makedirs(directory, exist_ok=True)
    sp = batch.points[point]
    ms = sp.ms
    om, _, pc = batch.observables()
    rho_beads = batch.get_field(abi.FIELD_RHO_BEADS, point)
    segs = [batch.segment_profile(u, point) for u in range(batch.nsp)]
    xp = batch.split(np.asarray(x, dtype=np.float64))[point]
    _, _, psi, _ = unflatten(xp, sp.NP, sp.NB, sp.MB)
    NP = (sp.NP,)
    write_cvt(pc[point], os.path.join(directory, "cvt.txt"))
    write_scalar_value(float(om[point]), os.path.join(directory, "grand_potential.txt"))
    heights = brush_heights(segs, sp)
    if heights:
        write_brush_height(heights, os.path.join(directory, "brush_height.txt"))
    write_field_array(psi, os.path.join(directory, "MEP.txt"), NP, names=["MEP"])
    write_field_array(rho_beads, os.path.join(directory, "rho_beads_K.txt"), NP, names=list(ms.monomers))
    write_segment_array(segs, ms, NP, directory)
    write_state_fraction_profiles(segs, ms, NP, directory)
    write_total_state_fractions(segs, ms, NP, directory)
    return dict(omega=float(om[point]), cvt=tuple(pc[point]), brush_heights=heights, rho_segments=segs)
