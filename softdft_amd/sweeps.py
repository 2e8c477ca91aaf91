"""Continuation sweeps on top of the batched solver.

The reference's slit-width continuation (routines/adjust_distanceZ.jl:15-155) steps the wall
separation H, warm-starts each width from the previous (narrower) converged system and records per
width the contact pressure P (contact_value_theorem, first wall), the disjoining pressure
Pi = P - P_bulk and the energy Omega + P_bulk * H (adjust_distanceZ.jl:95-110); its summary file
(adjust_distanceZ.jl:138-148) lists every step against the last one.

The reference driver is stale: it calls eval_inhomogeneous(dft_system, 1), a method that does not
exist (src/iDFT/evaluation/eval_inhomogeneous.jl:4 defines only the one-argument form), after
copying the previous densities into the new system. Here the warm start is done on the solution
vector itself: the excess fields mu_ex, lng and Psi of the new width take the previous width's
values on the left half (1-based index k < round(NP_new / 2) + 2, Julia round-half-even) and the
mirror image on the right half, which is the reference's copy rule (adjust_distanceZ.jl:73-88), and
PsiC is copied (adjust_distanceZ.jl:90).

The pH continuation (routines/adjust_pH.jl:16-124) steps the pSpecies of H+ ("Y"), warm-starts each pH
from the previous one (on the fixed grid the solution vector is copied as is, see copy_solution), and
records per
step the brush height of the grafted species (brush_height.jl) and the average degree of ionization
(the `avg` state-2 column of total_state_fractions, degree_of_dissociation.jl), stopping once alpha
reaches 0.99 on an increasing sweep (0.01 on a decreasing one, adjust_pH.jl:107-113).

Batching: width i is warm-started from width i - 1. The widths are solved in `waves` rounds: round r
holds every width i with i % waves == r, so each round is one batched solve and every width except
those of round 0 (cold, from the bulk guess) starts from its immediate narrower neighbour, as in
the reference.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np

import math

from . import abi, analysis
from .problem import StatePoint, build_state_points, flatten, unflatten


def julia_round(x: float) -> int:
    """Julia's round(Int, x): nearest integer, ties to even (Python's round does the same)."""
    return int(round(x))


def remap_solution(sp_old: StatePoint, x_old: np.ndarray, sp_new: StatePoint) -> np.ndarray:
    """Warm start of width `sp_new` from the converged `x_old` of width `sp_old`
    (adjust_distanceZ.jl:64-90 applied to the solution vector)."""
    mu_o, lng_o, psi_o, psic_o = unflatten(x_old, sp_old.NP, sp_old.NB, sp_old.MB)
    mu, lng, psi, _ = unflatten(sp_new.x0(), sp_new.NP, sp_new.NB, sp_new.MB)
    mu, lng, psi = mu.copy(), lng.copy(), psi.copy()
    NP = sp_new.NP
    last = julia_round(NP / 2) + 2  # 1-based k < last
    for k in range(1, last):
        if k > sp_old.NP:
            raise ValueError(f"width step too large: the old grid (NP={sp_old.NP}) has no index {k}")
        k0, k1 = k - 1, NP - k  # 0-based index and its mirror (NP - k + 1 in 1-based)
        mu[k0] = mu_o[k0]
        mu[k1] = mu_o[k0]
        lng[k0] = lng_o[k0]
        lng[k1] = lng_o[k0]
        psi[k0] = psi_o[k0]
        psi[k1] = psi_o[k0]
    return flatten(mu, lng, psi, float(psic_o))


@dataclass
class WidthResult:
    step: int            # 1-based sweep index
    distance: float      # H
    pressure: float      # contact-value pressure at the first wall
    delta_p: float       # Pi(H) = P - P_bulk
    energy: float        # Omega + P_bulk * H
    omega: float
    iterations: int
    warm: bool
    residual: float = 0.0  # ||F(x) - x||_inf at the end (above tole: not converged within max_iters)


def width_sweep(datasets: Sequence[dict], *, waves: int = 8, precision: int = abi.PREC_MIXED,
                numerics: Optional[dict] = None, max_iters: int = 3000,
                on_round: Optional[Callable[[int, float], None]] = None, workers: int = 0,
                tune: bool = False) -> List[WidthResult]:
    """Solve the slit widths of `datasets` (increasing H, one chemistry) with neighbour warm starts;
    the bulk states of all widths are solved on the device in one launch (`workers` is not used:
    the rest of the host setup takes milliseconds per point); `tune` measures the launch
    configuration on the first round's context and replays it on the others.

    Returns one WidthResult per width, in sweep order."""
    sps = build_state_points(datasets, workers, device_bulk=True)
    n = len(sps)
    num = dict(numerics or sps[0].numerics)
    opts = dict(tole=num["tole"], damping=num["damping"], mixing_max=num["mixing_max"],
                switch_tole=num["switch_tole"], anderm=int(num["anderm"]), max_iters=max_iters,
                precision=precision)
    res: List[Optional[WidthResult]] = [None] * n

    def finish(batch, idx, x, it, err, warm):
        om, _, pc = batch.observables()
        for k, i in enumerate(idx):
            sp = sps[i]
            p_bulk = float(sp.bulk.pressure)
            res[i] = WidthResult(step=i + 1, distance=float(sp.H), pressure=float(pc[k, 0]),
                                 delta_p=float(pc[k, 0] - p_bulk), energy=float(om[k] + p_bulk * sp.H),
                                 omega=float(om[k]), iterations=int(it[k]), warm=warm, residual=float(err[k]))

    _waves_solve(sps, waves, opts, on_round, finish, tune=tune)
    return [r for r in res if r is not None]


def write_sweep_summary(results: Sequence[WidthResult], path: str) -> None:
    """adjust_distancesZ_results.csv (adjust_distanceZ.jl:138-148): every step against the last."""
    final_e, final_dp = results[-1].energy, results[-1].delta_p
    with open(path, "w") as f:
        f.write("step,distance,Pressure,deltaP,energy,deltaE\n")
        for r in results:
            f.write(f"{r.step},{r.distance!r},{r.pressure!r},{r.delta_p - final_dp!r},{r.energy!r},"
                    f"{r.energy - final_e!r}\n")


def write_run_info(r: WidthResult, directory: str) -> None:
    """Per-width run_info.csv (adjust_distanceZ.jl:112-116)."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "run_info.csv"), "w") as f:
        f.write("step,distance,Pressure,deltaP,energy\n")
        f.write(f"{r.step},{r.distance!r},{r.pressure!r},{r.delta_p!r},{r.energy!r}\n")


@dataclass
class PHResult:
    step: int             # 1-based sweep index
    pH: float
    brush_height: float   # NaN without exactly one grafted species (adjust_pH.jl:86-87)
    alpha: float          # avg fraction of state 2 over multistate segments, NaN without one
    omega: float
    iterations: int
    warm: bool
    residual: float = 0.0  # ||F(x) - x||_inf at the end (above tole: not converged within max_iters)


def copy_solution(sp_old: StatePoint, x_old: np.ndarray, sp_new: StatePoint) -> np.ndarray:
    """Warm start on an unchanged grid: the converged solution vector itself. The pH driver's mirror
    copy (adjust_pH.jl:50-80) assumes a profile symmetric about the slit centre, which a brush grafted
    to one wall is not, and the driver's stale solver call never reads the copied densities."""
    if sp_old.NP != sp_new.NP or sp_old.NB != sp_new.NB or sp_old.MB != sp_new.MB:
        return remap_solution(sp_old, x_old, sp_new)
    return np.array(x_old, dtype=np.float64, copy=True)


def _waves_solve(sps, waves, opts, on_round, finish, warm=remap_solution, tune=False):
    """Solve `sps` in `waves` batched rounds, point i warm-started from point i - 1 (cold in round 0);
    finish(batch, idx, x_parts, iters, warm) is called per round while the context is alive. With
    `tune`, round 0's context is autotuned and its decisions (keyed by species and grid-length bin,
    which every round shares) are replayed on the later rounds."""
    n = len(sps)
    xs: List[Optional[np.ndarray]] = [None] * n
    waves = max(1, min(waves, n))
    tuning = None
    for r in range(waves):
        idx = list(range(r, n, waves))
        if not idx:
            continue
        batch = abi.Batch([sps[i] for i in idx])
        if tune:
            if tuning is None:
                tuning = batch.autotune(reps=2)
            else:
                batch.set_tuning(tuning)
        x0 = [warm(sps[i - 1], xs[i - 1], sps[i]) if r > 0 else sps[i].x0() for i in idx]
        x, it, err = batch.solve(np.concatenate(x0), raise_on_noconv=False, **opts)
        parts = batch.split(x)
        for k, i in enumerate(idx):
            xs[i] = parts[k].copy()
        finish(batch, idx, x, it, err, r > 0)
        batch.close()
        if on_round is not None:
            on_round(r, float(np.max(err)))


def ph_sweep(datasets: Sequence[dict], pH_values: Sequence[float], *, waves: int = 8,
             precision: int = abi.PREC_MIXED, numerics: Optional[dict] = None, max_iters: int = 3000,
             on_round: Optional[Callable[[int, float], None]] = None, workers: int = 0,
             tune: bool = False) -> List[PHResult]:
    """pH continuation over `datasets` (one per pH in `pH_values`, same grid and chemistry) with
    neighbour warm starts; the result list ends at the reference's early-stop step (`tune` as in
    width_sweep)."""
    sps = build_state_points(datasets, workers, device_bulk=True)
    n = len(sps)
    num = dict(numerics or sps[0].numerics)
    opts = dict(tole=num["tole"], damping=num["damping"], mixing_max=num["mixing_max"],
                switch_tole=num["switch_tole"], anderm=int(num["anderm"]), max_iters=max_iters,
                precision=precision)
    res: List[Optional[PHResult]] = [None] * n

    def finish(batch, idx, x, it, err, warm):
        om, al, _ = batch.observables()
        for k, i in enumerate(idx):
            sp = sps[i]
            segs = [batch.segment_profile(u, k) for u in range(batch.nsp)]
            hs = analysis.brush_heights(segs, sp)
            multi = [u for u, s in enumerate(sp.ms.species) if any(len(f) > 1 for f in s.state_family)]
            alpha = float(al[k, multi[0], 1]) if multi else math.nan
            res[i] = PHResult(step=i + 1, pH=float(pH_values[i]),
                              brush_height=float(hs[0][1]) if len(hs) == 1 else math.nan,
                              alpha=alpha, omega=float(om[k]), iterations=int(it[k]), warm=warm,
                              residual=float(err[k]))

    _waves_solve(sps, waves, opts, on_round, finish, warm=copy_solution, tune=tune)
    out = [r for r in res if r is not None]
    increasing = n < 2 or pH_values[-1] >= pH_values[0]
    for k, r in enumerate(out):
        if (increasing and r.alpha >= 0.99) or (not increasing and r.alpha <= 0.01):
            return out[:k + 1]
    return out


def write_ph_summary(results: Sequence[PHResult], path: str) -> None:
    """summary_brush_alpha.csv (adjust_pH.jl:298-310)."""
    def g12(v: float) -> str:   # Julia's @printf "%.12g" (NaN / Inf spelled as Julia prints them)
        if math.isnan(v):
            return "NaN"
        if math.isinf(v):
            return "Inf" if v > 0 else "-Inf"
        return "%.12g" % v
    with open(path, "w") as f:
        f.write("step,pH,brush_height,alpha\n")
        for r in results:
            f.write(f"{r.step},{g12(r.pH)},{g12(r.brush_height)},{g12(r.alpha)}\n")
