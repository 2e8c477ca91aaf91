"""Pseudo-arclength continuation in pH, batched over independent branches.

The reference's pac_continuation (src/NonlinearSolvers/continuation/objective_pac.jl:1-228) traces a
solution branch in the pSpecies of H+ ("Y"):

* it solves the system at the starting pH, then at pH + ds_init (objective_pac.jl:64-101), and takes
  the arclength s of the second step as the Euclidean norm of the change of the whole vector;
* each later step (objective_pac.jl:126-224):
  - forms the tangent dY/ds from the stored history: a first-order difference at the first step,
    else the second-order backward difference on the non-uniform arclengths (objective_pac.jl:131-143);
  - normalises the tangent, predicts X_k + dS t, and corrects with the Anderson solver on the
    augmented map G(x, p) = (F(x; p), p_k + (ds - t_x . (F(x; p) - x_k)) / t_p), whose fixed point
    lies on the branch at arclength ds from X_k along t (objective_pac.jl:230-261);
  - then sets dS = (dS + ds_actual) / 2 (objective_pac.jl:172);
* the loop runs while 1.99 < pH < 12.01 and stops early once the first species' average degree of
  ionization reaches 0.99 while dS > 0 (or 0.01 while dS < 0, which its update never makes
  negative) (objective_pac.jl:125, 207-213).

Two details of the reference are kept as written:
* the ds of the constraint is captured once, when the continuation data are built
  (`Dict(:ds => dS)`, objective_pac.jl:104-106), so every corrector holds the arclength ds_init while
  the predictor uses the adapted dS;
* the step counter starts at 3, the history ring is indexed with mod1, and `num_keep` entries are kept.

What differs, and why:
* The unknowns. The reference continues the segment densities, bond densities and PsiC
  (objective_pac.jl:36), but objective_dft's inputs are the excess fields. objective_pac copies the
  densities into a freshly built system and returns the fields (objective_pac.jl:246-254,274-283), so
  its map does not depend on its own input. Here the unknowns are the solver's own vector x
  (mu_ex, lng, Psi, PsiC; flat_solution.jl:25-54) plus the pH. G above is then the objective_pac
  construction applied to a consistent map.
* The second start solve is warm-started from the first solution. The reference rebuilds the system
  from the bulk guess (establish_dft_system), which the solver then cannot read either.
* `max_steps` caps the number of predictor-corrector steps. The reference reads it (objective_pac.jl:55)
  but never uses it.

The grafted variant (objective_pac_graft.jl) continues the logarithm of the first species' grafting
density instead: update_cont_var_graft sets that density to exp(variable) (objective_pac_graft.jl:286-288),
the loop runs while the variable is below 1 (objective_pac_graft.jl:120), its steps report the density,
and it has no early stop. Here `variable="log_graft_density"` selects it; the grafting density is
per point in the library (sdft_set_graft_density), so branches may differ in it.

Every evaluation of the augmented map sets each branch's chemistry for its current pH:
* delta_muH, and the bulk mu_species from the batched device bulk solve, go to the context
  through sdft_set_chemistry;
* the fields are then evaluated in FP64 by sdft_objective.
The Anderson iteration on the augmented vectors runs on the host (numpy). It follows
AndersonMixing.jl:56-169:
* the Picard warm-up and the one-way switch;
* the damping rule;
* the stop test taken after the update.
The least-squares coefficients come from numpy's minimum-norm lstsq in place of IterativeSolvers'
lsqr!. Both solve the same problem, so only the path to the fixed point can differ, never the fixed
point itself.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import abi, analysis
from .molsys import build_molsys
from .problem import build_state_points


@dataclass
class PacOptions:
    """Continuation numerics (objective_pac.jl:46-55 defaults) and the corrector's Anderson numerics
    (AndersonMixing.jl:74-79 names)."""
    num_keep: int = 5
    ds_init: float = 0.05
    max_steps: int = 20000
    pH_min: float = 1.99
    pH_max: float = 12.01
    tole: float = 1e-8
    damping: float = 0.1
    mixing_max: float = 1.0
    switch_tole: float = 1e-3
    anderm: int = 5
    max_iters: int = 20000
    banded: bool = False     # FP64 objective on the banded bond operator (dense by default)


@dataclass
class PacStep:
    step: int                # the reference's step counter (1, 2: start solves; 3.. continuation)
    param: float             # as the reference reports it: the pH, or the grafting density
    cont: float              # the continuation variable itself: the pH, or log(grafting density)
    alpha: float             # average state-2 fraction of the first species (NaN if it has one state)
    brush_height: float      # first grafted species, NaN without one
    omega: float
    iterations: int
    residual: float
    arclength: float         # s after this step
    x: Optional[np.ndarray] = field(default=None, repr=False)


def pac_tangent(X_hist: np.ndarray, S_vec: np.ndarray, step: int, num_keep: int) -> np.ndarray:
    """Normalised tangent dY/dS from the history ring (objective_pac.jl:127-146): rows of X_hist are
    the stored solutions, slot mod1(k, num_keep) holds step k's."""
    i1 = (step - 2) % num_keep
    i2 = (step - 3) % num_keep
    i3 = (step - 4) % num_keep
    if step == 3 or num_keep < 3:
        t = X_hist[i1] - X_hist[i2]
    else:
        ds1 = S_vec[i1] - S_vec[i2]
        ds2 = S_vec[i2] - S_vec[i3]
        a = ds1 / (ds2 * (ds1 + ds2))
        b = -(ds1 + ds2) / (ds1 * ds2)
        c = (2.0 * ds1 + ds2) / (ds1 * (ds1 + ds2))
        t = a * X_hist[i3] + b * X_hist[i2] + c * X_hist[i1]
    return t / np.linalg.norm(t)


class HostAnderson:
    """One Anderson(m) iteration state (AndersonMixing.jl:56-169) on a host vector of length n."""

    def __init__(self, n: int, m: int, damping: float, mixing_max: float, switch_tole: float):
        self.m = m
        self.damping, self.mixing_max, self.switch_tole = damping, mixing_max, switch_tole
        self.X = np.zeros((m, n))
        self.G = np.zeros((m, n))
        self.iter = 0
        self.use_picard = True

    def update(self, x: np.ndarray, f: np.ndarray):
        """One iteration from x and f = F(x): returns (next x, err = max|F(x) - x|)."""
        self.iter += 1
        col = (self.iter - 1) % self.m
        g = f - x
        self.X[col] = x
        self.G[col] = g
        err = max(float(np.max(np.abs(g))), 1e-20)
        amix = min(self.mixing_max, self.damping / err)
        if self.iter < self.m or (err > self.switch_tole and self.use_picard):
            return (1.0 - amix) * x + amix * f, err
        self.use_picard = False
        idx = [j for j in range(self.m) if j != col]
        dX = self.X[idx] - x
        dG = self.G[idx] - g
        lam = np.linalg.lstsq(dG.T, -g, rcond=None)[0]
        return x + dX.T @ lam + amix * (g + dG.T @ lam), err


def _ph_species_index(dataset: dict) -> int:
    for k, sp in enumerate(dataset["species"]):
        if sp.get("sequence") == "Y":
            return k
    raise ValueError("pH continuation needs a species with sequence 'Y' (H+) given by pSpecies")


class PhChemistry:
    """delta_muH and mu_species of each branch's dataset at a given pH (update_cont_var +
    establish_dft_system, objective_pac.jl:263-270: a new molecular system and bulk state per pH),
    with the bulk states of all branches solved in one device launch."""

    def __init__(self, datasets: Sequence[dict], grafted: Sequence[Sequence[bool]]):
        self.datasets = [copy.deepcopy(d) for d in datasets]
        self.iy = [_ph_species_index(d) for d in self.datasets]
        self.grafted = [np.asarray(g, dtype=bool) for g in grafted]

    def __call__(self, pH: Sequence[float]):
        from .bulk_device import solve_bulk_batch
        mss = []
        for d, iy, p in zip(self.datasets, self.iy, pH):
            d["species"][iy]["pSpecies"] = float(p)
            mss.append(build_molsys(copy.deepcopy(d)))
        bulks = solve_bulk_batch(mss)
        dmu = np.stack([ms.delta_muH for ms in mss])
        mus = np.stack([b.mu_species.copy() for b in bulks])
        for k, g in enumerate(self.grafted):
            mus[k, g] = 0.0   # grafted species (handle_fixed.jl:43)
        return dmu, mus


class _PhVariable:
    """pSpecies of H+ (update_cont_var, objective_pac.jl:286-293)."""
    report = "pH"
    early_stop = True

    def __init__(self, datasets, sps, o: PacOptions):
        self.chem = PhChemistry(datasets, [[f.any for f in sp.fixed] for sp in sps])
        self.lo, self.hi = o.pH_min, o.pH_max

    @staticmethod
    def initial(datasets) -> np.ndarray:
        return np.array([float(d["species"][_ph_species_index(d)]["pSpecies"]) for d in datasets])

    def apply(self, batch: abi.Batch, v: np.ndarray) -> None:
        dmu, mus = self.chem(v)
        batch.set_chemistry(dmu, mus)

    def in_range(self, v: float) -> bool:
        return self.lo < v < self.hi

    @staticmethod
    def param(v: float) -> float:
        return float(v)


class _GraftVariable:
    """log of the first species' grafting density (update_cont_var_graft, objective_pac_graft.jl:286-288);
    the other species keep their densities."""
    report = "graft_dens"
    early_stop = False

    def __init__(self, datasets, sps, o: PacOptions):
        if not sps[0].fixed or not sps[0].fixed[0].any:
            raise ValueError("grafting-density continuation needs a grafted first species")
        self.base = np.array([[f.density for f in sp.fixed] for sp in sps], dtype=np.float64)

    @staticmethod
    def initial(datasets) -> np.ndarray:
        return np.array([math.log(float(d["species"][0]["fixed"]["density"])) for d in datasets])

    def apply(self, batch: abi.Batch, v: np.ndarray) -> None:
        dens = self.base.copy()
        dens[:, 0] = np.exp(v)
        batch.set_graft_density(dens)

    @staticmethod
    def in_range(v: float) -> bool:
        return v < 1.0

    @staticmethod
    def param(v: float) -> float:
        return math.exp(v)


_VARIABLES = {"pH": _PhVariable, "log_graft_density": _GraftVariable}


def _observe(batch: abi.Batch, sps, k: int, om, al):
    segs = [batch.segment_profile(u, k) for u in range(batch.nsp)]
    hs = analysis.brush_heights(segs, sps[k])
    # the first species' average state-2 fraction (objective_pac.jl:190-192)
    first = sps[k].ms.species[0]
    alpha = float(al[k, 0, 1]) if any(len(f) > 1 for f in first.state_family) else math.nan
    height = float(hs[0][1]) if len(hs) >= 1 else math.nan
    return alpha, height, float(om[k])


def pac_continuation(datasets: Sequence[dict], opts: Optional[PacOptions] = None, *,
                     variable: str = "pH",
                     on_step: Optional[Callable[[int, PacStep], None]] = None,
                     keep_x: bool = False) -> List[List[PacStep]]:
    """Trace one branch per dataset (all sharing one grid and molecular layout, e.g. one per salt),
    all branches stepping together, in `variable` ("pH" or "log_graft_density"). Returns per branch
    its steps; a branch stops on its own when it leaves the variable's range, reaches the pH
    variant's degree-of-ionization bound or max_steps."""
    o = opts or PacOptions()
    sps = build_state_points(datasets, device_bulk=True)
    B = len(sps)
    var = _VARIABLES[variable](datasets, sps, o)
    pH = var.initial(datasets)
    batch = abi.Batch(sps)
    batch.set_operator_band(fp32=True, fp64=o.banded)
    solve_opts = dict(tole=o.tole, damping=o.damping, mixing_max=o.mixing_max, switch_tole=o.switch_tole,
                      anderm=o.anderm, max_iters=o.max_iters, precision=abi.PREC_FP64, band_fp32=int(o.banded),
                      raise_on_noconv=False)
    out: List[List[PacStep]] = [[] for _ in range(B)]

    def record(k, step, x_aug, iters, err, s, om, al):
        alpha, height, omega = _observe(batch, sps, k, om, al)
        st = PacStep(step=step, param=var.param(x_aug[-1]), cont=float(x_aug[-1]), alpha=alpha,
                     brush_height=height, omega=omega,
                     iterations=int(iters), residual=float(err), arclength=float(s),
                     x=x_aug[:-1].copy() if keep_x else None)
        out[k].append(st)
        if on_step is not None:
            on_step(k, st)
        return st

    def evaluate_final(xs: np.ndarray, pHs: np.ndarray):
        # the fields' densities at the final point (objective_pac.jl:194-195 evaluates before analysis)
        var.apply(batch, pHs)
        batch.objective(np.concatenate(list(xs)), abi.PREC_FP64)
        return batch.observables()

    # steps 1 and 2: plain solves at pH and pH + ds_init (objective_pac.jl:64-101)
    x, it1, err1 = batch.solve(batch.x0(), **solve_opts)
    X_prev = np.stack([np.append(xp, p) for xp, p in zip(batch.split(x), pH)])
    om, al, _ = evaluate_final(X_prev[:, :-1], X_prev[:, -1])
    for k in range(B):
        record(k, 1, X_prev[k], it1[k], err1[k], 0.0, om, al)
    pH2 = pH + o.ds_init
    var.apply(batch, pH2)
    x, it2, err2 = batch.solve(x, **solve_opts)
    X_k = np.stack([np.append(xp, p) for xp, p in zip(batch.split(x), pH2)])
    s = np.linalg.norm(X_k - X_prev, axis=1)
    om, al, _ = evaluate_final(X_k[:, :-1], X_k[:, -1])
    for k in range(B):
        record(k, 2, X_k[k], it2[k], err2[k], s[k], om, al)

    n_aug = X_k.shape[1]
    X_hist = np.zeros((B, o.num_keep, n_aug))
    S_vec = np.zeros((B, o.num_keep))
    X_hist[:, 0], X_hist[:, 1] = X_prev, X_k
    S_vec[:, 1] = s
    dS = np.full(B, o.ds_init)
    ds_con = o.ds_init   # the constraint's ds, fixed at the start (objective_pac.jl:104-106)
    active = np.array([var.in_range(X_k[k, -1]) for k in range(B)])
    step = 3
    while active.any() and step - 2 <= o.max_steps:
        T = np.stack([pac_tangent(X_hist[k], S_vec[k], step, o.num_keep) for k in range(B)])
        X_prev = X_k.copy()
        Y = X_k + dS[:, None] * T          # predictor
        Y[~active] = X_k[~active]
        ands = [HostAnderson(n_aug, o.anderm, o.damping, o.mixing_max, o.switch_tole) for _ in range(B)]
        done = ~active
        iters = np.zeros(B, dtype=np.int64)
        errs = np.zeros(B)
        while not done.all():
            var.apply(batch, Y[:, -1])
            fx = batch.split(batch.objective(np.concatenate(list(Y[:, :-1])), abi.PREC_FP64))
            for k in range(B):
                if done[k]:
                    continue
                tx, tp = T[k, :-1], T[k, -1]
                p_new = X_prev[k, -1] + (ds_con - float(np.dot(tx, fx[k] - X_prev[k, :-1]))) / tp
                G = np.append(fx[k], p_new)
                Y[k], errs[k] = ands[k].update(Y[k], G)
                iters[k] = ands[k].iter
                if not np.all(np.isfinite(Y[k])):
                    raise abi.SdftError(-3, f"branch {k}: non-finite corrector iterate at step {step}")
                if errs[k] < o.tole:
                    done[k] = True
                elif ands[k].iter >= o.max_iters:
                    raise abi.SdftError(-4, f"branch {k}: corrector did not converge in {o.max_iters} "
                                            f"iterations at step {step} (err {errs[k]:.3e})")
        X_k = np.where(active[:, None], Y, X_k)
        ds_actual = np.linalg.norm(X_k - X_prev, axis=1)
        s = s + np.where(active, ds_actual, 0.0)
        dS = np.where(active, (dS + ds_actual) / 2.0, dS)
        slot = (step - 1) % o.num_keep
        X_hist[:, slot] = X_k
        S_vec[:, slot] = s
        om, al, _ = evaluate_final(X_k[:, :-1], X_k[:, -1])
        for k in range(B):
            if not active[k]:
                continue
            st = record(k, step, X_k[k], iters[k], errs[k], s[k], om, al)
            # the reference tests sign(dS), which its update (dS + ds_actual) / 2 keeps positive
            if var.early_stop and ((st.alpha >= 0.99 and dS[k] > 0.0) or (st.alpha <= 0.01 and dS[k] < 0.0)):
                active[k] = False
            elif not var.in_range(st.cont):
                active[k] = False
        step += 1
    batch.close()
    return out


def write_pac_summary(results: Sequence[PacStep], path: str, variable: str = "pH") -> None:
    """The reference's summary file (objective_pac.jl:383-390, objective_pac_graft.jl:373-380):
    step, the reported parameter (pH or graft_dens), brush height and alpha, printed %.12g."""
    def g12(v: float) -> str:
        if math.isnan(v):
            return "NaN"
        if math.isinf(v):
            return "Inf" if v > 0 else "-Inf"
        return "%.12g" % v
    with open(path, "w") as f:
        f.write(f"step,{_VARIABLES[variable].report},brush_height,alpha\n")
        for r in results:
            f.write(f"{r.step},{g12(r.param)},{g12(r.brush_height)},{g12(r.alpha)}\n")
