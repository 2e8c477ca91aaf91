"""GPU: two- and three-dimensional Cartesian grids (sdft_nd.hip) against the ND oracle.

The library evaluates every convolution as per-dimension cosine / sine transforms on the primary grid.
The oracle uses fftn on the mirrored domain, as the reference does. Checked:
* FP64 objective to 1e-11 relative at the bulk guess and at a non-uniform state (a few oracle
  Picard steps). Covered systems:
  - a 2-D slab of a monomeric weak acid with salt and solvent;
  - a 3-D square channel of weak-acid 4-mers, walls along x and y;
  - a 3-D slab between charged walls;
  - a two-point batch at two pH values;
* FP32 objective to 5e-5;
* a converged FP64 and mixed-precision solve of the channel against the oracle's solve
  (fields 1e-7, Omega 1e-9 relative plus the neutrality band).
Widths are chosen so that the map is well conditioned. In a channel only a bead diameter wide,
hundreds of wall points carry n3 between the FMT threshold 1e-10 and 1e-6, where the derivative of
the reference's free energy cancels to a few digits. There the oracle's own F(x) moves by 6e-6 when
x is perturbed by 1e-15 relative, so no two evaluation orders agree to 1e-11. The channel here
moves by 2e-13 under the same perturbation.
"""
import copy

import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.geometry_nd import build_state_point_nd

pytestmark = pytest.mark.gpu


def _nd(ds1, dims, bins, walls, charge=None):
    d = copy.deepcopy(ds1)
    d["geometry"]["dimensions"] = list(dims)
    d["geometry"]["bin_width"] = list(bins)
    ext = {"hard_walls": {"dims": list(walls)}}
    if charge is not None:
        ext["charged_walls"] = {"dims": [walls[-1]], "charge": [list(charge)]}
    d["geometry"]["external_field"] = ext
    return d


CASES = {
    "slab2d_monomer": lambda: _nd(configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=1, solvent=0.2, pH=5.0),
                                  [1.2, 3.0], [0.1, 0.1], [2]),
    "channel3d_4mer": lambda: _nd(configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=4, solvent=0.2, pH=5.0),
                                  [1.5, 1.6, 0.3], [0.1, 0.1, 0.1], [1, 2]),
    "slab3d_charged": lambda: _nd(configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=3, solvent=0.2, pH=6.0),
                                  [0.4, 0.3, 2.0], [0.1, 0.1, 0.1], [3], charge=(-0.05, -0.05)),
}


def _perturbed(o, x, n=3, a=0.3):
    for _ in range(n):
        x = (1 - a) * x + a * o.objective(x)
    return x


@pytest.mark.parametrize("name", sorted(CASES))
def test_nd_objective_matches_oracle(name):
    from oracle.dft_oracle_nd import OracleDFTND
    sp = build_state_point_nd(CASES[name]())
    print(f"case nd objective {name} grid {sp.grid}", flush=True)
    o = OracleDFTND(sp)
    b = abi.Batch([sp])
    for x in (sp.x0(), _perturbed(o, sp.x0())):
        ref = o.objective(x)
        scale = max(1.0, float(np.max(np.abs(ref))))
        f64 = b.objective(x, abi.PREC_FP64)
        assert np.max(np.abs(f64 - ref)) < 1e-11 * scale, name
        f32 = b.objective(x, abi.PREC_FP32)
        assert np.max(np.abs(f32 - ref)) < 5e-5 * scale, name
    b.close()


def test_nd_batch_of_two_points():
    from oracle.dft_oracle_nd import OracleDFTND
    base = configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=3, solvent=0.2, pH=5.0)
    dss = []
    for pH in (4.5, 6.5):
        d = copy.deepcopy(base)
        for s in d["species"]:
            if s.get("sequence") == "Y":
                s["pSpecies"] = pH
        dss.append(_nd(d, [0.5, 0.4, 1.6], [0.1, 0.1, 0.1], [3]))
    sps = [build_state_point_nd(d) for d in dss]
    print("case nd batch of two points", flush=True)
    b = abi.Batch(sps)
    xs = [_perturbed(OracleDFTND(sp), sp.x0(), n=2) for sp in sps]
    fx = b.split(b.objective(np.concatenate(xs), abi.PREC_FP64))
    for sp, x, f in zip(sps, xs, fx):
        ref = OracleDFTND(sp).objective(x)
        assert np.max(np.abs(f - ref)) < 1e-11 * max(1.0, float(np.max(np.abs(ref))))
    b.close()


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_nd_solve_matches_oracle_solve(precision):
    from oracle.dft_oracle import anderson_solve
    from oracle.dft_oracle_nd import OracleDFTND
    from parity_helpers import omega_band
    sp = build_state_point_nd(CASES["channel3d_4mer"]())
    num = dict(sp.numerics)
    num["tole"] = 1e-11
    print(f"case nd solve channel3d_4mer precision {precision}", flush=True)
    o = OracleDFTND(sp)
    r = anderson_solve(o.objective, sp.x0(), num, max_iters=20000)
    om_ref = o.grand_potential(r.x)
    band = omega_band(sp, o, r, om_ref, num)
    b = abi.Batch([sp])
    x, it, err = b.solve(sp.x0(), tole=1e-11, damping=num["damping"], mixing_max=num["mixing_max"],
                         switch_tole=num["switch_tole"], anderm=int(num["anderm"]), max_iters=20000,
                         precision=precision)
    om, _, _ = b.observables()
    assert err[0] < 1e-11
    assert np.max(np.abs(x - r.x)) < 1e-7
    assert abs(om[0] - om_ref) < 1e-9 * abs(om_ref) + band
    # the converged bead densities against the oracle's at the same point
    o.objective(x)
    rb = b.get_field(abi.FIELD_RHO_BEADS, 0).reshape(o.rho_beads.shape, order="F")
    assert np.max(np.abs(rb - o.rho_beads)) < 1e-8 * np.max(np.abs(o.rho_beads))
    b.close()
