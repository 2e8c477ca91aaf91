"""GPU: segments with three states (a diprotic acid, A -> B + Y, B -> C + Y; state_families.jl).

Species with more than two states per segment take the general-state instantiations:
* the direct propagator step (k_prop_step with 64-point blocks and LDS for 16 state pairs);
* the density accumulation with four-state register arrays.
The MFMA chain, banded-step and operator-GEMM propagators stay two-state. Against the oracle:
* free chains: FP64 objective 1e-11 and FP32 5e-5 at the bulk guess and a non-uniform state;
  converged FP64 and mixed solves (fields 1e-7, Omega 1e-9 relative plus the neutrality band);
  per-state ionization fractions;
* a grafted brush: objectives;
* a three-dimensional channel: objectives.
"""
import copy

import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point

pytestmark = pytest.mark.gpu


def _perturbed(o, x, n=3, a=0.3):
    for _ in range(n):
        x = (1 - a) * x + a * o.objective(x)
    return x


def _check_objectives(sp, o, b):
    for x in (sp.x0(), _perturbed(o, sp.x0())):
        ref = o.objective(x)
        scale = max(1.0, float(np.max(np.abs(ref))))
        assert np.max(np.abs(b.objective(x, abi.PREC_FP64) - ref)) < 1e-11 * scale
        assert np.max(np.abs(b.objective(x, abi.PREC_FP32) - ref)) < 5e-5 * scale


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_three_state_chains_match_oracle(precision):
    from oracle.dft_oracle import OracleDFT, anderson_solve
    from parity_helpers import omega_band
    sp = build_state_point(configs.diprotic_polyelectrolyte(H=5.0, dz=0.05, n_seg=6, solvent=0.2, pH=5.5))
    assert max(len(f) for f in sp.ms.species[0].state_family) == 3
    print(f"case three-state chains precision {precision}", flush=True)
    o = OracleDFT(sp)
    b = abi.Batch([sp])
    _check_objectives(sp, o, b)
    num = dict(sp.numerics)
    num["tole"] = 1e-11
    r = anderson_solve(o.objective, sp.x0(), num, max_iters=20000)
    om_ref = o.grand_potential(r.x)
    band = omega_band(sp, o, r, om_ref, num)
    alpha_ref = o.state_fractions()[0]
    x, it, err = b.solve(sp.x0(), tole=1e-11, damping=num["damping"], mixing_max=num["mixing_max"],
                         switch_tole=num["switch_tole"], anderm=int(num["anderm"]), max_iters=20000,
                         precision=precision)
    om, al, _ = b.observables()
    assert err[0] < 1e-11
    assert np.max(np.abs(x - r.x)) < 1e-7
    assert abs(om[0] - om_ref) < 1e-9 * abs(om_ref) + band
    assert np.max(np.abs(al[0, 0, :3] - alpha_ref)) < 1e-8
    b.close()


def test_three_state_grafted_brush_objective():
    from oracle.dft_oracle import OracleDFT
    sp = build_state_point(configs.diprotic_polyelectrolyte(H=6.0, dz=0.05, n_seg=8, pH=5.5, pKa=5.0,
                                                            grafted=0.05))
    print("case three-state grafted brush", flush=True)
    b = abi.Batch([sp])
    _check_objectives(sp, OracleDFT(sp), b)
    b.close()


def test_three_state_channel_3d_objective():
    from oracle.dft_oracle_nd import OracleDFTND
    from softdft_amd.geometry_nd import build_state_point_nd
    d = configs.diprotic_polyelectrolyte(H=2.0, dz=0.1, n_seg=3, solvent=0.2, pH=5.5)
    d["geometry"]["dimensions"] = [1.5, 1.6, 0.3]
    d["geometry"]["bin_width"] = [0.1, 0.1, 0.1]
    d["geometry"]["external_field"] = {"hard_walls": {"dims": [1, 2]}}
    sp = build_state_point_nd(d)
    print(f"case three-state 3-D channel grid {sp.grid}", flush=True)
    b = abi.Batch([sp])
    _check_objectives(sp, OracleDFTND(sp), b)
    b.close()
