"""Full-size checks at the benchmark configurations, through the C ABI.

Where the oracle finishes in about a second per evaluation (the config-2 system: NP = 3001,
100-mers, six bead types) the FP64 objective is compared with it directly (1e-11 relative to
max |F|). The FP32 objective's error grows with the grid: at NP = 3001 it is 6.6e-5 relative,
largest in Psi, whose Coulomb operator sums 3001 long-range terms that nearly cancel (bound 2e-4;
the mixed solve only uses FP32 down to a residual of 1e-5, then polishes in FP64).

The batched pH sweep on the config-1 grid (NP = 801, 20-mers) is checked through properties that
do not depend on a CPU solve:
  * every point converges (||F(x) - x||_inf < 1e-8) in the mixed-precision solve;
  * the mixed solve's grand potential agrees with the batched FP64 solve to 1e-8 relative (the
    benchmark's acceptance bound);
  * the density profiles are mirror-symmetric about the slit centre (identical hard walls), to the
    solver tolerance;
  * the degree of ionization of the weak acid does not decrease with pH and goes from mostly
    neutral at pH 2 to mostly charged at pH 12.
"""
import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point
from oracle.dft_oracle import OracleDFT

pytestmark = pytest.mark.gpu


def test_config2_objective_matches_oracle_at_full_size():
    sp = build_state_point(configs.config2())
    assert sp.NP == 3001
    b = abi.Batch([sp])
    o = OracleDFT(sp)
    rng = np.random.default_rng(23)
    x0 = b.x0()
    for x in (x0, x0 + 0.02 * rng.standard_normal(len(x0))):
        print(f"case config2 objective n={len(x)}", flush=True)
        ref = o.objective(x)
        scale = max(1.0, float(np.max(np.abs(ref))))
        fx = b.objective(x, abi.PREC_FP64)
        assert np.max(np.abs(fx - ref)) / scale < 1e-11
        fx32 = b.objective(x, abi.PREC_FP32)
        assert np.max(np.abs(fx32 - ref)) / scale < 2e-4
    b.close()


def test_config3_batch_properties_at_full_size():
    pHs = np.linspace(2.0, 12.0, 64)
    dss = [configs.weak_polyelectrolyte(H=16.0, dz=0.02, n_seg=20, rho_poly=configs.RHO_POLY_SEG,
                                        solvent=configs.RHO_SOLVENT, pKa=4.5, pH=float(pH), salt_M=0.010)
           for pH in pHs]
    sps = [build_state_point(d) for d in dss]
    assert all(sp.NP == 801 for sp in sps)
    b = abi.Batch(sps)
    b.autotune(reps=1)
    num = dict(tole=1e-8, damping=0.1, mixing_max=0.02, switch_tole=100.0, anderm=5, max_iters=3000)
    print("case config3 mixed batch", flush=True)
    x, it, err = b.solve(b.x0(), precision=abi.PREC_MIXED, **num)
    assert np.all(err < 1e-8)
    om_m, al_m, _ = b.observables()
    rho = [b.get_field(abi.FIELD_RHO_BEADS, p) for p in range(len(sps))]
    print("case config3 FP64 batch", flush=True)
    b.solve(b.x0(), precision=abi.PREC_FP64, **num)
    om_d, _, _ = b.observables()
    assert np.max(np.abs(om_m - om_d) / np.abs(om_d)) < 1e-8
    for r in rho:
        scale = float(np.max(np.abs(r)))
        assert np.max(np.abs(r - r[::-1])) < 1e-6 * scale
    # species 0 is the weak-acid chain; state 1 is its charged form
    alpha = al_m[:, 0, 1]
    assert np.all(np.diff(alpha) > -1e-9)
    assert alpha[0] < 0.05 and alpha[-1] > 0.9
    b.close()
