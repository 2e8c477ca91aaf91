"""GPU: slit-width continuation with neighbour warm starts against independent cold solves."""
import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point
from softdft_amd.sweeps import width_sweep

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tune", [(abi.PREC_FP64, False), (abi.PREC_MIXED, False), (abi.PREC_MIXED, True)])
def test_width_sweep_matches_cold_solves(precision, tune):
    Hs = [4.0 + 0.05 * i for i in range(6)]
    dss = [configs.weak_polyelectrolyte(H=H, dz=0.05, n_seg=4, solvent=0.2, pH=7.0) for H in Hs]
    print(f"case width_sweep precision={precision} tune={tune}", flush=True)
    res = width_sweep(dss, waves=3, precision=precision, tune=tune)
    # round r of 3 solves widths i % 3 == r; round 0 starts cold, the others from width i - 1
    assert [r.warm for r in res] == [i % 3 != 0 for i in range(6)]
    sps = [build_state_point(d) for d in dss]
    b = abi.Batch(sps)  # cold: every width from its bulk guess
    num = sps[0].numerics
    x, it, err = b.solve(b.x0(), tole=num["tole"], damping=num["damping"], mixing_max=num["mixing_max"],
                         switch_tole=num["switch_tole"], anderm=int(num["anderm"]), max_iters=3000,
                         precision=abi.PREC_FP64)
    om, _, pc = b.observables()
    for k, (r, sp) in enumerate(zip(res, sps)):
        # the warm start changes the path, not the fixed point: both stop at ||F(x) - x|| < 1e-8
        assert abs(r.omega - om[k]) < 1e-8 * max(1.0, abs(om[k]))
        assert abs(r.delta_p - (pc[k, 0] - sp.bulk.pressure)) < 1e-7
        assert abs(r.energy - (om[k] + sp.bulk.pressure * sp.H)) < 1e-7 * max(1.0, abs(r.energy))
    b.close()


def test_ph_sweep_matches_cold_solves():
    from softdft_amd import analysis
    from softdft_amd.sweeps import ph_sweep
    pHs = [4.0, 4.5, 5.0, 5.5, 6.0, 6.5]
    dss = [configs.weak_polyelectrolyte(H=8.0, dz=0.05, n_seg=10, pH=pH, pKa=5.0, grafted=0.05) for pH in pHs]
    print("case ph_sweep brush", flush=True)
    res = ph_sweep(dss, pHs, waves=3, precision=abi.PREC_MIXED)
    assert [r.step for r in res] == list(range(1, len(res) + 1))
    assert [r.warm for r in res] == [i % 3 != 0 for i in range(len(res))]
    sps = [build_state_point(d) for d in dss]
    b = abi.Batch(sps)
    num = sps[0].numerics
    x, it, err = b.solve(b.x0(), tole=num["tole"], damping=num["damping"], mixing_max=num["mixing_max"],
                         switch_tole=num["switch_tole"], anderm=int(num["anderm"]), max_iters=3000,
                         precision=abi.PREC_FP64)
    om, al, _ = b.observables()
    for k, r in enumerate(res):
        segs = [b.segment_profile(u, k) for u in range(b.nsp)]
        (_, h), = analysis.brush_heights(segs, sps[k])
        assert abs(r.omega - om[k]) < 1e-8 * max(1.0, abs(om[k]))
        assert abs(r.alpha - al[k, 0, 1]) < 1e-7
        assert abs(r.brush_height - h) < 1e-6 * h
    # ionization grows with pH on a weak acid
    assert all(res[k + 1].alpha > res[k].alpha for k in range(len(res) - 1))
    b.close()


def test_width_sweep_against_oracle_solves():
    """Pi(H) = P_cvt - P_bulk and the energy Omega + P_bulk H of a warm-started width sweep
    (routines/adjust_distanceZ.jl:96-108) against oracle solves of the same widths, every point
    solved to 1e-11: Pi and the energy within 1e-9 (plus the neutrality band of Omega)."""
    from oracle.dft_oracle import OracleDFT, anderson_solve
    from parity_helpers import omega_band
    Hs = [4.0, 4.5, 5.0]
    dss = [configs.weak_polyelectrolyte(H=H, dz=0.05, n_seg=4, solvent=0.2, pH=7.0) for H in Hs]
    num = dict(dss[0]["numerical"])
    num["tole"] = 1e-11
    print("case width_sweep vs oracle", flush=True)
    res = width_sweep(dss, waves=3, precision=abi.PREC_MIXED, numerics=num, max_iters=6000)
    assert [r.warm for r in res] == [False, True, True]
    for r, d in zip(res, dss):
        sp = build_state_point(d)
        o = OracleDFT(sp)
        rr = anderson_solve(o.objective, sp.x0(), num, max_iters=20000)
        om = o.grand_potential(rr.x)
        band = omega_band(sp, o, rr, om, num)
        pc = o.contact_pressure()
        assert r.residual < 1e-11
        assert abs(r.delta_p - (pc[0] - sp.bulk.pressure)) < 1e-9
        assert abs(r.energy - (om + sp.bulk.pressure * sp.H)) < 1e-9 * max(1.0, abs(r.energy)) + band
        assert r.distance == pytest.approx(sp.H)
