"""Monomerbead species (independent beads: src/iDFT/coordinate_system/Cartesian/species_model/
monomerbead/*.jl) with several beads and with a grafted first bead, beside a freely jointed chain,
against the FP64 oracle: objective (FP64 1e-11, FP32 5e-5), segment profiles, and solves to 1e-12
(densities 1e-8, Omega 1e-10 plus the neutrality band, degree of ionization 1e-9)."""
import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point
from oracle.dft_oracle import OracleDFT, anderson_solve
from parity_helpers import omega_band

pytestmark = pytest.mark.gpu


def _free():
    ds = configs.weak_polyelectrolyte(H=6.0, dz=0.05, n_seg=4, solvent=0.2, pH=5.0)
    ds["species"].append({"sequence": "AAA", "density": 0.03, "model": "monomerbead"})
    return ds


def _grafted():
    ds = configs.weak_polyelectrolyte(H=6.0, dz=0.05, n_seg=4, solvent=0.2, pH=5.5)
    ds["species"].append({"sequence": "AA", "density": 0.0, "model": "monomerbead",
                          "fixed": {"density": 0.03, "segment": [1], "position": ["left"]}})
    return ds


CASES = {"free": _free, "grafted": _grafted}


@pytest.mark.parametrize("name", list(CASES))
def test_monomerbead_objective_matches_oracle(name):
    sp = build_state_point(CASES[name]())
    b = abi.Batch([sp])
    u = b.nsp - 1
    rng = np.random.default_rng(13)
    for x in (b.x0(), b.x0() + 0.03 * rng.standard_normal(b.n_total)):
        print(f"case monomerbead objective {name}", flush=True)
        o = OracleDFT(sp)
        ref = o.objective(x)
        scale = max(1.0, float(np.max(np.abs(ref))))
        fx = b.objective(x, abi.PREC_FP64)
        assert np.max(np.abs(fx - ref)) / scale < 1e-11
        assert np.allclose(b.get_field(abi.FIELD_RHO_BEADS, 0), o.rho_beads, rtol=1e-11, atol=1e-14)
        seg = b.segment_profile(u, 0)
        assert np.max(np.abs(seg - o.rho_seg[u])) <= 1e-11 * float(np.max(np.abs(o.rho_seg[u])))
        fx32 = b.objective(x, abi.PREC_FP32)
        assert np.max(np.abs(fx32 - ref)) / scale < 5e-5
    b.close()


@pytest.mark.parametrize("name", list(CASES))
def test_monomerbead_solve_matches_oracle(name):
    sp = build_state_point(CASES[name]())
    b = abi.Batch([sp])
    num = dict(sp.numerics)
    num["tole"] = 1e-12
    kw = dict(tole=1e-12, damping=num["damping"], mixing_max=num["mixing_max"], switch_tole=num["switch_tole"],
              anderm=int(num["anderm"]), max_iters=8000)
    o = OracleDFT(sp)
    r = anderson_solve(o.objective, sp.x0(), num, max_iters=20000)
    rb_ref = o.rho_beads.copy()
    om_ref = o.grand_potential(r.x)
    band = omega_band(sp, o, r, om_ref, num)
    fr = o.state_fractions()
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        print(f"case monomerbead solve {name} precision={prec}", flush=True)
        x, it, err = b.solve(b.x0(), precision=prec, **kw)
        assert err[0] < 1e-12
        om, al, pc = b.observables()
        rb = b.get_field(abi.FIELD_RHO_BEADS, 0)
        assert np.max(np.abs(rb - rb_ref)) / np.max(np.abs(rb_ref)) < 1e-8
        assert abs(om[0] - om_ref) < 1e-10 * max(1.0, abs(om_ref)) + band
        for u, f in enumerate(fr):
            if f is not None:
                assert np.allclose(al[0, u, :len(f)], f, atol=1e-9)
    b.close()
