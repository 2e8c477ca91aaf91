"""Square-well, point-charge and file-given walls through the library (their potentials enter as Ext;
the square-well walls add contact-theorem terms in sdft_observables). Tolerances as in the parity
suite: converged Omega 1e-8 relative to max(1, |Omega|), contact pressures 1e-8 absolute."""
import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point
from oracle.dft_oracle import OracleDFT, anderson_solve

pytestmark = pytest.mark.gpu


def _ideal():
    return {"procedure": "iDFT",
            "geometry": {"coordinate_system": "Cartesian", "dimensions": [4.0], "bin_width": [0.05],
                         "external_field": {"hard_walls": {"dims": [1]},
                                            "sw_walls": {"dims": [1], "monomers": {
                                                "A": {"energy": [[-0.7, 0.4]], "lambda": [[1.5, 1.3]]}}}}},
            "species": [{"sequence": "A", "density": 0.3}, {"sequence": "S", "density": 0.2}],
            "monomers": {"A": {"diameter": 1.0, "valence": 0.0}, "S": {"diameter": 0.8, "valence": 0.0}},
            "numerical": dict(configs.SCF_NUMERICS), "model": ["ideal"]}


def test_square_well_walls_ideal_gas_contact_theorem_on_device():
    sp = build_state_point(_ideal())
    b = abi.Batch([sp])
    print("case sw walls ideal gas", flush=True)
    fx = b.objective(b.x0(), abi.PREC_FP64)
    assert np.max(np.abs(fx - b.x0())) < 1e-14
    _, _, pc = b.observables()
    assert pc[0, 0] == pytest.approx(sp.bulk.pressure, rel=1e-12)
    assert pc[0, 1] == pytest.approx(sp.bulk.pressure, rel=1e-12)
    b.close()


def _chains_sw_walls():
    ds = configs.hard_sphere_chains(H=6.0, dz=0.05, n_seg=5)
    ds["geometry"]["external_field"]["sw_walls"] = {
        "dims": [1], "monomers": {"A": {"energy": [[-0.5, -0.2]], "lambda": [[1.5, 1.5]]},
                                  "S": {"energy": [[0.3, 0.3]], "lambda": [[1.2, 1.2]]}}}
    return ds


def _pe_pointcharge():
    ds = configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=4, solvent=0.2, pH=6.0)
    ds["geometry"]["external_field"]["pointcharge_walls"] = {"dims": [1], "charge": [[-0.04, -0.01]]}
    return ds


@pytest.mark.parametrize("name,builder", [("chains_sw_walls", _chains_sw_walls), ("pe_pointcharge", _pe_pointcharge)])
def test_wall_systems_solve_matches_oracle(name, builder):
    sp = build_state_point(builder())
    b = abi.Batch([sp])
    o = OracleDFT(sp)
    num = sp.numerics
    print(f"case walls solve {name}", flush=True)
    r = anderson_solve(o.objective, sp.x0(), num)
    om_ref = o.grand_potential(r.x)
    pc_ref = o.contact_pressure()
    opts = dict(tole=num["tole"], damping=num["damping"], mixing_max=num["mixing_max"],
                switch_tole=num["switch_tole"], anderm=int(num["anderm"]), max_iters=3000)
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        x, it, err = b.solve(b.x0(), precision=prec, **opts)
        om, _, pc = b.observables()
        assert err[0] < num["tole"]
        assert abs(om[0] - om_ref) < 1e-8 * max(1.0, abs(om_ref))
        assert np.allclose(pc[0], pc_ref, atol=1e-8)
    b.close()
