"""sdft_observables_noalloc (no device allocation; partial sums in the library's static per-GPU
scratch array) against sdft_observables, on two calls in a row, the second with a smaller batch."""
import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point
from parity_helpers import hip_free_memory

pytestmark = pytest.mark.gpu


def _points(pHs):
    return [build_state_point(configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=4, solvent=0.2, pH=float(pH),
                                                           wall_charge=(-0.05, -0.02))) for pH in pHs]


def test_observables_noalloc_matches_observables():
    for pHs in ((4.0, 5.0, 6.0), (7.0,)):
        b = abi.Batch(_points(pHs))
        print(f"case observables_noalloc P={len(pHs)}", flush=True)
        x, it, err = b.solve(b.x0(), max_iters=3000)
        free0 = hip_free_memory()
        om_n, al_n, pc_n = b.observables_noalloc()
        free1 = hip_free_memory()
        om, al, pc = b.observables()
        assert free1 == free0
        assert np.array_equal(om_n, om)
        assert np.array_equal(pc_n, pc)
        assert np.array_equal(al_n, al, equal_nan=True)
        assert np.all(np.isfinite(om)) and np.all(pc != 0.0)
        assert np.isfinite(al[:, 0, :]).all()
        b.close()
