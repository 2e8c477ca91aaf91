"""Writes brush500_dz0.05_pH3_1mM_x.npy: the converged solution vector (reference flatten order) of the
500-mer brush chemistry (grafting 0.05, pKa 5, pH 3, 1 mM salt, H = 81.9149) on a dz = 0.05 grid
(NP 1640), solved by libsoftdft_hip in mixed precision (banded bond operator) to ||F(x)-x|| < 1e-8.
Needs a HIP device. It is an input state for the oracle pins of the propagator evaluation at a
physical brush state, where the propagators of one segment span more than 16 decades."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from softdft_amd import abi, configs  # noqa: E402
from softdft_amd.problem import build_state_point  # noqa: E402


def main():
    ds = configs.weak_polyelectrolyte(H=81.9149, dz=0.05, n_seg=500, pKa=5.0, pH=3.0, salt_M=0.001, grafted=0.05,
                                      numerics=configs.SCF_NUMERICS_BRUSH)
    sp = build_state_point(ds)
    b = abi.Batch([sp])
    num = sp.numerics
    x, it, err = b.solve(b.x0(), tole=1e-8, damping=num["damping"], mixing_max=num["mixing_max"],
                         switch_tole=num["switch_tole"], anderm=int(num["anderm"]), max_iters=4000,
                         precision=abi.PREC_MIXED)
    print(f"iterations {it[0]}, residual {err[0]:.3e}")
    np.save(os.path.join(HERE, "brush500_dz0.05_pH3_1mM_x.npy"), x)


if __name__ == "__main__":
    main()
