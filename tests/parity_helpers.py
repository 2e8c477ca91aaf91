"""Helpers shared by the GPU parity tests (not collected by pytest)."""
import ctypes as C

import numpy as np

from oracle.dft_oracle import OracleDFT, anderson_solve


def omega_band(sp, o, r, om_ref, num):
    """Width in Omega of the reference algorithm's own family of fixed points at a converged state.

    compute_neutrality_shift (neutrality_shift.jl:60-69) returns no shift once the charge imbalance
    |Q + sum_u N_u q_u| is at most 1e-10, so every state whose imbalance lies in that band is a fixed
    point: Psi_C is defined only to about 2e-10 / |f'| (f' = -sum_u N_u q_u^2 over the free charged
    species), and Omega to that times dOmega/dPsi_C. The slope is measured here by re-solving the
    oracle from its converged state with Psi_C moved across the band. 0 for systems without free
    charged species."""
    Q, N, q = o.neutrality_inputs
    fp = float(np.sum(N * q * q))
    if not fp > 0.0:
        return 0.0
    band = 2e-10 / fp
    x1 = r.x.copy()
    x1[-1] += 2.0 * band
    o1 = type(o)(sp)
    r1 = anderson_solve(o1.objective, x1, num, max_iters=20000)
    dpsi = float(r1.x[-1] - r.x[-1])
    if abs(dpsi) < 1e-15:
        return 0.0
    slope = abs(o1.grand_potential(r1.x) - om_ref) / abs(dpsi)
    return slope * band


def hip_free_memory() -> int:
    """Free device memory from hipMemGetInfo of the HIP runtime the library itself links."""
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def truncated_oracle(sp):
    """Oracle whose bond operators keep only |j| <= Be of the filtered bond kernel c, where B is the
    last offset with |c| > 1e-9 max|c| over the bond types and Be is B rounded up to even: the
    operator the banded chain kernel applies."""
    o = OracleDFT(sp)
    L = o.L
    d = np.arange(L)
    dd = np.minimum(d, L - d)
    cs = [np.fft.ifft(o.wbond_filt[:, u]).real for u in range(sp.MB)]
    B = max(int(dd[np.abs(c) > 1e-9 * np.abs(c).max()].max()) for c in cs)
    Be = B + (B & 1)
    for u, c in enumerate(cs):
        c = c.copy()
        c[dd > Be] = 0.0
        o.wbond_filt[:, u] = np.fft.fft(c).real
    return o, Be
