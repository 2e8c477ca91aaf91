"""The chain kernel's two launch regimes against each other and against the FP64 oracle.

With more active points than half the CUs, both sweeps of a point share one 16-wave workgroup
(the sweeps offset by half a step); with fewer, each sweep runs in its own workgroup. The regime is
decided on the device from the number of points evaluated at the precision, so the same state
points evaluated in a large batch (paired sweeps) and in a small one (one sweep per workgroup) run
the two code paths over the same arithmetic: the chain species' segment densities (formed from the
propagators alone, per point) must agree bitwise, and F(x) must match the oracle with the same
truncated bond operator at 1e-11 (FP64; same arithmetic, different summation order). The banded bond
operator is selected at both precisions, and the chain kernel's profiling entry point, which
rejects species that do not run on it, confirms that it does.

Cases:
  * one-state segments (a neutral chain): one accumulation chain per unit;
  * two-state segments whose states have different bead diameters: the (a, b) columns apply
    different bond operators, one accumulation chain per column;
  * two-state segments with one bead diameter: both input states of a unit run as two chains over
    one operator window (the benchmark's chains).
"""
import copy

import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_points
from parity_helpers import truncated_oracle

pytestmark = pytest.mark.gpu

N_LARGE = 160   # > half the CUs of an MI355X (256): paired sweeps
SAMPLE = [0, 77, 159]


def _neutral(n):
    return [configs.hard_sphere_chains(H=6.0 + 0.01 * k, dz=0.05, n_seg=8) for k in range(n)]


def _two_diameters(n):
    out = []
    for k in range(n):
        d = configs.weak_polyelectrolyte(H=6.0, dz=0.05, n_seg=6, solvent=0.2, pH=4.0 + 2.0 * k / n)
        d = copy.deepcopy(d)
        d["monomers"]["B"]["diameter"] = 1.2   # the dissociated state is larger: distinct bond operators
        out.append(d)
    return out


def _one_diameter(n):
    return [configs.weak_polyelectrolyte(H=6.0, dz=0.05, n_seg=6, solvent=0.2, pH=4.0 + 2.0 * k / n)
            for k in range(n)]


@pytest.mark.parametrize("make", [_neutral, _two_diameters, _one_diameter],
                         ids=["one_state", "two_states_two_operators", "two_states_one_operator"])
def test_paired_and_single_sweep_regimes_agree(make):
    sps = build_state_points(make(N_LARGE))
    big = abi.Batch(sps)
    big.set_operator_band(fp32=True, fp64=True)
    x0 = big.x0()
    rng = np.random.default_rng(11)
    x = x0 + 0.02 * rng.standard_normal(len(x0))
    print(f"case chain regimes {make.__name__}: {N_LARGE} points", flush=True)
    f_big = big.split(big.objective(x, abi.PREC_FP64))
    seg_big = [big.segment_profile(0, i) for i in SAMPLE]
    big.kernel_profile(abi.KP_CHAIN_FP64, reps=1)   # raises unless the chain species runs on the kernel
    xs = big.split(x)
    big.close()
    small = abi.Batch([sps[i] for i in SAMPLE])
    small.set_operator_band(fp32=True, fp64=True)
    print(f"case chain regimes {make.__name__}: {len(SAMPLE)} points", flush=True)
    small.objective(np.concatenate([xs[i] for i in SAMPLE]), abi.PREC_FP64)
    for k, i in enumerate(SAMPLE):
        seg = small.segment_profile(0, k)
        assert np.array_equal(seg_big[k], seg), f"point {i}: paired and single-sweep segment densities differ"
    small.close()
    for i in SAMPLE[:2]:
        print(f"case chain regimes {make.__name__}: oracle point {i}", flush=True)
        o, Be = truncated_oracle(sps[i])
        assert 2 * Be + 1 <= sps[i].NP   # the band fits the mirror construction: the chain kernel runs
        ref = o.objective(xs[i])
        err = np.max(np.abs(f_big[i] - ref)) / max(1.0, float(np.max(np.abs(ref))))
        assert err < 1e-11, f"point {i}: |F - F_oracle| / max|F_oracle| = {err:.3e}"


def test_paired_regime_fp32_close_to_fp64():
    sps = build_state_points(_two_diameters(N_LARGE))
    b = abi.Batch(sps)
    b.set_operator_band(fp32=True, fp64=True)
    x = b.x0()
    print("case chain regimes fp32 vs fp64", flush=True)
    f64 = b.objective(x, abi.PREC_FP64)
    f32 = b.objective(x, abi.PREC_FP32)
    b.close()
    err = np.max(np.abs(f32 - f64)) / np.max(np.abs(f64))
    assert err < 5e-5, f"FP32 vs FP64 objective: {err:.3e}"
