"""Launch configurations chosen by sdft_autotune / sdft_set_tuning, through the C ABI.

Every tile shape of the convolution GEMMs and both propagator mappings must reproduce the FP64
oracle (1e-11 relative, as the parity tests), and the FP32 objective within 5e-5; a saved tuning
replayed on a fresh context gives bitwise the solve of the tuned context.
"""
import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point
from oracle.dft_oracle import OracleDFT

pytestmark = pytest.mark.gpu

CONV_KINDS = (abi.TUNE_CONV_FWD, abi.TUNE_CONV_PRE, abi.TUNE_CONV_BACK)


def _points():
    # two grid lengths (one kernel set each) with two and one points
    dss = [configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=6, solvent=0.2, pH=pH) for pH in (4.0, 7.0)]
    dss.append(configs.weak_polyelectrolyte(H=14.0, dz=0.05, n_seg=6, solvent=0.2, pH=6.0))
    return [build_state_point(d) for d in dss]


def _banded_oracle(sp, half_width):
    """Oracle whose bond operators keep |j| <= half_width(B) of the filtered bond kernel c, B the last
    offset with |c| > 1e-9 max|c| over the bond types."""
    o = OracleDFT(sp)
    L = o.L
    d = np.arange(L)
    dd = np.minimum(d, L - d)
    cs = [np.fft.ifft(o.wbond_filt[:, u]).real for u in range(sp.MB)]
    B = max(int(dd[np.abs(c) > 1e-9 * np.abs(c).max()].max()) for c in cs)
    W = half_width(B)
    for u, c in enumerate(cs):
        c = c.copy()
        c[dd > W] = 0.0
        o.wbond_filt[:, u] = np.fft.fft(c).real
    return o


@pytest.mark.parametrize("shape", [abi.CONV_TALL, abi.CONV_WIDE64, abi.CONV_WIDE128, abi.CONV_MID, abi.CONV_NARROW])
def test_every_conv_tile_shape_matches_oracle(shape):
    sps = _points()
    b = abi.Batch(sps)
    b.set_tuning([dict(kind=k, key=0, choice=shape) for k in CONV_KINDS])
    rng = np.random.default_rng(5)
    x0 = b.x0()
    for x in (x0, x0 + 0.05 * rng.standard_normal(len(x0))):
        print(f"case conv shape {shape} n={len(x)}", flush=True)
        fx64 = b.split(b.objective(x, abi.PREC_FP64))
        fx32 = b.split(b.objective(x, abi.PREC_FP32))
        for p, sp in enumerate(sps):
            ref = OracleDFT(sp).objective(b.split(x)[p])
            scale = max(1.0, float(np.max(np.abs(ref))))
            assert np.max(np.abs(fx64[p] - ref)) / scale < 1e-11
            assert np.max(np.abs(fx32[p] - ref)) / scale < 5e-5
    b.close()


@pytest.mark.parametrize("choice", [0, 1, 2])
def test_propagator_mappings_match_banded_oracle(choice):
    # choices 1, 2: persistent chain kernel, weights in registers / re-read; choice 0: one banded
    # step kernel launch per step; all apply the half-width Be = B rounded up to even
    sp = _points()[0]
    b = abi.Batch([sp])
    b.set_operator_band(fp32=True, fp64=True)
    b.set_tuning([dict(kind=abi.TUNE_PROPAGATOR, key=0, precision=0, choice=choice)])
    hw = lambda B: B + (B & 1)
    o = _banded_oracle(sp, hw)
    rng = np.random.default_rng(9)
    x0 = b.x0()
    for x in (x0, x0 + 0.05 * rng.standard_normal(len(x0))):
        print(f"case propagator mapping {choice}", flush=True)
        fx = b.objective(x, abi.PREC_FP64)
        ref = o.objective(x)
        # fields to 1e-11; PsiC (the last entry) is a Newton root found to the reference's tolerance
        # 1e-10 (neutrality_shift.jl), so summation-order differences in its inputs move it by up to that
        scale = max(1.0, float(np.max(np.abs(ref))))
        assert np.max(np.abs(fx[:-1] - ref[:-1])) / scale < 1e-11
        assert abs(fx[-1] - ref[-1]) < 1e-10
    b.close()


def test_autotune_decisions_and_replay():
    sps = _points()
    b = abi.Batch(sps)
    print("case autotune", flush=True)
    ent = b.autotune(reps=2)
    kinds = {e["kind"] for e in ent}
    assert abi.TUNE_PROPAGATOR in kinds and abi.TUNE_CONV_FWD in kinds
    for e in ent:
        cands = [u for u in e["us"] if u >= 0]
        assert len(cands) >= 2 and all(np.isfinite(cands))
        assert e["us"][e["choice"]] == min(cands)
    # conv entries: one per grid-length bin (NP 101 and 281 -> bounds 256, 512) and stage
    assert sorted({e["key"] for e in ent if e["kind"] == abi.TUNE_CONV_FWD}) == [256, 512]
    x, it, err = b.solve(b.x0(), precision=abi.PREC_MIXED, max_iters=3000)
    om, al, _ = b.observables()
    assert np.all(err < 1e-8)
    # the same decisions on a fresh context: bitwise the same solve
    b2 = abi.Batch(sps)
    b2.set_tuning(ent)
    x2, it2, err2 = b2.solve(b2.x0(), precision=abi.PREC_MIXED, max_iters=3000)
    om2, _, _ = b2.observables()
    assert np.array_equal(it, it2) and np.array_equal(x, x2) and np.array_equal(om, om2)
    # against the fixed rules: the same fixed point
    b3 = abi.Batch(sps)
    x3, it3, err3 = b3.solve(b3.x0(), precision=abi.PREC_MIXED, max_iters=3000)
    om3, _, _ = b3.observables()
    assert np.max(np.abs(om - om3) / np.maximum(1.0, np.abs(om3))) < 1e-8
    for bb in (b, b2, b3):
        bb.close()


def test_kernel_profiling_leaves_the_context_unchanged():
    """Every kernel class can be timed after an evaluation, and the timing launches (which write the
    library's own buffers and a scratch F(x)) do not change what the next evaluation returns."""
    sps = _points()
    b = abi.Batch(sps)
    x = b.x0()
    before = {p: b.objective(x, p) for p in (abi.PREC_FP32, abi.PREC_FP64)}
    print("case kernel profiling", flush=True)
    for kc in range(abi.KP_ACCUM_FP64 + 1):
        try:
            ms, fl = b.kernel_profile(kc, reps=2)
        except abi.SdftError as e:  # a class the context does not run (e.g. no chain-kernel species)
            assert "chain" in str(e)
            continue
        assert ms > 0.0 and np.isfinite(fl) and fl >= 0.0
    for p, fx in before.items():
        assert np.array_equal(b.objective(x, p), fx)
    b.close()
