"""GPU parity of libsoftdft_hip against the FP64 oracle, through the C ABI.

Tolerances (relative to max |oracle|):
  * FP64 objective F(x):                1e-11  (same arithmetic, different summation order)
  * FP32 objective F(x):                5e-5   (single-precision convolutions and propagators)
  * converged densities, FP64 and mixed: 1e-8 relative (max norm per field family), both sides solved
                                        to ||F(x)-x||_inf < 1e-12 so the comparison resolves the bound
  * Omega at the same state (objective at x, then observables): 1e-12 relative
  * converged Omega, FP64 and mixed:    1e-10 relative to max(1, |Omega|), plus the width of the
                                        reference's fixed-point family where the neutrality shift's
                                        1e-10 charge tolerance leaves Psi_C a band (omega_band)
  * degree of ionization / CVT:         1e-9 absolute
"""
import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point
from oracle.dft_oracle import OracleDFT, anderson_solve
from parity_helpers import hip_free_memory, omega_band, truncated_oracle

pytestmark = pytest.mark.gpu


def _systems():
    return {
        "chains": configs.hard_sphere_chains(H=6.0, dz=0.05, n_seg=5),
        "weak_pe": configs.weak_polyelectrolyte(H=6.0, dz=0.05, n_seg=6, solvent=0.2, pH=5.0),
        "weak_pe_charged_walls": configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=4, solvent=0.2, pH=6.0,
                                                              wall_charge=(-0.05, -0.02)),
        "mfmt_msa": configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=4, solvent=0.2, pH=4.0, msa="msa"),
        "brush": configs.weak_polyelectrolyte(H=8.0, dz=0.05, n_seg=10, pH=6.0, pKa=5.0, grafted=0.05),
        "square_well": configs.square_well_chains(H=6.0, dz=0.05, n_seg=5),
        "square_well_monomers": configs.square_well_chains(H=6.0, dz=0.05, n_seg=1, rho_poly=0.2, base="mfmt"),
    }


SYSTEMS = _systems()


def _mfmt_variant(ds):
    ds = dict(ds)
    ds["model"] = ["ideal", "mfmt", "msa", "coul"]
    return ds


SYSTEMS["mfmt_msa"] = _mfmt_variant(SYSTEMS["mfmt_msa"])


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_objective_fp64_matches_oracle(name):
    sp = build_state_point(SYSTEMS[name])
    b = abi.Batch([sp])
    o = OracleDFT(sp)
    rng = np.random.default_rng(7)
    x0 = b.x0()
    for x in (x0, x0 + 0.05 * rng.standard_normal(len(x0))):
        print(f"case objective_fp64 {name} n={len(x)}", flush=True)
        fx = b.objective(x, abi.PREC_FP64)
        ref = o.objective(x)
        scale = max(1.0, float(np.max(np.abs(ref))))
        assert np.max(np.abs(fx - ref)) / scale < 1e-11
        rb = b.get_field(abi.FIELD_RHO_BEADS, 0)
        assert np.allclose(rb, o.rho_beads, rtol=1e-11, atol=1e-14)
    b.close()


@pytest.mark.parametrize("name", ["chains", "weak_pe", "brush", "square_well"])
def test_objective_fp32_is_close(name):
    sp = build_state_point(SYSTEMS[name])
    b = abi.Batch([sp])
    x0 = b.x0()
    print(f"case objective_fp32 {name}", flush=True)
    fx = b.objective(x0, abi.PREC_FP32)
    ref = OracleDFT(sp).objective(x0)
    assert np.max(np.abs(fx - ref)) / max(1.0, float(np.max(np.abs(ref)))) < 5e-5
    b.close()


TIGHT = 1e-12


def _rel_maxdiff(a, ref):
    return float(np.max(np.abs(a - ref))) / max(float(np.max(np.abs(ref))), 1e-300)


@pytest.mark.parametrize("name", ["chains", "weak_pe", "weak_pe_charged_walls", "brush", "square_well",
                                  "square_well_monomers"])
def test_solve_matches_oracle(name):
    """Both sides solved to ||F(x)-x||_inf < 1e-12, four decades below the 1e-8 bound being checked:
    converged bead and bond densities within 1e-8 relative, Omega within 1e-10 relative (plain FP64 and
    mixed precision) plus the width of the reference's own fixed-point family (the neutrality shift's
    1e-10 charge tolerance leaves Psi_C free in a narrow band: ~1e-10 in Omega for the brush, measured
    by omega_band), degree of ionization and contact pressure within 1e-9."""
    sp = build_state_point(SYSTEMS[name])
    b = abi.Batch([sp])
    o = OracleDFT(sp)
    num = dict(sp.numerics)
    num["tole"] = TIGHT
    kw = dict(tole=TIGHT, damping=num["damping"], mixing_max=num["mixing_max"], switch_tole=num["switch_tole"],
              anderm=int(num["anderm"]), max_iters=6000)
    print(f"case solve {name}", flush=True)
    x, it, err = b.solve(b.x0(), **kw)
    om, al, pc = b.observables()
    rb = b.get_field(abi.FIELD_RHO_BEADS, 0)
    rbo = b.get_field(abi.FIELD_RHO_BONDS, 0)
    r = anderson_solve(o.objective, sp.x0(), num, max_iters=20000)
    rb_ref, rbo_ref = o.rho_beads.copy(), o.rho_bonds.copy()
    om_ref = o.grand_potential(r.x)
    assert err[0] < TIGHT and r.err < TIGHT
    assert _rel_maxdiff(rb, rb_ref) < 1e-8
    assert np.allclose(rb, rb_ref, rtol=1e-8, atol=1e-12 * float(np.max(rb_ref)))
    if sp.MB:
        assert _rel_maxdiff(rbo, rbo_ref) < 1e-8
    band = omega_band(sp, o, r, om_ref, num)
    assert abs(om[0] - om_ref) < 1e-10 * max(1.0, abs(om_ref)) + band
    fr = o.state_fractions()
    for u, f in enumerate(fr):
        if f is not None:
            assert np.allclose(al[0, u, :len(f)], f, atol=1e-9)
    assert np.allclose(pc[0], o.contact_pressure(), rtol=1e-8, atol=1e-9)
    print(f"case solve mixed {name}", flush=True)
    x2, it2, err2 = b.solve(b.x0(), precision=abi.PREC_MIXED, **kw)
    om2, _, _ = b.observables()
    rb2 = b.get_field(abi.FIELD_RHO_BEADS, 0)
    assert err2[0] < TIGHT
    assert _rel_maxdiff(rb2, rb_ref) < 1e-8
    assert abs(om2[0] - om_ref) < 1e-10 * max(1.0, abs(om_ref)) + band
    b.close()


@pytest.mark.parametrize("name", ["weak_pe_charged_walls", "brush", "square_well", "mfmt_msa"])
def test_omega_at_the_same_state_matches_oracle(name):
    """grand_potential (grand_potential.jl:4-28) with rho from the objective at x and mu_ex, lng from x."""
    sp = build_state_point(SYSTEMS[name])
    b = abi.Batch([sp])
    o = OracleDFT(sp)
    rng = np.random.default_rng(5)
    x = b.x0() + 0.01 * rng.standard_normal(sp.n)
    print(f"case omega same state {name}", flush=True)
    b.objective(x, abi.PREC_FP64)
    om, al, pc = b.observables()
    o.objective(x)
    om_ref = o.grand_potential(x)
    assert abs(om[0] - om_ref) < 1e-12 * max(1.0, abs(om_ref))
    assert np.allclose(pc[0], o.contact_pressure(), rtol=1e-12, atol=1e-14)
    b.close()


def test_observables_reuse_their_buffers():
    """sdft_observables allocates its work buffers once per context: 100 calls leave the device's
    free memory unchanged."""
    sp = build_state_point(SYSTEMS["weak_pe_charged_walls"])
    b = abi.Batch([sp, sp])
    b.objective(b.x0(), abi.PREC_FP64)
    first = b.observables()
    free0 = hip_free_memory()
    for _ in range(100):
        om, al, pc = b.observables()
    free1 = hip_free_memory()
    assert free1 == free0
    assert np.array_equal(om, first[0]) and np.array_equal(pc, first[2])
    assert np.array_equal(al, first[1], equal_nan=True)
    b.close()


@pytest.mark.parametrize("name", ["chains", "weak_pe", "weak_pe_charged_walls", "mfmt_msa"])
def test_banded_fp64_objective_matches_truncated_oracle(name):
    sp = build_state_point(SYSTEMS[name])
    o, Be = truncated_oracle(sp)
    assert 2 * Be + 1 <= sp.NP  # mirror-extended band fits: the persistent chain kernel runs
    b = abi.Batch([sp])
    b.set_operator_band(fp32=True, fp64=True)
    rng = np.random.default_rng(11)
    x0 = b.x0()
    for x in (x0, x0 + 0.05 * rng.standard_normal(len(x0))):
        print(f"case banded_fp64 {name} Be={Be}", flush=True)
        fx = b.objective(x, abi.PREC_FP64)
        ref = o.objective(x)
        assert np.max(np.abs(fx - ref)) / max(1.0, float(np.max(np.abs(ref)))) < 1e-11
    # the dense operator is back after resetting the band choice
    b.set_operator_band(fp32=True, fp64=False)
    fx = b.objective(x0, abi.PREC_FP64)
    ref = OracleDFT(sp).objective(x0)
    assert np.max(np.abs(fx - ref)) / max(1.0, float(np.max(np.abs(ref)))) < 1e-11
    b.close()


def test_banded_fp64_ragged_batch():
    dss = [configs.weak_polyelectrolyte(H=H, dz=0.05, n_seg=5, solvent=0.2, pH=6.0) for H in (4.1, 5.55, 7.0)]
    sps = [build_state_point(d) for d in dss]
    b = abi.Batch(sps)
    b.set_operator_band(fp32=True, fp64=True)
    print("case banded_fp64 ragged", flush=True)
    x0 = b.x0()
    fx = b.objective(x0, abi.PREC_FP64)
    for p, sp in enumerate(sps):
        o, Be = truncated_oracle(sp)
        assert 2 * Be + 1 <= sp.NP
        ref = o.objective(sp.x0())
        assert np.max(np.abs(b.split(fx)[p] - ref)) / max(1.0, np.max(np.abs(ref))) < 1e-11
    b.close()


def test_batched_points_equal_individual_solves():
    dss = [configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=6, solvent=0.2, pH=pH) for pH in (3.0, 4.5, 6.0, 9.0)]
    sps = [build_state_point(d) for d in dss]
    b = abi.Batch(sps)
    print("case batch-vs-single", flush=True)
    xb, itb, errb = b.solve(b.x0(), max_iters=3000)
    omb, alb, _ = b.observables()
    for p, sp in enumerate(sps):
        bi = abi.Batch([sp])
        xi, iti, erri = bi.solve(bi.x0(), max_iters=3000)
        omi, ali, _ = bi.observables()
        # reductions use a fixed number of blocks per point and a fixed summation order, so a point
        # solved inside a batch follows exactly the same trajectory as when solved alone
        assert iti[0] == itb[p]
        assert np.array_equal(xi, b.split(xb)[p])
        assert omi[0] == omb[p]
        bi.close()
    b.close()


def test_mixed_batch_equals_individual_mixed_solves():
    dss = [configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=6, solvent=0.2, pH=pH) for pH in (3.0, 6.0, 9.0)]
    sps = [build_state_point(d) for d in dss]
    b = abi.Batch(sps)
    print("case mixed batch-vs-single", flush=True)
    xb, itb, errb = b.solve(b.x0(), max_iters=3000, precision=abi.PREC_MIXED)
    omb, _, _ = b.observables()
    for p, sp in enumerate(sps):
        bi = abi.Batch([sp])
        xi, iti, erri = bi.solve(bi.x0(), max_iters=3000, precision=abi.PREC_MIXED)
        omi, _, _ = bi.observables()
        assert iti[0] == itb[p]
        assert np.array_equal(xi, b.split(xb)[p])
        assert omi[0] == omb[p]
        bi.close()
    b.close()


def test_ragged_batch_of_slit_widths():
    dss = [configs.weak_polyelectrolyte(H=H, dz=0.05, n_seg=4, solvent=0.2, pH=7.0) for H in (3.0, 4.1, 5.55)]
    sps = [build_state_point(d) for d in dss]
    assert len({sp.NP for sp in sps}) == 3
    b = abi.Batch(sps)
    print("case ragged", flush=True)
    x0 = b.x0()
    fx = b.objective(x0)
    for p, sp in enumerate(sps):
        ref = OracleDFT(sp).objective(sp.x0())
        assert np.max(np.abs(b.split(fx)[p] - ref)) / max(1.0, np.max(np.abs(ref))) < 1e-11
    x, it, err = b.solve(x0, max_iters=3000, precision=abi.PREC_MIXED)
    assert np.all(err < 1e-8)
    b.close()


def test_nonconvergence_and_nonfinite_inputs_are_reported():
    sp = build_state_point(SYSTEMS["weak_pe"])
    b = abi.Batch([sp])
    print("case max_iters", flush=True)
    with pytest.raises(abi.SdftError) as e:
        b.solve(b.x0(), max_iters=3)
    assert e.value.code == -4
    x = b.x0()
    x[5] = np.nan
    print("case nan", flush=True)
    with pytest.raises(abi.SdftError) as e:
        b.solve(x, max_iters=50)
    assert e.value.code == -4
    b.close()


def _chains_not_first():
    ds = configs.hard_sphere_chains(H=5.0, dz=0.05, n_seg=5, rho_poly=0.06, rho_solvent=0.2)
    ds["species"] = [{"sequence": "S", "density": 0.2}, {"sequence": "AAAAA", "density": 0.06},
                     {"sequence": "AAA", "density": 0.04}]
    return ds


@pytest.mark.parametrize("mapping", [0, 1, 2])
def test_chain_species_after_other_species(mapping):
    """Chain species whose propagator fields do not start at offset 0 (solvent listed first, two chain
    species): every propagator mapping (per-step kernels 0, chain kernel 1 and 2), FP64 dense, FP64
    banded (against the truncated oracle) and FP32."""
    sp = build_state_point(_chains_not_first())
    b = abi.Batch([sp])
    b.set_tuning([dict(kind=abi.TUNE_PROPAGATOR, key=u, precision=0, choice=mapping) for u in range(b.nsp)])
    rng = np.random.default_rng(3)
    x = b.x0() + 0.03 * rng.standard_normal(b.n_total)
    print(f"case chains after other species mapping={mapping}", flush=True)
    ref = OracleDFT(sp).objective(x)
    scale = max(1.0, float(np.max(np.abs(ref))))
    fx = b.objective(x, abi.PREC_FP64)
    assert np.max(np.abs(fx - ref)) / scale < 1e-11
    fx32 = b.objective(x, abi.PREC_FP32)
    assert np.max(np.abs(fx32 - ref)) / scale < 5e-5
    b.set_operator_band(fp32=True, fp64=True)
    o, _ = truncated_oracle(sp)
    refb = o.objective(x)
    fxb = b.objective(x, abi.PREC_FP64)
    assert np.max(np.abs(fxb - refb)) / scale < 1e-11
    for u in (1, 2):
        seg = b.segment_profile(u, 0)
        assert np.max(np.abs(seg - o.rho_seg[u])) <= 1e-11 * max(float(np.max(np.abs(o.rho_seg[u]))), 1e-300)
    b.close()
