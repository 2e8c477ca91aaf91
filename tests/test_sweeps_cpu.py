"""Host logic of the slit-width continuation (softdft_amd/sweeps.py) on the CPU."""
import numpy as np

from softdft_amd import configs
from softdft_amd.problem import build_state_point, flatten, unflatten
from softdft_amd.sweeps import julia_round, remap_solution, write_sweep_summary, WidthResult


def _symmetric_x(sp, seed):
    """A solution-like vector whose fields are mirror-symmetric about the slit centre, as the
    converged fields of a symmetric slit are."""
    rng = np.random.default_rng(seed)
    NP, NB, MB = sp.NP, sp.NB, sp.MB
    half = rng.standard_normal((NP, NB + 2 * MB + 1))
    sym = 0.5 * (half + half[::-1])
    mu = sym[:, :NB]
    lng = sym[:, NB:NB + 2 * MB].reshape(NP, 2, MB)
    psi = sym[:, -1]
    return flatten(mu, lng, psi, 0.123)


def test_julia_round_is_half_even():
    assert [julia_round(v) for v in (0.5, 1.5, 2.5, 30.5, 31.0)] == [0, 2, 2, 30, 31]


def test_width_warm_start_keeps_the_left_half_and_mirrors_it():
    for H_old, H_new in ((3.0, 3.05), (3.0, 3.1), (4.0, 4.02)):
        so = build_state_point(configs.weak_polyelectrolyte(H=H_old, dz=0.05, n_seg=4, solvent=0.2, pH=6.0))
        sn = build_state_point(configs.weak_polyelectrolyte(H=H_new, dz=0.05, n_seg=4, solvent=0.2, pH=6.0))
        assert sn.NP >= so.NP
        xo = _symmetric_x(so, 3)
        xn = remap_solution(so, xo, sn)
        mo, lo, po, co = unflatten(xo, so.NP, so.NB, so.MB)
        mn, ln, pn, cn = unflatten(xn, sn.NP, sn.NB, sn.MB)
        # the new profile is the old one near the left wall and mirror-symmetric as a whole; the copy
        # runs one index past the centre, so the mirrored writes replace the last copied index
        nkeep = sn.NP // 2 - 1
        assert np.array_equal(mn[:nkeep], mo[:nkeep])
        assert np.array_equal(ln[:nkeep], lo[:nkeep])
        assert np.array_equal(pn[:nkeep], po[:nkeep])
        assert np.array_equal(mn, mn[::-1]) and np.array_equal(pn, pn[::-1]) and np.array_equal(ln, ln[::-1])
        assert cn == co


def test_sweep_summary_is_relative_to_the_last_width(tmp_path):
    rs = [WidthResult(1, 4.0, 1.5, 0.25, -3.0, -10.0, 100, False),
          WidthResult(2, 4.1, 1.4, 0.15, -2.5, -9.9, 60, True)]
    path = tmp_path / "adjust_distancesZ_results.csv"
    write_sweep_summary(rs, str(path))
    lines = path.read_text().splitlines()
    assert lines[0] == "step,distance,Pressure,deltaP,energy,deltaE"
    f = lines[1].split(",")
    assert float(f[3]) == 0.25 - 0.15 and float(f[5]) == -3.0 - (-2.5)
    assert lines[2].split(",")[3:6:2] == ["0.0", "0.0"]


def test_ph_summary_format(tmp_path):
    from softdft_amd.sweeps import PHResult, write_ph_summary
    rs = [PHResult(step=1, pH=3.0, brush_height=float("nan"), alpha=0.123456789012345, omega=-1.0,
                   iterations=10, warm=False),
          PHResult(step=2, pH=3.05, brush_height=4.25, alpha=1.0 / 3.0, omega=-1.0, iterations=9, warm=True)]
    p = tmp_path / "summary_brush_alpha.csv"
    write_ph_summary(rs, str(p))
    assert p.read_text() == ("step,pH,brush_height,alpha\n1,3,NaN,0.123456789012\n"
                             "2,3.05,4.25,0.333333333333\n")
