"""Analysis writers (softdft_amd/analysis.py) on the CPU: Julia number printing, file layouts of
src/iDFT/evaluation/analysis/{print_array,degree_of_dissociation,brush_height}.jl, and the brush
height against closed forms."""
import os
from types import SimpleNamespace

import numpy as np
import pytest

from softdft_amd.analysis import (brush_heights, julia_repr, write_field_array, write_segment_array,
                                  write_state_fraction_profiles, write_total_state_fractions,
                                  write_scalar_value, write_brush_height, write_cvt)


@pytest.mark.parametrize("x,text", [
    (1.0, "1.0"), (0.1, "0.1"), (-2.5, "-2.5"), (100000.0, "100000.0"), (999999.0, "999999.0"),
    (1e6, "1.0e6"), (1234567.0, "1.234567e6"), (123456.789, "123456.789"), (0.001, "0.001"),
    (0.0001, "0.0001"), (1e-5, "1.0e-5"), (-0.00012, "-0.00012"), (1.5e-7, "1.5e-7"),
    (1e16, "1.0e16"), (5e-324, "5.0e-324"), (1.7976931348623157e308, "1.7976931348623157e308"),
    (0.0, "0.0"), (-0.0, "-0.0"), (float("nan"), "NaN"), (float("inf"), "Inf"), (float("-inf"), "-Inf"),
    (1 / 3, "0.3333333333333333"), (6.666666666666667e-10, "6.666666666666667e-10"), (12.0, "12.0"),
    (3, "3"),
])
def test_julia_float_printing(x, text):
    assert julia_repr(x) == text


def test_julia_repr_round_trips():
    rng = np.random.default_rng(0)
    for v in np.concatenate([rng.standard_normal(200) * 10.0 ** rng.integers(-12, 12, 200),
                             rng.random(50)]):
        assert float(julia_repr(float(v))) == float(v)


def test_field_array_layout(tmp_path):
    f = tmp_path / "rho.txt"
    arr = np.array([[1.0, 0.5], [2.0, 1e-7], [3.0, 0.0]])
    write_field_array(arr, str(f), (3,), names=["A", "B"])
    assert f.read_text() == "Index_1  A B\n1  1.0 0.5\n2  2.0 1.0e-7\n3  3.0 0.0\n"
    g = tmp_path / "mep.txt"
    write_field_array(np.array([0.25, -1.0]), str(g), (2,), names=["MEP"])
    assert g.read_text() == "Index_1  MEP\n1  0.25\n2  -1.0\n"
    h = tmp_path / "s.txt"
    write_scalar_value(-12.5, str(h))
    assert h.read_text() == "-12.5\n"
    write_scalar_value(2.0, str(h), label="Omega")
    assert h.read_text() == "Omega: 2.0\n"
    c = tmp_path / "cvt.txt"
    write_cvt((0.5, 0.25), str(c))
    assert c.read_text() == "# Contact Value Theorem results\ndim1  0.5  0.25\n"


def _toy_molsys():
    # species "AB": segment 1 in states {A, AH}, segment 2 in state {B}; species "C": one segment
    s1 = SimpleNamespace(name="AB", state_family=[[0, 1], [2]])
    s2 = SimpleNamespace(name="C", state_family=[[3]])
    return SimpleNamespace(species=[s1, s2], monomers=["A", "AH", "B", "C"])


def test_segment_and_state_fraction_files(tmp_path):
    ms = _toy_molsys()
    NP = 3
    r1 = np.zeros((NP, 2, 2))
    r1[:, 0, 0] = [1.0, 3.0, 0.0]
    r1[:, 1, 0] = [1.0, 1.0, 0.0]
    r1[:, 0, 1] = [2.0, 2.0, 2.0]
    r2 = np.ones((NP, 1, 1))
    d = str(tmp_path)
    write_segment_array([r1, r2], ms, (NP,), d)
    assert not os.path.exists(os.path.join(d, "rho_segments_C.txt"))   # one segment: skipped
    assert (tmp_path / "rho_segments_AB.txt").read_text() == (
        "Index_1  seg1_A seg1_AH seg2_B\n1  1.0 1.0 2.0\n2  3.0 1.0 2.0\n3  0.0 0.0 2.0\n")
    write_state_fraction_profiles([r1, r2], ms, (NP,), d)
    assert not os.path.exists(os.path.join(d, "state_fraction_C.txt"))
    assert (tmp_path / "state_fraction_AB.txt").read_text() == (
        "Index_1  seg1_A seg1_AH seg2_B\n1  0.5 0.5 1.0\n2  0.75 0.25 1.0\n3  0.0 0.0 1.0\n")
    write_total_state_fractions([r1, r2], ms, (NP,), d)
    # segment 1: A sums to 4, AH to 2; segment 2 has one state and is skipped
    assert (tmp_path / "total_state_fractions_AB.txt").read_text() == (
        "segment state1 state2\n1 0.6666666666666666 0.3333333333333333\n\n"
        "# Average state fractions across all dissociating segments:\navg 0.6666666666666666 0.3333333333333333\n")


def test_total_state_fraction_padding(tmp_path):
    # segments with 3 and 2 states: the shorter row is padded with empty columns
    s = SimpleNamespace(name="X", state_family=[[0, 1, 2], [0, 1]])
    ms = SimpleNamespace(species=[s], monomers=["a", "b", "c"])
    r = np.zeros((2, 3, 2))
    r[:, :, 0] = 1.0
    r[:, 0, 1] = 1.0
    r[:, 1, 1] = 3.0
    write_total_state_fractions([r], ms, (2,), str(tmp_path))
    lines = (tmp_path / "total_state_fractions_X.txt").read_text().split("\n")
    third = julia_repr(1.0 / 3.0)
    assert lines[1] == f"1 {third} {third} {third}"
    assert lines[2] == "2 0.25 0.75 "
    assert lines[5] == f"avg {julia_repr((1.0 / 3.0 + 0.25) / 2)} {julia_repr((1.0 / 3.0 + 0.75) / 2)} {third}"


def _toy_brush(NP, dz):
    spec = SimpleNamespace(name="G", state_family=[[0], [0]])
    free = SimpleNamespace(name="S", state_family=[[1]])
    ms = SimpleNamespace(species=[spec, free], monomers=["g", "s"])
    fixed = [SimpleNamespace(any=True, segments=np.array([True, False])),
             SimpleNamespace(any=False, segments=np.array([False]))]
    return SimpleNamespace(ms=ms, fixed=fixed, NP=NP, dz=dz, trapez=np.ones((NP, 2)))


def test_brush_height_closed_forms(tmp_path):
    NP, dz = 101, 0.05
    sp = _toy_brush(NP, dz)
    z = np.arange(NP) * dz
    # grafted segment at k = 10 only, free segment uniform over the slit: numerator and
    # denominator add the two parts with weights rho/dz and rho*trapez
    r = np.zeros((NP, 1, 2))
    r[10, 0, 0] = 0.2
    r[:, 0, 1] = 0.5
    (name, h), = brush_heights([r, np.zeros((NP, 1, 1))], sp)
    num = 0.2 / dz * z[10] + 0.5 * z.sum()
    den = 0.2 / dz + 0.5 * NP
    assert name == "G"
    assert h == pytest.approx(2.0 * num / den, rel=1e-14)
    # uniform free density only (grafted weight zero): 2<z> is the slit's far grid point
    r[10, 0, 0] = 0.0
    (_, h), = brush_heights([r, np.zeros((NP, 1, 1))], sp)
    assert h == pytest.approx((NP - 1) * dz, rel=1e-14)
    # a half-weighted wall point (trapez 0.5) shifts the mean by the closed-form amount
    sp.trapez[NP - 1, 0] = 0.5
    (_, h2), = brush_heights([r, np.zeros((NP, 1, 1))], sp)
    w = np.ones(NP)
    w[-1] = 0.5
    assert h2 == pytest.approx(2.0 * (w * z).sum() / w.sum(), rel=1e-14)
    write_brush_height([("G", 1.5)], str(tmp_path / "bh.txt"))
    assert (tmp_path / "bh.txt").read_text() == (
        "# Brush heights per species (computed as 2 * <z>)\n# Format: species brush_height\nG 1.5\n")
