"""CPU-side checks of the C-ABI library: it loads, exports every entry point the
header declares, validates descriptors before touching a device, and fails
loudly (no silent CPU path) when no HIP device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "softdft_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sdft_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = abi.load_library()
    names = _declared()
    assert set(abi.EXPORTED) <= set(names)
    for n in names:
        assert hasattr(lib, n), n
    assert lib.sdft_abi_version() == abi.SDFT_ABI_VERSION


def test_descriptor_validation_happens_before_device_use():
    lib = abi.load_library()
    d = abi.SdftDesc()
    d.abi_version = 999
    h = C.c_void_p()
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -1
    assert b"abi_version" in lib.sdft_last_error(None)
    d.abi_version = abi.SDFT_ABI_VERSION
    d.n_points = 0
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -1
    d.n_points, d.NB, d.MB, d.n_species = 1, 2, 1, 1
    d.functionals = abi.F_MFMT | abi.F_TPT1MFMT
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -5
    assert lib.sdft_solution_length(None) == -1


def test_batch_requires_a_shared_layout():
    a = build_state_point(configs.hard_sphere_chains(H=3.0, dz=0.05, n_seg=3))
    b = build_state_point(configs.weak_polyelectrolyte(H=3.0, dz=0.05, n_seg=3))
    with pytest.raises(ValueError):
        abi.Batch([a, b])


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a HIP device is present")
def test_no_device_fails_loudly():
    sp = build_state_point(configs.hard_sphere_chains(H=3.0, dz=0.05, n_seg=3))
    with pytest.raises(abi.SdftError) as e:
        abi.Batch([sp])
    assert e.value.code == -2


def test_hard_walls_need_contact_tables_and_indices_inside_the_grid():
    """hard_walls = 1 with a NULL contact-index table, or an index outside [0, NP), is rejected
    before any device call (the contact-value theorem would otherwise drop or misread the hard-wall
    term, hard_walls.jl:49-79)."""
    lib = abi.load_library()
    sp = build_state_point(configs.hard_sphere_chains(H=3.0, dz=0.05, n_seg=3))
    assert sp.hard_walls
    d, keep, _, _ = abi.build_descriptor([sp])
    h = C.c_void_p()
    d.wall_lo = None
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -1
    assert b"wall_lo and wall_hi" in lib.sdft_last_error(None)
    d.wall_lo = abi._ptr(keep["lo"], C.c_int32)
    d.wall_hi = None
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -1
    d.wall_hi = abi._ptr(keep["hi"], C.c_int32)
    bad = keep["hi"].copy()
    bad[0, 0] = sp.NP
    d.wall_hi = abi._ptr(bad, C.c_int32)
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -1
    assert b"contact indices" in lib.sdft_last_error(None)
    bad[0, 0] = -1
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -1
    d.wall_hi = abi._ptr(keep["hi"], C.c_int32)
    if not os.path.exists("/dev/kfd"):
        # a valid descriptor passes validation and reaches the device check
        assert lib.sdft_create(C.byref(d), C.byref(h)) == -2


def test_multi_dimensional_descriptor_validation():
    """2-D / 3-D descriptors are checked before any device call: NP must be the grid's point count
    and dz its volume element, every point must share the grid, 2..128 points per dimension, no
    contact-value tables, and grafted segments only in 3-D (determine_fixed has no 2-D method)."""
    import copy
    from softdft_amd.geometry_nd import build_state_point_nd
    lib = abi.load_library()
    d1 = configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=3, solvent=0.2, pH=5.0)
    d3 = copy.deepcopy(d1)
    d3["geometry"]["dimensions"] = [0.5, 0.4, 2.0]
    d3["geometry"]["bin_width"] = [0.1, 0.1, 0.1]
    d3["geometry"]["external_field"] = {"hard_walls": {"dims": [3]}}
    sp = build_state_point_nd(d3)
    d, keep, _, _ = abi.build_descriptor([sp])
    assert d.ndim == 3 and d.hard_walls == 0
    h = C.c_void_p()
    keep["NP"][0] += 1
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -1
    assert b"NP" in lib.sdft_last_error(None)
    keep["NP"][0] -= 1
    keep["grid"][0, 0] = 200
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -1
    keep["grid"][0, 0] = sp.grid[0]
    d.hard_walls = 1
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -5
    d.hard_walls = 0
    if not os.path.exists("/dev/kfd"):
        assert lib.sdft_create(C.byref(d), C.byref(h)) == -2
    # a grafted species on a 2-D grid
    fseg = keep["fseg"].copy()
    keep["fseg"][0] = 0
    d.ndim = 2
    keep["grid"][0, :] = [sp.grid[0], sp.grid[1] * sp.grid[2], 1]
    keep["bins"][0, :] = [0.1, 0.1, 1.0]
    keep["dz"][0] = 0.01
    assert lib.sdft_create(C.byref(d), C.byref(h)) == -5
    assert b"grafted" in lib.sdft_last_error(None)
    keep["fseg"][:] = fseg
