"""Bulk states on the device (sdft_bulk_solve) against the host restatement bulk.solve_bulk, which
the oracle pins fix (Carnahan-Starling pressure, Henderson-Hasselbalch dissociation, the
bulk-consistency KAT): mu_species, mu_ex, lng, bead and bond densities and pressure to 1e-12
relative to max(1, |value|), on points of every benchmark configuration."""
import copy
import time

import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.bulk import solve_bulk
from softdft_amd.bulk_device import solve_bulk_batch
from softdft_amd.molsys import build_molsys
from softdft_amd.problem import build_state_point, build_state_points

pytestmark = pytest.mark.gpu


def _cases():
    c3 = configs.config3(256)
    c4 = configs.config4(512)
    c5 = configs.config5()
    return {
        "C1": [configs.config1()],
        "C2": [configs.config2()],
        "C3": [c3[i] for i in (0, 40, 100, 128, 200, 255)],
        "C4": [c4[i] for i in (0, 255, 511)],
        "C5": [c5[i] for i in (0, 7, 60, 127)],
        "square_well": [configs.square_well_chains(H=6.0, dz=0.05, n_seg=5)],
        "mfmt_msa": [dict(configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=4, solvent=0.2, pH=4.0, msa="msa"),
                          model=["ideal", "mfmt", "msa", "coul"])],
        "diprotic": [configs.diprotic_polyelectrolyte(H=5.0, dz=0.05, n_seg=4, solvent=0.2, pH=pH)
                     for pH in (3.0, 5.5, 8.0)],
    }


def _close(a, b, tol=1e-12):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.all(np.abs(a - b) <= tol * np.maximum(1.0, np.abs(b)))


@pytest.mark.parametrize("name", list(_cases()))
def test_device_bulk_matches_host_bulk(name):
    dss = _cases()[name]
    mss = [build_molsys(copy.deepcopy(d)) for d in dss]
    print(f"case bulk {name} points={len(dss)}", flush=True)
    dev = solve_bulk_batch(mss)
    for d, bd in zip(dss, dev):
        ref = solve_bulk(build_molsys(copy.deepcopy(d)))
        assert _close(bd.mu_ex, ref.mu_ex)
        assert _close(bd.lng, ref.lng)
        assert _close(bd.rho_beads, ref.rho_beads)
        assert _close(bd.rho_bonds, ref.rho_bonds)
        assert _close(bd.rho_species, ref.rho_species)
        assert _close(bd.pressure, ref.pressure)
        assert _close(bd.mu_species, ref.mu_species)


def test_device_bulk_setup_gives_the_same_objective():
    dss = configs.config3(256)[::64]
    t0 = time.perf_counter()
    host = [build_state_point(d) for d in dss]
    t1 = time.perf_counter()
    dev = build_state_points(dss, device_bulk=True)
    t2 = time.perf_counter()
    print(f"case bulk setup: host {t1 - t0:.2f} s, device bulk {t2 - t1:.2f} s for {len(dss)} points", flush=True)
    bh, bd = abi.Batch(host), abi.Batch(dev)
    xh, xd = bh.x0(), bd.x0()
    assert np.max(np.abs(xh - xd)) < 1e-12 * max(1.0, float(np.max(np.abs(xh))))
    fh = bh.objective(xh, abi.PREC_FP64)
    fd = bd.objective(xd, abi.PREC_FP64)
    assert np.max(np.abs(fh - fd)) < 1e-11 * max(1.0, float(np.max(np.abs(fh))))
    bh.close()
    bd.close()
