"""CPU pins of the multi-dimensional oracle (oracle/dft_oracle_nd.py) and of the host setup
(softdft_amd/geometry_nd.py).

* Slab reduction: a grid uniform in every dimension but the walled one is the one-dimensional system.
  At equal bin widths the 3-D weights restricted to k_x = k_y = 0 are the 1-D weights. The 3-D bond
  kernel sin(k sigma) / (k sigma), the Coulomb Laplacian and the filter cut-off (from the largest
  bin) also coincide. The neutrality shift is invariant under scaling every species amount by the
  cross-section. So F(x) of the 3-D (and 2-D) slab equals the 1-D oracle's F(x) broadcast over the
  cross-section, including at non-uniform x. Omega scales with the cross-section area.
* Permutation: walls along x instead of z on the permuted grid give the permuted result. This pins
  the vector weights and the external field along the other axes, which the slab leaves at zero.
* Square channel (walls along x and y): F(x) is invariant under the x <-> y transposition.
* Bulk: without walls the bulk fields are a fixed point, F(x0) = x0 (all but PsiC).
"""
import copy

import numpy as np
import pytest

from oracle.dft_oracle import OracleDFT
from oracle.dft_oracle_nd import OracleDFTND
from softdft_amd import configs
from softdft_amd.geometry_nd import build_state_point_nd
from softdft_amd.problem import build_state_point, unflatten


def _nd(ds1, dims, bins, walls):
    d = copy.deepcopy(ds1)
    d["geometry"]["dimensions"] = list(dims)
    d["geometry"]["bin_width"] = list(bins)
    d["geometry"]["external_field"] = {"hard_walls": {"dims": list(walls)}}
    return d


def _picard(o, x, n=3, a=0.3):
    for _ in range(n):
        x = (1 - a) * x + a * o.objective(x)
    return x


def _broadcast(x1, sp1, grid, axis):
    """1-D solution fields broadcast over the other dimensions, the 1-D grid along `axis`."""
    mu, lng, psi, psic = unflatten(x1, sp1.NP, sp1.NB, sp1.MB)
    D = len(grid)

    def bc(a):
        if a.size == 0:
            return np.zeros((int(np.prod(grid)),) + a.shape[1:])
        shape = [1] * D + list(a.shape[1:])
        shape[axis] = a.shape[0]
        A = a.reshape(shape)
        return np.broadcast_to(A, tuple(grid) + a.shape[1:]).reshape((-1,) + a.shape[1:], order="F")
    return np.concatenate([bc(mu).ravel(order="F"), bc(lng).ravel(order="F"), bc(psi).ravel(order="F"), [psic]])


@pytest.mark.parametrize("D,n_seg", [(3, 4), (2, 1)])
def test_slab_reduces_to_the_one_dimensional_oracle(D, n_seg):
    ds1 = configs.weak_polyelectrolyte(H=3.0, dz=0.1, n_seg=n_seg, solvent=0.2, pH=5.0)
    sp1 = build_state_point(ds1)
    dims = [0.3, 0.4, 3.0][-D:]
    spn = build_state_point_nd(_nd(ds1, dims, [0.1] * D, [D]))
    o1, on = OracleDFT(sp1), OracleDFTND(spn)
    x1 = _picard(o1, sp1.x0())
    xn = _broadcast(x1, sp1, spn.grid, D - 1)
    f1 = o1.objective(x1)
    fn = on.objective(xn)
    assert np.max(np.abs(fn - _broadcast(f1, sp1, spn.grid, D - 1))) < 1e-12 * max(1.0, np.max(np.abs(f1)))
    area = np.prod(spn.grid[:-1]) * np.prod(spn.bins[:-1])
    om1, omn = o1.grand_potential(x1), on.grand_potential(xn)
    assert omn / area == pytest.approx(om1, rel=1e-12)


def test_walls_along_x_give_the_permuted_result():
    ds1 = configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=3, solvent=0.2, pH=5.0)
    spz = build_state_point_nd(_nd(ds1, [0.3, 0.4, 2.0], [0.1] * 3, [3]))
    spx = build_state_point_nd(_nd(ds1, [2.0, 0.4, 0.3], [0.1] * 3, [1]))
    oz, ox = OracleDFTND(spz), OracleDFTND(spx)
    xz = _picard(oz, spz.x0(), n=2)
    NB, MB = spz.NB, spz.MB

    def perm(x, grid, order):
        mu, lng, psi, psic = unflatten(x, int(np.prod(grid)), NB, MB)
        g = tuple(grid)

        def p(a):
            A = a.reshape(g + a.shape[1:], order="F")
            A = np.transpose(A, order + tuple(range(3, A.ndim)))
            return A.reshape((-1,) + a.shape[1:], order="F")
        return np.concatenate([p(mu).ravel(order="F"), p(lng).ravel(order="F"), p(psi).ravel(order="F"), [psic]])
    xx = perm(xz, spz.grid, (2, 1, 0))
    fz = oz.objective(xz)
    fx = ox.objective(xx)
    assert np.max(np.abs(fx - perm(fz, spz.grid, (2, 1, 0)))) < 1e-12 * np.max(np.abs(fz))
    assert ox.grand_potential(xx) == pytest.approx(oz.grand_potential(xz), rel=1e-12)


def test_square_channel_is_transpose_invariant():
    ds1 = configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=3, solvent=0.2, pH=5.0)
    sp = build_state_point_nd(_nd(ds1, [1.2, 1.2, 0.3], [0.1] * 3, [1, 2]))
    o = OracleDFTND(sp)
    x = _picard(o, sp.x0(), n=2)
    NB, MB = sp.NB, sp.MB
    mu, lng, psi, psic = unflatten(x, sp.NP, NB, MB)
    f = o.objective(x)
    fm, fl, fp, fc = unflatten(f, sp.NP, NB, MB)
    g = sp.grid
    T = lambda a: np.swapaxes(a.reshape(g + a.shape[1:], order="F"), 0, 1)
    A = T(fm)
    assert np.max(np.abs(A - np.swapaxes(A, 0, 1))) < 1e-11 * np.max(np.abs(A))
    P = T(fp)
    assert np.max(np.abs(P - np.swapaxes(P, 0, 1))) < 1e-11 * max(1e-300, np.max(np.abs(P)))
    # the channel is not a slab: the profile varies along both walled axes
    assert np.ptp(A[:, g[1] // 2, 0, 0]) > 1e-3 and np.ptp(A[g[0] // 2, :, 0, 0]) > 1e-3


def test_wall_free_bulk_is_a_fixed_point():
    ds1 = configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=3, solvent=0.2, pH=5.0)
    d = _nd(ds1, [0.6, 0.5, 0.7], [0.1] * 3, [])
    d["geometry"]["external_field"] = {}
    sp = build_state_point_nd(d)
    o = OracleDFTND(sp)
    x0 = sp.x0()
    # the fields; PsiC = 0 of the bulk guess is not the neutrality shift of a charged bulk (the 1-D
    # bulk consistency check leaves it out for the same reason)
    assert np.max(np.abs(o.objective(x0) - x0)[:-1]) < 1e-12


def test_nd_setup_rejects_what_the_path_does_not_build():
    from softdft_amd.molsys import SetupError
    ds1 = configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=3, solvent=0.2, pH=5.0)
    with pytest.raises(SetupError):
        build_state_point_nd(_nd(ds1, [0.4, 2.0], [0.1, 0.1], [2]))   # chains need 3-D (get_renorm)
    g = configs.weak_polyelectrolyte(H=2.0, dz=0.1, n_seg=1, pH=5.0, grafted=0.05)
    with pytest.raises(SetupError):
        build_state_point_nd(_nd(g, [0.4, 2.0], [0.1] * 2, [2]))     # grafting needs 3-D (determine_fixed)


def test_grafted_chain_placement_and_distances_in_3d():
    """determine_fixed (CartesianXYZ, determine_fixed.jl:31-86): centre of the wall in x and y,
    round(d / h_z / 2) steps from the left wall, density 1 / (L_x L_y); the oracle's distances are the
    Euclidean ones of compute_distance_squared, and its bond renormalisation is 4 pi D^2."""
    d = configs.weak_polyelectrolyte(H=3.0, dz=0.125, n_seg=4, pH=3.0, pKa=5.0, grafted=0.05)
    sp = build_state_point_nd(_nd(d, [2.0, 2.0, 3.0], [0.125, 0.125, 0.125], [3]))
    g = sp.grid
    fx = sp.fixed[0]
    assert fx.any and fx.segments[0] and not fx.segments[1:].any()
    i, j, k = np.unravel_index(int(fx.coords[0]), g, order="F")
    assert (i, j, k) == (round(g[0] / 2), round(g[1] / 2), round(1.0 / 0.125 / 2))   # 0-based
    assert fx.density == pytest.approx(1.0 / (2.0 * 2.0))
    assert sp.mu_species[0] == 0.0
    o = OracleDFTND(sp)
    K = np.arange(sp.NP)
    a = np.array(np.unravel_index(K, g, order="F")).T * 0.125
    c = np.array([i, j, k]) * 0.125
    assert np.allclose(o._dist2(K, int(fx.coords[0])), ((a - c) ** 2).sum(1), rtol=0, atol=1e-14)
    assert o._renorm(1.0) == pytest.approx(4.0 * np.pi)
