"""sdft_solve's launch modes through the C ABI: captured poll windows (hipGraph) against
kernel-by-kernel launches, and the FP32 / FP64 pipelines on two streams against one after the other.

The SCF kernels are deterministic (fixed reduction orders, no atomics on values that are read back
in the same iteration), so both launch modes must give bitwise-identical solutions, iteration
counts, residuals and grand potentials. Each solve replays its graphs many times with the host's
polls (copies and synchronisations) between the replays; the module also interleaves eager work
(objective evaluations, eager solves) between graph solves on one context and checks every one.
"""
import numpy as np
import pytest

from softdft_amd import abi, configs
from softdft_amd.problem import build_state_point

pytestmark = pytest.mark.gpu


def _solve(b, graphs, concurrent=True, **kw):
    b.set_launch_mode(graph=graphs, concurrent=concurrent)
    x, it, err = b.solve(b.x0(), **kw)
    om, alpha, pc = b.observables()
    return x, it, err, om, alpha


def _assert_same(a, b):
    xa, ita, erra, oma, ala = a
    xb, itb, errb, omb, alb = b
    assert np.array_equal(ita, itb)
    assert np.array_equal(xa, xb)
    assert np.array_equal(erra, errb)
    assert np.array_equal(oma, omb)
    assert np.array_equal(ala, alb, equal_nan=True)


@pytest.mark.parametrize("precision", [abi.PREC_MIXED, abi.PREC_FP64])
def test_graph_windows_equal_eager_launches(precision):
    dss = [configs.weak_polyelectrolyte(H=5.0, dz=0.05, n_seg=6, solvent=0.2, pH=pH) for pH in (3.0, 6.0, 9.0)]
    b = abi.Batch([build_state_point(d) for d in dss])
    print(f"case graphs vs eager, precision {precision}", flush=True)
    n0 = b.set_graph_capture(None)
    eager = _solve(b, False, max_iters=3000, precision=precision)
    assert b.set_graph_capture(None) == n0  # no graph in eager mode
    g1 = _solve(b, True, max_iters=3000, precision=precision)
    built = b.set_graph_capture(None) - n0
    # one graph per pair of active pipelines: FP32, FP32+FP64, FP64 (mixed) or FP64 alone
    assert 1 <= built <= 3
    assert built >= (2 if precision == abi.PREC_MIXED else 1)
    _assert_same(eager, g1)
    # eager work on the same context between graph solves, then replay again
    fx = b.objective(b.x0(), abi.PREC_FP64)
    assert np.all(np.isfinite(fx))
    g2 = _solve(b, True, max_iters=3000, precision=precision)
    _assert_same(eager, g2)
    eager2 = _solve(b, False, max_iters=3000, precision=precision)
    _assert_same(eager, eager2)
    # the FP32 and FP64 pipelines in turn instead of on two streams: the same solve
    for graphs in (False, True):
        seq = _solve(b, graphs, concurrent=False, max_iters=3000, precision=precision)
        _assert_same(eager, seq)
    b.close()


def test_graph_windows_with_a_partial_last_window_and_per_step_propagators():
    # a grafted brush takes the per-step propagator path (many launches per iteration); max_iters
    # that is not a multiple of the poll interval ends with a window launched kernel by kernel
    ds = configs.weak_polyelectrolyte(H=8.0, dz=0.05, n_seg=10, pH=6.0, pKa=5.0, grafted=0.05)
    b = abi.Batch([build_state_point(ds)])
    print("case graphs vs eager, brush, max_iters 23", flush=True)
    kw = dict(max_iters=23, precision=abi.PREC_MIXED, raise_on_noconv=False, check_every=5)
    eager = _solve(b, False, **kw)
    graph = _solve(b, True, **kw)
    assert eager[1][0] == 23  # still iterating: the partial window ran
    _assert_same(eager, graph)
    kw["max_iters"] = 3000
    eager = _solve(b, False, **kw)
    graph = _solve(b, True, **kw)
    assert eager[2][0] < 1e-8
    _assert_same(eager, graph)
    b.close()
