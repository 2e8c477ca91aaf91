"""CPU: the pseudo-arclength driver's pure parts (src/NonlinearSolvers/continuation/objective_pac.jl).

* pac_tangent: on a branch X(s) that is a quadratic in s the second-order backward difference on
  non-uniform arclengths is exact, so the returned direction equals X'(s_k) / |X'(s_k)|; the first
  step uses the plain difference (objective_pac.jl:131-143). The ring is read at mod1 slots, so a
  history written in slot order (step - 1) % num_keep gives the same tangent after wrapping.
* HostAnderson follows AndersonMixing.jl:56-169: the first m - 1 iterations are exact Picard steps
  x <- (1 - amix) x + amix F(x) with amix = min(mixing_max, damping / err); on a linear contraction
  F(x) = A x + b it reaches the closed-form fixed point (I - A)^-1 b.
"""
import numpy as np

from softdft_amd.continuation import HostAnderson, pac_tangent


def _quad_branch(s):
    a = np.array([1.0, -2.0, 0.5, 3.0])
    b = np.array([0.3, 0.1, -0.7, 1.1])
    c = np.array([0.05, -0.02, 0.04, 0.01])
    return a + b * s + c * s * s, b + 2.0 * c * s


def test_tangent_first_step_is_the_plain_difference():
    X = np.zeros((5, 4))
    X[0], _ = _quad_branch(0.0)
    X[1], _ = _quad_branch(0.7)
    t = pac_tangent(X, np.array([0.0, 0.7, 0, 0, 0]), 3, 5)
    d = X[1] - X[0]
    assert np.allclose(t, d / np.linalg.norm(d), atol=1e-15)


def test_tangent_is_exact_on_a_quadratic_branch_with_uneven_steps_and_ring_wrap():
    num_keep = 3
    s_steps = [0.0, 0.4, 1.1, 1.35, 2.2, 2.9]
    X = np.zeros((num_keep, 4))
    S = np.zeros(num_keep)
    for k, s in enumerate(s_steps, start=1):      # step k stored at mod1(k, num_keep)
        X[(k - 1) % num_keep], _ = _quad_branch(s)
        S[(k - 1) % num_keep] = s
        if k >= 3:
            step = k + 1                          # the tangent formed before step k + 1
            t = pac_tangent(X, S, step, num_keep)
            _, d = _quad_branch(s)
            assert np.allclose(t, d / np.linalg.norm(d), atol=1e-12), k
    # a dropped term (first-order difference instead) is not exact on the same data
    X1 = X[(len(s_steps) - 1) % num_keep] - X[(len(s_steps) - 2) % num_keep]
    _, d = _quad_branch(s_steps[-1])
    assert not np.allclose(X1 / np.linalg.norm(X1), d / np.linalg.norm(d), atol=1e-6)


def test_host_anderson_picard_warmup_and_linear_fixed_point():
    rng = np.random.default_rng(0)
    n = 12
    Q = rng.standard_normal((n, n))
    A = 0.6 * Q / np.linalg.norm(Q, 2)
    b = rng.standard_normal(n)
    F = lambda x: A @ x + b
    x_star = np.linalg.solve(np.eye(n) - A, b)
    an = HostAnderson(n, 5, damping=0.5, mixing_max=0.7, switch_tole=1e3)
    x = np.zeros(n)
    for it in range(1, 5):                        # iter < m: Picard
        f = F(x)
        err = np.max(np.abs(f - x))
        amix = min(0.7, 0.5 / err)
        x_new, e = an.update(x, f)
        assert e == err
        assert np.array_equal(x_new, (1.0 - amix) * x + amix * f)
        x = x_new
    for _ in range(200):
        x, err = an.update(x, F(x))
        if err < 1e-13:
            break
    assert not an.use_picard
    assert np.max(np.abs(x - x_star)) < 1e-12
