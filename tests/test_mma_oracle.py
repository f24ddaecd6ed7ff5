"""Pins the CPU oracle of the method of moving asymptotes (oracle/mma_oracle.c).

The reference's MMA needs PETSc; the oracle is checked against what the method must deliver,
with the optima computed independently of its formulas:
* a separable quadratic under a linear budget converges to the KKT point found by bisection
  on the budget's multiplier;
* the robust (min-max) form with a_j = 1 converges to the point that minimises the largest
  of two squared distances (their midpoint);
* every update stays inside the move limits and [0, 1];
* the first update places the asymptotes at x -/+ 0.5 (upper - lower);
* the dual residual at exit is below the interior-point tolerance.
"""
import numpy as np
import pytest

import oracle


def budget_kkt(t, V):
    """argmin sum (x - t)^2 s.t. sum x <= V, 0 <= x <= 1: x = clip(t - lam / 2) with the
    smallest lam >= 0 meeting the budget (bisection)."""
    x = np.clip(t, 0, 1)
    if x.sum() <= V:
        return x
    lo, hi = 0.0, 4.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.clip(t - mid / 2, 0, 1).sum() > V:
            lo = mid
        else:
            hi = mid
    return np.clip(t - hi / 2, 0, 1)


def run_budget(mma, t, V, iters, x0=0.5, movlim=0.2):
    n = t.size
    x = np.full(n, x0)
    for _ in range(iters):
        x = mma.update(x, 2 * (x - t), [x.sum() / n - V / n], np.full((1, n), 1.0 / n), movlim)
    return x


@pytest.mark.parametrize("n,frac", [(500, 0.3), (3000, 0.42)])
def test_budget_problem_converges_to_kkt_point(n, frac):
    """The budget is active (mean target 0.5 > frac): with an interior optimum and an inactive
    budget the approximation's curvature vanishes with the gradient and the iterates settle
    into a bounded two-cycle, as they do with the reference's formulas."""
    t = np.random.default_rng(n).random(n)
    V = frac * n
    x = run_budget(oracle.MMA(n, 1), t, V, 60)
    xs = budget_kkt(t, V)
    assert np.max(np.abs(x - xs)) < 1e-6
    assert x.sum() <= V * (1 + 1e-9)


def test_robust_minmax_converges_to_midpoint():
    """min z s.t. |x - t_j|^2 / n <= z (a_j = 1, zero objective): the minimiser of the larger
    distance between two points is their midpoint."""
    n = 400
    rng = np.random.default_rng(7)
    t1, t2 = 0.2 + 0.6 * rng.random(n), 0.2 + 0.6 * rng.random(n)
    mma = oracle.MMA(n, 2, a=[1.0, 1.0])
    x = np.full(n, 0.5)
    for _ in range(80):
        g = np.array([np.sum((x - t1) ** 2), np.sum((x - t2) ** 2)]) / n
        dg = np.stack([2 * (x - t1), 2 * (x - t2)]) / n
        x = mma.update(x, np.zeros(n), g, dg, 0.1)
    assert np.max(np.abs(x - 0.5 * (t1 + t2))) < 1e-4


def test_move_limits_bounds_and_first_asymptotes():
    n, movlim = 1000, 0.15
    rng = np.random.default_rng(3)
    x0 = rng.random(n)
    mma = oracle.MMA(n, 2, a=[1.0, 0.0])
    x1 = mma.update(x0, rng.standard_normal(n), [0.3, -0.1], rng.standard_normal((2, n)), movlim)
    assert np.all(np.abs(x1 - x0) <= movlim + 1e-15) and np.all(x1 >= 0) and np.all(x1 <= 1)
    st = mma.state()
    up, lo = np.minimum(1.0, x0 + movlim), np.maximum(0.0, x0 - movlim)
    assert np.max(np.abs(st["L"] - (x0 - 0.5 * (up - lo)))) < 1e-15
    assert np.max(np.abs(st["U"] - (x0 + 0.5 * (up - lo)))) < 1e-15
    assert st["k"] == 1 and 0 < st["newton_iters"] < 500
    assert st["residual"] < 1e-7 * np.sqrt(n + 2)
    assert abs(mma.design_change(x1) - np.max(np.abs(x1 - x0))) == 0.0


def test_asymptotes_follow_the_iterate_history():
    """From the third update on, asymptotes widen where the last two steps kept their sign
    and narrow where they oscillated, within [x - 10 span, x - 0.01 span] (span = upper - lower)."""
    n = 300
    t = np.random.default_rng(9).random(n)
    mma = oracle.MMA(n, 1)
    x = np.full(n, 0.5)
    hist = [x]
    for _ in range(4):
        x = mma.update(x, 2 * (x - t), [x.mean() - 0.4], np.full((1, n), 1.0 / n), 0.2)
        hist.append(x)
    st = mma.state()
    xk = hist[-2]                 # the design the last subproblem was built at
    span = np.maximum(1e-5, np.minimum(1, xk + 0.2) - np.maximum(0, xk - 0.2))
    assert np.all(st["L"] <= xk - 0.01 * span + 1e-15) and np.all(st["L"] >= xk - 10 * span - 1e-15)
    assert np.all(st["U"] >= xk + 0.01 * span - 1e-15) and np.all(st["U"] <= xk + 10 * span + 1e-15)
