"""Method of moving asymptotes on the device (lbmx_mma_*) against the CPU oracle.

The device sums over the design variables in a fixed blocked order, the oracle sums
sequentially, and pow/sqrt come from different math libraries, so the designs agree to the
interior-point tolerance rather than bit for bit: 1e-9 on x in [0, 1] per update, over ten
updates in a row. The robust case has the shape of applications/RobustPump2d/robustPump2d.cc:
three realisations with a_j = 1 and a volume constraint with a_j = 0.
"""
import numpy as np
import pytest

import lbmx
import oracle

pytestmark = pytest.mark.gpu


def budget_inputs(x, t, V):
    n = x.size
    return 2 * (x - t), np.array([x.sum() / n - V]), np.full((1, n), 1.0 / n)


def robust_inputs(x, targets, weights, vol):
    n = x.size
    g = [w * np.sum((x - t) ** 2) / n for t, w in zip(targets, weights)] + [x.mean() - vol]
    dg = [2 * w * (x - t) / n for t, w in zip(targets, weights)] + [np.full(n, 1.0 / n)]
    return np.zeros(n), np.array(g), np.stack(dg)


@pytest.mark.parametrize("n", [1000, 524288])
def test_budget_updates_match_oracle(lbmx_mod, n):
    t = np.random.default_rng(1).random(n)
    dev, ref = lbmx.MMA(n, 1), oracle.MMA(n, 1)
    xd = xr = np.full(n, 0.5)
    for it in range(10):
        xd = dev.update(xd, *budget_inputs(xd, t, 0.35), 0.2)
        xr = ref.update(xr, *budget_inputs(xr, t, 0.35), 0.2)
        assert np.max(np.abs(xd - xr)) < 1e-9, (it, float(np.max(np.abs(xd - xr))))
        assert abs(dev.design_change() - ref.design_change(xr)) < 1e-9
    sd, sr = dev.state(), ref.state()
    assert np.max(np.abs(sd["L"] - sr["L"])) < 1e-9 and np.max(np.abs(sd["U"] - sr["U"])) < 1e-9
    assert abs(sd["lam"][0] - sr["lam"][0]) <= 1e-6 * abs(sr["lam"][0])
    dev.close()


def test_robust_minmax_updates_match_oracle(lbmx_mod):
    n = 20000
    rng = np.random.default_rng(2)
    targets = [0.2 + 0.6 * rng.random(n) for _ in range(3)]
    weights = [1.0, 1.3, 0.8]
    a = [1.0, 1.0, 1.0, 0.0]
    dev, ref = lbmx.MMA(n, 4, a=a), oracle.MMA(n, 4, a=a)
    xd = xr = np.full(n, 0.5)
    for it in range(10):
        xd = dev.update(xd, *robust_inputs(xd, targets, weights, 0.45), 0.1)
        xr = ref.update(xr, *robust_inputs(xr, targets, weights, 0.45), 0.1)
        assert np.max(np.abs(xd - xr)) < 1e-9, (it, float(np.max(np.abs(xd - xr))))
    sd, sr = dev.state(), ref.state()
    assert np.max(np.abs(sd["lam"] - sr["lam"])) <= 1e-6 * np.max(np.abs(sr["lam"]))
    assert sd["newton_iters"] > 0
    dev.close()


def test_options_asymptote_type_one_and_constraint_modification(lbmx_mod):
    n = 3000
    t = np.random.default_rng(3).random(n)
    dev, ref = lbmx.MMA(n, 1), oracle.MMA(n, 1)
    for h in (dev, ref):
        h.set_asymptotes(0.3, 0.6, 1.1)
    dev.set_options(True, 1)
    ref.set_options(True, 1)
    xd = xr = np.full(n, 0.5)
    for it in range(6):
        xd = dev.update(xd, *budget_inputs(xd, t, 0.3), 0.25)
        xr = ref.update(xr, *budget_inputs(xr, t, 0.3), 0.25)
        assert np.max(np.abs(xd - xr)) < 1e-9, it
    dev.close()


def test_mma_argument_errors(lbmx_mod):
    for m in (0, 9):
        with pytest.raises(lbmx.LbmxError):
            lbmx.MMA(100, m)
    h = lbmx.MMA(100, 2)
    with pytest.raises(lbmx.LbmxError):
        h.set_options(False, 2)
    h.close()
