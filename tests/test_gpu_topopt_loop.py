"""The reference's optimisation loop with every stage on the device
(applications/RobustPump2d/robustPump2d.cc:217-310, one realisation): design-box mapping and
double projection filter into the solver's design in HBM, forward + adjoint with checkpoints,
the filtered sensitivities of the objective and of a fluid-volume constraint, and an MMA update.

A 2D channel with a Brinkman design box, objective = pressure drop, fluid volume <= 50% of the
box: from the uniform grey start the loop must lower the pressure drop at every iteration (by 11% over ten on this
200-step transient) while the volume constraint holds, and the filtered gradient must agree with central differences of the loop's
own objective at the final design.
"""
import numpy as np
import pytest

import lbmx
import workloads

pytestmark = pytest.mark.gpu

B = lbmx.make_box


def test_filtered_mma_loop_lowers_pressure_drop(lbmx_mod):
    nx, ny, steps = 96, 40, 200
    dbox = (16, 79, 1, 38, 0, 0)
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN,
                       omega=1.25, interp_q=0.1, alpha_max=0.1)
    s = lbmx.Solver(d)
    workloads.channel2d(s, nx, ny, objective=lbmx.OBJ_PRESSURE_DROP, time_const=1.0 / steps)
    s.checkpoint_config(16)
    filt = lbmx.Filter(lbmx.make_filter_desc(nx, ny, 1, B(*dbox), B(*dbox), ((3, True, 2.0, 0.5), (2, True, 1.0, 0.5))))
    nd = filt.n_design
    mask = np.zeros(nx * ny)
    mask.reshape(ny, nx)[dbox[2]:dbox[3] + 1, dbox[0]:dbox[1] + 1] = 1.0
    mma = lbmx.MMA(nd, 1)
    x = np.full(nd, 0.5)
    vol_frac = 0.5
    history = []
    for it in range(10):
        filt.design_to_solver(x, s)
        J, _ = s.gradient(steps, want_sens=False)
        dJ = filt.solver_sensitivities(s)
        phys = s.get_design()
        vol = np.sum(phys * mask) / nd - vol_frac                 # fluid fraction of the box - 0.5
        dvol = filt.sensitivities(mask / nd)
        history.append((J, vol))
        scale = 1.0 / abs(history[0][0])
        x = mma.update(x, dJ * scale, [vol], [dvol], 0.2)
        assert np.all(x >= 0) and np.all(x <= 1)
    J0, J_end = history[0][0], history[-1][0]
    assert J_end < 0.93 * J0, history
    assert all(b[0] <= a[0] for a, b in zip(history, history[1:])), history
    assert history[-1][1] <= 0.02, history
    # the loop's gradient at the last design against central differences of J(filter(x))
    filt.design_to_solver(x, s)
    s.gradient(steps, want_sens=False)
    g = filt.solver_sensitivities(s)
    h = 1e-5
    for i in np.argsort(-np.abs(g))[:3]:
        Jpm = []
        for sgn in (1, -1):
            xx = x.copy()
            xx[i] += sgn * h
            filt.design_to_solver(xx, s)
            s.init_equilibrium()
            Jpm.append(s.forward(steps))
        fd = (Jpm[0] - Jpm[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * np.max(np.abs(g)), (i, fd, g[i])
    mma.close()
    filt.close()
    s.close()
