"""The optimisation loop with the design and its sensitivities kept in HBM
(applications/RobustPump2d/robustPump2d.cc:217-310, one realisation): the *_device entry points
of the filter, the solver and the MMA take GPU arrays (torch tensors here), so no per-variable
array crosses PCIe between the stages of an iteration.

The device loop runs the same kernels on the same values as the host-array loop, so every
iterate (design, objective, sensitivities) must be bitwise equal to it. The borrowed
sensitivity view and the device design setter are checked against their host counterparts.
"""
import numpy as np
import pytest

import lbmx
import workloads

pytestmark = pytest.mark.gpu

B = lbmx.make_box
NX, NY, STEPS = 80, 32, 120
DBOX = (12, 67, 1, 30, 0, 0)


def _setup():
    d = lbmx.make_desc(NX, NY, lattice=lbmx.D2Q9, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN,
                       omega=1.25, interp_q=0.1, alpha_max=0.1)
    s = lbmx.Solver(d)
    workloads.channel2d(s, NX, NY, objective=lbmx.OBJ_PRESSURE_DROP, time_const=1.0 / STEPS)
    s.checkpoint_config(12)
    filt = lbmx.Filter(lbmx.make_filter_desc(NX, NY, 1, B(*DBOX), B(*DBOX), ((3, True, 2.0, 0.5), (2, True, 1.0, 0.5))))
    mask = np.zeros(NX * NY)
    mask.reshape(NY, NX)[DBOX[2]:DBOX[3] + 1, DBOX[0]:DBOX[1] + 1] = 1.0
    return s, filt, mask


def test_device_loop_equals_host_loop(lbmx_mod):
    import torch
    iters = 4
    # host arrays
    s, filt, mask = _setup()
    nd = filt.n_design
    mma = lbmx.MMA(nd, 1)
    x = np.full(nd, 0.5)
    host = []
    for _ in range(iters):
        filt.design_to_solver(x, s)
        J, _ = s.gradient(STEPS, want_sens=False)
        dJ = filt.solver_sensitivities(s)
        vol = np.sum(s.get_design() * mask) / nd - 0.5
        dvol = filt.sensitivities(mask / nd)
        x = mma.update(x, dJ * (1.0 / abs(host[0][0] if host else J)), [vol], [dvol], 0.2)
        host.append((J, x.copy(), dJ.copy(), mma.design_change()))
    # device arrays
    s, filt, mask = _setup()
    mma = lbmx.MMA(nd, 1)
    xt = torch.full((nd,), 0.5, dtype=torch.float64, device="cuda")
    mask_t = torch.from_numpy(mask / nd).to("cuda")
    J0 = None
    for it in range(iters):
        filt.design_to_solver_device(xt, s)
        J, _ = s.gradient(STEPS, want_sens=False)
        dJ = filt.solver_sensitivities_device(s)
        phys = filt.design_device(xt)
        vol = np.sum(phys.cpu().numpy() * mask) / nd - 0.5
        dvol = filt.sensitivities_device(mask_t)
        J0 = J if J0 is None else J0
        mma.update_device(xt, dJ * (1.0 / abs(J0)), [vol], [dvol], 0.2)
        Jh, xh, dJh, chh = host[it]
        assert J == Jh, it
        assert np.array_equal(dJ.cpu().numpy(), dJh), it
        assert np.array_equal(xt.cpu().numpy(), xh), it
        assert mma.design_change_device() == chh, it
    # the device loop lowered the pressure drop as the host loop did
    assert host[-1][0] < host[0][0]


def test_sensitivity_view_and_device_design(lbmx_mod):
    import torch
    s, _, _ = _setup()
    g = workloads.design_field(NX, NY, 1, seed=5)
    s.set_design(g)
    J, sens = s.gradient(STEPS)
    view = s.sensitivity_device()
    assert view.device.type == "cuda" and view.numel() == NX * NY
    assert np.array_equal(view.cpu().numpy(), sens)
    # the same design set from the device gives the same gradient
    s.set_design(np.full(NX * NY, 0.3))
    s.set_design_device(torch.from_numpy(g).to("cuda"))
    assert np.array_equal(s.get_design(), g)
    J2, sens2 = s.gradient(STEPS)
    assert J2 == J
    assert np.array_equal(sens2, sens)
