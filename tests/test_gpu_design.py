"""Design pipeline on the device (lbmx_filter_*) against the CPU oracle, through the C ABI.

* The density filter and its transpose sum their cone taps in the order of the reference's
  compressed-row products with separately rounded operations, so they equal the oracle bit
  for bit, in 2D and 3D, for radii 1..16, filter boxes reaching outside the grid and grids
  that are not multiples of the tile.
* The projection uses the device tanh (within an ulp of the host's): the two-stage chain of
  applications/RobustPump2d/robustPump2d.cc:140-146 (R = 8 / 4, beta = 2 / 1, eta = 0.5 / eta_r)
  matches the oracle to 1e-14 on fields in [0, 1] and 1e-13 relative on sensitivities.
* Writing a solver's design on the device, and reading its dJ/dgamma there, give the same
  numbers as the host paths; the chained gradient dJ/d(design) matches central differences.
"""
import numpy as np
import pytest

import lbmx
import oracle
import workloads

pytestmark = pytest.mark.gpu

B = lbmx.make_box


def chain_pair(dims, dbox, fbox, stages, background=None):
    desc = lbmx.make_filter_desc(*dims, B(*dbox), B(*fbox), stages)
    dev = lbmx.Filter(desc)
    ref = oracle.FilterChain(*dims, dbox, fbox, stages=stages, background=background)
    if background is not None:
        dev.set_background(background)
    return dev, ref


DENSITY_CASES = [
    ((100, 37, 1), (0, 99, 0, 36, 0, 0), 1),
    ((100, 37, 1), (5, 90, 3, 30, 0, 0), 2),
    ((130, 70, 1), (-4, 200, 10, 80, 0, 0), 5),     # filter box reaches outside the grid
    ((256, 128, 1), (16, 239, 8, 119, 0, 0), 8),
    ((70, 45, 1), (0, 69, 0, 44, 0, 0), 16),        # radius larger than the tile height
    ((33, 21, 17), (2, 30, 1, 19, 3, 14), 3),
    ((24, 18, 12), (0, 23, 0, 17, 0, 11), 5),
]


@pytest.mark.parametrize("dims,fbox,R", DENSITY_CASES)
def test_density_filter_bitwise(lbmx_mod, dims, fbox, R):
    nx, ny, nz = dims
    full = (0, nx - 1, 0, ny - 1, 0, nz - 1)
    dev, ref = chain_pair(dims, full, fbox, ((R, False, 1.0, 0.5),))
    rng = np.random.default_rng(R)
    x = rng.random(ref.n)
    y = dev.design(x)
    assert np.array_equal(y, ref.density(x, R)), float(np.max(np.abs(y - ref.density(x, R))))
    s = rng.standard_normal(ref.n)
    g = dev.sensitivities(s)
    assert np.array_equal(g, ref.density_transpose(s, R)), float(np.max(np.abs(g - ref.density_transpose(s, R))))
    dev.close()


@pytest.mark.parametrize("eta2", [0.3, 0.5, 0.7])
def test_robust_pump_double_projection(lbmx_mod, eta2):
    nx, ny = 160, 96
    dbox = (20, 139, 12, 83, 0, 0)
    stages = ((8, True, 2.0, 0.5), (4, True, 1.0, eta2))
    rng = np.random.default_rng(11)
    bg = (rng.random(nx * ny) > 0.3).astype(np.float64)   # solid/fluid layout outside the design box
    dev, ref = chain_pair((nx, ny, 1), dbox, dbox, stages, background=bg)
    d = rng.random(ref.nd)
    p = dev.design(d)
    pr = ref.design(d)
    assert np.max(np.abs(p - pr)) < 1e-14
    for k in range(2):
        assert np.max(np.abs(dev.field(2 * k) - ref.dens[k])) < 1e-14
    s = rng.standard_normal(ref.n)
    g, gr = dev.sensitivities(s), ref.sensitivities(s)
    assert np.max(np.abs(g - gr)) < 1e-13 * np.max(np.abs(gr))
    # continuation: beta grows on both stages (DoubleProjectionFilter2d::increaseBeta)
    dev.increase_beta(1.5)
    ref.stages[0][2] *= 1.5
    ref.stages[1][2] *= 1.5
    assert np.max(np.abs(dev.design(d) - ref.design(d))) < 1e-14
    dev.close()


def test_projection_3d_chain(lbmx_mod):
    dims = (28, 20, 16)
    dbox = (3, 24, 2, 17, 2, 13)
    stages = ((3, True, 4.0, 0.5),)
    rng = np.random.default_rng(12)
    dev, ref = chain_pair(dims, dbox, dbox, stages, background=rng.random(28 * 20 * 16))
    d = rng.random(ref.nd)
    assert np.max(np.abs(dev.design(d) - ref.design(d))) < 1e-14
    s = rng.standard_normal(ref.n)
    gr = ref.sensitivities(s)
    assert np.max(np.abs(dev.sensitivities(s) - gr)) < 1e-13 * np.max(np.abs(gr))
    dev.close()


def test_variable_threshold_double_projection(lbmx_mod):
    """VariableDoubleProjectionFilter2d as in applications/diffuser/diffuser.cc:224 (radii 4 / 2):
    a per-node eta field on the second stage."""
    nx, ny = 120, 64
    dbox = (30, 90, 1, 62, 0, 0)
    stages = ((4, True, 2.0, 0.5), (2, True, 3.0, 0.5))
    rng = np.random.default_rng(15)
    dev, ref = chain_pair((nx, ny, 1), dbox, dbox, stages)
    etaf = rng.uniform(0.3, 0.7, nx * ny)
    dev.set_eta_field(1, etaf)
    ref.eta_fields[1] = etaf
    d = rng.random(ref.nd)
    assert np.max(np.abs(dev.design(d) - ref.design(d))) < 1e-14
    s = rng.standard_normal(ref.n)
    gr = ref.sensitivities(s)
    assert np.max(np.abs(dev.sensitivities(s) - gr)) < 1e-13 * np.max(np.abs(gr))
    dev.set_eta_field(1, None)      # back to the constant threshold
    ref.eta_fields[1] = None
    assert np.max(np.abs(dev.design(d) - ref.design(d))) < 1e-14
    dev.close()


def channel_solver(nx, ny, prec=lbmx.FP64):
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, precision=prec,
                       omega=1.25, interp_q=0.1, alpha_max=0.1)
    s = lbmx.Solver(d)
    workloads.channel2d(s, nx, ny, objective=lbmx.OBJ_PRESSURE_DROP, time_const=1.0 / 60)
    return s


@pytest.mark.parametrize("prec", [lbmx.FP64, lbmx.FP32])
def test_design_to_solver_and_device_sensitivities(lbmx_mod, prec):
    nx, ny = 96, 40
    dbox = (16, 79, 1, 38, 0, 0)
    stages = ((4, True, 2.0, 0.5), (2, True, 1.0, 0.5))
    dev = lbmx.Filter(lbmx.make_filter_desc(nx, ny, 1, B(*dbox), B(*dbox), stages))
    s = channel_solver(nx, ny, prec)
    d = np.random.default_rng(13).random(dev.n_design)
    phys = dev.design(d)
    dev.design_to_solver(d, s)
    assert np.array_equal(s.get_design(), phys)
    J, sens = s.gradient(60)
    g_dev = dev.solver_sensitivities(s)
    g_host = dev.sensitivities(sens)
    assert np.array_equal(g_dev, g_host)
    # the device-set design drives the solver exactly as the same design set from the host
    s2 = channel_solver(nx, ny, prec)
    s2.set_design(phys)
    J2, sens2 = s2.gradient(60)
    assert J2 == J and np.array_equal(sens2, sens)
    s.close()
    s2.close()
    dev.close()


def test_chained_gradient_central_differences(lbmx_mod):
    """dJ/d(design) through the mapping, both projection stages and the solver's adjoint,
    against central differences of J(filter(design)) on design cells in the flow path."""
    nx, ny, steps = 64, 24, 60
    dbox = (12, 51, 1, 22, 0, 0)
    stages = ((3, True, 2.0, 0.5), (2, True, 1.0, 0.4))
    dev = lbmx.Filter(lbmx.make_filter_desc(nx, ny, 1, B(*dbox), B(*dbox), stages))
    s = channel_solver(nx, ny)
    s.checkpoint_config(0)
    d = 0.6 + 0.3 * np.random.default_rng(14).random(dev.n_design)
    dev.design_to_solver(d, s)
    s.gradient(steps)
    g = dev.solver_sensitivities(s)
    h = 1e-5
    nxd = dbox[1] - dbox[0] + 1
    scale = np.max(np.abs(g))
    for (xd, yd) in ((5, 11), (20, 4), (33, 17), (19, 11)):
        i = yd * nxd + xd
        Jpm = []
        for sign in (1, -1):
            dd = d.copy()
            dd[i] += sign * h
            dev.design_to_solver(dd, s)
            s.init_equilibrium()
            Jpm.append(s.forward(steps))
        fd = (Jpm[0] - Jpm[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * scale, (xd, yd, fd, g[i])
    s.close()
    dev.close()


def test_filter_argument_errors(lbmx_mod):
    with pytest.raises(lbmx.LbmxError):
        lbmx.Filter(lbmx.make_filter_desc(20, 10, 1, B(0, 19, 0, 9), stages=((0, False, 1.0, 0.5),)))
    with pytest.raises(lbmx.LbmxError):
        lbmx.Filter(lbmx.make_filter_desc(20, 10, 1, B(0, 20, 0, 9), stages=((2, False, 1.0, 0.5),)))
    with pytest.raises(lbmx.LbmxError):
        lbmx.Filter(lbmx.make_filter_desc(20, 10, 1, B(0, 19, 0, 9), stages=((17, False, 1.0, 0.5),)))
    f = lbmx.Filter(lbmx.make_filter_desc(20, 10, 1, B(2, 17, 1, 8), stages=((2, True, 1.0, 0.5),)))
    with pytest.raises(lbmx.LbmxError) as e:   # projection derivative needs a filtered design
        f.sensitivities(np.ones(200))
    assert e.value.code == lbmx.ERR_STATE
    with pytest.raises(lbmx.LbmxError):
        f.design(np.ones(5))
    f.close()
