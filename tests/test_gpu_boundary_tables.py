"""Tabulated boundary values (lbmx_add_boundary_table) on the device against the oracle: a
velocity inlet and pressure outlet whose profiles and time factors no built-in form gives, in
2D (per-component and scalar tables) and 3D, single instances and a batch."""
import numpy as np
import pytest

import lbmx
import oracle
from boundary_tables import inlet_profile_2d, ramp_table, table_channel2d, table_duct3d
from helpers import relmax

pytestmark = pytest.mark.gpu


def _pair(nx, ny, nz=1, lattice=lbmx.D2Q9, coll=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, precision=lbmx.FP64,
          omega=1.25):
    kw = dict(lattice=lattice, collision=coll, design=design, omega=omega, interp_q=0.1,
              alpha_max=1.0 if design == lbmx.DESIGN_PBB else 0.1)
    return lbmx.Solver(lbmx.make_desc(nx, ny, nz, precision=precision, **kw)), oracle.Oracle(lbmx.make_desc(nx, ny, nz, **kw))


@pytest.mark.parametrize("ncomp", [2, 1])
@pytest.mark.parametrize("coll,design", [(lbmx.MRT, lbmx.DESIGN_BRINKMAN), (lbmx.CASCADED, lbmx.DESIGN_PBB)])
def test_table_channel_matches_the_oracle(lbmx_mod, ncomp, coll, design):
    nx, ny, steps = 60, 30, 80
    s, o = _pair(nx, ny, coll=coll, design=design)
    g = np.random.default_rng(5).uniform(0.2, 1.0, nx * ny)
    for obj in (s, o):
        table_channel2d(obj, nx, ny, steps, ncomp=ncomp, time_const=1.0 / steps)
        obj.set_design(g)
    s.init_equilibrium(); o.init_equilibrium()
    Js, Jo = s.forward(steps), o.forward(steps)
    assert relmax(s.get_state(), o.f) < 1e-12
    assert abs(Js - Jo) <= 1e-10 * abs(Jo)
    s.checkpoint_config(7)
    Jg, gs = s.gradient(steps)
    Jog, go, _ = o.gradient(steps)
    assert abs(Jg - Jog) <= 1e-10 * abs(Jog)
    assert relmax(gs, go) < 1e-10


def test_table_duct_matches_the_oracle(lbmx_mod):
    nx, ny, nz, steps = 30, 14, 12, 30
    s, o = _pair(nx, ny, nz, lattice=lbmx.D3Q19, omega=1.6)
    g = np.random.default_rng(6).uniform(0.2, 1.0, nx * ny * nz)
    for obj in (s, o):
        table_duct3d(obj, nx, ny, nz, steps, time_const=1.0 / steps)
        obj.set_design(g)
    s.init_equilibrium(); o.init_equilibrium()
    Js, Jo = s.forward(steps), o.forward(steps)
    assert relmax(s.get_state(), o.f) < 1e-12
    assert abs(Js - Jo) <= 1e-10 * abs(Jo)
    s.checkpoint_config(0)
    Jg, gs = s.gradient(steps)
    Jog, go, _ = o.gradient(steps)
    assert abs(Jg - Jog) <= 1e-10 * abs(Jog)
    assert relmax(gs, go) < 1e-10


def test_table_members_in_a_batch(lbmx_mod):
    steps = 40
    members = []
    for nx, ny in ((64, 30), (90, 36)):
        s = lbmx.Solver(lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.CASCADED, design=lbmx.DESIGN_PBB,
                                       precision=lbmx.FP32, omega=1.0 / 0.6, interp_q=0.1, alpha_max=1.0))
        table_channel2d(s, nx, ny, steps, time_const=1.0 / steps)
        s.set_design(np.random.default_rng(nx).uniform(0.2, 1.0, nx * ny))
        members.append(s)
    ref = []
    for m in members:
        m.checkpoint_config(5)
        ref.append(m.gradient(steps))
    b = lbmx.Batch(members)
    b.checkpoint_config(5)
    J, sens = b.gradient(steps)
    for i, (Js, gs) in enumerate(ref):
        assert J[i] == Js
        assert np.array_equal(sens[i], gs)
    b.close()


def test_table_arguments_are_checked(lbmx_mod):
    nx, ny = 20, 12
    s, _ = _pair(nx, ny)
    B = lbmx.make_box
    inlet = lbmx.make_bc(lbmx.BC_ZOUHE_VELOCITY, 0.05)
    with pytest.raises(lbmx.LbmxError) as e:  # one value short
        s.add_boundary_table(B(0, 0, 1, ny - 2), inlet, np.ones(ny - 3), 1, None)
    assert e.value.code == lbmx.ERR_ARG
    with pytest.raises(lbmx.LbmxError) as e:  # three components on a 2D lattice
        s.add_boundary_table(B(0, 0, 1, ny - 2), inlet, np.ones(3 * (ny - 2)), 3, None)
    assert e.value.code == lbmx.ERR_ARG
    with pytest.raises(lbmx.LbmxError) as e:  # bounce-back has no value
        s.add_boundary_table(B(0, nx - 1, 0, 0), lbmx.make_bc(lbmx.BC_BOUNCEBACK), None, 1, np.ones(5))
    assert e.value.code == lbmx.ERR_ARG
    with pytest.raises(lbmx.LbmxError) as e:  # a NaN in the profile
        s.add_boundary_table(B(0, 0, 1, ny - 2), inlet, np.full(ny - 2, np.nan), 1, None)
    assert e.value.code == lbmx.ERR_ARG
    # a horizon past the end of the time table
    s.add_boundary_table(B(0, 0, 1, ny - 2), inlet, inlet_profile_2d(ny, 2), 2, ramp_table(10))
    s.init_equilibrium()
    s.forward(10)
    with pytest.raises(lbmx.LbmxError) as e:
        s.forward(1)
    assert e.value.code == lbmx.ERR_ARG
    with pytest.raises(lbmx.LbmxError) as e:
        s.gradient(11)
    assert e.value.code == lbmx.ERR_ARG
    s.gradient(10)
