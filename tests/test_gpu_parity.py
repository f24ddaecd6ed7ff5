"""Parity of the HIP engine (through the C ABI) with the CPU oracle: FP64 plain form.

Bars (written here, from the engine's contract): populations 1e-12 relative max-norm,
objective 1e-10 relative, gradient 1e-10 relative max-norm.
"""
import numpy as np
import pytest

import lbmx
from helpers import build_pair, relmax

pytestmark = pytest.mark.gpu

POP_TOL = 1e-12
OBJ_TOL = 1e-10
GRAD_TOL = 1e-10

CASES_2D = [(lbmx.BGK, lbmx.DESIGN_PBB), (lbmx.MRT, lbmx.DESIGN_PBB),
            (lbmx.BGK, lbmx.DESIGN_BRINKMAN), (lbmx.MRT, lbmx.DESIGN_BRINKMAN),
            (lbmx.BGK, lbmx.DESIGN_NONE), (lbmx.CASCADED, lbmx.DESIGN_PBB), (lbmx.CASCADED, lbmx.DESIGN_NONE)]
CASES_3D = [(lbmx.MRT, lbmx.DESIGN_PBB), (lbmx.MRT, lbmx.DESIGN_BRINKMAN), (lbmx.BGK, lbmx.DESIGN_BRINKMAN)]


@pytest.mark.parametrize("coll,design", CASES_2D)
def test_forward_2d(lbmx_mod, coll, design):
    s, o, _ = build_pair(lbmx.D2Q9, coll, design, 70, 34)
    s.init_equilibrium(); o.init_equilibrium()
    Js = s.forward(150); Jo = o.forward(150)
    assert s.timestep() == 150
    assert relmax(s.get_state(), o.f) < POP_TOL
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)


@pytest.mark.parametrize("coll,design", CASES_3D)
def test_forward_3d(lbmx_mod, coll, design):
    s, o, _ = build_pair(lbmx.D3Q19, coll, design, 40, 18, 14, setup="duct", omega=1.6)
    s.init_equilibrium(); o.init_equilibrium()
    Js = s.forward(60); Jo = o.forward(60)
    assert relmax(s.get_state(), o.f) < POP_TOL
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)


def test_forward_standard_mrt3d(lbmx_mod):
    s, o, _ = build_pair(lbmx.D3Q19, lbmx.MRT_STANDARD, lbmx.DESIGN_NONE, 30, 14, 12, setup="duct", omega=1.6)
    s.init_equilibrium(); o.init_equilibrium()
    s.forward(40); o.forward(40)
    assert relmax(s.get_state(), o.f) < POP_TOL


@pytest.mark.parametrize("coll,design", CASES_2D[:4] + [(lbmx.CASCADED, lbmx.DESIGN_PBB)])
def test_gradient_2d(lbmx_mod, coll, design):
    s, o, _ = build_pair(lbmx.D2Q9, coll, design, 50, 26, time_const=1.0 / 80)
    s.checkpoint_config(0)
    Js, gs = s.gradient(80)
    Jo, go, ao = o.gradient(80)
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL
    assert relmax(s.get_adjoint(), ao) < GRAD_TOL


@pytest.mark.parametrize("coll,design", CASES_3D)
def test_gradient_3d(lbmx_mod, coll, design):
    s, o, _ = build_pair(lbmx.D3Q19, coll, design, 30, 14, 12, setup="duct", omega=1.6, time_const=1.0 / 30)
    s.checkpoint_config(0)
    Js, gs = s.gradient(30)
    Jo, go, _ = o.gradient(30)
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL


def test_checkpointed_gradient_equals_stored(lbmx_mod):
    s, _, _ = build_pair(lbmx.D2Q9, lbmx.MRT, lbmx.DESIGN_PBB, 40, 20, time_const=1.0 / 100)
    s.checkpoint_config(0)
    J0, g0 = s.gradient(100)
    for mode in (lbmx.CHECKPOINT_DYNAMIC, lbmx.CHECKPOINT_STATIC, lbmx.CHECKPOINT_BINOMIAL):
        for slots in (2, 5, 17):
            before = s.stats()["recompute_sweeps"]
            s.checkpoint_config(slots, mode)
            J1, g1 = s.gradient(100)
            assert J1 == J0
            assert np.array_equal(g1, g0), (mode, slots)
            assert s.stats()["recompute_sweeps"] - before == lbmx.checkpoint_simulate(slots, 100, mode)[0]


def test_thermal_forward_and_gradient(lbmx_mod):
    u0 = (0.03, 0.0, 0.0)
    s, o, _ = build_pair(lbmx.D2Q9_D2Q5, lbmx.MRT, lbmx.DESIGN_PBB, 30, 14, setup="thermal", time_const=1.0 / 60)
    s.init_equilibrium(u0=u0, T0=0.5); o.init_equilibrium(u0=u0, T0=0.5)
    s.forward(60); o.forward(60)
    assert relmax(s.get_state(), o.f) < POP_TOL
    s.checkpoint_config(0)
    Js, gs = s.gradient(60, u0=u0, T0=0.5)
    Jo, go, _ = o.gradient(60, u0=u0, T0=0.5)
    assert abs(Jo) > 1e-4
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL


def test_set_state_roundtrip(lbmx_mod):
    s, o, _ = build_pair(lbmx.D2Q9, lbmx.BGK, lbmx.DESIGN_PBB, 33, 17)
    o.init_equilibrium()
    o.forward(25)
    s.set_state(o.f, 25)
    assert np.array_equal(s.get_state(), o.f)
    s.forward(10); o.forward(10)
    assert s.timestep() == 35
    assert relmax(s.get_state(), o.f) < POP_TOL


def test_forward_darcy_mrt(lbmx_mod):
    """DarcyMRT2d (forward only, as the reference): drag plus a constant body force."""
    import oracle
    import workloads
    nx, ny = 70, 34
    kw = dict(lattice=lbmx.D2Q9, collision=lbmx.MRT, design=lbmx.DESIGN_DARCY, omega=1.25, interp_q=0.1,
              alpha_max=0.05, body_force=(2e-6, -1e-6, 0.0))
    s = lbmx.Solver(lbmx.make_desc(nx, ny, **kw))
    o = oracle.Oracle(lbmx.make_desc(nx, ny, **kw))
    g = workloads.design_field(nx, ny, 1, seed=21)
    for obj in (s, o):
        workloads.channel2d(obj, nx, ny, time_const=1.0 / 120)
        obj.set_design(g)
    s.init_equilibrium(); o.init_equilibrium()
    Js = s.forward(120); Jo = o.forward(120)
    assert relmax(s.get_state(), o.f) < POP_TOL
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    with pytest.raises(lbmx.LbmxError):
        s.gradient(5)



@pytest.mark.parametrize("coll,design", [(lbmx.MRT, lbmx.DESIGN_BRINKMAN), (lbmx.BGK, lbmx.DESIGN_PBB),
                                         (lbmx.CASCADED, lbmx.DESIGN_PBB)])
def test_zouhe_neumann_outlet_forward_and_gradient(lbmx_mod, coll, design):
    """IncompressibleZouHeNeumannEast2d outlet (u_x extrapolated from the two nodes inwards):
    the forward reads the neighbours' post-stream values inside the sweep, the adjoint sends
    the neighbours' share through the side buffer and the fix-up launch."""
    s, o, _ = build_pair(lbmx.D2Q9, coll, design, 46, 22, setup="neumann", time_const=1.0 / 70)
    s.init_equilibrium(); o.init_equilibrium()
    s.forward(70); o.forward(70)
    assert relmax(s.get_state(), o.f) < POP_TOL
    s.checkpoint_config(0)
    Js, gs = s.gradient(70)
    Jo, go, ao = o.gradient(70)
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL
    assert relmax(s.get_adjoint(), ao) < GRAD_TOL
    # checkpointed with recomputation: bitwise the stored-trajectory result
    s.checkpoint_config(5)
    J2, g2 = s.gradient(70)
    assert J2 == Js and np.array_equal(g2, gs)


def test_thermal_neumann_outlets_forward_and_gradient(lbmx_mod):
    """ThermalZouHeNeumannEast2d<.,D2Q5> with a flow Neumann outlet on the thermal lattice."""
    u0 = (0.03, 0.0, 0.0)
    s, o, _ = build_pair(lbmx.D2Q9_D2Q5, lbmx.MRT, lbmx.DESIGN_PBB, 30, 14, setup="thermal-neumann",
                         time_const=1.0 / 60)
    s.init_equilibrium(u0=u0, T0=0.5); o.init_equilibrium(u0=u0, T0=0.5)
    s.forward(60); o.forward(60)
    assert relmax(s.get_state(), o.f) < POP_TOL
    s.checkpoint_config(0)
    Js, gs = s.gradient(60, u0=u0, T0=0.5)
    Jo, go, ao = o.gradient(60, u0=u0, T0=0.5)
    assert abs(Jo) > 1e-4
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL
    assert relmax(s.get_adjoint(), ao) < GRAD_TOL


def test_throughflow_objective_gradient(lbmx_mod):
    """ThroughFlow2d (src/functionals/throughFlow2d.hh:16-88): J and dJ/dgamma against the oracle."""
    s, o, _ = build_pair(lbmx.D2Q9, lbmx.MRT, lbmx.DESIGN_BRINKMAN, 44, 20, time_const=1.0 / 60)
    inlet, outlet = lbmx.make_box(0, 0, 1, 18), lbmx.make_box(43, 43, 1, 18)
    for obj in (s, o):
        obj.set_objective(lbmx.make_objective(lbmx.OBJ_THROUGHFLOW, inlet, outlet, 1.0 / 60))
    s.checkpoint_config(0)
    Js, gs = s.gradient(60)
    Jo, go, _ = o.gradient(60)
    assert abs(Jo) > 1e-3
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL
