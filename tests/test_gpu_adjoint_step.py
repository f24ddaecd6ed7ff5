"""The reverse loop step by step (lbmx_adjoint_reset / lbmx_adjoint_step), as the reference's
application drives it (applications/RobustPump2d/robustPump2d.cc:263-277): resetAdjoints,
setCurrentTimestep(N), then per step getForwardSolution, adjointCollide, computeCollideSource,
adjointStream.

Two ways of providing the forward state of step t are checked against the fused gradient:
- the state handed back through lbmx_set_state (a caller-held checkpoint, AoS FP64);
- the solver advanced to step t by lbmx_forward (the sweep recomputes s_t from p_{t-1}).
Both run the same sweep on the same values, so dJ/dgamma and the adjoint populations must be
bitwise equal to lbmx_gradient's, in the FP64 form and in the FP16 storage form (whose state
values are FP32 in registers and exact in FP64). The FP64 results are also held to the oracle's
gradient at the parity bar (1e-10).
"""
import numpy as np
import pytest

import lbmx
from helpers import build_pair, relmax

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-10

CASES = [
    ("d2q9 mrt brinkman", dict(lattice=lbmx.D2Q9, coll=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, nx=48, ny=22), 40),
    ("d2q9 bgk pbb", dict(lattice=lbmx.D2Q9, coll=lbmx.BGK, design=lbmx.DESIGN_PBB, nx=40, ny=20), 30),
    ("d3q19 mrt pbb", dict(lattice=lbmx.D3Q19, coll=lbmx.MRT, design=lbmx.DESIGN_PBB, nx=20, ny=10, nz=8,
                           setup="duct", omega=1.6), 16),
    ("thermal", dict(lattice=lbmx.D2Q9_D2Q5, coll=lbmx.MRT, design=lbmx.DESIGN_PBB, nx=30, ny=14,
                     setup="thermal"), 20),
    ("d2q9 mrt neumann outlet", dict(lattice=lbmx.D2Q9, coll=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, nx=40, ny=18,
                                     setup="neumann"), 30),
]


def _pair(kw, steps, precision=lbmx.FP64):
    kw = dict(kw)
    return build_pair(kw.pop("lattice"), kw.pop("coll"), kw.pop("design"), kw.pop("nx"), kw.pop("ny"),
                      kw.pop("nz", 1), precision=precision, time_const=1.0 / steps, **kw)


def _init(s, thermal):
    if thermal:
        s.init_equilibrium(u0=(0.03, 0.0, 0.0), T0=0.5)
    else:
        s.init_equilibrium()


def _fused(s, steps, thermal):
    s.checkpoint_config(0)
    if thermal:
        J, g = s.gradient(steps, u0=(0.03, 0.0, 0.0), T0=0.5)
    else:
        J, g = s.gradient(steps)
    return J, g, s.get_adjoint()


@pytest.mark.parametrize("precision", [lbmx.FP64, lbmx.FP16_SHIFTED], ids=["fp64", "fp16"])
@pytest.mark.parametrize("name,kw,steps", CASES, ids=[c[0] for c in CASES])
def test_stepwise_reverse_loop_equals_fused_gradient(lbmx_mod, name, kw, steps, precision):
    thermal = kw["lattice"] == lbmx.D2Q9_D2Q5
    s, o, _ = _pair(kw, steps, precision)
    J, g, adj = _fused(s, steps, thermal)

    # forward loop with every state kept by the caller (makeCheckpoint after each stream)
    _init(s, thermal)
    states = [s.get_state()]
    for _ in range(steps):
        s.forward(1)
        states.append(s.get_state())
    s.adjoint_reset(0)
    s.adjoint_set_timestep(steps)
    for t in range(steps, 0, -1):
        s.set_state(states[t], t)
        s.adjoint_step(with_source=True)
    assert np.array_equal(s.get_sensitivity(), g)
    assert np.array_equal(s.get_adjoint(), adj)

    # the same loop, each state recomputed by advancing from the initial state
    s.adjoint_reset(steps)
    for t in range(steps, 0, -1):
        _init(s, thermal)
        s.forward(t)
        assert s.timestep() == t
        s.adjoint_step(with_source=True)
    assert np.array_equal(s.get_sensitivity(), g)
    assert np.array_equal(s.get_adjoint(), adj)

    if precision == lbmx.FP64:
        if thermal:
            _, go, ao = o.gradient(steps, u0=(0.03, 0.0, 0.0), T0=0.5)
        else:
            _, go, ao = o.gradient(steps)
        assert relmax(s.get_sensitivity(), go) < GRAD_TOL
        assert relmax(s.get_adjoint(), ao) < GRAD_TOL


def test_stepwise_reverse_without_source_is_zero(lbmx_mod):
    """With no objective source the adjoint stays zero (the reverse loop is linear in it)."""
    s, _, _ = _pair(CASES[0][1], 10)
    s.init_equilibrium()
    s.forward(10)
    s.adjoint_reset(10)
    s.adjoint_step(with_source=False)
    assert not np.any(s.get_sensitivity())
    assert not np.any(s.get_adjoint())


def test_stepwise_reverse_state_checks(lbmx_mod):
    s, _, _ = _pair(CASES[0][1], 10)
    s.init_equilibrium()
    s.forward(5)
    with pytest.raises(lbmx.LbmxError) as e:
        s.adjoint_step()                      # no reset yet
    assert e.value.code == lbmx.ERR_STATE
    s.adjoint_reset(4)
    with pytest.raises(lbmx.LbmxError) as e:
        s.adjoint_step()                      # the forward state is at step 5, the reverse step is 4
    assert e.value.code == lbmx.ERR_STATE
    s.adjoint_set_timestep(5)
    s.adjoint_step()
    s.adjoint_set_timestep(0)
    with pytest.raises(lbmx.LbmxError) as e:
        s.adjoint_step()                      # step 0 (the initial state) is never reversed
    assert e.value.code == lbmx.ERR_STATE
