"""Functionals given by explicit terms (lbmx_set_objective_terms) on the GPU: forward J and
dJ/dgamma against the oracle at the FP64 bars, batched members bitwise equal to single ones."""
import numpy as np
import pytest

import lbmx
import oracle
import workloads
from helpers import build_pair, relmax

pytestmark = pytest.mark.gpu

OBJ_TOL = 1e-10
GRAD_TOL = 1e-10
B = lbmx.make_box


def _set_both(s, o, terms, tc):
    s.set_objective_terms(terms, time_const=tc)
    o.set_objective_terms(terms, time_const=tc)


def test_domain_temperature_thermal(lbmx_mod):
    """DomainTemperature2d (src/functionals/domainTemperature2d.hh) on the thermal channel."""
    nx, ny, steps = 30, 14, 60
    u0 = (0.03, 0.0, 0.0)
    s, o, _ = build_pair(lbmx.D2Q9_D2Q5, lbmx.MRT, lbmx.DESIGN_PBB, nx, ny, setup="thermal")
    box = B(1, nx - 2, 1, ny - 2)
    _set_both(s, o, [(lbmx.TERM_TEMPERATURE, 1.0 / ((nx - 2) * (ny - 2)), box)], 1.0 / steps)
    s.checkpoint_config(0)
    Js, gs = s.gradient(steps, u0=u0, T0=0.5)
    Jo, go, _ = o.gradient(steps, u0=u0, T0=0.5)
    print(f"domain temperature J {Js:.12e} (oracle {Jo:.12e}), gradient rel. {relmax(gs, go):.1e}")
    assert abs(Jo) > 1e-3
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL
    s.init_equilibrium(u0=u0, T0=0.5); o.init_equilibrium(u0=u0, T0=0.5)
    assert abs(s.forward(steps) - o.forward(steps)) <= OBJ_TOL * abs(Jo)


@pytest.mark.parametrize("lattice,coll,design", [(lbmx.D2Q9, lbmx.MRT, lbmx.DESIGN_BRINKMAN),
                                                 (lbmx.D2Q9, lbmx.CASCADED, lbmx.DESIGN_PBB),
                                                 (lbmx.D2Q9_D2Q5, lbmx.MRT, lbmx.DESIGN_BRINKMAN)])
def test_mixed_terms_2d(lbmx_mod, lattice, coll, design):
    nx, ny, steps = 40, 20, 80
    setup = "thermal" if lattice == lbmx.D2Q9_D2Q5 else "channel"
    s, o, _ = build_pair(lattice, coll, design, nx, ny, setup=setup)
    terms = [(lbmx.TERM_DENSITY, 0.3, B(0, 0, 1, ny - 2)), (lbmx.TERM_VELOCITY_X, -2.0 / 100, B(10, 19, 5, 14)),
             (lbmx.TERM_TOTAL_PRESSURE_FLUX, 0.5, B(nx - 1, nx - 1, 1, ny - 2))]
    if lattice == lbmx.D2Q9_D2Q5:
        terms.append((lbmx.TERM_HEAT_FLUX_X, -1.0, B(nx - 1, nx - 1, 1, ny - 2)))
    _set_both(s, o, terms, 1.0 / steps)
    kw = dict(u0=(0.03, 0.0, 0.0), T0=0.5) if lattice == lbmx.D2Q9_D2Q5 else {}
    s.checkpoint_config(0)
    Js, gs = s.gradient(steps, **kw)
    Jo, go, _ = o.gradient(steps, **kw)
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL


def test_mixed_terms_3d(lbmx_mod):
    nx, ny, nz, steps = 24, 12, 10, 30
    s, o, _ = build_pair(lbmx.D3Q19, lbmx.MRT, lbmx.DESIGN_BRINKMAN, nx, ny, nz, setup="duct", omega=1.6)
    terms = [(lbmx.TERM_DENSITY, 1.0, B(0, 0, 1, ny - 2, 1, nz - 2)),
             (lbmx.TERM_DENSITY, -1.0, B(nx - 1, nx - 1, 1, ny - 2, 1, nz - 2)),
             (lbmx.TERM_VELOCITY_X, 0.25, B(6, 17, 2, 9, 2, 7))]
    _set_both(s, o, terms, 1.0 / steps)
    s.checkpoint_config(0)
    Js, gs = s.gradient(steps)
    Jo, go, _ = o.gradient(steps)
    assert abs(Js - Jo) <= OBJ_TOL * abs(Jo)
    assert relmax(gs, go) < GRAD_TOL


def test_terms_in_a_batch(lbmx_mod):
    """Members with different explicit functionals in one batch: bitwise their single runs."""
    grids = [(48, 20), (64, 24), (40, 18)]
    members = []
    for i, (nx, ny) in enumerate(grids):
        s = lbmx.Solver(lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.CASCADED, design=lbmx.DESIGN_PBB,
                                       precision=lbmx.FP32, omega=1.0 / 0.6, interp_q=0.1, alpha_max=1.0))
        workloads.channel2d(s, nx, ny, objective=lbmx.OBJ_NONE)
        s.set_design(workloads.design_field(nx, ny, 1, seed=40 + i))
        s.set_objective_terms([(lbmx.TERM_DENSITY, 1.0 / (ny - 2), B(0, 0, 1, ny - 2)),
                               (lbmx.TERM_VELOCITY_X, -0.5 * (i + 1), B(5, nx - 6, 3, ny - 4))], time_const=1.0 / 40)
        members.append(s)
    ref = []
    for m in members:
        m.checkpoint_config(5)
        ref.append(m.gradient(40))
    b = lbmx.Batch(members)
    b.checkpoint_config(5)
    J, sens = b.gradient(40)
    for i, (Js, gs) in enumerate(ref):
        assert J[i] == Js, (i, J[i], Js)
        assert np.array_equal(sens[i], gs), i
    b.close()


def test_term_errors_engine(lbmx_mod):
    s = lbmx.Solver(lbmx.make_desc(16, 8, lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_NONE))
    with pytest.raises(lbmx.LbmxError):
        s.set_objective_terms([(lbmx.TERM_TEMPERATURE, 1.0, B(0, 3, 0, 3))])
    with pytest.raises(lbmx.LbmxError):
        s.set_objective_terms([(lbmx.TERM_DENSITY, 1.0, B(0, 3, 0, 3))] * 5)
    with pytest.raises(lbmx.LbmxError):
        s.set_objective_terms([(lbmx.TERM_DENSITY, 1.0, B(0, 16, 0, 3))])
    s.close()
