"""Linear box functionals given by explicit terms (lbmx_set_objective_terms), the extension
point that stands in for application-defined LBMacroFunctional subclasses
(src/functionals/LBMacroFunctional.hh:6-14). The oracle's term values are pinned by sums
computed here from the populations; its adjoint sources by central differences of those
values; the engine's enum functionals equal their explicit-term forms.
"""
import numpy as np
import pytest

import lbmx
import oracle

RNG = np.random.default_rng(424242)
B = lbmx.make_box


def _oracle(lattice, nx, ny):
    d = lbmx.make_desc(nx, ny, lattice=lattice, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, omega=1.25,
                       interp_q=0.1, alpha_max=0.1, thermal_diffusivity=0.02, ramp_p=1.0, ramp_ck=0.1)
    return oracle.Oracle(d)


def _state(o, T=0.6):
    """A random state near a moving, warm equilibrium (AoS, natural node order)."""
    ex, ey, _, _, w = o.lattice()
    f = np.empty((o.nodes, o.Q))
    for n in range(o.nodes):
        rho = 1 + 0.02 * RNG.standard_normal()
        u = 0.05 * RNG.standard_normal(2)
        cu = ex[:9] * u[0] + ey[:9] * u[1]
        f[n, :9] = w[:9] * (rho + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u) * (1 + 0.01 * RNG.standard_normal(9))
        if o.Q > 9:
            f[n, 9:] = w[9:] * T * (1 + 0.05 * RNG.standard_normal(o.Q - 9))
    return f


def _node_quantities(o, f):
    ex, ey, _, _, _ = o.lattice()
    rho = f[:, :9].sum(1)
    jx = (f[:, :9] * ex[:9]).sum(1)
    jy = (f[:, :9] * ey[:9]).sum(1)
    T = f[:, 9:].sum(1) if o.Q > 9 else np.zeros(len(f))
    return rho, jx, jy, T


def _box_mask(nx, ny, b):
    m = np.zeros((ny, nx), bool)
    m[b.y0:b.y1 + 1, b.x0:b.x1 + 1] = True
    return m.ravel()


def test_temperature_term_is_pinned():
    """DomainTemperature2d (src/functionals/domainTemperature2d.hh): timeConst / area * sum T over
    the domain box, and its source timeConst / area on every D2Q5 population."""
    nx, ny = 12, 9
    o = _oracle(lbmx.D2Q9_D2Q5, nx, ny)
    box = B(2, 9, 1, 6)
    area = 8 * 6
    o.set_objective_terms([(lbmx.TERM_TEMPERATURE, 1.0 / area, box)], time_const=0.25)
    f = _state(o)
    _, _, _, T = _node_quantities(o, f)
    m = _box_mask(nx, ny, box)
    J = o.objective(f.ravel())
    assert abs(J - 0.25 * T[m].sum() / area) < 1e-14
    g = o.objective_grad(f.ravel()).reshape(o.nodes, o.Q)
    assert np.all(g[m, 9:] == 0.25 / area)
    assert np.all(g[m, :9] == 0) and np.all(g[~m] == 0)


@pytest.mark.parametrize("lattice", [lbmx.D2Q9, lbmx.D2Q9_D2Q5])
def test_mixed_terms_value_and_gradient(lattice):
    nx, ny = 14, 10
    o = _oracle(lattice, nx, ny)
    terms = [(lbmx.TERM_DENSITY, 0.3, B(0, 0, 1, 8)), (lbmx.TERM_VELOCITY_X, -2.0, B(3, 10, 2, 7)),
             (lbmx.TERM_TOTAL_PRESSURE_FLUX, 0.5, B(13, 13, 1, 8))]
    if lattice == lbmx.D2Q9_D2Q5:
        terms.append((lbmx.TERM_HEAT_FLUX_X, -1.5, B(13, 13, 1, 8)))
    o.set_objective_terms(terms, time_const=0.5)
    f = _state(o)
    rho, jx, jy, T = _node_quantities(o, f)
    per_kind = {lbmx.TERM_DENSITY: rho, lbmx.TERM_VELOCITY_X: jx,
                lbmx.TERM_TOTAL_PRESSURE_FLUX: (rho / 3 + 0.5 * (jx ** 2 + jy ** 2)) * jx,
                lbmx.TERM_HEAT_FLUX_X: jx * T}
    expect = 0.5 * sum(wgt * per_kind[k][_box_mask(nx, ny, b)].sum() for k, wgt, b in terms)
    J = o.objective(f.ravel())
    assert abs(J - expect) < 1e-13 * max(1.0, abs(expect))
    # adjoint source against central differences of the value, on random entries
    g = o.objective_grad(f.ravel())
    h = 1e-6
    for idx in RNG.choice(f.size, 40, replace=False):
        fp, fm = f.ravel().copy(), f.ravel().copy()
        fp[idx] += h
        fm[idx] -= h
        fd = (o.objective(fp) - o.objective(fm)) / (2 * h)
        assert abs(fd - g[idx]) < 1e-8, idx


@pytest.mark.parametrize("kind,terms", [
    (lbmx.OBJ_PRESSURE_DROP, lambda a, b, na, nb: [(lbmx.TERM_DENSITY, 1 / na, a), (lbmx.TERM_DENSITY, -1 / nb, b)]),
    (lbmx.OBJ_THROUGHFLOW, lambda a, b, na, nb: [(lbmx.TERM_VELOCITY_X, -1 / na, a), (lbmx.TERM_VELOCITY_X, -1 / nb, b)]),
    (lbmx.OBJ_DISSIPATED_POWER, lambda a, b, na, nb: [(lbmx.TERM_TOTAL_PRESSURE_FLUX, 1.0, a),
                                                      (lbmx.TERM_TOTAL_PRESSURE_FLUX, -1.0, b)]),
])
def test_enum_functionals_equal_their_terms(kind, terms):
    nx, ny = 16, 8
    a, b = B(0, 0, 1, 6), B(15, 15, 1, 6)
    o = _oracle(lbmx.D2Q9, nx, ny)
    f = _state(o).ravel()
    o.set_objective(lbmx.make_objective(kind, a, b, 0.125))
    J1, g1 = o.objective(f), o.objective_grad(f)
    o.set_objective_terms(terms(a, b, 6, 6), time_const=0.125)
    J2, g2 = o.objective(f), o.objective_grad(f)
    assert abs(J1 - J2) <= 1e-15 * abs(J1)
    assert np.max(np.abs(g1 - g2)) <= 1e-15 * np.max(np.abs(g1))


def test_term_errors():
    o = _oracle(lbmx.D2Q9, 8, 8)
    with pytest.raises(lbmx.LbmxError):
        o.set_objective_terms([(lbmx.TERM_TEMPERATURE, 1.0, B(0, 1, 0, 1))])
    with pytest.raises(lbmx.LbmxError):
        o.set_objective_terms([(lbmx.TERM_DENSITY, 1.0, B(0, 1, 0, 1))] * 5)
