"""Boundary values from tables (lbmx_add_boundary_table) in the oracle: pinned against the
oracle's own built-in profiles and time modes (themselves pinned by the closure tests), and
node by node against the prescribed values."""
import numpy as np

import lbmx
import oracle
from boundary_tables import inlet_profile_2d, ramp_table, table_channel2d


def _oracle(nx, ny, nz=1, lattice=lbmx.D2Q9, coll=lbmx.MRT):
    return oracle.Oracle(lbmx.make_desc(nx, ny, nz, lattice=lattice, collision=coll, design=lbmx.DESIGN_BRINKMAN,
                                        omega=1.25, interp_q=0.1, alpha_max=0.1))


def test_table_reproduces_the_builtin_profile_and_sine():
    """A table holding the parabolic profile and the sine modulation gives the built-in
    form's run to rounding."""
    nx, ny, steps = 40, 22, 60
    R, c0 = (ny - 1) / 2.0, (ny - 1) / 2.0
    period, amp = 23.0, 0.4
    y = np.arange(1, ny - 1, dtype=float)
    prof = (R * R - (y - c0) ** 2) / (R * R)
    t = np.arange(steps + 1, dtype=float)
    tf = 1.0 + amp * np.sin(2 * np.pi * t / period)
    B = lbmx.make_box
    runs = []
    for tabled in (False, True):
        o = _oracle(nx, ny)
        o.add_boundary(B(0, nx - 1, 0, 0), lbmx.make_bc(lbmx.BC_BOUNCEBACK))
        o.add_boundary(B(0, nx - 1, ny - 1, ny - 1), lbmx.make_bc(lbmx.BC_BOUNCEBACK))
        if tabled:
            o.add_boundary_table(B(0, 0, 1, ny - 2), lbmx.make_bc(lbmx.BC_ZOUHE_VELOCITY, 0.05), prof, 1, tf)
        else:
            o.add_boundary(B(0, 0, 1, ny - 2),
                           lbmx.make_bc(lbmx.BC_ZOUHE_VELOCITY, 0.05, profile=lbmx.PROFILE_PARABOLIC, center=(c0, 0),
                                        radius=(R, 1), time_mode=lbmx.TIME_SINE, t_period=period, t_amp=amp,
                                        t_offset=1.0))
        o.add_boundary(B(nx - 1, nx - 1, 1, ny - 2), lbmx.make_bc(lbmx.BC_ZOUHE_PRESSURE, 1.0))
        o.set_objective(lbmx.make_objective(lbmx.OBJ_PRESSURE_DROP, B(1, 1, 1, ny - 2), B(nx - 2, nx - 2, 1, ny - 2)))
        o.init_equilibrium()
        J = o.forward(steps)
        runs.append((J, o.f.copy()))
    (J0, f0), (J1, f1) = runs
    assert abs(J1 - J0) <= 1e-13 * abs(J0)
    assert np.max(np.abs(f1 - f0)) < 1e-14


def test_table_values_prescribed_node_by_node():
    """After the tabulated Zou-He inlet at step t, each inlet node's velocity is
    value * dir * a(y) * b(t), component by component."""
    nx, ny, steps = 12, 10, 30
    o = _oracle(nx, ny, coll=lbmx.BGK)
    table_channel2d(o, nx, ny, steps, ncomp=2)
    ex, ey, _, _, w = o.lattice()
    c = np.stack([ex, ey], 1).astype(float)
    pieces = o.boundaries()
    k = [i for i, p in enumerate(pieces) if p[0] == lbmx.BC_ZOUHE_VELOCITY][0]
    prof = inlet_profile_2d(ny, 2).reshape(-1, 2)
    tf = ramp_table(steps)
    rng = np.random.default_rng(3)
    for t in (0, 7, steps):
        f = (w[None, :] * (1 + 0.01 * rng.standard_normal((ny * nx, 9)))).ravel()
        out = o.apply_bc(k, t, f).reshape(ny, nx, 9)
        for j in range(1, ny - 1):
            u = c.T @ out[j, 0]
            assert np.allclose(u, 0.05 * prof[j - 1] * tf[t], rtol=0, atol=2e-16), (t, j)
