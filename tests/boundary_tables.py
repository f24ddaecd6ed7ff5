"""Problem builders with boundary values given as tables (lbmx_add_boundary_table): a velocity
inlet and a pressure outlet whose per-node profiles and time factors no built-in profile or
time mode produces."""
import numpy as np

import lbmx


def inlet_profile_2d(ny, ncomp):
    """A skewed, non-polynomial inlet over y = 1..ny-2: u_x ~ (1 - |s|^1.5)(1 + 0.2 sin(3 pi y / ny)),
    u_y a small cross flow (ncomp = 2); ncomp = 1 gives the u_x factor only."""
    y = np.arange(1, ny - 1, dtype=float)
    s = 2.0 * (y - 0.5) / (ny - 2) - 1.0
    ax = (1.0 - np.abs(s) ** 1.5) * (1.0 + 0.2 * np.sin(3 * np.pi * y / ny))
    if ncomp == 1:
        return ax
    ay = 0.15 * np.sin(2 * np.pi * y / ny) * (1.0 - s * s)
    return np.stack([ax, ay], 1).ravel()


def outlet_profile_2d(ny):
    y = np.arange(1, ny - 1, dtype=float)
    return 1.0 + 0.002 * np.cos(np.pi * y / ny)


def ramp_table(steps):
    """b(t) = 1 - exp(-t / 40) + 0.1 sin(t / 7), t = 0..steps (no built-in time mode)."""
    t = np.arange(steps + 1, dtype=float)
    return 1.0 - np.exp(-t / 40.0) + 0.1 * np.sin(t / 7.0)


def table_channel2d(obj, nx, ny, steps, ncomp=2, time_const=1.0):
    """Channel with bounce-back walls, the tabulated inlet (u_max 0.05) on x = 0 and the
    tabulated pressure outlet on x = nx - 1; objective: pressure drop."""
    B = lbmx.make_box
    obj.add_boundary(B(0, nx - 1, 0, 0), lbmx.make_bc(lbmx.BC_BOUNCEBACK))
    obj.add_boundary(B(0, nx - 1, ny - 1, ny - 1), lbmx.make_bc(lbmx.BC_BOUNCEBACK))
    inlet = lbmx.make_bc(lbmx.BC_ZOUHE_VELOCITY, 0.05, (1.0, 1.0, 0.0))
    obj.add_boundary_table(B(0, 0, 1, ny - 2), inlet, inlet_profile_2d(ny, ncomp), ncomp, ramp_table(steps))
    outlet = lbmx.make_bc(lbmx.BC_ZOUHE_PRESSURE, 1.0)
    obj.add_boundary_table(B(nx - 1, nx - 1, 1, ny - 2), outlet, outlet_profile_2d(ny), 1, None)
    obj.set_objective(lbmx.make_objective(lbmx.OBJ_PRESSURE_DROP, B(1, 1, 1, ny - 2), B(nx - 2, nx - 2, 1, ny - 2),
                                          time_const))


def table_duct3d(obj, nx, ny, nz, steps, time_const=1.0):
    """Duct with a three-component tabulated Hecht inlet (a swirl on a bump) and a uniform
    pressure outlet with a tabulated time factor."""
    B = lbmx.make_box
    bb = lbmx.make_bc(lbmx.BC_BOUNCEBACK)
    obj.add_boundary(B(0, nx - 1, 0, 0, 0, nz - 1), bb)
    obj.add_boundary(B(0, nx - 1, ny - 1, ny - 1, 0, nz - 1), bb)
    obj.add_boundary(B(0, nx - 1, 1, ny - 2, 0, 0), bb)
    obj.add_boundary(B(0, nx - 1, 1, ny - 2, nz - 1, nz - 1), bb)
    y = np.arange(1, ny - 1, dtype=float)[None, :]
    z = np.arange(1, nz - 1, dtype=float)[:, None]
    sy = 2.0 * (y - 0.5) / (ny - 2) - 1.0
    sz = 2.0 * (z - 0.5) / (nz - 2) - 1.0
    bump = np.cos(0.5 * np.pi * sy) * np.cos(0.5 * np.pi * sz)
    prof = np.stack([bump, 0.3 * sz * bump, -0.3 * sy * bump], -1)  # (z, y, 3): natural order z, y, x=0
    inlet = lbmx.make_bc(lbmx.BC_HECHT_VELOCITY, 0.04, (1.0, 1.0, 1.0))
    obj.add_boundary_table(B(0, 0, 1, ny - 2, 1, nz - 2), inlet, prof.ravel(), 3, ramp_table(steps))
    outlet = lbmx.make_bc(lbmx.BC_HECHT_PRESSURE, 1.0)
    t = np.arange(steps + 1, dtype=float)
    obj.add_boundary_table(B(nx - 1, nx - 1, 1, ny - 2, 1, nz - 2), outlet, None, 1, 1.0 + 0.001 * np.sin(t / 5.0))
    obj.set_objective(lbmx.make_objective(lbmx.OBJ_PRESSURE_DROP, B(1, 1, 1, ny - 2, 1, nz - 2),
                                          B(nx - 2, nx - 2, 1, ny - 2, 1, nz - 2), time_const))
