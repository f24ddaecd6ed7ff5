"""Synthetic workloads: design fields and the channel / duct problem setups.

Every setup function takes any object with ``add_boundary``, ``set_design`` and
``set_objective`` (the engine's ``lbmx.Solver``), so the same problem can be built on
other implementations of that interface too.

Design field: gamma ~ U[0,1) per node from SplitMix64 (seeded), smoothed with a 5x5 (5x5x5)
box mean truncated at the domain edges (mean over the in-domain neighbours).
Porosity map: alpha = alpha_max * q (1 - gamma) / (q + gamma), alpha_max = 0.1, q = 0.1
(the reference's BPInterpolation, src/core/interpolationFunctions.hh:18-21).
"""
import numpy as np

import lbmx

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_uniform(n, seed):
    """n doubles in [0,1) from the SplitMix64 stream with the given seed."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _box_sum_axis(a, r, axis):
    c = np.cumsum(a, axis=axis)
    pad = [(0, 0)] * a.ndim
    pad[axis] = (1, 0)
    c = np.pad(c, pad)
    n = a.shape[axis]
    idx = np.arange(n)
    hi = np.minimum(idx + r, n - 1) + 1
    lo = np.maximum(idx - r, 0)
    return np.take(c, hi, axis=axis) - np.take(c, lo, axis=axis)


def box_filter(a, r=2):
    s = a.astype(np.float64)
    cnt = np.ones_like(s)
    for ax in range(a.ndim):
        if a.shape[ax] > 1:
            s = _box_sum_axis(s, r, ax)
            cnt = _box_sum_axis(cnt, r, ax)
    return s / cnt


def design_field(nx, ny, nz=1, seed=0, smooth=True):
    """gamma in natural order (k*ny + j)*nx + i, shape (nz, ny, nx) flattened."""
    g = splitmix64_uniform(nx * ny * nz, seed).reshape(nz, ny, nx)
    if smooth:
        g = box_filter(g, 2)
    return np.ascontiguousarray(g).ravel()


def channel2d(obj, nx, ny, u_max=0.05, rho_out=1.0, objective=lbmx.OBJ_PRESSURE_DROP, time_const=1.0,
              modulation=None, outlet="pressure"):
    """Channel (the 256 x 128 and 1024 x 512 workloads): Zou-He parabolic velocity inlet (x=0, 1<=y<=ny-2), Zou-He pressure
    outlet (x=nx-1) or, with outlet="neumann", the Zou-He Neumann outflow of
    applications/pump2d/pump2d.cc:82 (u_x extrapolated from the two nodes inwards), on-grid
    bounce-back walls at y=0 and y=ny-1 (corners included)."""
    B = lbmx.make_box
    mid = (ny - 1) / 2.0
    kw = {}
    if modulation is not None:
        kw = dict(time_mode=lbmx.TIME_SINE, t_period=modulation[0], t_amp=modulation[1], t_offset=1.0)
    obj.add_boundary(B(0, nx - 1, 0, 0), lbmx.make_bc(lbmx.BC_BOUNCEBACK))
    obj.add_boundary(B(0, nx - 1, ny - 1, ny - 1), lbmx.make_bc(lbmx.BC_BOUNCEBACK))
    obj.add_boundary(B(0, 0, 1, ny - 2), lbmx.make_bc(lbmx.BC_ZOUHE_VELOCITY, u_max, (1.0, 0.0, 0.0),
                                                      lbmx.PROFILE_PARABOLIC, (mid, 0.0), (mid, 1.0), **kw))
    if outlet == "neumann":
        obj.add_boundary(B(nx - 1, nx - 1, 1, ny - 2), lbmx.make_bc(lbmx.BC_ZOUHE_NEUMANN))
    else:
        obj.add_boundary(B(nx - 1, nx - 1, 1, ny - 2), lbmx.make_bc(lbmx.BC_ZOUHE_PRESSURE, rho_out))
    if objective != lbmx.OBJ_NONE:
        inlet = B(0, 0, 1, ny - 2)
        outlet = B(nx - 1, nx - 1, 1, ny - 2)
        obj.set_objective(lbmx.make_objective(objective, inlet, outlet, time_const))


def duct3d(obj, nx, ny, nz, u_in=0.04, rho_out=1.0, objective=lbmx.OBJ_PRESSURE_DROP, time_const=1.0):
    """Duct (the 384 x 192 x 192 workload): Hecht velocity inlet on the west face, Hecht pressure outlet on the east
    face, on-grid bounce-back on the four side walls (edges and corners included), as
    applications/Diffuser3d/diffuser3d.cc:60-99 arranges its walls."""
    B = lbmx.make_box
    bb = lbmx.make_bc(lbmx.BC_BOUNCEBACK)
    obj.add_boundary(B(0, nx - 1, 0, 0, 0, nz - 1), bb)
    obj.add_boundary(B(0, nx - 1, ny - 1, ny - 1, 0, nz - 1), bb)
    obj.add_boundary(B(0, nx - 1, 1, ny - 2, 0, 0), bb)
    obj.add_boundary(B(0, nx - 1, 1, ny - 2, nz - 1, nz - 1), bb)
    obj.add_boundary(B(0, 0, 1, ny - 2, 1, nz - 2), lbmx.make_bc(lbmx.BC_HECHT_VELOCITY, u_in, (1.0, 0.0, 0.0)))
    obj.add_boundary(B(nx - 1, nx - 1, 1, ny - 2, 1, nz - 2), lbmx.make_bc(lbmx.BC_HECHT_PRESSURE, rho_out))
    if objective != lbmx.OBJ_NONE:
        inlet = B(0, 0, 1, ny - 2, 1, nz - 2)
        outlet = B(nx - 1, nx - 1, 1, ny - 2, 1, nz - 2)
        obj.set_objective(lbmx.make_objective(objective, inlet, outlet, time_const))


def thermal_channel2d(obj, nx, ny, u_max=0.05, T_in=1.0, T_wall=0.0, time_const=1.0, thermal_outlet="adiabatic",
                      outlet="pressure"):
    """Thermal channel (the 2048 x 1024 workload): flow as channel2d; hot inlet T=1 (Dirichlet, west), cold
    walls T=0 (Dirichlet, north/south), adiabatic outlet (bounce-back of the D2Q5 part) or,
    with thermal_outlet="neumann", the temperature extrapolated from the two nodes inwards
    (ThermalZouHeNeumann2d, as applications/Regenerator2d/regenerator2d.cc:429 uses it);
    objective: outlet heat flux (ThermalOutflowEast2d)."""
    B = lbmx.make_box
    channel2d(obj, nx, ny, u_max, objective=lbmx.OBJ_NONE, outlet=outlet)
    obj.add_boundary(B(1, nx - 2, 0, 0), lbmx.make_bc(lbmx.BC_THERMAL_DIRICHLET, T_wall))
    obj.add_boundary(B(1, nx - 2, ny - 1, ny - 1), lbmx.make_bc(lbmx.BC_THERMAL_DIRICHLET, T_wall))
    obj.add_boundary(B(0, 0, 1, ny - 2), lbmx.make_bc(lbmx.BC_THERMAL_DIRICHLET, T_in))
    if thermal_outlet == "neumann":
        obj.add_boundary(B(nx - 1, nx - 1, 1, ny - 2), lbmx.make_bc(lbmx.BC_THERMAL_NEUMANN))
    else:
        obj.add_boundary(B(nx - 1, nx - 1, 1, ny - 2), lbmx.make_bc(lbmx.BC_THERMAL_BOUNCEBACK))
    obj.set_objective(lbmx.make_objective(lbmx.OBJ_HEAT_FLUX, B(nx - 1, nx - 1, 1, ny - 2), None, time_const))


def heaviside_project(g, beta=8.0, eta=0.5):
    """Smoothed Heaviside (tanh) projection of a filtered design,
    (tanh(beta eta) + tanh(beta (g - eta))) / (tanh(beta eta) + tanh(beta (1 - eta)))
    (ProjectionFilter2d::filterDesign, src/topOpt/projectionFilter2d.cc:27-47)."""
    g = np.asarray(g, dtype=np.float64)
    return (np.tanh(beta * eta) + np.tanh(beta * (g - eta))) / (np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta)))


def heaviside_project_grad(g, beta=8.0, eta=0.5):
    """d(projection)/dg, the factor filterSensitivities applies (projectionFilter2d.cc:65-67)."""
    g = np.asarray(g, dtype=np.float64)
    return beta * (1.0 - np.tanh(beta * (g - eta)) ** 2) / (np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta)))


# Robust ensemble of the batched form: 3 projection realisations (eroded / nominal / dilated)
# x 8 design candidates on uneven grids, two of each size.
ENSEMBLE_ETAS = (0.3, 0.5, 0.7)
ENSEMBLE_GRIDS = ((256, 128), (256, 128), (512, 256), (512, 256), (768, 384), (768, 384), (1024, 512), (1024, 512))


def ensemble_candidate(nx, ny, seed):
    """A synthetic design candidate of the robust ensemble: the smoothed random field stretched
    to [0, 1], joined with a fluid channel (|y - centre| below about ny/8) that runs from the
    inlet to the outlet, as any candidate of a pump optimisation keeps one. Without the channel
    the dilated realisation (eta = 0.7) of the bare field (values near 0.5) is solid almost
    everywhere, and the velocity inlet then pumps mass into a closed domain until the state
    overflows (FP32, after about 2500 steps)."""
    g = design_field(nx, ny, 1, seed=seed).reshape(ny, nx)
    g = (g - g.min()) / (g.max() - g.min())
    y = np.arange(ny, dtype=np.float64)[:, None]
    band = 1.0 / (1.0 + np.exp((np.abs(y - (ny - 1) / 2.0) - ny / 8.0) / 2.0))
    return np.maximum(g, band).ravel()


def ensemble_members(make_solver, grids=ENSEMBLE_GRIDS, etas=ENSEMBLE_ETAS, beta=8.0, steps=3000, seed0=1000):
    """Builds the robust ensemble: for candidate c (grid grids[c], design ensemble_candidate with
    seed seed0 + c) and realisation eta, a channel problem whose design is the projected field.
    make_solver(nx, ny) returns an object with the Solver setup interface. Returns
    [(solver, c, eta), ...]."""
    out = []
    for c, (nx, ny) in enumerate(grids):
        raw = ensemble_candidate(nx, ny, seed0 + c)
        for eta in etas:
            s = make_solver(nx, ny)
            channel2d(s, nx, ny, time_const=1.0 / steps)
            s.set_design(heaviside_project(raw, beta, eta))
            out.append((s, c, eta))
    return out


def modulated_channel(obj, nx=1024, ny=512, steps=5000, period=1000.0, amp=0.5):
    """Unsteady channel: the inlet amplitude is modulated by 1 + amp sin(2 pi t / period)
    (the functor pattern of applications/RobustPump2d/robustPump2d.cc:14-19); objective: the
    time-integrated dissipated power (total-pressure flux in minus out), timeConst = 1/steps."""
    channel2d(obj, nx, ny, objective=lbmx.OBJ_DISSIPATED_POWER, time_const=1.0 / steps, modulation=(period, amp))
