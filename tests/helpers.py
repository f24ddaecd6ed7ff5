"""Shared problem builders for the tests."""
import numpy as np

import lbmx
import oracle
import workloads


def build_pair(lattice, coll, design, nx, ny, nz=1, precision=lbmx.FP64, omega=1.25, alpha_max=None,
               setup="channel", seed=3, time_const=1.0, modulation=None, thermal=False):
    if alpha_max is None:
        alpha_max = 1.0 if design == lbmx.DESIGN_PBB else 0.1
    d = lbmx.make_desc(nx, ny, nz, lattice=lattice, collision=coll, design=design, precision=precision,
                       omega=omega, interp_q=0.1, alpha_max=alpha_max, thermal_diffusivity=0.02,
                       ramp_p=1.0, ramp_ck=0.1)
    do = lbmx.make_desc(nx, ny, nz, lattice=lattice, collision=coll, design=design, precision=lbmx.FP64,
                        omega=omega, interp_q=0.1, alpha_max=alpha_max, thermal_diffusivity=0.02,
                        ramp_p=1.0, ramp_ck=0.1)
    s = lbmx.Solver(d)
    o = oracle.Oracle(do)
    g = workloads.design_field(nx, ny, nz, seed=seed)
    for obj in (s, o):
        if setup == "duct":
            workloads.duct3d(obj, nx, ny, nz, time_const=time_const)
        elif setup == "thermal":
            workloads.thermal_channel2d(obj, nx, ny, time_const=time_const)
        elif setup == "thermal-neumann":
            workloads.thermal_channel2d(obj, nx, ny, time_const=time_const, thermal_outlet="neumann",
                                        outlet="neumann")
        elif setup == "neumann":
            workloads.channel2d(obj, nx, ny, time_const=time_const, outlet="neumann")
        else:
            workloads.channel2d(obj, nx, ny, time_const=time_const, modulation=modulation)
        obj.set_design(g)
    return s, o, g


def relmax(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
