"""Writes the golden vectors under tests/golden/ from the CPU oracle.

The reference cannot run (its solver needs PETSc, which is absent), so these vectors are
the oracle's outputs on small problems, frozen after the oracle was pinned by the closed
forms, invariants and finite differences of tests/test_oracle_pinning.py. They let the
device parity tests compare against fixed numbers and catch any later drift of the oracle.

Each case: the problem (lattice, operator, coupling, grid, setup, steps, seed) and the
outputs J (objective summed over the steps), sens (dJ/dgamma, natural node order), and the
state after the forward steps (array-of-structures [node][q], FP64).

Run:  python tests/golden/make_golden.py [case ...]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "lbm-topopt_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import lbmx  # noqa: E402
import oracle  # noqa: E402
import workloads  # noqa: E402

CASES = {
    # reference-exact operator of the 2D channel: PBB<IncompressibleBGK2d, BP(q=0.1)>
    "d2q9_bgk_pbb_channel": dict(lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_PBB, nx=32, ny=16,
                                 nz=1, setup="channel", steps=40, seed=0, omega=1.25, alpha_max=1.0),
    "d2q9_mrt_pbb_channel_sine": dict(lattice=lbmx.D2Q9, collision=lbmx.MRT, design=lbmx.DESIGN_PBB, nx=32,
                                      ny=16, nz=1, setup="channel_sine", steps=40, seed=1, omega=1.25,
                                      alpha_max=1.0),
    "d2q9_bgk_brinkman_channel": dict(lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_BRINKMAN, nx=32,
                                      ny=16, nz=1, setup="channel", steps=40, seed=0, omega=1.25, alpha_max=0.1),
    # the ensemble operator: PBB<IncompressibleCascaded2d, BP(q=0.1)> at tau = 0.6
    "d2q9_cascaded_pbb_channel": dict(lattice=lbmx.D2Q9, collision=lbmx.CASCADED, design=lbmx.DESIGN_PBB, nx=32,
                                      ny=16, nz=1, setup="channel", steps=40, seed=2, omega=1.0 / 0.6,
                                      alpha_max=1.0),
    "d3q19_mrt_brinkman_duct": dict(lattice=lbmx.D3Q19, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, nx=14,
                                    ny=10, nz=8, setup="duct", steps=20, seed=0, omega=1.6, alpha_max=0.1),
    "d2q9_d2q5_mrt_brinkman_thermal": dict(lattice=lbmx.D2Q9_D2Q5, collision=lbmx.MRT,
                                           design=lbmx.DESIGN_BRINKMAN, nx=24, ny=12, nz=1, setup="thermal",
                                           steps=40, seed=0, omega=1.25, alpha_max=0.1),
}


def make_problem(impl_cls, c, precision=lbmx.FP64):
    d = lbmx.make_desc(c["nx"], c["ny"], c["nz"], lattice=c["lattice"], collision=c["collision"],
                       design=c["design"], precision=precision, omega=c["omega"], interp_q=0.1,
                       alpha_max=c["alpha_max"], thermal_diffusivity=0.02, ramp_p=1.0, ramp_ck=0.1)
    obj = impl_cls(d)
    tc = 1.0 / c["steps"]
    if c["setup"] == "duct":
        workloads.duct3d(obj, c["nx"], c["ny"], c["nz"], time_const=tc)
    elif c["setup"] == "thermal":
        workloads.thermal_channel2d(obj, c["nx"], c["ny"], time_const=tc)
    elif c["setup"] == "channel_sine":
        workloads.channel2d(obj, c["nx"], c["ny"], time_const=tc, modulation=(20.0, 0.5))
    else:
        workloads.channel2d(obj, c["nx"], c["ny"], time_const=tc)
    obj.set_design(workloads.design_field(c["nx"], c["ny"], c["nz"], seed=c["seed"]))
    return obj


def init_kwargs(c):
    return dict(u0=(0.03, 0.0, 0.0), T0=0.5) if c["setup"] == "thermal" else {}


def main(names=None):
    for name, c in CASES.items():
        if names and name not in names:
            continue
        o = make_problem(oracle.Oracle, c)
        kw = init_kwargs(c)
        J, sens, _ = o.gradient(c["steps"], **kw)
        o.init_equilibrium(1.0, kw.get("u0", (0, 0, 0)), kw.get("T0", 0.0))
        Jf = o.forward(c["steps"])
        assert abs(Jf - J) <= 1e-14 * abs(J)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), J=np.array([J]), sens=sens, state=o.f,
                            case=np.array(json.dumps(c)))
        print(f"{name}: J={J:.15e} |sens|max={np.max(np.abs(sens)):.3e}")


if __name__ == "__main__":
    main(sys.argv[1:])
