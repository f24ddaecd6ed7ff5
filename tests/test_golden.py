"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces the frozen vectors (detects any drift of the checker).
GPU: the engine reproduces them through the C ABI -- FP64 plain form to the parity bars,
and the reduced-storage forms to their stated accuracy.
"""
import glob
import json
import os
import sys

import numpy as np
import pytest

import lbmx
import oracle
from helpers import relmax

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
NAMES = [os.path.splitext(os.path.basename(f))[0] for f in FILES]

# reduced-storage forms on these short runs (20-40 steps): relative error of J and max-norm
# relative error of dJ/dgamma against the FP64 vectors (FP32 collide, FP64 objective and
# sensitivity accumulation, stochastic rounding of the FP16/BF16 storage). Measured: FP32
# J 1e-6 .. 4e-6, gradient <= 1e-6; FP16 J <= 7e-5, gradient <= 1.3e-4; BF16 J <= 7e-5,
# gradient <= 8.4e-4. The 1e-4 gradient target of the FP16 form is asserted on the
# 384 x 192 x 192, 3000-step workload itself (tests/test_gpu_fullsize.py).
FORM_TOL = {lbmx.FP32: (1e-5, 5e-6), lbmx.FP16_SHIFTED: (2e-4, 3e-4), lbmx.BF16_SHIFTED: (1e-3, 2e-3)}
# the temperature populations sit further from their storage shift than the flow ones
FORM_TOL_THERMAL = {lbmx.FP32: (1e-5, 5e-6), lbmx.FP16_SHIFTED: (2e-4, 4e-4), lbmx.BF16_SHIFTED: (1e-3, 4e-3)}


def load(name):
    z = np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)
    return json.loads(str(z["case"])), float(z["J"][0]), z["sens"], z["state"]


def test_golden_files_present():
    assert len(FILES) >= 5


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    c, J, sens, state = load(name)
    o = make_golden.make_problem(oracle.Oracle, c)
    kw = make_golden.init_kwargs(c)
    Jo, so, _ = o.gradient(c["steps"], **kw)
    o.init_equilibrium(1.0, kw.get("u0", (0, 0, 0)), kw.get("T0", 0.0))
    o.forward(c["steps"])
    assert abs(Jo - J) <= 1e-14 * abs(J)
    assert relmax(so, sens) <= 1e-13
    assert relmax(o.f, state) <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_engine_reproduces_golden_fp64(lbmx_mod, name):
    c, J, sens, state = load(name)
    s = make_golden.make_problem(lbmx.Solver, c)
    kw = make_golden.init_kwargs(c)
    s.init_equilibrium(1.0, kw.get("u0", (0, 0, 0)), kw.get("T0", 0.0))
    Jf = s.forward(c["steps"])
    assert relmax(s.get_state(), state) < 1e-12
    assert abs(Jf - J) <= 1e-10 * abs(J)
    Jg, sg = s.gradient(c["steps"], **kw)
    assert abs(Jg - J) <= 1e-10 * abs(J)
    assert relmax(sg, sens) < 1e-10
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [lbmx.FP32, lbmx.FP16_SHIFTED, lbmx.BF16_SHIFTED])
@pytest.mark.parametrize("name", ["d2q9_bgk_pbb_channel", "d3q19_mrt_brinkman_duct", "d2q9_bgk_brinkman_channel",
                                  "d2q9_d2q5_mrt_brinkman_thermal"])
def test_engine_reduced_storage_forms(lbmx_mod, name, prec):
    c, J, sens, _ = load(name)
    s = make_golden.make_problem(lbmx.Solver, c, precision=prec)
    Jg, sg = s.gradient(c["steps"], **make_golden.init_kwargs(c))
    s.close()
    ej, es = abs(Jg - J) / abs(J), relmax(sg, sens)
    print(f"{name} precision {prec}: J rel err {ej:.2e}, sensitivity rel err {es:.2e}")
    tj, ts = (FORM_TOL_THERMAL if "thermal" in name else FORM_TOL)[prec]
    assert ej <= tj and es <= ts
