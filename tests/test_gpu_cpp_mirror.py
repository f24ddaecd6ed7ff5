"""The C++ host mirror (lbm-topopt_amd/host/lbm_topopt.hpp) driven by a program written in the
reference application's style (tests/cpp/robust_pump_miniature.cpp): the optimisation loop of
RobustPump2d on a small channel through the reference's class names. The program checks the
per-step objective (collideAndStream + computeMacroFunctional, evaluated from the state's
macros on the host) against the fused forward pass to 1e-12, and that the pressure drop falls
at every iteration with the volume constraint met."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "robust_pump_miniature")


def test_robust_pump_miniature_through_the_mirror(lbmx_mod, tmp_path):
    if not os.path.exists(EXE):
        raise RuntimeError(f"{EXE} not built (run __graft_entry__.build())")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "RESULT OK" in r.stdout
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("iteration")]
    assert len(lines) == 10
    assert (tmp_path / "robust_pump_miniature_design.bin").exists()
    assert (tmp_path / "robust_pump_miniature_fields.vts").exists()


EXE_UNFUSED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "unfused_reverse_loop")


def test_unfused_reverse_loop_with_caller_checkpointing(lbmx_mod, tmp_path):
    """The reference's unfused forward and reverse loops through the mirror's AdjointLBSolver
    (resetAdjoints, setCurrentTimestep, adjointCollide, computeCollideSource, adjointStream) and
    two CheckpointingBase implementations (tests/cpp/unfused_reverse_loop.cpp): dJ/dgamma and the
    adjoint populations bitwise equal to the fused pass, J within 1e-12."""
    if not os.path.exists(EXE_UNFUSED):
        raise RuntimeError(f"{EXE_UNFUSED} not built (run __graft_entry__.build())")
    r = subprocess.run([EXE_UNFUSED], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "RESULT OK" in r.stdout
