"""torch.autograd binding: the gradient from backward() equals central finite differences of
the engine's own forward objective (the check the reference's FiniteDifferenceCheck2d makes,
src/topOpt/finiteDifferenceCheck2d.cc:68-92, here central and with an explicit tolerance)."""
import numpy as np
import pytest
import torch

import lbmx
import lbmx_torch
import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("coll,design", [(lbmx.BGK, lbmx.DESIGN_BRINKMAN), (lbmx.CASCADED, lbmx.DESIGN_PBB)])
def test_autograd_matches_central_differences(lbmx_mod, coll, design):
    nx, ny, steps = 48, 20, 60
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=coll, design=design, omega=1.25, interp_q=0.1,
                       alpha_max=0.1 if design == lbmx.DESIGN_BRINKMAN else 1.0)
    s = lbmx.Solver(d)
    workloads.channel2d(s, nx, ny, time_const=1.0 / steps)
    g0 = workloads.design_field(nx, ny, 1, seed=9)
    gamma = torch.tensor(g0, dtype=torch.float64, requires_grad=True)
    J = lbmx_torch.objective(gamma, s, steps)
    J.backward()
    grad = gamma.grad.numpy()
    rng = np.random.default_rng(4)
    cells = rng.choice(nx * ny, 16, replace=False)  # 16 seed-chosen cells, as in the 1024x512 gradient check
    fd = lbmx_torch.central_difference(s, g0, steps, cells, h=1e-5)
    scale = np.max(np.abs(grad))
    assert np.max(np.abs(fd - grad[cells])) <= 1e-6 * scale, (fd, grad[cells])
    s.close()


def test_gpu_design_tensor_stays_on_device(lbmx_mod):
    """A design tensor on the GPU: same J and bitwise the same gradient as the host path."""
    nx, ny, steps = 40, 18, 40
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, omega=1.25,
                       interp_q=0.1, alpha_max=0.1)
    s = lbmx.Solver(d)
    workloads.channel2d(s, nx, ny, time_const=1.0 / steps)
    g0 = workloads.design_field(nx, ny, 1, seed=2)
    gh = torch.tensor(g0, dtype=torch.float64, requires_grad=True)
    Jh = lbmx_torch.objective(gh, s, steps)
    Jh.backward()
    gd = torch.tensor(g0, dtype=torch.float64, device="cuda", requires_grad=True)
    Jd = lbmx_torch.objective(gd, s, steps)
    Jd.backward()
    assert gd.grad.is_cuda
    assert Jd.item() == Jh.item()
    assert np.array_equal(gd.grad.cpu().numpy(), gh.grad.numpy())
    s.close()


def test_engine_on_a_torch_stream(lbmx_mod):
    """The engine's launches on a torch stream (one shared HIP runtime): same results."""
    nx, ny, steps = 40, 18, 30
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_PBB, omega=1.25)
    s = lbmx.Solver(d)
    workloads.channel2d(s, nx, ny, time_const=1.0 / steps)
    s.set_design(workloads.design_field(nx, ny, 1, seed=3))
    J0, g0 = s.gradient(steps)
    st = torch.cuda.Stream()
    s.set_stream(st.cuda_stream)
    J1, g1 = s.gradient(steps)
    assert J1 == J0 and np.array_equal(g1, g0)
    s.set_stream(0)
    s.close()
