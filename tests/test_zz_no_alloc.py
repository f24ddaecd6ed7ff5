"""lbmx_design_change_noalloc (the no-allocation variant of lbmx_mma_design_change_device)
against that entry point, through the C ABI: two calls in a row on one device, the second on a
smaller design, share the library's static scratch array."""
import ctypes as C

import numpy as np
import pytest

import lbmx

pytestmark = pytest.mark.gpu


def _noalloc(xnew, xold):
    import torch
    ch = torch.full((1,), -1.0, dtype=torch.float64, device=xnew.device)
    lib = lbmx.library()
    n = xnew.numel()
    code = lib.lbmx_design_change_noalloc(lbmx._tptr(xnew, n), lbmx._tptr(xold, n), n, lbmx._tptr(ch, 1), None)
    lbmx._check(code)
    torch.cuda.synchronize()
    return float(ch.cpu()[0])


def _handle_change(n, seed):
    """An MMA update on device arrays, then the handle's design change and the two designs."""
    import torch
    rng = np.random.default_rng(seed)
    t = rng.random(n)
    h = lbmx.MMA(n, 1)
    x = torch.full((n,), 0.5, dtype=torch.float64, device="cuda")
    xold = x.clone()
    xc = x.cpu().numpy()
    dfdx = torch.tensor(2 * (xc - t), device="cuda")
    dg = torch.full((n,), 1.0 / n, dtype=torch.float64, device="cuda")
    h.update_device(x, dfdx, [xc.mean() - 0.35], [dg], 0.2)
    ch = h.design_change_device()
    h.close()
    return ch, x, xold


def test_noalloc_design_change_matches_the_handle_entry_point(lbmx_mod):
    import torch
    results = []
    for n, seed in ((700_001, 1), (4_099, 2)):  # the second call is smaller than the first
        ch, x, xold = _handle_change(n, seed)
        got = _noalloc(x, xold)
        results.append((n, ch, got))
        assert got == ch, (n, got, ch)  # a max-abs reduction is exact: the same value bit for bit
        assert got == float(torch.max(torch.abs(x - xold)).cpu())
    assert results[0][1] != results[1][1]


def test_noalloc_rejects_bad_sizes(lbmx_mod):
    import torch
    lib = lbmx.library()
    a = torch.zeros(4, dtype=torch.float64, device="cuda")
    p = lbmx._tptr(a, 4)
    for n in (0, -3, (1 << 40) + 1):
        assert lib.lbmx_design_change_noalloc(p, p, n, p, None) == lbmx.ERR_ARG
    assert lib.lbmx_design_change_noalloc(None, p, 4, p, None) == lbmx.ERR_ARG
