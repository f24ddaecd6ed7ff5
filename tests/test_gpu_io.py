"""Output of the current state as a VTK StructuredGrid (lbmx_write_vts, the counterpart of
LBSolver::outputFields): the file parses as XML, its extents are the grid's, and its fields are
the state's macros and the design, node for node in natural order."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import lbmx
import workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lattice,dims", [(lbmx.D2Q9, (24, 12, 1)), (lbmx.D3Q19, (10, 8, 6)),
                                          (lbmx.D2Q9_D2Q5, (20, 10, 1))])
def test_write_vts_matches_macros(lbmx_mod, tmp_path, lattice, dims):
    nx, ny, nz = dims
    d = lbmx.make_desc(nx, ny, nz, lattice=lattice, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN, omega=1.25,
                       interp_q=0.1, alpha_max=0.1, thermal_diffusivity=0.02, ramp_p=1.0, ramp_ck=0.1)
    s = lbmx.Solver(d)
    if lattice == lbmx.D3Q19:
        workloads.duct3d(s, nx, ny, nz)
    elif lattice == lbmx.D2Q9_D2Q5:
        workloads.thermal_channel2d(s, nx, ny)
    else:
        workloads.channel2d(s, nx, ny)
    s.set_design(workloads.design_field(nx, ny, nz, seed=3))
    s.init_equilibrium(1.0, (0.02, 0.0, 0.0), 0.5)
    s.forward(15)
    path = str(tmp_path / "fields.vts")
    s.write_vts(path)
    root = ET.parse(path).getroot()
    grid = root.find("StructuredGrid")
    assert grid.get("WholeExtent") == f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    arrays = {a.get("Name"): np.array(a.text.split(), dtype=float) for a in grid.iter("DataArray") if a.get("Name")}
    mac = s.get_macros()
    assert np.array_equal(arrays["rho"], mac[:, 0]) and np.array_equal(arrays["ux"], mac[:, 1])
    assert np.array_equal(arrays["uy"], mac[:, 2])
    assert ("uz" in arrays) == (nz > 1) and ("T" in arrays) == (lattice == lbmx.D2Q9_D2Q5)
    if nz > 1:
        assert np.array_equal(arrays["uz"], mac[:, 3])
    if lattice == lbmx.D2Q9_D2Q5:
        assert np.array_equal(arrays["T"], mac[:, 4])
    assert np.array_equal(arrays["gamma"], s.get_design())
    pts = np.array(next(a for a in grid.iter("DataArray") if a.get("NumberOfComponents") == "3").text.split(),
                   dtype=float).reshape(-1, 3)
    assert pts.shape[0] == nx * ny * nz and tuple(pts[1]) == (1.0, 0.0, 0.0) and tuple(pts[nx]) == (0.0, 1.0, 0.0)
    s.close()
