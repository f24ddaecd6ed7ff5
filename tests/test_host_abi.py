"""C-ABI library and host logic, without a device.

* liblbmx.so loads and exports every function that include/lbmx.h declares;
* the Python binding's symbol list is the header's;
* boundary boxes split into the same pieces, in the same order, as the oracle's restatement
  of the reference's splitting, with the documented corner/edge rule;
* the checkpoint schedule replay (Wang dynamic checkpointing) delivers the right state at
  every reverse step, never exceeds its slots, and meets closed forms and bounds.
"""
import os
import re
from math import comb

import numpy as np
import pytest

import lbmx
import oracle
import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbmx.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbmx_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_binding_symbols():
    assert header_functions() == sorted(lbmx.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = lbmx.library()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_host_mirror_header_compiles():
    """The C++ mirror of the reference's classes (lbm-topopt_amd/host/lbm_topopt.hpp) and the
    program written against it compile with the system C++ compiler (no device needed)."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "robust_pump_miniature.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "lbm-topopt_amd", "host"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_petsc_binary_vec_format(tmp_path):
    """writeVectorToBinary / loadVectorFromBinary: PETSc's binary Vec layout, big-endian int32
    class id 1211214, int32 length, big-endian FP64 values (the byte string is built here
    from that description, not by the library)."""
    import struct
    v = np.random.default_rng(8).standard_normal(37)
    path = str(tmp_path / "Design_variables_iter_0.bin")
    lbmx.write_petsc_vec(path, v)
    raw = open(path, "rb").read()
    assert raw == struct.pack(">ii", 1211214, v.size) + v.astype(">f8").tobytes()
    assert open(path + ".info").read() == "-vecload_block_size 1\n"
    assert np.array_equal(lbmx.load_petsc_vec(path, v.size), v)
    with pytest.raises(lbmx.LbmxError):
        lbmx.load_petsc_vec(path, v.size + 1)                     # length mismatch
    open(path, "wb").write(struct.pack(">ii", 1211215, 2) + b"\0" * 16)
    with pytest.raises(lbmx.LbmxError):
        lbmx.load_petsc_vec(path, 2)                              # wrong class id
    open(path, "wb").write(struct.pack(">ii", 1211214, 4) + b"\0" * 16)
    with pytest.raises(lbmx.LbmxError):
        lbmx.load_petsc_vec(path, 4)                              # truncated data


def test_dmda_orderings():
    """DMDA global (PETSc) order <-> natural order. Worked example from PETSc's ordering rule:
    a 3 x 2 grid over 2 x 1 ranks owning 2 and 1 columns holds rank 0's nodes (0,0), (1,0),
    (0,1), (1,1), then rank 1's (2,0), (2,1), i.e. natural indices 0, 1, 3, 4, 2, 5."""
    natural = np.arange(6, dtype=float) * 10
    glob = lbmx.dmda_natural_to_global(natural, (3, 2), [2, 1], [2])
    assert np.array_equal(glob, [0, 10, 30, 40, 20, 50])
    assert np.array_equal(lbmx.dmda_global_to_natural(glob, (3, 2), [2, 1], [2]), natural)
    # one rank: the orders coincide
    v = np.random.default_rng(3).standard_normal(7 * 5 * 3 * 2)
    assert np.array_equal(lbmx.dmda_natural_to_global(v, (7, 5, 3), [7], [5], [3], dof=2), v)
    # 3D, dof = 19, 3 x 2 x 2 ranks with uneven ranges: the node at natural (i, j, k) lands in
    # the block of the rank owning it, at its natural position inside that block
    dims, lx, ly, lz, dof = (7, 5, 4), [3, 2, 2], [2, 3], [1, 3], 19
    nat = np.random.default_rng(4).standard_normal(7 * 5 * 4 * dof)
    glob = lbmx.dmda_natural_to_global(nat, dims, lx, ly, lz, dof=dof)
    sx, sy, sz = np.cumsum([0] + lx), np.cumsum([0] + ly), np.cumsum([0] + lz)
    for (i, j, k) in [(0, 0, 0), (6, 4, 3), (3, 2, 1), (4, 1, 2), (2, 4, 0)]:
        ri, rj, rk = (np.searchsorted(sx, i, "right") - 1, np.searchsorted(sy, j, "right") - 1,
                      np.searchsorted(sz, k, "right") - 1)
        blocks_before = 0
        for kk in range(len(lz)):
            for jj in range(len(ly)):
                for ii in range(len(lx)):
                    if (kk, jj, ii) < (rk, rj, ri):
                        blocks_before += lx[ii] * ly[jj] * lz[kk]
        local = ((k - sz[rk]) * ly[rj] + (j - sy[rj])) * lx[ri] + (i - sx[ri])
        g = blocks_before + local
        n = (k * 5 + j) * 7 + i
        assert np.array_equal(glob[g * dof:(g + 1) * dof], nat[n * dof:(n + 1) * dof])
    assert np.array_equal(lbmx.dmda_global_to_natural(glob, dims, lx, ly, lz, dof=dof), nat)
    with pytest.raises(lbmx.LbmxError):
        lbmx.dmda_natural_to_global(nat, dims, [3, 2, 1], ly, lz, dof=dof)  # ranges do not add up


def test_error_paths_without_device():
    lib = lbmx.library()
    import ctypes as C
    assert lib.lbmx_create(None, None) == lbmx.ERR_ARG
    assert lib.lbmx_destroy(None) == lbmx.LBMX_OK
    assert lbmx.library().lbmx_last_error() is not None
    bad = lbmx.make_desc(0, 4)
    h = C.c_void_p()
    assert lib.lbmx_create(C.byref(bad), C.byref(h)) != lbmx.LBMX_OK
    with pytest.raises(lbmx.LbmxError):
        lbmx.checkpoint_simulate(1, 10)
    # the copy diagnostics reject their arguments before touching a device, with a message
    for call in (lambda: lbmx.plane_copy_gbs(7, 4096), lambda: lbmx.plane_copy_gbs(19, 16),
                 lambda: lbmx.stream_copy_gbs(nbytes=100)):
        with pytest.raises(lbmx.LbmxError) as e:
            call()
        assert e.value.code == lbmx.ERR_ARG and "needs" in str(e.value)


def test_host_side_checks_of_the_newer_entry_points():
    """Argument and state checks that fail before any device work (no GPU needed)."""
    lib = lbmx.library()
    s = lbmx.Solver(lbmx.make_desc(12, 8, lattice=lbmx.D2Q9, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN,
                                   alpha_max=0.1))
    for call, code in ((lambda: s.adjoint_step(), lbmx.ERR_STATE),            # no lbmx_adjoint_reset
                       (lambda: s.adjoint_set_timestep(3), lbmx.ERR_STATE),
                       (lambda: s.get_sensitivity(), lbmx.ERR_STATE),         # no gradient yet
                       (lambda: s.checkpoint_config(8, mode=5), lbmx.ERR_ARG)):
        with pytest.raises(lbmx.LbmxError) as e:
            call()
        assert e.value.code == code
    assert lib.lbmx_set_design_device(s.h, None, 96) == lbmx.ERR_ARG
    assert lib.lbmx_get_sensitivity_device(s.h, None, None) == lbmx.ERR_ARG
    assert lib.lbmx_mma_update_device(None, None, None, None, None, 0.1) == lbmx.ERR_ARG
    assert lib.lbmx_filter_design_device(None, None, 0, None, 0) == lbmx.ERR_ARG
    # Zou-He Neumann: east/west of a 2D incompressible grid with nx >= 4; the thermal one needs D2Q5
    B = lbmx.make_box
    s.add_boundary(B(11, 11, 1, 6), lbmx.make_bc(lbmx.BC_ZOUHE_NEUMANN))
    with pytest.raises(lbmx.LbmxError) as e:
        s.add_boundary(B(0, 11, 7, 7), lbmx.make_bc(lbmx.BC_ZOUHE_NEUMANN))  # north: not defined
    assert e.value.code == lbmx.ERR_SUP
    with pytest.raises(lbmx.LbmxError) as e:
        s.add_boundary(B(0, 0, 1, 6), lbmx.make_bc(lbmx.BC_THERMAL_NEUMANN))
    assert e.value.code == lbmx.ERR_SUP
    narrow = lbmx.Solver(lbmx.make_desc(3, 8, lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_PBB))
    with pytest.raises(lbmx.LbmxError) as e:
        narrow.add_boundary(B(2, 2, 1, 6), lbmx.make_bc(lbmx.BC_ZOUHE_NEUMANN))  # the neighbours would be boundary nodes
    assert e.value.code == lbmx.ERR_SUP
    cube = lbmx.Solver(lbmx.make_desc(8, 6, 6, lattice=lbmx.D3Q19, collision=lbmx.MRT, design=lbmx.DESIGN_PBB))
    with pytest.raises(lbmx.LbmxError) as e:
        cube.add_boundary(B(7, 7, 1, 4, 1, 4), lbmx.make_bc(lbmx.BC_ZOUHE_NEUMANN))
    assert e.value.code == lbmx.ERR_SUP
    for h in (s, narrow, cube):
        h.close()


def _pieces_engine_vs_oracle(desc, setup, *dims):
    s = lbmx.Solver(desc)
    o = oracle.Oracle(desc)
    for obj in (s, o):
        setup(obj, *dims)
    got = s.boundaries()
    exp = o.boundaries()
    s.close()
    return got, exp


@pytest.mark.parametrize("case", ["channel", "duct", "thermal"])
def test_boundary_splitting_matches_oracle(case):
    if case == "duct":
        d = lbmx.make_desc(9, 7, 6, lattice=lbmx.D3Q19, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN)
        got, exp = _pieces_engine_vs_oracle(d, workloads.duct3d, 9, 7, 6)
        # 4 side walls: faces, 4 edges along x, inlet/outlet corners handled by their pieces
        assert len(got) > 8
    elif case == "thermal":
        d = lbmx.make_desc(9, 7, lattice=lbmx.D2Q9_D2Q5, collision=lbmx.MRT, design=lbmx.DESIGN_BRINKMAN)
        got, exp = _pieces_engine_vs_oracle(d, workloads.thermal_channel2d, 9, 7)
    else:
        d = lbmx.make_desc(9, 7, lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_PBB)
        got, exp = _pieces_engine_vs_oracle(d, workloads.channel2d, 9, 7)
    assert [(t, tuple(n), tuple(b)) for t, n, b in got] == [(t, tuple(n), tuple(b)) for t, n, b in exp]
    # every boundary node of a piece lies on the face/edge/corner its normal names
    nx, ny, nz = d.nx, d.ny, d.nz
    for t, n, b in got:
        for a, (lo, hi, N) in enumerate(((b[0], b[1], nx), (b[2], b[3], ny), (b[4], b[5], nz))):
            if n[a] == -1:
                assert lo == hi == 0
            elif n[a] == 1:
                assert lo == hi == N - 1
            elif N > 1:
                assert lo >= 1 and hi <= N - 2, "edge boxes exclude the corners"


def test_unsupported_boundary_is_rejected_without_side_effects():
    d = lbmx.make_desc(8, 6, lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_PBB)
    s = lbmx.Solver(d)
    n0 = len(s.boundaries())
    with pytest.raises(lbmx.LbmxError):  # Hecht conditions are 3D only
        s.add_boundary(lbmx.make_box(0, 0, 1, 4), lbmx.make_bc(lbmx.BC_HECHT_VELOCITY, 0.01, (1.0, 0.0, 0.0)))
    assert len(s.boundaries()) == n0
    s.close()


def test_design_roundtrip_and_validation():
    d = lbmx.make_desc(8, 6, lattice=lbmx.D2Q9, collision=lbmx.BGK, design=lbmx.DESIGN_PBB)
    s = lbmx.Solver(d)
    g = workloads.design_field(8, 6, 1, seed=11)
    s.set_design(g)
    assert np.array_equal(s.get_design(), g)
    with pytest.raises(lbmx.LbmxError):
        s.set_design(np.zeros(5))
    s.close()


# ----------------------------------------------------------------------- checkpointing
def test_checkpoint_single_slot_closed_form():
    """Two slots leave one stored state: every reverse step restarts from state 0, so the
    recomputation is sum_{i=1}^{n-2} i = (n-1)(n-2)/2 forward steps."""
    for n in (3, 5, 10, 37, 100):
        adv, mx = lbmx.checkpoint_simulate(2, n)
        assert adv == comb(n - 1, 2)
        assert mx <= 2


def test_checkpoint_enough_slots_no_recompute():
    for n in (2, 7, 30):
        adv, mx = lbmx.checkpoint_simulate(n, n)
        assert adv == 0 and mx <= n


@pytest.mark.parametrize("slots", [3, 4, 6, 10, 48, 64])
def test_checkpoint_invariants_and_bound(slots):
    prev = None
    for n in (10, 50, 200, 1000):
        adv, mx = lbmx.checkpoint_simulate(slots, n)  # raises if a wrong state is delivered
        assert mx <= slots
        # repetition bound: with s - 1 free checkpoints every state is recomputed at most
        # t - 1 times, t the least repetition number with C(s - 1 + t, t) >= n (Griewank)
        t = 1
        while comb(slots - 1 + t, t) < n:
            t += 1
        assert adv <= (t - 1) * n + n
        if prev is not None:
            assert adv >= prev[0]
        prev = (adv, n)
    # more slots never cost more recomputation
    a1, _ = lbmx.checkpoint_simulate(slots, 500)
    a2, _ = lbmx.checkpoint_simulate(slots + 1, 500)
    assert a2 <= a1


@pytest.mark.parametrize("slots", [2, 3, 5, 16, 48, 64])
def test_static_checkpointing_schedule(slots):
    """LBStaticCheckpointing: the forward pass keeps the checkpoints of a plan over the known
    horizon; the replay checks that every reverse step still gets the right state. Planning
    the horizon at once with the level rule ends in the placement the online rule reaches after
    the same steps, so the recompute count equals the dynamic schedule's."""
    for n in (2, 9, 64, 500, 3000):
        adv_s, mx_s = lbmx.checkpoint_simulate(slots, n, lbmx.CHECKPOINT_STATIC)
        adv_d, _ = lbmx.checkpoint_simulate(slots, n, lbmx.CHECKPOINT_DYNAMIC)
        assert mx_s <= slots
        assert adv_s == adv_d, (slots, n)
    with pytest.raises(lbmx.LbmxError):
        lbmx.checkpoint_simulate(slots, 10, 7)                   # unknown mode


def _fewest_recomputes(n_max, free_max):
    """Fewest reverse-phase forward steps over every checkpoint placement, by exhaustive search
    over the first split: with state 0 held and s free slots, reversing m states costs
    R(m, s) = min_k [k + R(m-k, s-1) + R(k, s)], R(1, s) = 0, R(m, 0) = m(m-1)/2. The forward
    pass itself supplies the first descent's n-1 steps."""
    R = [[m * (m - 1) // 2 for m in range(n_max + 1)]]
    for s in range(1, free_max + 1):
        row = [0, 0]
        for m in range(2, n_max + 1):
            row.append(min(k + R[s - 1][m - k] + row[k] for k in range(1, m)))
        R.append(row)
    return lambda n, slots: R[min(slots, max(2, n)) - 1][n] - (n - 1) if n > 1 else 0


def test_binomial_checkpointing_is_optimal():
    """The binomial placement delivers every state (the replay checks each one) with the
    fewest forward steps any placement allows, never more than the level-based schedules."""
    best = _fewest_recomputes(120, 7)
    for slots in (2, 3, 4, 5, 8):
        for n in list(range(1, 41)) + [77, 120]:
            adv, mx = lbmx.checkpoint_simulate(slots, n, lbmx.CHECKPOINT_BINOMIAL)
            assert mx <= slots
            assert adv == best(n, slots), (slots, n)
            assert adv <= lbmx.checkpoint_simulate(slots, n)[0]
    # the benchmark workloads; 4726 from the same exhaustive search run to n = 3000
    assert lbmx.checkpoint_simulate(48, 3000, lbmx.CHECKPOINT_BINOMIAL) == (4726, 48)
    assert lbmx.checkpoint_simulate(48, 3000, lbmx.CHECKPOINT_DYNAMIC)[0] == 5161
    for slots, n in ((64, 3000), (64, 5000), (64, 4000)):
        adv_b, mx_b = lbmx.checkpoint_simulate(slots, n, lbmx.CHECKPOINT_BINOMIAL)
        assert mx_b <= slots and adv_b < lbmx.checkpoint_simulate(slots, n)[0]
    assert lbmx.checkpoint_simulate(30, 30, lbmx.CHECKPOINT_BINOMIAL)[0] == 0


def test_checkpoint_headline_schedules():
    """The checkpoint counts of the benchmark workloads stay within twice the offline optimum."""
    def revolve_recompute(l, s):
        t = 0
        while comb(s + t, t) < l:
            t += 1
        return t * l - comb(s + t, t - 1) - l
    for slots, n in ((64, 5000), (48, 3000)):
        adv, mx = lbmx.checkpoint_simulate(slots, n)
        assert mx <= slots
        assert adv <= 2 * revolve_recompute(n - 1, slots - 1) + n


# ----------------------------------------------------------------------- robust ensemble inputs
def test_heaviside_projection_closed_forms():
    g = np.linspace(0.0, 1.0, 101)
    for eta in workloads.ENSEMBLE_ETAS:
        p = workloads.heaviside_project(g, 8.0, eta)
        assert abs(p[0]) < 1e-15 and abs(p[-1] - 1.0) < 1e-15
        assert np.all(np.diff(p) > 0)
        # p(eta) = tanh(beta eta) / (tanh(beta eta) + tanh(beta (1 - eta)))
        pe = workloads.heaviside_project(np.array([eta]), 8.0, eta)[0]
        assert abs(pe - np.tanh(8 * eta) / (np.tanh(8 * eta) + np.tanh(8 * (1 - eta)))) < 1e-15
        d = workloads.heaviside_project_grad(g, 8.0, eta)
        fd = (workloads.heaviside_project(g + 1e-6, 8.0, eta) - workloads.heaviside_project(g - 1e-6, 8.0, eta)) / 2e-6
        assert np.max(np.abs(fd - d)) < 1e-8


def test_ensemble_members_layout():
    class Stub:
        def __init__(self, nx, ny):
            self.nx, self.ny, self.bcs, self.obj, self.design = nx, ny, [], None, None

        def add_boundary(self, box, bc):
            self.bcs.append((box, bc))

        def set_objective(self, o):
            self.obj = o

        def set_design(self, g):
            self.design = np.asarray(g)

    ms = workloads.ensemble_members(Stub, steps=3000)
    assert len(ms) == 24
    assert sum(s.nx * s.ny for s, _, _ in ms) == 5898240
    for s, c, eta in ms:
        assert (s.nx, s.ny) == workloads.ENSEMBLE_GRIDS[c] and eta in workloads.ENSEMBLE_ETAS
        assert s.design.size == s.nx * s.ny and len(s.bcs) == 4 and s.obj is not None
        assert abs(s.obj.time_const - 1.0 / 3000) < 1e-18
    # the three realisations of one candidate project the same raw field: a larger threshold
    # erodes (dilated eta = 0.3 >= nominal 0.5 >= eroded 0.7 pointwise)
    a, b, c = ms[0][0].design, ms[1][0].design, ms[2][0].design
    assert np.all(a >= b) and np.all(b >= c) and not np.array_equal(a, c)
