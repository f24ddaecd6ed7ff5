"""Batched robust-ensemble form against single-instance gradients, through the C ABI.

Each member of a batch must get exactly the gradient it gets alone: the batched sweep runs
the same node code on the same data, so dJ/dgamma is bitwise equal. J is bitwise equal too,
for every tile shape: the objective is summed per 64-node row segment, whatever the tile.
"""
import numpy as np
import pytest

import lbmx
import workloads

pytestmark = pytest.mark.gpu

STEPS = 40


def member(nx, ny, coll, design, prec, seed, eta, sine=False, objective=lbmx.OBJ_PRESSURE_DROP):
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=coll, design=design, precision=prec,
                       omega=1.0 / 0.6, interp_q=0.1, alpha_max=1.0 if design == lbmx.DESIGN_PBB else 0.1)
    s = lbmx.Solver(d)
    workloads.channel2d(s, nx, ny, objective=objective, time_const=1.0 / STEPS,
                        modulation=(17.0, 0.4) if sine else None)
    s.set_design(workloads.heaviside_project(workloads.design_field(nx, ny, 1, seed=seed), 8.0, eta))
    return s


def build(coll, design, prec):
    # uneven grids (not multiples of the tile), different designs, boundary time modes and functionals
    spec = [((64, 32), 0.3, False, lbmx.OBJ_PRESSURE_DROP), ((96, 40), 0.5, True, lbmx.OBJ_PRESSURE_DROP),
            ((130, 50), 0.7, False, lbmx.OBJ_OUTFLOW), ((70, 21), 0.5, True, lbmx.OBJ_DISSIPATED_POWER)]
    return [member(nx, ny, coll, design, prec, 1000 + i, eta, sine, obj)
            for i, ((nx, ny), eta, sine, obj) in enumerate(spec)]


def singles(members, slots):
    out = []
    for m in members:
        m.checkpoint_config(slots)
        out.append(m.gradient(STEPS))
    return out


SHAPES = [(1, False), (2, True), (4, True), (2, False), (4, False)]


def check_J(J, Js, rows):
    assert J == Js, (rows, J, Js)


@pytest.mark.parametrize("rows,paired", SHAPES, ids=[f"{r}{'p' if p else ''}" for r, p in SHAPES])
@pytest.mark.parametrize("prec", [lbmx.FP64, lbmx.FP32])
def test_batch_equals_single_instances(lbmx_mod, prec, rows, paired):
    members = build(lbmx.CASCADED, lbmx.DESIGN_PBB, prec)
    ref = singles(members, 6)
    b = lbmx.Batch(members)
    b.set_tile(rows, paired)
    assert b.tile() == (rows, paired)
    b.checkpoint_config(6)
    J, sens = b.gradient(STEPS)
    for i, (Js, gs) in enumerate(ref):
        check_J(J[i], Js, rows)
        assert np.array_equal(sens[i], gs), (i, float(np.max(np.abs(sens[i] - gs))))
    st = b.stats()
    assert st["adj_sweeps"] == STEPS
    b.close()


@pytest.mark.parametrize("coll,design", [(lbmx.BGK, lbmx.DESIGN_BRINKMAN), (lbmx.MRT, lbmx.DESIGN_PBB),
                                         (lbmx.MRT, lbmx.DESIGN_NONE)])
def test_batch_other_operators(lbmx_mod, coll, design):
    members = build(coll, design, lbmx.FP64)
    ref = singles(members, 0)
    b = lbmx.Batch(members)
    b.checkpoint_config(0)
    J, sens = b.gradient(STEPS)
    for i, (Js, gs) in enumerate(ref):
        check_J(J[i], Js, 4)
        assert np.array_equal(sens[i], gs)
    b.close()


def test_batch_autotune_and_repeat(lbmx_mod):
    members = build(lbmx.CASCADED, lbmx.DESIGN_PBB, lbmx.FP32)
    ref = singles(members, 5)
    b = lbmx.Batch(members)
    b.checkpoint_config(5)
    J0, _ = b.gradient(STEPS)
    rows, ms = b.autotune(4)
    assert rows in (1, 2, 4) and all(0 < v < 1e3 for v in ms)
    assert b.tile()[0] == rows
    # the timed sweeps overwrote the members' dJ/dgamma arrays: reading them is an error now
    for m in members:
        with pytest.raises(lbmx.LbmxError) as e:
            m.get_sensitivity()
        assert e.value.code == lbmx.ERR_STATE
    for _ in range(2):  # the batch is reusable; results do not depend on the tile shape
        J, sens = b.gradient(STEPS)
        for i, (Js, gs) in enumerate(ref):
            check_J(J[i], Js, rows)
            assert np.array_equal(sens[i], gs)
            assert np.array_equal(members[i].get_sensitivity(), gs)
    b.close()


def test_batch_rejects_mixed_members(lbmx_mod):
    a = member(64, 32, lbmx.CASCADED, lbmx.DESIGN_PBB, lbmx.FP32, 1, 0.5)
    c = member(64, 32, lbmx.MRT, lbmx.DESIGN_PBB, lbmx.FP32, 2, 0.5)
    with pytest.raises(lbmx.LbmxError):
        lbmx.Batch([a, c])
    with pytest.raises(lbmx.LbmxError):
        lbmx.Batch([a, a])


def test_batch_graph_replays_match_eager(lbmx_mod):
    """The captured gradient pass replays to exactly the eager results: several replays with
    eager work and synchronisations of the program in between, and a design change."""
    members = build(lbmx.CASCADED, lbmx.DESIGN_PBB, lbmx.FP32)
    b = lbmx.Batch(members)
    b.checkpoint_config(5)
    J0, s0 = b.gradient(STEPS)                      # eager
    b.set_graph(True)
    for rep in range(3):
        J, sens = b.gradient(STEPS)                 # capture on the first call, then replays
        assert np.array_equal(J, J0), rep
        for i in range(len(members)):
            assert np.array_equal(sens[i], s0[i]), (rep, i)
        members[rep % len(members)].gradient(7)     # the program's own eager work in between
    # new designs are read from the same device arrays by the replay
    g2 = [workloads.heaviside_project(workloads.design_field(m.desc.nx, m.desc.ny, 1, seed=77 + i), 8.0, 0.5)
          for i, m in enumerate(members)]
    for m, g in zip(members, g2):
        m.set_design(g)
    Jg, sg = b.gradient(STEPS)
    b.set_graph(False)
    Je, se = b.gradient(STEPS)
    assert np.array_equal(Jg, Je)
    for i in range(len(members)):
        assert np.array_equal(sg[i], se[i])
    assert not np.array_equal(Jg, J0)
    b.close()


def test_batch_checkpoint_placements_agree(lbmx_mod):
    """The static and binomial placements of the shared schedule give the dynamic schedule's
    results bit for bit, eager and replayed from a captured graph, with fewer or equal
    recomputations."""
    members = build(lbmx.CASCADED, lbmx.DESIGN_PBB, lbmx.FP32)
    b = lbmx.Batch(members)
    b.checkpoint_config(5, lbmx.CHECKPOINT_DYNAMIC)
    J0, s0 = b.gradient(STEPS)
    for mode in (lbmx.CHECKPOINT_STATIC, lbmx.CHECKPOINT_BINOMIAL):
        b.checkpoint_config(5, mode)
        for graph in (False, True, True):
            b.set_graph(graph)
            J, sens = b.gradient(STEPS)
            assert np.array_equal(J, J0), (mode, graph)
            for i in range(len(members)):
                assert np.array_equal(sens[i], s0[i]), (mode, graph, i)
        b.set_graph(False)
    assert lbmx.checkpoint_simulate(5, STEPS, lbmx.CHECKPOINT_BINOMIAL)[0] <= lbmx.checkpoint_simulate(5, STEPS)[0]
    b.close()


def member3d(nx, ny, nz, coll, design, prec, seed, eta):
    d = lbmx.make_desc(nx, ny, nz, lattice=lbmx.D3Q19, collision=coll, design=design, precision=prec, omega=1.6,
                       interp_q=0.1, alpha_max=1.0 if design == lbmx.DESIGN_PBB else 0.1)
    s = lbmx.Solver(d)
    workloads.duct3d(s, nx, ny, nz, time_const=1.0 / STEPS)
    s.set_design(workloads.heaviside_project(workloads.design_field(nx, ny, nz, seed=seed), 8.0, eta))
    return s


@pytest.mark.parametrize("coll,design,prec,rows,paired", [(lbmx.MRT, lbmx.DESIGN_BRINKMAN, lbmx.FP64, 4, True),
                                                          (lbmx.MRT, lbmx.DESIGN_BRINKMAN, lbmx.FP64, 4, False),
                                                          (lbmx.MRT, lbmx.DESIGN_PBB, lbmx.FP32, 1, False),
                                                          (lbmx.BGK, lbmx.DESIGN_BRINKMAN, lbmx.FP64, 2, True)])
def test_batch_d3q19_equals_single_instances(lbmx_mod, coll, design, prec, rows, paired):
    """Robust 3D realisations (eta = 0.3/0.5/0.7) on uneven duct grids in one batch; each
    member's dJ/dgamma bitwise its single-instance gradient, also replayed from a HIP graph."""
    spec = [((40, 18, 14), 0.3), ((33, 15, 12), 0.5), ((48, 20, 10), 0.7)]
    members = [member3d(*g, coll, design, prec, 2000 + i, eta) for i, (g, eta) in enumerate(spec)]
    ref = singles(members, 6)
    b = lbmx.Batch(members)
    b.set_tile(rows, paired)
    b.checkpoint_config(6)
    J, sens = b.gradient(STEPS)
    for i, (Js, gs) in enumerate(ref):
        check_J(J[i], Js, rows)
        assert np.array_equal(sens[i], gs), (i, float(np.max(np.abs(sens[i] - gs))))
    b.set_graph(True)
    for _ in range(2):
        Jg, sg = b.gradient(STEPS)
        for i in range(len(members)):
            assert Jg[i] == J[i] and np.array_equal(sg[i], sens[i])
    b.close()


def test_batch_rejects_mixed_lattices(lbmx_mod):
    a = member(40, 20, lbmx.MRT, lbmx.DESIGN_PBB, lbmx.FP64, 1, 0.5)
    c = member3d(20, 10, 8, lbmx.MRT, lbmx.DESIGN_PBB, lbmx.FP64, 2, 0.5)
    with pytest.raises(lbmx.LbmxError) as e:
        lbmx.Batch([a, c])
    assert e.value.code == lbmx.ERR_ARG


def member_thermal(nx, ny, coll, design, prec, seed, eta):
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9_D2Q5, collision=coll, design=design, precision=prec, omega=1.25,
                       interp_q=0.1, alpha_max=1.0 if design == lbmx.DESIGN_PBB else 0.1, thermal_diffusivity=0.02,
                       ramp_p=1.0, ramp_ck=0.1)
    s = lbmx.Solver(d)
    workloads.thermal_channel2d(s, nx, ny, time_const=1.0 / STEPS)
    s.set_design(workloads.heaviside_project(workloads.design_field(nx, ny, 1, seed=seed), 8.0, eta))
    return s


@pytest.mark.parametrize("coll,design,prec", [(lbmx.MRT, lbmx.DESIGN_PBB, lbmx.FP64),
                                              (lbmx.BGK, lbmx.DESIGN_BRINKMAN, lbmx.FP32),
                                              (lbmx.MRT, lbmx.DESIGN_BRINKMAN, lbmx.FP16_SHIFTED),
                                              (lbmx.MRT, lbmx.DESIGN_PBB, lbmx.BF16_SHIFTED)])
def test_batch_thermal_equals_single_instances(lbmx_mod, coll, design, prec):
    """Thermal realisations (D2Q9 + D2Q5, heat-flux objective) in one batch: each member's
    dJ/dgamma bitwise its single-instance gradient from a moving, warm initial state."""
    spec = [((60, 26), 0.3), ((83, 30), 0.5), ((48, 21), 0.7)]
    if prec in (lbmx.FP16_SHIFTED, lbmx.BF16_SHIFTED):
        spec.append(((1025, 260), 0.5))  # wide enough for the single instance's two-node forward sweep
    members = [member_thermal(nx, ny, coll, design, prec, 3000 + i, eta) for i, ((nx, ny), eta) in enumerate(spec)]
    u0, T0 = (0.03, 0.0, 0.0), 0.5
    ref = []
    for m in members:
        m.checkpoint_config(6)
        ref.append(m.gradient(STEPS, u0=u0, T0=T0))
    for rows, paired in ((4, True), (4, False)):
        b = lbmx.Batch(members)
        b.set_tile(rows, paired)
        b.checkpoint_config(6)
        J, sens = b.gradient(STEPS, u0=u0, T0=T0)
        for i, (Js, gs) in enumerate(ref):
            assert abs(Js) > 0
            check_J(J[i], Js, rows)
            assert np.array_equal(sens[i], gs), (i, float(np.max(np.abs(sens[i] - gs))))
        b.close()


def test_batch_members_keep_their_sensitivities(lbmx_mod):
    """After a batched gradient each member hands out its own dJ/dgamma (getSensitivities, and
    the device array the filter's sensitivity map reads); its adjoint stays with the batch."""
    members = build(lbmx.MRT, lbmx.DESIGN_PBB, lbmx.FP64)[:2]
    b = lbmx.Batch(members)
    b.checkpoint_config(5)
    _, sens = b.gradient(STEPS)
    for m, g in zip(members, sens):
        assert np.array_equal(m.get_sensitivity(), g)
        assert np.array_equal(m.sensitivity_device().cpu().numpy(), g)
        with pytest.raises(lbmx.LbmxError) as e:
            m.get_adjoint()
        assert e.value.code == lbmx.ERR_STATE
    nx, ny = members[0].desc.nx, members[0].desc.ny
    box = lbmx.make_box(8, nx - 9, 1, ny - 2)
    filt = lbmx.Filter(lbmx.make_filter_desc(nx, ny, 1, box, box, ((2, False, 1.0, 0.5),)))
    assert np.array_equal(filt.solver_sensitivities(members[0]), filt.sensitivities(sens[0]))
    b.close()


@pytest.mark.parametrize("prec", [lbmx.FP16_SHIFTED, lbmx.BF16_SHIFTED], ids=["fp16", "bf16"])
def test_batch_shifted_storage_forms(lbmx_mod, prec):
    """The shifted-storage forms in a batch: the stochastic rounding is keyed by each member's
    own node offsets and steps, so every member's dJ/dgamma is bitwise its single-instance one.
    On grids of at least 1024 one-node workgroups a single instance sweeps forward with two
    x-adjacent nodes per thread, the batch always with one: the wide members take that path, and
    their odd widths leave the last thread of a row with one node."""
    members = build(lbmx.CASCADED, lbmx.DESIGN_PBB, prec)
    members3 = [member3d(30, 14, 12, lbmx.MRT, lbmx.DESIGN_BRINKMAN, prec, 4000 + i, eta)
                for i, eta in enumerate((0.4, 0.6))]
    # 17 x 64 and 3 x 11 x 32 one-node workgroups
    wide = [member(1025, 256, lbmx.CASCADED, lbmx.DESIGN_PBB, prec, 4100, 0.5, True),
            member(1023, 260, lbmx.CASCADED, lbmx.DESIGN_PBB, prec, 4101, 0.4, False, lbmx.OBJ_OUTFLOW)]
    wide3 = [member3d(129, 44, 32, lbmx.MRT, lbmx.DESIGN_BRINKMAN, prec, 4200, 0.5)]
    # a moment-space operator, whose doubled populations go straight to the codecs
    wide_mrt = [member(1025, 256, lbmx.MRT, lbmx.DESIGN_BRINKMAN, prec, 4300, 0.5)]
    for group in (members, members3, wide, wide3, wide_mrt):
        ref = singles(group, 6)
        b = lbmx.Batch(group)
        b.checkpoint_config(6)
        J, sens = b.gradient(STEPS)
        for i, (Js, gs) in enumerate(ref):
            check_J(J[i], Js, 4)
            assert np.array_equal(sens[i], gs), (i, float(np.max(np.abs(sens[i] - gs))))
        b.close()
