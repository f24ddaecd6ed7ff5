"""The batched robust ensemble at its real configuration, and a small FP64 ensemble against the
oracle.

The robust optimisation of applications/RobustPump2d/robustPump2d.cc:40-47,146-166 runs its
eroded / nominal / dilated realisations (Heaviside thresholds eta = 0.3 / 0.5 / 0.7) and its
design candidates on separate MPISplit communicators (src/core/MPISplit.cc:150-234); here they
are members of one batch.
"""
import numpy as np
import pytest

import lbmx
import oracle
import workloads

pytestmark = pytest.mark.gpu


def _cascaded(nx, ny, precision):
    return lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.CASCADED, design=lbmx.DESIGN_PBB,
                          precision=precision, omega=1.0 / 0.6, interp_q=0.1, alpha_max=1.0)


def test_robust_ensemble_24_members_3000_steps(lbmx_mod):
    """24 members (eta 0.3/0.5/0.7, beta 8, x 8 candidates on 256x128, 512x256, 768x384 and
    1024x512, two of each), PBB<IncompressibleCascaded2d> at tau 0.6, FP32, 3000 steps, 64
    checkpoints, autotuned tile: every member's batched dJ/dgamma is bitwise its
    single-instance gradient, and so is J."""
    steps, slots = 3000, 64
    members = [s for s, _, _ in workloads.ensemble_members(lambda nx, ny: lbmx.Solver(_cascaded(nx, ny, lbmx.FP32)),
                                                           steps=steps)]
    assert len(members) == 24
    assert sum(m.nodes for m in members) == 5898240
    b = lbmx.Batch(members)
    b.checkpoint_config(slots, lbmx.CHECKPOINT_BINOMIAL)
    rows, _ = b.autotune(4)
    J, sens = b.gradient(steps)
    rows_used, paired = b.tile()
    assert rows_used == rows
    print(f"ensemble: tile {rows} rows{' paired' if paired else ''}")
    ref = []
    for i, m in enumerate(members):
        m.checkpoint_config(slots, lbmx.CHECKPOINT_BINOMIAL)
        ref.append(m.gradient(steps))
    bad = [i for i, (Js, gs) in enumerate(ref) if not (np.isfinite(Js) and np.all(np.isfinite(gs)))]
    if bad:
        print("members with a non-finite single-instance gradient:", bad)
    assert not bad and np.all(np.isfinite(J))
    for i, (Js, gs) in enumerate(ref):
        assert J[i] == Js, (i, J[i], Js)
        assert np.array_equal(sens[i], gs), (i, float(np.max(np.abs(sens[i] - gs))))
    b.close()
    for m in members:
        m.close()


@pytest.mark.parametrize("rows,paired", [(4, True), (1, False)])
def test_small_fp64_ensemble_matches_the_oracle(lbmx_mod, rows, paired):
    """3 realisations x 2 uneven grids in one FP64 batch, 60 steps, 7 checkpoints: every
    member's J and dJ/dgamma against the oracle at the plain form's bars (1e-10 relative;
    dJ/dgamma in max-norm)."""
    steps = 60
    grids = [(64, 32), (90, 41)]
    etas = (0.3, 0.5, 0.7)
    members, refs = [], []
    for c, (nx, ny) in enumerate(grids):
        raw = workloads.design_field(nx, ny, 1, seed=500 + c)
        for eta in etas:
            g = workloads.heaviside_project(raw, 8.0, eta)
            s = lbmx.Solver(_cascaded(nx, ny, lbmx.FP64))
            o = oracle.Oracle(_cascaded(nx, ny, lbmx.FP64))
            for obj in (s, o):
                workloads.channel2d(obj, nx, ny, time_const=1.0 / steps)
                obj.set_design(g)
            Jo, so, _ = o.gradient(steps)
            members.append(s)
            refs.append((Jo, so))
    b = lbmx.Batch(members)
    b.set_tile(rows, paired)
    b.checkpoint_config(7, lbmx.CHECKPOINT_BINOMIAL)
    J, sens = b.gradient(steps)
    for i, (Jo, so) in enumerate(refs):
        assert abs(J[i] - Jo) <= 1e-10 * abs(Jo), (i, J[i], Jo)
        err = float(np.max(np.abs(sens[i] - so)) / np.max(np.abs(so)))
        assert err <= 1e-10, (i, err)
    b.close()
    for m in members:
        m.close()


def test_reference_adjoint_on_an_ensemble_member(lbmx_mod):
    """One member of the robust ensemble at its full horizon (256 x 128, eta 0.5, 3000 steps),
    FP64: the engine's dJ/dgamma (true transpose) against the oracle and central differences
    of J, and the reference's as-coded cascaded adjoint (oracle only) against both."""
    nx, ny, steps = 256, 128, 3000
    d = _cascaded(nx, ny, lbmx.FP64)
    g = workloads.heaviside_project(workloads.ensemble_candidate(nx, ny, 1000), 8.0, 0.5)
    s = lbmx.Solver(d)
    workloads.channel2d(s, nx, ny, time_const=1.0 / steps)
    s.set_design(g)
    s.checkpoint_config(64, lbmx.CHECKPOINT_BINOMIAL)
    J, sens = s.gradient(steps)
    o = oracle.Oracle(d, threads=16)
    workloads.channel2d(o, nx, ny, time_const=1.0 / steps)
    o.set_design(g)
    o.set_reference_reverse(True)
    Jr, sens_ref, _ = o.gradient(steps)
    assert abs(J - Jr) <= 1e-10 * abs(Jr)  # the forward pass is the same
    scale = np.max(np.abs(sens))
    dev = float(np.max(np.abs(sens_ref - sens)) / scale)
    cells = np.concatenate([np.argsort(-np.abs(sens_ref - sens))[:4],
                            np.random.default_rng(0).choice(g.size, 4, replace=False)])
    h = 1e-5
    fd = []
    for c in cells:
        vals = []
        for sgn in (1, -1):
            gg = g.copy()
            gg[c] += sgn * h
            s.set_design(gg)
            s.init_equilibrium()
            vals.append(s.forward(steps))
        fd.append((vals[0] - vals[1]) / (2 * h))
    fd = np.array(fd)
    e_true = float(np.max(np.abs(sens[cells] - fd)) / scale)
    e_ref = float(np.max(np.abs(sens_ref[cells] - fd)) / scale)
    print(f"ensemble member 256x128 eta 0.5, {steps} steps: reference adjoint vs true {dev:.2e} (max-norm rel.); "
          f"vs central FD on {len(cells)} cells: true {e_true:.1e}, reference {e_ref:.1e}")
    assert e_true < 1e-6
    assert e_ref > 10 * e_true
    s.close()
