"""The reference's own adjoint of the cascaded collision, and how far it is from the true one.

ReverseIncompressibleCascaded2d (src/adjointDynamics/reverseCascaded2d.hh:27-133) is not the
transpose of IncompressibleCascaded2d (src/dynamics/cascaded2d.hh:40-95): where it uses its own
recomputation of k3 and k4 (:45-52, without the forward's /4) as the coefficients of du_x and
du_y in the adjoint of k5 (:65, :69), its result departs from the transpose. The oracle holds
that as-coded variant for this report only (Oracle.set_reference_reverse); the engine's
reverse is the true transpose. Its parity with the reference is unpinned: the checks below
show its structure, not its values.
"""
import numpy as np
import pytest

import lbmx
import oracle
import workloads

RNG = np.random.default_rng(7)
W9 = np.array([4 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9])


def _node_oracle(reference):
    o = oracle.Oracle(lbmx.make_desc(4, 4, lattice=lbmx.D2Q9, collision=lbmx.CASCADED, design=lbmx.DESIGN_NONE,
                                     omega=1.0 / 0.6))
    o.set_reference_reverse(reference)
    return o


def _state(u_scale=0.05):
    c = _c()
    rho = 1 + 0.02 * RNG.standard_normal()
    u = u_scale * RNG.standard_normal(2)
    cu = c @ u
    return W9 * (rho + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u) + 0.01 * W9 * RNG.standard_normal(9)


def _c():
    o = _node_oracle(False)
    ex, ey, _, _, _ = o.lattice()
    return np.stack([ex, ey], 1).astype(float)


def _dot_mismatch(o, f, lam, d, h=1e-6):
    """|lam.(dC d) - (dC^T lam).d| / |lam.(dC d)| with dC d by central differences."""
    Jd = (o.collide_node(f + h * d, 1.0)[0] - o.collide_node(f - h * d, 1.0)[0]) / (2 * h)
    lhs = lam @ Jd
    rhs = o.reverse_collide_node(f, 1.0, lam)[0] @ d
    return abs(lhs - rhs) / abs(lhs)


def test_true_reverse_is_the_transpose_and_the_reference_reverse_is_not():
    true_o, ref_o = _node_oracle(False), _node_oracle(True)
    mism_true, mism_ref = [], []
    for _ in range(20):
        f, lam, d = _state(), RNG.standard_normal(9), RNG.standard_normal(9)
        mism_true.append(_dot_mismatch(true_o, f, lam, d))
        mism_ref.append(_dot_mismatch(ref_o, f, lam, d))
    print(f"per-node dot test at |u| ~ 0.05: true transpose median {np.median(mism_true):.1e}, "
          f"reference as coded median {np.median(mism_ref):.1e}")
    assert max(mism_true) < 1e-7
    assert np.median(mism_ref) > 1e-4


def test_reference_deviation_structure():
    """The departure lies in span{c_x, c_y} (it enters only through the adjoints of u_x and
    u_y), is linear in lam through kappa_5 = sum_q T_q5 lam_q alone (the k5 column of the
    moment-to-distribution map, cascaded2d.hh:77-95), and vanishes at an equilibrium at rest."""
    true_o, ref_o = _node_oracle(False), _node_oracle(True)
    c = _c()
    T5 = np.array([4, 1, -2, 1, -2, 1, -2, 1, -2], float)
    f = _state()
    diffs = []
    for _ in range(6):
        lam = RNG.standard_normal(9)
        dlt = ref_o.reverse_collide_node(f, 1.0, lam)[0] - true_o.reverse_collide_node(f, 1.0, lam)[0]
        coef, res, *_ = np.linalg.lstsq(c, dlt, rcond=None)
        assert np.max(np.abs(c @ coef - dlt)) < 1e-15
        diffs.append(coef / (T5 @ lam))
    assert np.max(np.abs(np.array(diffs) - diffs[0])) < 1e-12 * np.max(np.abs(diffs[0]))
    # lam orthogonal to the k5 column: no departure
    lam = RNG.standard_normal(9)
    lam -= (T5 @ lam) / (T5 @ T5) * T5
    dlt = ref_o.reverse_collide_node(f, 1.0, lam)[0] - true_o.reverse_collide_node(f, 1.0, lam)[0]
    assert np.max(np.abs(dlt)) < 1e-16
    # rest equilibrium: the recomputed k3, k4 agree with the forward's
    feq0 = W9 * 1.01
    lam = RNG.standard_normal(9)
    dlt = ref_o.reverse_collide_node(feq0, 1.0, lam)[0] - true_o.reverse_collide_node(feq0, 1.0, lam)[0]
    assert np.max(np.abs(dlt)) < 1e-16


def test_reference_gradient_on_an_ensemble_member_grid():
    """One member of the robust ensemble (256 x 128, eta = 0.5, PBB<cascaded> at tau 0.6), on a
    300-step horizon: dJ/dgamma of the true transpose against central differences of J, and
    the reference's as-coded adjoint against both."""
    nx, ny, steps = 256, 128, 300
    d = lbmx.make_desc(nx, ny, lattice=lbmx.D2Q9, collision=lbmx.CASCADED, design=lbmx.DESIGN_PBB,
                       omega=1.0 / 0.6, interp_q=0.1, alpha_max=1.0)
    g = workloads.heaviside_project(workloads.ensemble_candidate(nx, ny, 1000), 8.0, 0.5)
    res = {}
    for ref in (False, True):
        o = oracle.Oracle(d, threads=8)
        workloads.channel2d(o, nx, ny, time_const=1.0 / steps)
        o.set_design(g)
        o.set_reference_reverse(ref)
        J, sens, _ = o.gradient(steps)
        res[ref] = (J, sens, o)
    J, s_true, o = res[False]
    _, s_ref, _ = res[True]
    scale = np.max(np.abs(s_true))
    dev = np.max(np.abs(s_ref - s_true)) / scale
    cells = np.argsort(-np.abs(s_ref - s_true))[:3]
    cells = np.concatenate([cells, RNG.choice(g.size, 3, replace=False)])
    h = 1e-5
    fd = []
    for cidx in cells:
        vals = []
        for sgn in (1, -1):
            gg = g.copy()
            gg[cidx] += sgn * h
            o.set_design(gg)
            o.init_equilibrium()
            vals.append(o.forward(steps))
        fd.append((vals[0] - vals[1]) / (2 * h))
    fd = np.array(fd)
    e_true = np.max(np.abs(s_true[cells] - fd)) / scale
    e_ref = np.max(np.abs(s_ref[cells] - fd)) / scale
    print(f"256x128, {steps} steps: reference adjoint vs true transpose {dev:.2e} (max-norm rel.); "
          f"vs central FD on {len(cells)} cells: true {e_true:.1e}, reference {e_ref:.1e}")
    assert e_true < 1e-6
    assert e_ref > 10 * e_true
    assert dev > 1e-5
