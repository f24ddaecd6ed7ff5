"""Pins the CPU oracle of the design pipeline (oracle/design_oracle.c) without a reference run.

The reference's filters need PETSc, which is absent; these checks fix what a plausible
mistake would change:
* closed forms of the cone: R = 1 is the identity (its axis neighbours sit at distance R and
  weigh 0), the R = 2 weights and denominators at interior, edge and corner rows in 2D and 3D;
* structure: rows outside the filter box are the identity, columns outside the box still feed
  rows inside it, constants are preserved inside the box;
* the transpose: <F x, s> = <x, F^T s> for the forward map F = D^-1 W and the sensitivity map;
* the projection: p(0) = 0, p(1) = 1, p(eta) = tanh(beta eta) / (tanh(beta eta) + tanh(beta (1 - eta))),
  its derivative against central differences;
* the whole chain (mapping, two stages, extraction) against central differences.
"""
import math

import numpy as np
import pytest

import oracle

SQ2, SQ3 = math.sqrt(2.0), math.sqrt(3.0)


def delta(n, i):
    e = np.zeros(n)
    e[i] = 1.0
    return e


def test_radius_one_is_identity():
    f = oracle.FilterChain(9, 7, 1, (0, 8, 0, 6, 0, 0), stages=((1, False, 1.0, 0.5),))
    x = np.random.default_rng(1).random(f.n)
    assert np.array_equal(f.density(x, 1), x)
    assert np.array_equal(f.density_transpose(x, 1), x)
    assert np.array_equal(f.denominator(), np.ones(f.n))


def test_cone_weights_radius_two_2d():
    nx, ny = 11, 9
    f = oracle.FilterChain(nx, ny, 1, (0, nx - 1, 0, ny - 1, 0, 0), stages=((2, False, 1.0, 0.5),))
    den = f.denominator().reshape(ny, nx)
    # interior: centre 2, four axis neighbours 1, four diagonals 2 - sqrt2; (+-2, 0) weigh 0
    interior = 2 + 4 * 1 + 4 * (2 - SQ2)
    assert abs(den[4, 5] - interior) < 1e-14
    # corner (0, 0): centre, (1,0), (0,1), (1,1) in the grid
    assert abs(den[0, 0] - (2 + 1 + 1 + (2 - SQ2))) < 1e-14
    # edge (x = 5, y = 0): centre, three axis neighbours, two diagonals
    assert abs(den[0, 5] - (2 + 3 + 2 * (2 - SQ2))) < 1e-14
    # impulse response at an interior row: F[c, c + d] = w(d) / den(c)
    c = (4, 5)
    rows = {(0, 0): 2.0, (0, 1): 1.0, (1, 0): 1.0, (1, 1): 2 - SQ2, (0, 2): 0.0, (1, 2): None}
    for (dy, dx), w in rows.items():
        col = (c[0] + dy) * nx + c[1] + dx
        out = f.density(delta(f.n, col), 2).reshape(ny, nx)
        if w is None:  # (1, 2): distance sqrt5 > 2, not in the cone
            assert out[c] == 0.0
        else:
            assert abs(out[c] - w / interior) < 1e-15


def test_cone_weights_radius_two_3d():
    n = 7
    f = oracle.FilterChain(n, n, n, (0, n - 1, 0, n - 1, 0, n - 1), stages=((2, False, 1.0, 0.5),))
    den = f.denominator().reshape(n, n, n)
    # centre 2, 6 axis (1), 12 face diagonals (2 - sqrt2), 8 body diagonals (2 - sqrt3)
    assert abs(den[3, 3, 3] - (2 + 6 + 12 * (2 - SQ2) + 8 * (2 - SQ3))) < 1e-13
    # corner: centre, 3 axis, 3 face diagonals, 1 body diagonal
    assert abs(den[0, 0, 0] - (2 + 3 + 3 * (2 - SQ2) + (2 - SQ3))) < 1e-14


def test_rows_outside_the_box_are_identity_and_columns_outside_feed_the_box():
    nx, ny = 12, 10
    box = (3, 8, 2, 7, 0, 0)
    f = oracle.FilterChain(nx, ny, 1, box, stages=((3, False, 1.0, 0.5),))
    x = np.random.default_rng(2).random(f.n)
    y = f.density(x, 3).reshape(ny, nx)
    xr = x.reshape(ny, nx)
    inside = np.zeros((ny, nx), bool)
    inside[2:8, 3:9] = True
    assert np.array_equal(y[~inside], xr[~inside])
    assert np.array_equal(f.denominator().reshape(ny, nx)[~inside], np.ones((~inside).sum()))
    # a unit value just left of the box reaches the box row beside it
    col = 5 * nx + 2
    assert f.density(delta(f.n, col), 3).reshape(ny, nx)[5, 3] > 0


def test_constants_preserved():
    f = oracle.FilterChain(15, 11, 1, (2, 12, 1, 9, 0, 0), stages=((4, False, 1.0, 0.5),))
    y = f.density(np.full(f.n, 0.37), 4)
    assert np.max(np.abs(y - 0.37)) < 1e-15
    f3 = oracle.FilterChain(8, 7, 6, (1, 6, 0, 6, 1, 4), stages=((2, False, 1.0, 0.5),))
    assert np.max(np.abs(f3.density(np.full(f3.n, 0.81), 2) - 0.81)) < 1e-15


@pytest.mark.parametrize("dims,box,R", [((13, 9, 1), (2, 10, 1, 7, 0, 0), 3), ((16, 12, 1), (0, 15, 0, 11, 0, 0), 5),
                                        ((8, 7, 6), (1, 6, 1, 5, 0, 5), 2)])
def test_transpose_dot_product(dims, box, R):
    f = oracle.FilterChain(*dims, box, stages=((R, False, 1.0, 0.5),))
    rng = np.random.default_rng(3)
    x, s = rng.random(f.n), rng.standard_normal(f.n)
    lhs = np.dot(f.density(x, R), s)
    rhs = np.dot(x, f.density_transpose(s, R))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_projection_closed_forms_and_derivative():
    f = oracle.FilterChain(6, 5, 1, (0, 5, 0, 4, 0, 0), stages=((1, True, 8.0, 0.3),))
    for beta, eta in ((8.0, 0.3), (2.0, 0.5), (1.0, 0.7)):
        p = f.project(np.array([0.0, 1.0, eta] + [0.5] * 27), beta, eta)
        assert abs(p[0]) < 1e-15 and abs(p[1] - 1) < 1e-15
        pe = math.tanh(beta * eta) / (math.tanh(beta * eta) + math.tanh(beta * (1 - eta)))
        assert abs(p[2] - pe) < 1e-15
        x = np.linspace(0.01, 0.99, 30)
        s = f.project_sens(x, beta, eta, np.ones(30))
        h = 1e-6
        fd = (f.project(x + h, beta, eta) - f.project(x - h, beta, eta)) / (2 * h)
        assert np.max(np.abs(fd - s)) < 1e-7 * max(1.0, np.max(np.abs(s)))


@pytest.mark.parametrize("dims,dbox,fbox", [((24, 16, 1), (3, 20, 2, 13, 0, 0), (3, 20, 2, 13, 0, 0)),
                                            ((20, 14, 1), (4, 15, 3, 10, 0, 0), (2, 17, 1, 12, 0, 0)),
                                            ((9, 8, 7), (1, 7, 1, 6, 1, 5), (1, 7, 1, 6, 1, 5))])
def test_chain_gradient_matches_central_differences(dims, dbox, fbox):
    """Two projecting stages (DoubleProjectionFilter2d): the sensitivity map is the transpose
    of the chain's Jacobian (checked as directional derivatives)."""
    stages = ((3, True, 2.0, 0.5), (2, True, 1.5, 0.3)) if dims[2] == 1 else ((2, True, 2.0, 0.5), (1, True, 1.5, 0.4))
    rng = np.random.default_rng(4)
    bg = rng.random(dims[0] * dims[1] * dims[2])
    f = oracle.FilterChain(*dims, dbox, fbox, stages=stages, background=bg)
    d = rng.random(f.nd)
    s = rng.standard_normal(f.n)
    f.design(d)
    g = f.sensitivities(s)
    h = 1e-6
    for trial in range(3):
        v = rng.standard_normal(f.nd)
        fd = (np.dot(s, f.design(d + h * v)) - np.dot(s, f.design(d - h * v))) / (2 * h)
        f.design(d)
        assert abs(fd - np.dot(g, v)) < 1e-7 * max(1.0, abs(fd)), trial


def test_variable_threshold_projection():
    """VariableProjectionFilter2d: a per-node eta field; a constant field reproduces the
    constant-eta stage bit for bit, and each node follows its own threshold."""
    f = oracle.FilterChain(8, 6, 1, (0, 7, 0, 5, 0, 0), stages=((2, True, 4.0, 0.5),))
    rng = np.random.default_rng(6)
    x = rng.random(f.n)
    assert np.array_equal(f.project(x, 4.0, 0.5, np.full(f.n, 0.5)), f.project(x, 4.0, 0.5))
    etaf = rng.uniform(0.2, 0.8, f.n)
    p = f.project(x, 4.0, 0.5, etaf)
    for i in (0, 7, 23, 40):
        assert p[i] == f.project(np.full(f.n, x[i]), 4.0, etaf[i])[0]
    # the chain's gradient with a variable threshold
    f.eta_fields[0] = etaf
    d = rng.random(f.nd)
    s = rng.standard_normal(f.n)
    f.design(d)
    g = f.sensitivities(s)
    v = rng.standard_normal(f.nd)
    h = 1e-6
    fd = (np.dot(s, f.design(d + h * v)) - np.dot(s, f.design(d - h * v))) / (2 * h)
    assert abs(fd - np.dot(g, v)) < 1e-7 * max(1.0, abs(fd))


def test_mapping_round_trip_and_background():
    rng = np.random.default_rng(5)
    bg = rng.random(10 * 8)
    f = oracle.FilterChain(10, 8, 1, (2, 6, 1, 5, 0, 0), stages=((1, False, 1.0, 0.5),), background=bg)
    d = rng.random(f.nd)
    full = f.map_to_full(d)
    assert np.array_equal(f.extract(full), d)
    m = np.zeros((8, 10), bool)
    m[1:6, 2:7] = True
    assert np.array_equal(full.reshape(8, 10)[~m], bg.reshape(8, 10)[~m])
    # natural order over the box: x fastest
    assert full[1 * 10 + 3] == d[1] and full[2 * 10 + 2] == d[5]
