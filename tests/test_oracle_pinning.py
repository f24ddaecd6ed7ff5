"""Pins the CPU oracle against mathematics, independently of the oracle's own code.

Every check here derives its expected value from a closed form, a textbook basis built in
this file from its defining polynomials, an invariant, brute force, or central finite
differences of the oracle's forward map -- never from a copy of the oracle's expressions.
Reference constants that no published source fixes are cited at their file:line.
"""
from math import comb

import numpy as np
import pytest

import lbmx
import oracle

RNG = np.random.default_rng(20241202)


def mk(lattice=lbmx.D2Q9, coll=lbmx.BGK, design=lbmx.DESIGN_NONE, nx=4, ny=4, nz=1, omega=1.3,
       alpha_max=0.1, q=0.1, diff=0.02, ck=0.1):
    d = lbmx.make_desc(nx, ny, nz, lattice=lattice, collision=coll, design=design, omega=omega,
                       interp_q=q, alpha_max=alpha_max, thermal_diffusivity=diff, ramp_p=1.0, ramp_ck=ck)
    return oracle.Oracle(d)


def lat(o):
    ex, ey, ez, op, w = o.lattice()
    return np.stack([ex, ey, ez], 1).astype(float), op, w


def feq_text(c, w, rho, u):
    """Textbook incompressible second-order equilibrium w (rho + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2)."""
    cu = c @ u
    return w * (rho + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))


def near_eq(c, w, amp=0.01):
    rho = 1 + 0.02 * RNG.standard_normal()
    u = 0.05 * RNG.standard_normal(3)
    if not np.any(c[:, 2]):
        u[2] = 0
    return feq_text(c, w, rho, u) + amp * w * RNG.standard_normal(len(w))


def flow_dims(o):
    c, op, w = lat(o)
    Qf = {9: 9, 19: 19, 14: 9}[len(w)]
    return c[:Qf], op[:Qf], w[:Qf], Qf


def bp_closed(q, g):
    """Borrvall-Petersson interpolation q (1 - g) / (q + g)."""
    return q * (1 - g) / (q + g)


# ---------------------------------------------------------------- textbook moment bases
def basis_d2q9(c):
    """Lallemand & Luo (2000): rho, e, eps, jx, qx, jy, qy, pxx, pxy."""
    cx, cy = c[:, 0], c[:, 1]
    c2 = cx ** 2 + cy ** 2
    rows = dict(rho=np.ones_like(cx), e=-4 + 3 * c2, eps=4 - 10.5 * c2 + 4.5 * c2 ** 2, jx=cx,
                qx=(-5 + 3 * c2) * cx, jy=cy, qy=(-5 + 3 * c2) * cy, pxx=cx ** 2 - cy ** 2, pxy=cx * cy)
    return rows


def basis_d3q19(c):
    """d'Humieres, Ginzburg, Krafczyk, Lallemand, Luo (2002), D3Q19."""
    cx, cy, cz = c[:, 0], c[:, 1], c[:, 2]
    c2 = cx ** 2 + cy ** 2 + cz ** 2
    return dict(rho=np.ones_like(cx), e=19 * c2 - 30, eps=(21 * c2 ** 2 - 53 * c2 + 24) / 2,
                jx=cx, qx=(5 * c2 - 9) * cx, jy=cy, qy=(5 * c2 - 9) * cy, jz=cz, qz=(5 * c2 - 9) * cz,
                pxx3=3 * cx ** 2 - c2, pixx3=(3 * c2 - 5) * (3 * cx ** 2 - c2), pww=cy ** 2 - cz ** 2,
                piww=(3 * c2 - 5) * (cy ** 2 - cz ** 2), pxy=cx * cy, pyz=cy * cz, pxz=cx * cz,
                mx=(cy ** 2 - cz ** 2) * cx, my=(cz ** 2 - cx ** 2) * cy, mz=(cx ** 2 - cy ** 2) * cz)


def basis_d2q5(c):
    cx, cy = c[:, 0], c[:, 1]
    c2 = cx ** 2 + cy ** 2
    return dict(T=np.ones_like(cx), qx=cx, qy=cy, e=5 * c2 - 4, pxx=cx ** 2 - cy ** 2)


# Relaxation rates; viscous rows relax with omega. Constants as coded in the reference:
# incompressibleMRT2d.hh:88-90 (s_e = s_eps = 1.4, s_q = 1.2); MRT3d.hh:10-18
# (s1 = 1.19, s2 = 1.4, s4 = 1.2, s10 = 1.4, s16 = 1.4); thermalMRT2d.hh:105-106 (1.5, 1.5).
RATES_D2Q9 = dict(rho=0, jx=0, jy=0, e=1.4, eps=1.4, qx=1.2, qy=1.2, pxx="w", pxy="w")
RATES_D3Q19 = dict(rho=0, jx=0, jy=0, jz=0, e=1.19, eps=1.4, qx=1.2, qy=1.2, qz=1.2, pxx3="w", pww="w",
                   pxy="w", pyz="w", pxz="w", pixx3=1.4, piww=1.4, mx=1.4, my=1.4, mz=1.4)


def _rate(v, om):
    return om if v == "w" else float(v)


# ================================================================ lattices and equilibria
@pytest.mark.parametrize("lattice", [lbmx.D2Q9, lbmx.D3Q19, lbmx.D2Q9_D2Q5])
def test_lattice_isotropy(lattice):
    o = mk(lattice, nz=1)
    c, op, w = lat(o)
    parts = [(c[:9], op[:9], w[:9], 0)] if lattice != lbmx.D3Q19 else [(c, op, w, 0)]
    if lattice == lbmx.D2Q9_D2Q5:
        parts.append((c[9:], op[9:] - 9, w[9:], 9))
    for cc, oo, ww, _ in parts:
        D = 3 if lattice == lbmx.D3Q19 else 2
        I = np.eye(3)
        I[D:, D:] = 0
        assert abs(ww.sum() - 1) < 1e-15
        assert np.allclose(ww @ cc, 0, atol=1e-16)
        assert np.allclose(np.einsum("q,qa,qb->ab", ww, cc, cc), I / 3, atol=1e-15)
        assert np.all(cc[oo] == -cc)
        assert np.all(oo[oo] == np.arange(len(oo)))
        if len(ww) != 5:  # fourth-order isotropy of D2Q9 / D3Q19
            m4 = np.einsum("q,qa,qb,qc,qd->abcd", ww, cc, cc, cc, cc)
            ex = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I)
                  + np.einsum("ad,bc->abcd", I, I)) / 9
            assert np.allclose(m4, ex, atol=1e-15)


@pytest.mark.parametrize("lattice", [lbmx.D2Q9, lbmx.D3Q19])
def test_equilibrium_moments(lattice):
    nz = 2 if lattice == lbmx.D3Q19 else 1
    o = mk(lattice, nx=3, ny=2, nz=nz)
    c, _, w = lat(o)
    u = np.array([0.04, -0.03, 0.02 if lattice == lbmx.D3Q19 else 0.0])
    o.init_equilibrium(1.03, u)
    f = o.f.reshape(-1, len(w))[0]
    assert abs(f.sum() - 1.03) < 1e-15
    assert np.allclose(c.T @ f, u, atol=1e-16)  # incompressible: j = u (no density factor)
    Pi = np.einsum("q,qa,qb->ab", f, c, c)
    D = 3 if lattice == lbmx.D3Q19 else 2
    ex = 1.03 / 3 * np.eye(3) + np.outer(u, u)
    ex[D:, :] = 0
    ex[:, D:] = 0
    assert np.allclose(Pi, ex, atol=1e-15)


# ================================================================ BGK
@pytest.mark.parametrize("lattice", [lbmx.D2Q9, lbmx.D3Q19])
def test_bgk_closed_form(lattice):
    o = mk(lattice, lbmx.BGK, omega=1.37)
    c, _, w = lat(o)
    for _ in range(5):
        f = near_eq(c, w)
        out, mac = o.collide_node(f, 1.0)
        rho, j = f.sum(), c.T @ f
        fe = feq_text(c, w, rho, j)
        assert np.allclose(out, fe + (1 - 1.37) * (f - fe), rtol=0, atol=1e-15)
        assert abs(out.sum() - rho) < 1e-15 and np.allclose(c.T @ out, j, atol=1e-16)
        assert abs(mac[0] - rho) < 1e-15
    fe = feq_text(c, w, 1.0, np.zeros(3))
    assert np.allclose(o.collide_node(fe, 1.0)[0], fe, atol=1e-17)


@pytest.mark.parametrize("lattice", [lbmx.D2Q9, lbmx.D3Q19])
def test_standard_bgk_closed_form(lattice):
    """The compressible BGK: f' = f + omega (f_eq - f) with the textbook equilibrium
    w rho (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2) at the barycentric velocity u = j / rho."""
    om = 1.37
    o = mk(lattice, lbmx.BGK_STANDARD, omega=om, nz=2 if lattice == lbmx.D3Q19 else 1)
    c, _, w = lat(o)
    for scale in (0.9, 1.0, 1.15):
        f = near_eq(c, w, amp=0.02) * scale
        rho, j = f.sum(), c.T @ f
        u = j / rho
        cu = c @ u
        fe = w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))
        out, mac = o.collide_node(f, 1.0)
        assert np.allclose(out, fe + (1 - om) * (f - fe), rtol=0, atol=2e-15)
        assert abs(out.sum() - rho) < 2e-15 and np.allclose(c.T @ out, j, atol=1e-16)
        # differs from the incompressible operator once rho is away from 1
        if scale != 1.0:
            assert np.max(np.abs(out - mk(lattice, lbmx.BGK, omega=om, nz=o.desc.nz).collide_node(f, 1.0)[0])) > 1e-6


# ================================================================ MRT: basis, rates, equilibria
def _mrt_check(o, rows, rates, om, gamma=1.0):
    c, op, w, Qf = flow_dims(o)
    names = list(rows)
    M = np.array([rows[n] for n in names])
    G = M @ M.T
    assert np.allclose(G, np.diag(np.diag(G))), "basis rows must be orthogonal on this lattice"
    cons = [n for n in names if rates[n] == 0]
    ncons = [n for n in names if rates[n] != 0]
    s = np.array([_rate(rates[n], om) for n in names])
    for _ in range(4):
        f = near_eq(c, w)
        fo, _ = o.collide_node(np.concatenate([f, np.zeros(len(lat(o)[2]) - Qf)]), gamma)
        fo = fo[:Qf]
        # conserved moments unchanged
        for n in cons:
            assert abs(rows[n] @ fo - rows[n] @ f) < 1e-15
        # a perturbation orthogonal to the conserved rows relaxes row by row with rate s_k
        a = {n: 0.01 * RNG.standard_normal() for n in ncons}
        d = sum(a[n] * rows[n] / (rows[n] @ rows[n]) for n in ncons)
        f2 = f + d
        fo2, _ = o.collide_node(np.concatenate([f2, np.zeros(len(lat(o)[2]) - Qf)]), gamma)
        dm = M @ (fo2[:Qf] - fo)
        exp = np.array([(1 - s[k]) * a.get(n, 0.0) for k, n in enumerate(names)])
        assert np.allclose(dm, exp, atol=1e-15), (dm, exp)
    return M, s


def test_mrt2d_basis_rates_equilibrium():
    om = 1.31
    o = mk(lbmx.D2Q9, lbmx.MRT, omega=om)
    c, _, w, Qf = flow_dims(o)
    rows = basis_d2q9(c)
    M, s = _mrt_check(o, rows, RATES_D2Q9, om)
    names = list(rows)
    # equilibrium moments (Lallemand-Luo, incompressible, rho_0 = 1)
    for _ in range(3):
        f = near_eq(c, w)
        rho, (jx, jy, _) = f.sum(), c.T @ f
        fo, _ = o.collide_node(f, 1.0)
        m, mo = M @ f, M @ fo
        meq = dict(rho=rho, jx=jx, jy=jy, e=-2 * rho + 3 * (jx ** 2 + jy ** 2), eps=rho - 3 * (jx ** 2 + jy ** 2),
                   qx=-jx, qy=-jy, pxx=jx ** 2 - jy ** 2, pxy=jx * jy)
        for k, n in enumerate(names):
            assert abs(mo[k] - (m[k] - s[k] * (m[k] - meq[n]))) < 2e-15, n
    fe = feq_text(c, w, 0.98, np.array([0.03, -0.02, 0]))
    assert np.allclose(o.collide_node(fe, 1.0)[0], fe, atol=1e-16)


@pytest.mark.parametrize("coll", [lbmx.MRT, lbmx.MRT_STANDARD])
def test_mrt3d_basis_rates_equilibrium(coll):
    om = 1.6
    o = mk(lbmx.D3Q19, coll, omega=om, nz=2)
    c, _, w, Qf = flow_dims(o)
    rows = basis_d3q19(c)
    names = list(rows)
    M = np.array([rows[n] for n in names])
    s = np.array([_rate(RATES_D3Q19[n], om) for n in names])
    if coll == lbmx.MRT:
        _mrt_check(o, rows, RATES_D3Q19, om)
    for _ in range(3):
        f = near_eq(c, w)
        rho, j = f.sum(), c.T @ f
        fo, _ = o.collide_node(f, 1.0)
        m, mo = M @ f, M @ fo
        rs = 1.0 if coll == lbmx.MRT else rho  # StandardMRT3d divides quadratic terms by rho
        jj = j @ j / rs
        # d'Humieres 2002 equilibria with w_eps = 3, w_epsj = -11/2, w_xx = -1/2
        # (the values coded at MRT3d.hh:19-21)
        p3 = (3 * j[0] ** 2 - j @ j) / rs
        pw = (j[1] ** 2 - j[2] ** 2) / rs
        meq = dict(rho=rho, jx=j[0], jy=j[1], jz=j[2], e=-11 * rho + 19 * jj, eps=3 * rho - 5.5 * jj,
                   qx=-2 / 3 * j[0], qy=-2 / 3 * j[1], qz=-2 / 3 * j[2], pxx3=p3, pixx3=-0.5 * p3, pww=pw,
                   piww=-0.5 * pw, pxy=j[0] * j[1] / rs, pyz=j[1] * j[2] / rs, pxz=j[0] * j[2] / rs,
                   mx=0, my=0, mz=0)
        for k, n in enumerate(names):
            assert abs(mo[k] - (m[k] - s[k] * (m[k] - meq[n]))) < 4e-14, n
    if coll == lbmx.MRT:
        fe = feq_text(c, w, 1.01, np.array([0.02, -0.03, 0.01]))
        assert np.allclose(o.collide_node(fe, 1.0)[0], fe, atol=1e-16)


# ================================================================ cascaded (D2Q9)
def test_cascaded_closed_forms():
    """IncompressibleCascaded2d: conserves rho and j; the shear moments relax with omega to
    rho u_a u_b for any state (u = j, incompressible macros); at u = 0 every higher moment
    reaches its equilibrium with rate 1 (trace 2 rho/3, third order 0, fourth order rho/9).
    The u-dependent terms of k0, k3, k4, k5 are fixed by the central-moment test below."""
    om = 1.0 / 0.6
    o = mk(lbmx.D2Q9, lbmx.CASCADED, omega=om)
    c, _, w = lat(o)
    cx, cy = c[:, 0], c[:, 1]
    for _ in range(6):
        f = near_eq(c, w, amp=0.02)
        rho, j = f.sum(), c.T @ f
        out, _ = o.collide_node(f, 1.0)
        assert abs(out.sum() - rho) < 1e-15 and np.allclose(c.T @ out, j, atol=1e-16)
        pxx, pxy = (cx ** 2 - cy ** 2) @ f, (cx * cy) @ f
        assert abs((cx ** 2 - cy ** 2) @ out - (pxx - om * (pxx - rho * (j[0] ** 2 - j[1] ** 2)))) < 1e-15
        assert abs((cx * cy) @ out - (pxy - om * (pxy - rho * j[0] * j[1]))) < 1e-15
    # u = 0: remove the momentum of a random state with a combination that carries only momentum
    for _ in range(4):
        f = near_eq(c, w, amp=0.03)
        j = c.T @ f
        f = f - 3 * w * (c @ j)  # sum_q w c_q c_q = I/3 removes j exactly
        rho = f.sum()
        assert np.allclose(c.T @ f, 0, atol=1e-16)
        out, _ = o.collide_node(f, 1.0)
        assert abs((cx ** 2 + cy ** 2) @ out - 2 * rho / 3) < 1e-15
        assert abs((cx * cy ** 2) @ out) < 1e-15 and abs((cx ** 2 * cy) @ out) < 1e-15
        assert abs((cx ** 2 * cy ** 2) @ out - rho / 9) < 1e-15
        assert abs((cx ** 2 - cy ** 2) @ out - (1 - om) * ((cx ** 2 - cy ** 2) @ f)) < 1e-15


def _central_moment(g, c, rho, u, m, n):
    """Central moment about u of order (m, n) by the binomial transform of the raw moments, with
    the zeroth and first raw moments taken as rho and rho u: the consistency under which the
    cascaded collision is derived (Geier, Greiner & Korvink, Phys. Rev. E 73, 066705 (2006))."""
    cx, cy = c[:, 0], c[:, 1]
    total = 0.0
    for a in range(m + 1):
        for b in range(n + 1):
            raw = {(0, 0): rho, (1, 0): rho * u[0], (0, 1): rho * u[1]}.get((a, b))
            if raw is None:
                raw = (cx ** a * cy ** b) @ g
            total += comb(m, a) * comb(n, b) * (-u[0]) ** (m - a) * (-u[1]) ** (n - b) * raw
    return total


@pytest.mark.parametrize("rho_scale", [0.93, 1.0, 1.07, 1.2])
def test_cascaded_central_moment_relaxation(rho_scale):
    """IncompressibleCascaded2d is Geier's cascaded collision with the incompressible macros
    u = j: about that u, the post-collision central moments of orders 2-4 relax with the rates
    the reference assigns (shear with omega, the rest with 1) to the equilibria rho/3 per
    axis, 0 off-diagonal, 0 at third order and rho/9 at fourth. With rho and j conserved these
    six conditions fix all nine outputs, so every term of k0..k5 is pinned, u-dependent ones
    included, for any rho."""
    om = 1.0 / 0.6
    o = mk(lbmx.D2Q9, lbmx.CASCADED, omega=om)
    c, _, w = lat(o)
    for _ in range(4):
        f = near_eq(c, w, amp=0.03)
        f = f * (rho_scale / f.sum())
        rho, u = f.sum(), (c.T @ f)[:2]
        out, _ = o.collide_node(f, 1.0)

        def k(g, m, n):
            return _central_moment(g, c, rho, u, m, n)

        assert abs(k(out, 2, 0) + k(out, 0, 2) - 2 * rho / 3) < 1e-15
        assert abs(k(out, 2, 0) - k(out, 0, 2) - (1 - om) * (k(f, 2, 0) - k(f, 0, 2))) < 1e-15
        assert abs(k(out, 1, 1) - (1 - om) * k(f, 1, 1)) < 1e-15
        assert abs(k(out, 2, 1)) < 1e-15 and abs(k(out, 1, 2)) < 1e-15
        assert abs(k(out, 2, 2) - rho / 9) < 1e-15
        # the pre-collision state is not already relaxed, so the check has something to see
        assert abs(k(f, 2, 1)) > 1e-6 and abs(k(f, 2, 2) - rho / 9) > 1e-6


# ================================================================ design couplings
@pytest.mark.parametrize("lattice,coll", [(lbmx.D2Q9, lbmx.BGK), (lbmx.D2Q9, lbmx.MRT), (lbmx.D3Q19, lbmx.MRT)])
def test_pbb_pair_map(lattice, coll):
    nz = 2 if lattice == lbmx.D3Q19 else 1
    base = mk(lattice, coll, lbmx.DESIGN_NONE, nz=nz)
    pbb = mk(lattice, coll, lbmx.DESIGN_PBB, alpha_max=1.0, q=0.1, nz=nz)
    c, op, w = lat(base)
    half = (len(w) - 1) // 2
    for g in (1.0, 0.0, 0.37):
        f = near_eq(c, w)
        b, _ = base.collide_node(f, g)
        p, _ = pbb.collide_node(f, g)
        m = 0.5 * bp_closed(0.1, g)
        for d in range(1, len(w)):
            e = op[d]
            assert abs(p[d] + p[e] - (b[d] + b[e])) < 1e-15
            assert abs((p[d] - p[e]) - (1 - 2 * m) * (b[d] - b[e])) < 1e-15
        assert abs(p[0] - b[0]) == 0
        if g == 1.0:
            assert np.array_equal(p, b)
        if g == 0.0:
            assert np.allclose(p[1:1 + half], p[op[1:1 + half]], atol=1e-16)  # pair mean


@pytest.mark.parametrize("lattice,coll", [(lbmx.D2Q9, lbmx.BGK), (lbmx.D2Q9, lbmx.MRT), (lbmx.D3Q19, lbmx.BGK),
                                          (lbmx.D3Q19, lbmx.MRT)])
def test_brinkman_guo_forcing(lattice, coll):
    nz = 2 if lattice == lbmx.D3Q19 else 1
    om = 1.3
    base = mk(lattice, coll, lbmx.DESIGN_NONE, nz=nz, omega=om)
    br = mk(lattice, coll, lbmx.DESIGN_BRINKMAN, nz=nz, omega=om, alpha_max=0.8, q=0.1)
    c, _, w = lat(base)
    f = near_eq(c, w)
    # gamma = 1: alpha = 0, the base operator
    assert np.allclose(br.collide_node(f, 1.0)[0], base.collide_node(f, 1.0)[0], atol=1e-16)
    for g in (0.0, 0.3):
        f = near_eq(c, w)
        alpha = 0.8 * bp_closed(0.1, g)
        rho, j = f.sum(), c.T @ f
        v = j / (1 + alpha / 2)  # velocity with half the force, F = -alpha v
        F = -alpha * v
        out, _ = br.collide_node(f, g)
        assert abs(out.sum() - rho) < 1e-15
        assert np.allclose(c.T @ out, j + F, atol=1e-16), "momentum change equals the force"
        if coll == lbmx.BGK:  # Guo: second moment (1-w)Pi + w Pi_eq(v) + (1 - w/2)(vF + Fv)
            D = 3 if lattice == lbmx.D3Q19 else 2
            I = np.eye(3)
            I[D:, D:] = 0
            Pi = np.einsum("q,qa,qb->ab", f, c, c)
            Peq = rho / 3 * I + np.outer(v, v)
            ex = (1 - om) * Pi + om * Peq + (1 - om / 2) * (np.outer(v, F) + np.outer(F, v))
            assert np.allclose(np.einsum("q,qa,qb->ab", out, c, c), ex, atol=1e-15)


def test_darcy_mrt_forcing():
    """DarcyMRT2d (src/dynamics/darcyMRT2d.hh:33-59): the MRT collision with the unshifted
    macros u = j, plus a forcing term for F = -alpha_max BP_q(gamma) rho u + body force. Pinned
    in moment space against the textbook Guo source S = w (3 c.F + 9 (c.u)(c.F) - 3 u.F): the
    forcing changes mass by 0, momentum by F, and each other Lallemand-Luo row by (1 - s_k/2)
    times the source moment, except the energy-flux rows q, which receive the full source."""
    om, amax, q, body = 1.3, 0.4, 0.1, (2e-4, -1e-4, 0.0)
    d = lbmx.make_desc(4, 4, lattice=lbmx.D2Q9, collision=lbmx.MRT, design=lbmx.DESIGN_DARCY, omega=om,
                       interp_q=q, alpha_max=amax, body_force=body)
    o = oracle.Oracle(d)
    base = mk(lbmx.D2Q9, lbmx.MRT, omega=om)
    c, _, w = lat(o)
    rows = basis_d2q9(c)
    s = {k: _rate(v, om) for k, v in RATES_D2Q9.items()}
    for g in (0.0, 0.35, 1.0):
        f = near_eq(c, w)
        rho, j = f.sum(), c.T @ f
        add = o.collide_node(f, g)[0] - base.collide_node(f, g)[0]
        F = -amax * bp_closed(q, g) * rho * j + np.array(body)
        guo = w * (3 * (c @ F) + 9 * (c @ j) * (c @ F) - 3 * (j @ F))
        # add is a difference of two O(0.1) outputs: roundoff of a few 1e-17
        assert abs(add.sum()) < 5e-16
        assert np.allclose(c.T @ add, F, rtol=0, atol=5e-16)
        for name, row in rows.items():
            fac = 1.0 if name in ("qx", "qy") else 1.0 - s[name] / 2
            assert abs(row @ add - fac * (row @ guo)) < 5e-16, (g, name)
    with pytest.raises(lbmx.LbmxError):  # the reference has no reverse for this operator
        o.gradient(3)


# ================================================================ thermal (D2Q5 MRT, ramp)
def test_thermal_mrt_and_ramp():
    D, ck = 0.02, 0.1
    o = mk(lbmx.D2Q9_D2Q5, lbmx.MRT, lbmx.DESIGN_BRINKMAN, diff=D, ck=ck)
    c, _, w = lat(o)
    ct, wt = c[9:], w[9:]
    rows = basis_d2q5(ct)
    names = list(rows)
    M = np.array([rows[n] for n in names])
    assert np.allclose(M @ M.T, np.diag(np.diag(M @ M.T)))
    for g in (1.0, 0.0, 0.5):
        k_rel = (g * (ck * 2 - 1) + 1) / (ck * (1 + g))  # RAMP with p = 1: k(1) = 1, k(0) = 1/ck
        wT = 1 / (5 * D * k_rel + 0.5)                   # D2Q5 model with cs^2 = 1/5
        f = np.concatenate([near_eq(c[:9], w[:9]), 0.7 * wt + 0.01 * RNG.standard_normal(5) * wt])
        out, mac = o.collide_node(f, g)
        T, j = f[9:].sum(), c[:9].T @ f[:9]
        assert abs(out[9:].sum() - T) < 1e-15 and abs(mac[4] - T) < 1e-15
        m, mo = M @ f[9:], M @ out[9:]
        meq = dict(T=T, qx=T * j[0], qy=T * j[1], e=-2 * T, pxx=0.0)
        s = dict(T=0, qx=wT, qy=wT, e=1.5, pxx=1.5)  # thermalMRT2d.hh:105-106
        for k, n in enumerate(names):
            assert abs(mo[k] - (m[k] - s[n] * (m[k] - meq[n]))) < 1e-15, (g, n)


# ================================================================ reverse operators: FD
REV_CASES = [(lbmx.D2Q9, lbmx.BGK, lbmx.DESIGN_PBB), (lbmx.D2Q9, lbmx.MRT, lbmx.DESIGN_PBB),
             (lbmx.D2Q9, lbmx.BGK, lbmx.DESIGN_BRINKMAN), (lbmx.D2Q9, lbmx.MRT, lbmx.DESIGN_BRINKMAN),
             (lbmx.D2Q9, lbmx.MRT, lbmx.DESIGN_NONE), (lbmx.D3Q19, lbmx.BGK, lbmx.DESIGN_BRINKMAN),
             (lbmx.D3Q19, lbmx.MRT, lbmx.DESIGN_BRINKMAN), (lbmx.D3Q19, lbmx.MRT, lbmx.DESIGN_PBB),
             (lbmx.D2Q9_D2Q5, lbmx.MRT, lbmx.DESIGN_BRINKMAN), (lbmx.D2Q9_D2Q5, lbmx.BGK, lbmx.DESIGN_PBB),
             (lbmx.D2Q9, lbmx.CASCADED, lbmx.DESIGN_NONE), (lbmx.D2Q9, lbmx.CASCADED, lbmx.DESIGN_PBB)]


@pytest.mark.parametrize("lattice,coll,design", REV_CASES)
def test_reverse_collision_is_the_transpose(lattice, coll, design):
    nz = 2 if lattice == lbmx.D3Q19 else 1
    o = mk(lattice, coll, design, nz=nz, alpha_max=0.5 if design == lbmx.DESIGN_PBB else 0.3)
    c, _, w = lat(o)
    Qf = 19 if lattice == lbmx.D3Q19 else 9
    h = 1e-6
    for g in (0.35, 0.8):
        f = near_eq(c[:Qf], w[:Qf])
        if len(w) > Qf:
            f = np.concatenate([f, 0.6 * w[Qf:] * (1 + 0.05 * RNG.standard_normal(len(w) - Qf))])
        lam = RNG.standard_normal(len(w))
        df = RNG.standard_normal(len(w)) * w
        lpre, dg = o.reverse_collide_node(f, g, lam)
        fd = (o.collide_node(f + h * df, g)[0] - o.collide_node(f - h * df, g)[0]) / (2 * h)
        lhs, rhs = lam @ fd, lpre @ df
        assert abs(lhs - rhs) <= 1e-8 * max(1.0, abs(lhs)), (lhs, rhs)
        fdg = (o.collide_node(f, g + h)[0] - o.collide_node(f, g - h)[0]) / (2 * h)
        assert abs(lam @ fdg - dg) <= 1e-7 * max(1.0, abs(dg)), (lam @ fdg, dg)


# ================================================================ streaming
@pytest.mark.parametrize("lattice,shape", [(lbmx.D2Q9, (5, 4, 1)), (lbmx.D3Q19, (4, 3, 3)), (lbmx.D2Q9_D2Q5, (4, 3, 1))])
def test_stream_brute_force_and_transpose(lattice, shape):
    nx, ny, nz = shape
    o = mk(lattice, nx=nx, ny=ny, nz=nz)
    c, op, w = lat(o)
    Q = len(w)
    fp = RNG.standard_normal((nz, ny, nx, Q))
    out = o.stream(fp.ravel()).reshape(nz, ny, nx, Q)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for q in range(Q):
                    si, sj, sk = i - int(c[q, 0]), j - int(c[q, 1]), k - int(c[q, 2])
                    inside = 0 <= si < nx and 0 <= sj < ny and 0 <= sk < nz
                    exp = fp[sk, sj, si, q] if inside else fp[k, j, i, op[q]]
                    assert out[k, j, i, q] == exp
    lam = RNG.standard_normal(fp.size)
    x = RNG.standard_normal(fp.size)
    assert abs(lam @ o.stream(x) - o.adjoint_stream(lam) @ x) < 1e-12


# ================================================================ boundary conditions
def _channel(lattice, nx=7, ny=6, nz=1, **kw):
    import workloads
    o = mk(lattice, lbmx.BGK, lbmx.DESIGN_NONE, nx=nx, ny=ny, nz=nz)
    if lattice == lbmx.D3Q19:
        workloads.duct3d(o, nx, ny, nz, **kw)
    elif lattice == lbmx.D2Q9_D2Q5:
        workloads.thermal_channel2d(o, nx, ny, **kw)
    else:
        workloads.channel2d(o, nx, ny, **kw)
    return o


@pytest.mark.parametrize("lattice", [lbmx.D2Q9, lbmx.D3Q19, lbmx.D2Q9_D2Q5])
def test_boundary_conditions_prescribe_values(lattice):
    nz = 5 if lattice == lbmx.D3Q19 else 1
    o = _channel(lattice, nz=nz)
    c, op, w = lat(o)
    Q = len(w)
    Qf = 19 if lattice == lbmx.D3Q19 else 9
    nx, ny = o.desc.nx, o.desc.ny
    for piece, (typ, nrm, box) in enumerate(o.boundaries()):
        field = np.empty((nz, ny, nx, Q))
        for idx in np.ndindex(nz, ny, nx):
            ff = near_eq(c[:Qf], w[:Qf], 0.003)
            if Q > Qf:
                ff = np.concatenate([ff, 0.5 * w[Qf:] * (1 + 0.02 * RNG.standard_normal(Q - Qf))])
            field[idx] = ff
        t = 3
        out = o.apply_bc(piece, t, field.ravel()).reshape(field.shape)
        x0, x1, y0, y1, z0, z1 = box
        n = np.array(nrm, dtype=float)
        for k in range(z0, z1 + 1):
            for j in range(y0, y1 + 1):
                for i in range(x0, x1 + 1):
                    fi, fo = field[k, j, i], out[k, j, i]
                    if typ == lbmx.BC_BOUNCEBACK:
                        cn = c[:Qf] @ n
                        exp = fi[:Qf].copy()
                        sel = (cn <= 0) & (np.arange(Qf) > 0)
                        exp[sel] = fi[op[:Qf]][sel]
                        if lattice == lbmx.D3Q19 and tuple(nrm) == (-1, -1, -1):
                            exp = _corner_as_coded(fi[:Qf])
                        assert np.array_equal(fo[:Qf], exp), (piece, nrm)
                    elif typ in (lbmx.BC_ZOUHE_VELOCITY, lbmx.BC_HECHT_VELOCITY):
                        assert np.allclose(c[:Qf].T @ fo[:Qf], _inlet_velocity(o, piece, i, j, k, lattice), atol=1e-15)
                        known = (c[:Qf] @ n) >= 0
                        assert np.array_equal(fo[:Qf][known], fi[:Qf][known]), "only unknowns are written"
                    elif typ in (lbmx.BC_ZOUHE_PRESSURE, lbmx.BC_HECHT_PRESSURE):
                        assert abs(fo[:Qf].sum() - 1.0) < 1e-15
                        u = c[:Qf].T @ fo[:Qf]
                        assert np.allclose(u[1:], 0, atol=1e-15), "outlet velocity is normal"
                    elif typ == lbmx.BC_THERMAL_DIRICHLET:
                        T_set = 1.0 if nrm[0] == -1 else 0.0  # hot inlet, cold walls
                        assert abs(fo[Qf:].sum() - T_set) < 1e-15
                        assert np.array_equal(fo[:Qf], fi[:Qf])


def _corner_as_coded(fi):
    """The D3Q19 corner at (x0, y0, z0) as the reference codes it, in place on one node
    (bounceback3d.cc:434-451): six copies, then three swaps. The swap of 9 and 18 comes after
    18 = 9, so both end as the incoming f9, and f13 is never written; the generic rule would
    give f9 = f18 and f13 = f4."""
    f = np.array(fi, dtype=float)
    for dst, src in ((10, 1), (11, 2), (12, 3), (15, 6), (17, 8), (18, 9)):
        f[dst] = f[src]
    for a, b in ((5, 14), (7, 16), (9, 18)):
        f[a], f[b] = f[b], f[a]
    return f


def test_d3q19_corner_quirk_differs_from_the_rule_only_where_coded():
    """The as-coded corner map is a permutation-with-copy that differs from the generic
    bounce-back rule in exactly f9 and f13."""
    o = mk(lbmx.D3Q19)
    c, op, w = lat(o)
    fi = RNG.standard_normal(19)
    cn = c @ np.array([-1.0, -1.0, -1.0])
    rule = fi.copy()
    sel = (cn <= 0) & (np.arange(19) > 0)
    rule[sel] = fi[op][sel]
    quirk = _corner_as_coded(fi)
    assert list(np.nonzero(quirk != rule)[0]) == [9, 13]
    assert quirk[9] == quirk[18] == fi[9] and quirk[13] == fi[13]


# ---------------------------------------------------------------- Zou-He / Hecht closures
# Every velocity, pressure and Neumann piece the oracle accepts: (lattice, collision, box
# (as a function of the grid), bc type, value, direction). Zou-He (1997) and Hecht & Harting
# (2010) close the unknown populations of a face node by the prescribed moments plus the
# bounce-back of the non-equilibrium part along the normal: f_q - f_opp(q) equals its
# equilibrium difference 6 w_q c_q.j for the inward normal direction and, summed per pair,
# for the inward diagonal directions of each tangential axis. j is the node's momentum after
# the condition (incompressible: u = j; standard: u = j / rho, so 6 w rho c.u = 6 w c.j).
def _closure_cases():
    G2, G3 = (7, 6, 1), (5, 5, 5)
    X2, Y2 = G2[0] - 1, G2[1] - 1
    X3 = G3[0] - 1
    west2 = (0, 0, 1, Y2 - 1, 0, 0)
    east2 = (X2, X2, 1, Y2 - 1, 0, 0)
    north2 = (1, X2 - 1, Y2, Y2, 0, 0)
    west3 = (0, 0, 1, 3, 1, 3)
    east3 = (X3, X3, 1, 3, 1, 3)
    D2, D3 = lbmx.D2Q9, lbmx.D3Q19
    ZV, ZP, ZN = lbmx.BC_ZOUHE_VELOCITY, lbmx.BC_ZOUHE_PRESSURE, lbmx.BC_ZOUHE_NEUMANN
    HV, HP = lbmx.BC_HECHT_VELOCITY, lbmx.BC_HECHT_PRESSURE
    return [
        ("zh-vel-W-inc", D2, lbmx.BGK, G2, west2, ZV, 0.05, (1.0, 0.4, 0.0)),
        ("zh-vel-N-inc", D2, lbmx.BGK, G2, north2, ZV, 0.05, (0.3, -1.0, 0.0)),
        ("zh-vel-E-inc", D2, lbmx.BGK, G2, east2, ZV, 0.05, (-1.0, 0.25, 0.0)),
        ("zh-vel-W-std", D2, lbmx.BGK_STANDARD, G2, west2, ZV, 0.05, (1.0, 0.4, 0.0)),
        ("zh-prs-E-inc", D2, lbmx.BGK, G2, east2, ZP, 1.013, (1.0, 0.0, 0.0)),
        ("zh-prs-W-inc", D2, lbmx.BGK, G2, west2, ZP, 1.013, (1.0, 0.0, 0.0)),
        ("zh-prs-E-std", D2, lbmx.BGK_STANDARD, G2, east2, ZP, 1.013, (1.0, 0.0, 0.0)),
        ("zh-prs-W-std", D2, lbmx.BGK_STANDARD, G2, west2, ZP, 1.013, (1.0, 0.0, 0.0)),
        ("zh-neu-E", D2, lbmx.BGK, G2, east2, ZN, 0.0, (1.0, 0.0, 0.0)),
        ("zh-neu-W", D2, lbmx.BGK, G2, west2, ZN, 0.0, (1.0, 0.0, 0.0)),
        ("hecht-vel-W-inc", D3, lbmx.MRT, G3, west3, HV, 0.04, (1.0, 0.3, -0.2)),
        ("hecht-vel-W-std", D3, lbmx.MRT_STANDARD, G3, west3, HV, 0.04, (1.0, 0.3, -0.2)),
        ("hecht-prs-E-inc", D3, lbmx.MRT, G3, east3, HP, 1.013, (1.0, 0.0, 0.0)),
        ("hecht-prs-E-std", D3, lbmx.MRT_STANDARD, G3, east3, HP, 1.013, (1.0, 0.0, 0.0)),
    ]


def _closure_residuals(fo, c, w, op, n, typ, std, value, direction, m_neumann):
    """Residuals of the closure's defining equations at one node, as a function of its
    populations after the condition: prescribed moments, then the non-equilibrium
    bounce-back along the normal and per tangential axis."""
    D = 3 if len(w) == 19 else 2
    rho = fo.sum()
    j = c.T @ fo
    unknown = (c @ n) < 0
    res = []
    if typ in (lbmx.BC_ZOUHE_VELOCITY, lbmx.BC_HECHT_VELOCITY):
        u = j / rho if std else j
        res += list(u[:D] - value * np.asarray(direction)[:D])
    else:
        tang = [a for a in range(D) if n[a] == 0]
        res += [j[a] for a in tang]                  # tangential velocity 0
        if typ == lbmx.BC_ZOUHE_NEUMANN:
            res.append(j[0] - m_neumann)             # extrapolated normal velocity
        else:
            res.append(rho - value)                  # prescribed density
    inward = -n
    qn = int(np.nonzero(np.all(c == inward, axis=1))[0][0])
    res.append((fo[qn] - fo[op[qn]]) - 6 * w[qn] * (c[qn] @ j))
    for a in range(D):
        if n[a] != 0:
            continue
        S = np.nonzero(unknown & (c[:, a] != 0))[0]
        res.append(sum(fo[q] - fo[op[q]] for q in S) - sum(6 * w[q] * (c[q] @ j) for q in S))
    return np.array(res), unknown


@pytest.mark.parametrize("case", _closure_cases(), ids=lambda cs: cs[0])
def test_boundary_closures_are_pinned(case):
    """Zou-He / Hecht / Zou-He Neumann pieces (ZouHe2d.hh:29-44,69-84,109-124,149-174,252-270,
    440-527; Hecht3d.hh:69-89,182-204): only the unknown populations change, the prescribed
    moments hold, and the non-equilibrium bounce-back holds along the normal and per tangential
    axis. The rank check shows these equations leave no unknown free, so any other closure
    with the right moments fails."""
    name, lattice, coll, (nx, ny, nz), box, typ, value, direction = case
    o = mk(lattice, coll, lbmx.DESIGN_NONE, nx=nx, ny=ny, nz=nz)
    o.add_boundary(lbmx.make_box(*box), lbmx.make_bc(typ, value, direction))
    pieces = o.boundaries()
    assert len(pieces) == 1
    _, nrm, pbox = pieces[0]
    assert tuple(pbox) == tuple(box)
    c, op, w = lat(o)
    n = np.array(nrm, float)
    std = coll in (lbmx.BGK_STANDARD, lbmx.MRT_STANDARD)
    field = np.empty((nz, ny, nx, len(w)))
    for idx in np.ndindex(nz, ny, nx):
        field[idx] = near_eq(c, w, 0.01)
    out = o.apply_bc(0, 3, field.ravel()).reshape(field.shape)
    x0, x1, y0, y1, z0, z1 = box
    for k in range(z0, z1 + 1):
        for jj in range(y0, y1 + 1):
            for i in range(x0, x1 + 1):
                fi, fo = field[k, jj, i], out[k, jj, i]
                m_neu = 0.0
                if typ == lbmx.BC_ZOUHE_NEUMANN:
                    step = -int(nrm[0])
                    m1 = c[:, 0] @ field[k, jj, i + step]
                    m2 = c[:, 0] @ field[k, jj, i + 2 * step]
                    m_neu = (4 * m1 - m2) / 3
                res, unknown = _closure_residuals(fo, c, w, op, n, typ, std, value, direction, m_neu)
                assert np.array_equal(fo[~unknown], fi[~unknown]), (name, "a known population changed")
                assert np.max(np.abs(res)) < 2e-15, (name, res)
                # the residuals fix every unknown: their Jacobian over the unknowns has full rank
                U = np.nonzero(unknown)[0]
                Jm = np.empty((len(res), len(U)))
                for col, q in enumerate(U):
                    h = 1e-6
                    fp, fm = fo.copy(), fo.copy()
                    fp[q] += h
                    fm[q] -= h
                    rp = _closure_residuals(fp, c, w, op, n, typ, std, value, direction, m_neu)[0]
                    rm = _closure_residuals(fm, c, w, op, n, typ, std, value, direction, m_neu)[0]
                    Jm[:, col] = (rp - rm) / (2 * h)
                sv = np.linalg.svd(Jm, compute_uv=False)
                assert len(U) == (5 if lattice == lbmx.D3Q19 else 3)
                assert sv[len(U) - 1] > 1e-3 * sv[0], (name, sv)
    # nothing outside the piece moves
    rest = np.ones((nz, ny, nx), bool)
    rest[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1] = False
    assert np.array_equal(out[rest], field[rest])


def test_closure_check_rejects_a_wrong_closure():
    """A closure with the right moments but another normal/diagonal split (the Zou-He west
    velocity with 0.1 u_x moved from f8 to f1 and f7, which keeps j) fails the check."""
    o = mk(lbmx.D2Q9)
    c, op, w = lat(o)
    fi = near_eq(c, w, 0.01)
    ux, uy = 0.05, 0.02
    g = fi.copy()
    g[1] = 0.5 * (g[6] - g[2]) + g[5] + ux / 6 - 0.5 * uy + 0.05 * ux
    g[7] = 0.5 * (g[2] - g[6]) + g[3] + ux / 6 + 0.5 * uy + 0.05 * ux
    g[8] = g[4] + 2 / 3 * ux - 0.1 * ux
    assert np.allclose(c.T @ g, [ux, uy, 0], atol=1e-15)
    res, _ = _closure_residuals(g, c, w, op, np.array([-1.0, 0, 0]), lbmx.BC_ZOUHE_VELOCITY, False, 1.0,
                                (ux, uy, 0), 0.0)
    assert np.max(np.abs(res)) > 1e-3


def _inlet_velocity(o, piece, i, j, k, lattice):
    """Inlet velocities of the workloads (closed forms)."""
    if lattice == lbmx.D3Q19:  # uniform Hecht inlet
        return np.array([0.04, 0, 0])
    ny = o.desc.ny
    R = (ny - 1) / 2
    return np.array([0.05 * (R * R - (j - R) ** 2) / (R * R), 0, 0])


@pytest.mark.parametrize("lattice", [lbmx.D2Q9, lbmx.D3Q19, lbmx.D2Q9_D2Q5])
def test_adjoint_boundary_conditions_are_transposes(lattice):
    nz = 5 if lattice == lbmx.D3Q19 else 1
    o = _channel(lattice, nz=nz)
    Q = len(lat(o)[2])
    N = o.desc.nx * o.desc.ny * nz * Q
    for piece in range(len(o.boundaries())):
        f = RNG.standard_normal(N)
        d = RNG.standard_normal(N)
        lam = RNG.standard_normal(N)
        Ld = o.apply_bc(piece, 2, f + d) - o.apply_bc(piece, 2, f)
        lhs = lam @ Ld
        rhs = o.apply_bc_adjoint(piece, 2, lam) @ d
        assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs)), piece


def _neumann_channel(thermal, nx=9, ny=6):
    import workloads
    if thermal:
        o = mk(lbmx.D2Q9_D2Q5, lbmx.BGK, lbmx.DESIGN_NONE, nx=nx, ny=ny)
        workloads.thermal_channel2d(o, nx, ny, thermal_outlet="neumann", outlet="neumann")
    else:
        o = mk(lbmx.D2Q9, lbmx.BGK, lbmx.DESIGN_NONE, nx=nx, ny=ny)
        workloads.channel2d(o, nx, ny, outlet="neumann")
    return o


@pytest.mark.parametrize("thermal", [False, True], ids=["flow", "temperature"])
def test_neumann_conditions_extrapolate(thermal):
    """Zou-He Neumann outlets (ZouHe2d.hh:440-527, thermalZouHe2d.hh:229-300): after the
    condition, the node's u_x (flow) or T (temperature) equals (4 m_1 - m_2)/3 of the two nodes
    inwards, the flow's u_y is 0, and only the unknown populations change."""
    o = _neumann_channel(thermal)
    c, op, w = lat(o)
    Q, Qf = len(w), 9
    nx, ny = o.desc.nx, o.desc.ny
    pieces = [(k, p) for k, p in enumerate(o.boundaries())
              if p[0] in (lbmx.BC_ZOUHE_NEUMANN, lbmx.BC_THERMAL_NEUMANN)]
    assert {p[0] for _, p in pieces} == ({lbmx.BC_ZOUHE_NEUMANN, lbmx.BC_THERMAL_NEUMANN} if thermal
                                         else {lbmx.BC_ZOUHE_NEUMANN})
    for piece, (typ, nrm, box) in pieces:
        field = np.empty((ny, nx, Q))
        for idx in np.ndindex(ny, nx):
            ff = near_eq(c[:Qf], w[:Qf], 0.02)
            if Q > Qf:
                ff = np.concatenate([ff, 0.5 * w[Qf:] * (1 + 0.2 * RNG.standard_normal(Q - Qf))])
            field[idx] = ff
        out = o.apply_bc(piece, 3, field.ravel()).reshape(field.shape)
        x0, x1, y0, y1 = box[:4]
        assert x0 == x1 == nx - 1 and (y0, y1) == (1, ny - 2)
        changed = np.zeros(Q, bool)
        for j in range(y0, y1 + 1):
            fi, fo = field[j, x0], out[j, x0]
            n1, n2 = field[j, x0 - 1], field[j, x0 - 2]
            if typ == lbmx.BC_ZOUHE_NEUMANN:
                m1, m2 = c[:Qf, 0] @ n1[:Qf], c[:Qf, 0] @ n2[:Qf]
                u = c[:Qf].T @ fo[:Qf]
                assert abs(u[0] - (4 * m1 - m2) / 3) < 1e-15
                assert abs(u[1]) < 1e-15
                known = (c[:Qf] @ np.array(nrm, float)) >= 0
                assert np.array_equal(fo[:Qf][known], fi[:Qf][known])
            else:
                m1, m2 = n1[Qf:].sum(), n2[Qf:].sum()
                assert abs(fo[Qf:].sum() - (4 * m1 - m2) / 3) < 1e-15
                assert np.array_equal(fo[:Qf], fi[:Qf])
            changed |= fo != fi
        # the unknowns: the three directions entering from the east (flow) or the one D2Q5 direction
        expect = np.zeros(Q, bool)
        if typ == lbmx.BC_ZOUHE_NEUMANN:
            expect[:Qf] = c[:Qf, 0] < 0
        else:
            expect[Qf:] = c[Qf:, 0] < 0
        assert np.array_equal(changed, expect)
        # nothing else in the field moves
        rest = np.ones((ny, nx), bool)
        rest[y0:y1 + 1, x0] = False
        assert np.array_equal(out[rest], field[rest])


@pytest.mark.parametrize("thermal", [False, True], ids=["flow", "temperature"])
def test_neumann_adjoints_are_transposes(thermal):
    """The dot test of the whole-field map, which includes the neighbour blocks."""
    o = _neumann_channel(thermal)
    Q = len(lat(o)[2])
    N = o.desc.nx * o.desc.ny * Q
    for piece, (typ, _, _) in enumerate(o.boundaries()):
        if typ not in (lbmx.BC_ZOUHE_NEUMANN, lbmx.BC_THERMAL_NEUMANN):
            continue
        f, d, lam = RNG.standard_normal(N), RNG.standard_normal(N), RNG.standard_normal(N)
        Ld = o.apply_bc(piece, 2, f + d) - o.apply_bc(piece, 2, f)
        lhs = lam @ Ld
        rhs = o.apply_bc_adjoint(piece, 2, lam) @ d
        assert abs(lhs - rhs) < 1e-11 * max(1, abs(lhs)), piece


def test_boundary_piece_split_2d():
    """Edges exclude corners; corners get their own diagonal orientation (geometry2d.cc:41-49)."""
    o = _channel(lbmx.D2Q9, nx=7, ny=6)
    pieces = o.boundaries()
    got = {(t, n[:2], b[:4]) for t, n, b in pieces}
    X, Y = 6, 5
    walls = {(0, (0, 1), (1, X - 1, Y, Y)), (0, (-1, 1), (0, 0, Y, Y)), (0, (1, 1), (X, X, Y, Y)),
             (0, (0, -1), (1, X - 1, 0, 0)), (0, (-1, -1), (0, 0, 0, 0)), (0, (1, -1), (X, X, 0, 0))}
    assert walls <= got
    assert (lbmx.BC_ZOUHE_VELOCITY, (-1, 0), (0, 0, 1, Y - 1)) in got
    assert (lbmx.BC_ZOUHE_PRESSURE, (1, 0), (X, X, 1, Y - 1)) in got
    assert len(pieces) == 8


# ================================================================ functionals
@pytest.mark.parametrize("kind", [lbmx.OBJ_PRESSURE_DROP, lbmx.OBJ_OUTFLOW, lbmx.OBJ_DISSIPATED_POWER,
                                  lbmx.OBJ_THROUGHFLOW])
def test_functional_closed_form_and_gradient(kind):
    nx, ny = 6, 5
    o = mk(lbmx.D2Q9, nx=nx, ny=ny)
    a = lbmx.make_box(0, 0, 1, ny - 2)
    b = lbmx.make_box(nx - 1, nx - 1, 1, ny - 2)
    o.set_objective(lbmx.make_objective(kind, a, b, 0.5))
    u = np.array([0.03, 0.01, 0.0])
    o.init_equilibrium(1.02, u)
    J = o.objective(o.f)
    p = 1.02 / 3 + 0.5 * (u @ u)
    # ThroughFlow2d: -(mean u_x over the inflow) - (mean u_x over the outflow), times timeConst
    exp = {lbmx.OBJ_PRESSURE_DROP: 0.0, lbmx.OBJ_OUTFLOW: -0.5 * 0.03,
           lbmx.OBJ_DISSIPATED_POWER: 0.0, lbmx.OBJ_THROUGHFLOW: -0.5 * 0.06}[kind]
    assert abs(J - exp) < 1e-15
    # non-uniform state: central differences of J
    f = o.f + 0.001 * RNG.standard_normal(o.f.size)
    g = o.objective_grad(f)
    for idx in RNG.choice(f.size, 12, replace=False):
        e = np.zeros_like(f)
        e[idx] = 1e-5
        fd = (o.objective(f + e) - o.objective(f - e)) / 2e-5
        assert abs(fd - g[idx]) < 1e-9
    # total-pressure flux on a box of one node: (rho/3 + u^2/2) u_x
    if kind == lbmx.OBJ_DISSIPATED_POWER:
        one = lbmx.make_box(2, 2, 2, 2)
        empty = lbmx.make_box(0, -1, 0, -1)
        o.set_objective(lbmx.make_objective(kind, one, empty, 1.0))
        o.init_equilibrium(1.02, u)
        assert abs(o.objective(o.f) - p * 0.03) < 1e-16


# ================================================================ end-to-end gradient: FD
FD_CASES = [
    (lbmx.D2Q9, lbmx.BGK, lbmx.DESIGN_PBB, (12, 8, 1), "channel"),
    (lbmx.D2Q9, lbmx.MRT, lbmx.DESIGN_BRINKMAN, (12, 8, 1), "channel"),
    (lbmx.D2Q9, lbmx.CASCADED, lbmx.DESIGN_PBB, (12, 8, 1), "channel"),
    (lbmx.D3Q19, lbmx.MRT, lbmx.DESIGN_BRINKMAN, (8, 6, 6), "duct"),
    (lbmx.D2Q9_D2Q5, lbmx.MRT, lbmx.DESIGN_BRINKMAN, (10, 7, 1), "thermal"),
    (lbmx.D2Q9, lbmx.MRT, lbmx.DESIGN_BRINKMAN, (12, 8, 1), "neumann"),
    (lbmx.D2Q9_D2Q5, lbmx.MRT, lbmx.DESIGN_PBB, (10, 7, 1), "thermal-neumann"),
]


@pytest.mark.parametrize("lattice,coll,design,shape,setup", FD_CASES)
def test_gradient_matches_central_differences(lattice, coll, design, shape, setup):
    import workloads
    nx, ny, nz = shape
    steps = 25
    o = mk(lattice, coll, design, nx=nx, ny=ny, nz=nz, omega=1.4,
           alpha_max=1.0 if design == lbmx.DESIGN_PBB else 0.2)
    if setup == "duct":
        workloads.duct3d(o, nx, ny, nz, time_const=1.0 / steps)
    elif setup == "thermal":
        workloads.thermal_channel2d(o, nx, ny, time_const=1.0 / steps)
    elif setup == "thermal-neumann":
        workloads.thermal_channel2d(o, nx, ny, time_const=1.0 / steps, thermal_outlet="neumann", outlet="neumann")
    elif setup == "neumann":
        workloads.channel2d(o, nx, ny, time_const=1.0 / steps, outlet="neumann")
    else:
        workloads.channel2d(o, nx, ny, time_const=1.0 / steps)
    g = workloads.design_field(nx, ny, nz, seed=5)
    o.set_design(g)
    thermal = setup.startswith("thermal")
    kw = dict(u0=(0.03, 0, 0), T0=0.5) if thermal else {}
    J, sens, lam0 = o.gradient(steps, **kw)
    # The reference's reverse loop runs while step > 0, so the first collision (of the initial
    # equilibrium) is never reverse-collided (applications/RobustPump2d/robustPump2d.cc:268-276).
    # Its design term vanishes from the rest state; from a moving initial state it is added
    # here from the final adjoint so that the completed gradient equals the derivative of J.
    o.init_equilibrium(1.0, kw.get("u0", (0, 0, 0)), kw.get("T0", 0.0))
    Q = o.Q
    step0 = np.array([o.reverse_collide_node(o.f[n * Q:(n + 1) * Q], g[n], lam0[n * Q:(n + 1) * Q])[1]
                      for n in range(nx * ny * nz)])
    if not thermal:
        assert np.max(np.abs(step0)) <= 1e-13 * np.max(np.abs(sens))
    full = sens + step0

    def Jof(gg):
        o.set_design(gg)
        o.init_equilibrium(1.0, kw.get("u0", (0, 0, 0)), kw.get("T0", 0.0))
        return o.forward(steps)

    h = 1e-5
    cells = RNG.choice(nx * ny * nz, 4, replace=False)
    for cidx in cells:
        gp, gm = g.ravel().copy(), g.ravel().copy()
        gp[cidx] += h
        gm[cidx] -= h
        fd = (Jof(gp) - Jof(gm)) / (2 * h)
        scale = np.max(np.abs(full))
        assert abs(fd - full[cidx]) <= 1e-6 * scale, (cidx, fd, full[cidx])
    o.set_design(g)
