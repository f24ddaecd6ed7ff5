"""ORACLE wrapper -- test infrastructure only.

ctypes access to liblbm_oracle.so (the plain-C restatement of the reference's hot path).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "lbm-topopt_amd"))
import lbmx  # noqa: E402  (struct definitions of include/lbmx.h)

LIB = os.path.join(HERE, "liblbm_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    P = C.POINTER
    dp = P(C.c_double)
    ip = P(C.c_int)
    vp = C.c_void_p
    sig = {
        "orc_create": ([P(lbmx.Desc)], vp),
        "orc_destroy": ([vp], None),
        "orc_add_boundary": ([vp, P(lbmx.Box), P(lbmx.BCDesc)], C.c_int),
        "orc_add_boundary_table": ([vp, P(lbmx.Box), P(lbmx.BCDesc), dp, C.c_longlong, C.c_int, dp, C.c_int], C.c_int),
        "orc_num_boundaries": ([vp], C.c_int),
        "orc_boundary_info": ([vp, C.c_int, ip, ip, ip], C.c_int),
        "orc_set_design": ([vp, dp], C.c_int),
        "orc_set_objective": ([vp, P(lbmx.Objective)], C.c_int),
        "orc_set_objective_terms": ([vp, P(lbmx.Term), C.c_int, C.c_double], C.c_int),
        "orc_num_dof": ([vp], C.c_int),
        "orc_set_threads": ([C.c_int], None),
        "orc_set_reference_reverse": ([vp, C.c_int], None),
        "orc_lattice": ([vp, ip, ip, ip, ip, ip, dp], None),
        "orc_collide_node": ([vp, dp, dp, dp, C.c_double], None),
        "orc_reverse_collide_node": ([vp, dp, C.c_double, dp, dp, dp], None),
        "orc_apply_bc": ([vp, C.c_int, C.c_int, dp], None),
        "orc_apply_bc_adjoint": ([vp, C.c_int, C.c_int, dp], None),
        "orc_stream": ([vp, dp, dp], None),
        "orc_adjoint_stream": ([vp, dp, dp], None),
        "orc_objective": ([vp, dp, C.c_int], C.c_double),
        "orc_objective_grad": ([vp, dp, C.c_int, dp], None),
        "orc_init_equilibrium": ([vp, dp, C.c_double, dp, C.c_double], None),
        "orc_forward": ([vp, dp, C.c_int, ip, dp], None),
        "orc_gradient": ([vp, C.c_int, C.c_double, dp, C.c_double, dp, dp, dp], C.c_int),
        "orc_macros": ([vp, dp, dp], None),
        "odf_denominator": ([C.c_int, C.c_int, C.c_int, C.c_int, ip, dp], None),
        "odf_density_forward": ([C.c_int, C.c_int, C.c_int, C.c_int, ip, dp, dp], None),
        "odf_density_transpose": ([C.c_int, C.c_int, C.c_int, C.c_int, ip, dp, dp], None),
        "odf_project": ([C.c_int, C.c_int, C.c_int, ip, C.c_double, C.c_double, dp, dp, dp], None),
        "odf_project_sens": ([C.c_int, C.c_int, C.c_int, ip, C.c_double, C.c_double, dp, dp, dp], None),
        "odf_map_to_full": ([C.c_int, C.c_int, C.c_int, ip, dp, dp, dp], None),
        "odf_extract": ([C.c_int, C.c_int, C.c_int, ip, dp, dp], None),
        "omma_create": ([C.c_int, C.c_int], vp),
        "omma_destroy": ([vp], None),
        "omma_set_a": ([vp, dp], None),
        "omma_set_asymptotes": ([vp, C.c_double, C.c_double, C.c_double], None),
        "omma_set_options": ([vp, C.c_int, C.c_int], None),
        "omma_update": ([vp, dp, dp, dp, dp, C.c_double], None),
        "omma_design_change": ([vp, dp], C.c_double),
        "omma_state": ([vp, dp, dp, dp, ip, ip, dp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """Same construction calls as lbmx.Solver, on the CPU restatement."""

    def __init__(self, desc, threads=1):
        self.lib = library()
        self.lib.orc_set_threads(threads)
        self.desc = desc
        self.h = self.lib.orc_create(C.byref(desc))
        if not self.h:
            raise ValueError("oracle: unsupported configuration")
        self.nodes = desc.nx * desc.ny * desc.nz
        self.Q = self.lib.orc_num_dof(self.h)
        self.t = 0
        self.f = None

    def __del__(self):
        try:
            if self.h:
                self.lib.orc_destroy(self.h)
        except Exception:
            pass

    def set_reference_reverse(self, on=True):
        """Cascaded reverse collision as the reference codes it (not the true transpose)."""
        self.lib.orc_set_reference_reverse(self.h, 1 if on else 0)

    def lattice(self):
        Q = C.c_int()
        ex = (C.c_int * 24)(); ey = (C.c_int * 24)(); ez = (C.c_int * 24)(); op = (C.c_int * 24)()
        w = (C.c_double * 24)()
        self.lib.orc_lattice(self.h, C.byref(Q), ex, ey, ez, op, w)
        n = Q.value
        return (np.array(ex[:n]), np.array(ey[:n]), np.array(ez[:n]), np.array(op[:n]),
                np.array(w[:n]))

    def add_boundary(self, box, bc):
        code = self.lib.orc_add_boundary(self.h, C.byref(box), C.byref(bc))
        if code:
            raise lbmx.LbmxError(code, "oracle add_boundary")

    def add_boundary_table(self, box, bc, profile=None, ncomp=1, time_factors=None):
        prof = None if profile is None else np.ascontiguousarray(profile, dtype=np.float64).ravel()
        tf = None if time_factors is None else np.ascontiguousarray(time_factors, dtype=np.float64).ravel()
        code = self.lib.orc_add_boundary_table(self.h, C.byref(box), C.byref(bc),
                                               None if prof is None else _d(prof), 0 if prof is None else prof.size,
                                               ncomp, None if tf is None else _d(tf), 0 if tf is None else tf.size)
        if code:
            raise lbmx.LbmxError(code, "oracle add_boundary_table")

    def boundaries(self):
        out = []
        for k in range(self.lib.orc_num_boundaries(self.h)):
            t = C.c_int(); n = (C.c_int * 3)(); b = (C.c_int * 6)()
            self.lib.orc_boundary_info(self.h, k, C.byref(t), n, b)
            out.append((t.value, tuple(n), tuple(b)))
        return out

    def set_design(self, gamma):
        g = np.ascontiguousarray(gamma, dtype=np.float64).ravel()
        assert g.size == self.nodes
        self.lib.orc_set_design(self.h, _d(g))

    def set_objective(self, obj):
        code = self.lib.orc_set_objective(self.h, C.byref(obj))
        if code:
            raise lbmx.LbmxError(code, "oracle set_objective")

    def set_objective_terms(self, terms, time_const=1.0):
        code = self.lib.orc_set_objective_terms(self.h, lbmx.make_terms(terms), len(terms), time_const)
        if code:
            raise lbmx.LbmxError(code, "oracle set_objective_terms")

    # node primitives
    def collide_node(self, fin, gamma):
        fin = np.ascontiguousarray(fin, dtype=np.float64)
        out = np.empty(self.Q); mac = np.empty(5)
        self.lib.orc_collide_node(self.h, _d(fin), _d(out), _d(mac), gamma)
        return out, mac

    def reverse_collide_node(self, fin, gamma, lam_post):
        fin = np.ascontiguousarray(fin, dtype=np.float64)
        lp = np.ascontiguousarray(lam_post, dtype=np.float64)
        out = np.empty(self.Q); dg = np.zeros(1)
        self.lib.orc_reverse_collide_node(self.h, _d(fin), gamma, _d(lp), _d(out), _d(dg))
        return out, dg[0]

    def apply_bc(self, piece, t, f):
        f = np.ascontiguousarray(f, dtype=np.float64).copy()
        self.lib.orc_apply_bc(self.h, piece, t, _d(f))
        return f

    def apply_bc_adjoint(self, piece, t, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64).copy()
        self.lib.orc_apply_bc_adjoint(self.h, piece, t, _d(lam))
        return lam

    def stream(self, fpost):
        fpost = np.ascontiguousarray(fpost, dtype=np.float64)
        out = np.empty_like(fpost)
        self.lib.orc_stream(self.h, _d(fpost), _d(out))
        return out

    def adjoint_stream(self, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        out = np.empty_like(lam)
        self.lib.orc_adjoint_stream(self.h, _d(lam), _d(out))
        return out

    def objective(self, f, t=0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        return self.lib.orc_objective(self.h, _d(f), t)

    def objective_grad(self, f, t=0):
        f = np.ascontiguousarray(f, dtype=np.float64)
        g = np.zeros_like(f)
        self.lib.orc_objective_grad(self.h, _d(f), t, _d(g))
        return g

    # loops
    def init_equilibrium(self, rho0=1.0, u0=(0.0, 0.0, 0.0), T0=0.0):
        self.f = np.empty(self.nodes * self.Q)
        u = np.asarray(u0, dtype=np.float64)
        self.lib.orc_init_equilibrium(self.h, _d(self.f), rho0, _d(u), T0)
        self.t = 0

    def forward(self, nsteps):
        t = C.c_int(self.t)
        J = np.zeros(1)
        self.lib.orc_forward(self.h, _d(self.f), nsteps, C.byref(t), _d(J))
        self.t = t.value
        return J[0]

    def gradient(self, nsteps, rho0=1.0, u0=(0.0, 0.0, 0.0), T0=0.0):
        u = np.asarray(u0, dtype=np.float64)
        J = np.zeros(1)
        sens = np.empty(self.nodes)
        adj = np.empty(self.nodes * self.Q)
        code = self.lib.orc_gradient(self.h, nsteps, rho0, _d(u), T0, _d(J), _d(sens), _d(adj))
        if code:
            raise lbmx.LbmxError(code, "oracle gradient")
        return J[0], sens, adj

    def macros(self, f=None):
        f = self.f if f is None else np.ascontiguousarray(f, dtype=np.float64)
        m = np.empty(self.nodes * 5)
        self.lib.orc_macros(self.h, _d(f), _d(m))
        return m.reshape(self.nodes, 5)


def _box6(b):
    """(x0, x1, y0, y1, z0, z1) from an lbmx.Box or a 6-sequence, as a C int array."""
    if hasattr(b, "x0"):
        b = (b.x0, b.x1, b.y0, b.y1, b.z0, b.z1)
    return (C.c_int * 6)(*[int(v) for v in b])


class FilterChain:
    """Design pipeline restated on the CPU: FullToDesignMapping2d + DensityFilter2d/3d and
    ProjectionFilter2d stages (src/topOpt/fullToDesignMapping2d.cc, densityFilter2d.cc,
    projectionFilter2d.cc, doubleProjectionFilter2d.cc). stages: (radius, project, beta, eta)."""

    def __init__(self, nx, ny, nz, design_box, filter_box=None, stages=((8, True, 1.0, 0.5),), background=None):
        self.lib = library()
        self.dims = (nx, ny, nz)
        self.n = nx * ny * nz
        self.dbox = _box6(design_box)
        self.fbox = _box6(filter_box if filter_box is not None else design_box)
        self.stages = [list(s) for s in stages]
        self.background = np.ones(self.n) if background is None else np.array(background, dtype=np.float64)
        d = list(self.dbox)
        self.nd = (d[1] - d[0] + 1) * (d[3] - d[2] + 1) * (d[5] - d[4] + 1)
        self.dens = [None] * len(self.stages)
        self.eta_fields = [None] * len(self.stages)   # per-node thresholds (variable projection)

    def denominator(self, stage=0):
        out = np.empty(self.n)
        self.lib.odf_denominator(*self.dims, self.stages[stage][0], self.fbox, _d(out))
        return out

    def density(self, x, R):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(self.n)
        self.lib.odf_density_forward(*self.dims, R, self.fbox, _d(x), _d(out))
        return out

    def density_transpose(self, s, R):
        s = np.ascontiguousarray(s, dtype=np.float64)
        out = np.empty(self.n)
        self.lib.odf_density_transpose(*self.dims, R, self.fbox, _d(s), _d(out))
        return out

    def _etaf(self, etaf):
        return None if etaf is None else _d(np.ascontiguousarray(etaf, dtype=np.float64))

    def project(self, x, beta, eta, etaf=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(x.size)
        self.lib.odf_project(*self.dims, self.fbox, beta, eta, self._etaf(etaf), _d(x), _d(out))
        return out

    def project_sens(self, x, beta, eta, s, etaf=None):
        """s * projDiff(x) inside the box (a new array)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        s = np.array(s, dtype=np.float64)
        self.lib.odf_project_sens(*self.dims, self.fbox, beta, eta, self._etaf(etaf), _d(x), _d(s))
        return s

    def map_to_full(self, design):
        design = np.ascontiguousarray(design, dtype=np.float64)
        out = np.empty(self.n)
        self.lib.odf_map_to_full(*self.dims, self.dbox, _d(design), _d(self.background), _d(out))
        return out

    def extract(self, full):
        full = np.ascontiguousarray(full, dtype=np.float64)
        out = np.empty(self.nd)
        self.lib.odf_extract(*self.dims, self.dbox, _d(full), _d(out))
        return out

    def design(self, design):
        """mapDesignToFull + filterDesign of every stage; caches the density-filtered fields."""
        x = self.map_to_full(design)
        for k, (R, proj, beta, eta) in enumerate(self.stages):
            x = self.density(x, R)
            self.dens[k] = x
            if proj:
                x = self.project(x, beta, eta, self.eta_fields[k])
        return x

    def sensitivities(self, full_sens):
        """filterSensitivities (stages reversed) + mapFullToDesign."""
        s = np.array(full_sens, dtype=np.float64)
        for k in range(len(self.stages) - 1, -1, -1):
            R, proj, beta, eta = self.stages[k]
            if proj:
                s = self.project_sens(self.dens[k], beta, eta, s, self.eta_fields[k])
            s = self.density_transpose(s, R)
        return self.extract(s)


class MMA:
    """Method of moving asymptotes restated on the CPU (src/topOpt/MMA.cc): n design
    variables, m constraints; update() overwrites x with the subproblem's solution."""

    def __init__(self, n, m, a=None):
        self.lib = library()
        self.n, self.m = n, m
        self.h = self.lib.omma_create(n, m)
        if a is not None:
            self.set_robust_constants(a)

    def __del__(self):
        try:
            if self.h:
                self.lib.omma_destroy(self.h)
        except Exception:
            pass

    def set_robust_constants(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.lib.omma_set_a(self.h, _d(a))

    def set_asymptotes(self, init, dec, inc):
        self.lib.omma_set_asymptotes(self.h, init, dec, inc)

    def set_options(self, constraint_modification=False, robust_asymptotes_type=0):
        self.lib.omma_set_options(self.h, int(constraint_modification), int(robust_asymptotes_type))

    def update(self, x, dfdx, gx, dgdx, movlim):
        x = np.array(x, dtype=np.float64)
        dfdx = np.ascontiguousarray(dfdx, dtype=np.float64)
        gx = np.ascontiguousarray(gx, dtype=np.float64)
        dgdx = np.ascontiguousarray(np.asarray(dgdx, dtype=np.float64).reshape(self.m, self.n))
        self.lib.omma_update(self.h, _d(x), _d(dfdx), _d(gx), _d(dgdx), movlim)
        return x

    def design_change(self, xnew):
        xnew = np.ascontiguousarray(xnew, dtype=np.float64)
        return self.lib.omma_design_change(self.h, _d(xnew))

    def state(self):
        L, U, lam = np.empty(self.n), np.empty(self.n), np.empty(self.m)
        k, it, res = C.c_int(), C.c_int(), C.c_double()
        self.lib.omma_state(self.h, _d(L), _d(U), _d(lam), C.byref(k), C.byref(it), C.byref(res))
        return {"L": L, "U": U, "lam": lam, "k": k.value, "newton_iters": it.value, "residual": res.value}
