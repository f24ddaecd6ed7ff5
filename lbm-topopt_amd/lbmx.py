"""Python host binding of the lbmx C ABI (``include/lbmx.h``) via ctypes.

The engine itself is C++/HIP (``csrc/``) built into ``liblbmx.so`` next to this file.
This module only marshals arguments; it fails loudly when the library is missing.
Array conventions follow the reference's single-rank PETSc DMDA order:
design/sensitivity ``(k*ny+j)*nx+i``; distributions array-of-structures ``[node][q]``.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbmx.so")

# ---- enums (include/lbmx.h) ----
D2Q9, D3Q19, D2Q9_D2Q5 = 0, 1, 2
BGK, MRT, CASCADED, MRT_STANDARD, BGK_STANDARD = 0, 1, 2, 3, 4
DESIGN_NONE, DESIGN_PBB, DESIGN_BRINKMAN, DESIGN_DARCY = 0, 1, 2, 3
FP64, FP32, FP16_SHIFTED, BF16_SHIFTED = 0, 1, 2, 3
BC_BOUNCEBACK, BC_ZOUHE_VELOCITY, BC_ZOUHE_PRESSURE, BC_ZOUHE_NEUMANN = 0, 1, 2, 3
BC_HECHT_VELOCITY, BC_HECHT_PRESSURE, BC_THERMAL_BOUNCEBACK, BC_THERMAL_DIRICHLET = 4, 5, 6, 7
BC_THERMAL_NEUMANN = 8
PROFILE_UNIFORM, PROFILE_PARABOLIC, PROFILE_PARABOLIC_DISK, PROFILE_PARABOLIC_DUCT, PROFILE_TABLE = 0, 1, 2, 3, 4
TIME_CONSTANT, TIME_SINE, TIME_TABLE = 0, 1, 2
OBJ_NONE, OBJ_PRESSURE_DROP, OBJ_OUTFLOW, OBJ_DISSIPATED_POWER, OBJ_HEAT_FLUX, OBJ_THROUGHFLOW = 0, 1, 2, 3, 4, 5
TERM_DENSITY, TERM_VELOCITY_X, TERM_TOTAL_PRESSURE_FLUX, TERM_HEAT_FLUX_X, TERM_TEMPERATURE = 0, 1, 2, 3, 4

LBMX_OK, ERR_ARG, ERR_SUP, ERR_MEM, ERR_STATE, ERR_DEVICE = 0, 62, 56, 55, 73, 76
CHECKPOINT_DYNAMIC, CHECKPOINT_STATIC, CHECKPOINT_BINOMIAL = 0, 1, 2

Q_OF = {D2Q9: 9, D3Q19: 19, D2Q9_D2Q5: 14}


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lattice", C.c_int), ("collision", C.c_int), ("design", C.c_int),
                ("precision", C.c_int), ("reserved0", C.c_int),
                ("omega", C.c_double), ("interp_q", C.c_double), ("alpha_max", C.c_double),
                ("body_force", C.c_double * 3),
                ("thermal_diffusivity", C.c_double), ("ramp_p", C.c_double), ("ramp_ck", C.c_double)]


class Box(C.Structure):
    _fields_ = [("x0", C.c_int), ("x1", C.c_int), ("y0", C.c_int), ("y1", C.c_int),
                ("z0", C.c_int), ("z1", C.c_int)]


class BCDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("profile", C.c_int), ("time_mode", C.c_int), ("reserved0", C.c_int),
                ("value", C.c_double), ("dir", C.c_double * 3), ("center", C.c_double * 2),
                ("radius", C.c_double * 2), ("t_period", C.c_double), ("t_amp", C.c_double),
                ("t_offset", C.c_double), ("t_phase", C.c_double)]


class Objective(C.Structure):
    _fields_ = [("kind", C.c_int), ("reserved0", C.c_int), ("time_const", C.c_double),
                ("box_a", Box), ("box_b", Box)]


class Term(C.Structure):
    _fields_ = [("kind", C.c_int), ("reserved0", C.c_int), ("weight", C.c_double), ("box", Box)]


class FilterDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("design_box", Box), ("filter_box", Box), ("nstages", C.c_int),
                ("radius", C.c_int * 2), ("project", C.c_int * 2),
                ("beta", C.c_double * 2), ("eta", C.c_double * 2)]


FILTER_MAX_RADIUS = 16


class Stats(C.Structure):
    _fields_ = [("fwd_kernel_ms", C.c_double), ("adj_kernel_ms", C.c_double),
                ("forward_phase_ms", C.c_double), ("reverse_phase_ms", C.c_double),
                ("fwd_sweeps", C.c_longlong), ("adj_sweeps", C.c_longlong),
                ("checkpoint_copies", C.c_longlong), ("recompute_sweeps", C.c_longlong)]


def make_desc(nx, ny, nz=1, lattice=D2Q9, collision=BGK, design=DESIGN_PBB, precision=FP64,
              omega=1.25, interp_q=0.1, alpha_max=1.0, body_force=(0.0, 0.0, 0.0),
              thermal_diffusivity=0.01, ramp_p=1.0, ramp_ck=0.1):
    d = Desc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.lattice, d.collision, d.design, d.precision = lattice, collision, design, precision
    d.omega, d.interp_q, d.alpha_max = omega, interp_q, alpha_max
    for a in range(3):
        d.body_force[a] = body_force[a]
    d.thermal_diffusivity, d.ramp_p, d.ramp_ck = thermal_diffusivity, ramp_p, ramp_ck
    return d


def make_box(x0, x1, y0, y1, z0=0, z1=0):
    return Box(x0, x1, y0, y1, z0, z1)


def make_bc(type, value=0.0, dir=(1.0, 0.0, 0.0), profile=PROFILE_UNIFORM, center=(0.0, 0.0),
            radius=(1.0, 1.0), time_mode=TIME_CONSTANT, t_period=1.0, t_amp=0.0, t_offset=1.0,
            t_phase=0.0):
    b = BCDesc()
    b.type, b.profile, b.time_mode = type, profile, time_mode
    b.value = value
    for a in range(3):
        b.dir[a] = dir[a]
    b.center[0], b.center[1] = center
    b.radius[0], b.radius[1] = radius
    b.t_period, b.t_amp, b.t_offset, b.t_phase = t_period, t_amp, t_offset, t_phase
    return b


def make_objective(kind, box_a, box_b=None, time_const=1.0):
    o = Objective()
    o.kind, o.time_const = kind, time_const
    o.box_a = box_a
    o.box_b = box_b if box_b is not None else Box(0, -1, 0, -1, 0, -1)
    return o


def make_terms(terms):
    """(kind, weight, box) tuples as a Term array."""
    arr = (Term * max(1, len(terms)))()
    for i, (kind, weight, box) in enumerate(terms):
        arr[i].kind, arr[i].weight, arr[i].box = kind, weight, box
    return arr


_lib = None


def library():
    """Load liblbmx.so (the HIP engine). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"lbmx engine library not built: {LIB_PATH} (run __graft_entry__.build())")
    # One HIP runtime per process. The torch wheel ships its own HIP runtime and loads it by
    # path; loaded after the engine, it would be a second runtime beside the system one, fail
    # to open the GPU, and device pointers could not pass between them. Importing torch first
    # makes the engine bind to torch's runtime (same soname), so torch tensors and the engine
    # share device memory.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    dp = P(C.c_double)
    sig = {
        "lbmx_create": [P(Desc), P(C.c_void_p)],
        "lbmx_destroy": [C.c_void_p],
        "lbmx_set_stream": [C.c_void_p, C.c_void_p],
        "lbmx_add_boundary": [C.c_void_p, P(Box), P(BCDesc)],
        "lbmx_add_boundary_table": [C.c_void_p, P(Box), P(BCDesc), dp, C.c_size_t, C.c_int, dp, C.c_size_t],
        "lbmx_set_design": [C.c_void_p, dp, C.c_size_t],
        "lbmx_get_design": [C.c_void_p, dp, C.c_size_t],
        "lbmx_set_objective": [C.c_void_p, P(Objective)],
        "lbmx_set_objective_terms": [C.c_void_p, P(Term), C.c_int, C.c_double],
        "lbmx_init_equilibrium": [C.c_void_p, C.c_double, dp, C.c_double],
        "lbmx_init_equilibrium_field": [C.c_void_p, dp, C.c_size_t],
        "lbmx_forward": [C.c_void_p, C.c_int, dp],
        "lbmx_get_state": [C.c_void_p, dp, C.c_size_t],
        "lbmx_set_state": [C.c_void_p, dp, C.c_size_t, C.c_int],
        "lbmx_get_macros": [C.c_void_p, dp, C.c_size_t],
        "lbmx_get_timestep": [C.c_void_p, P(C.c_int)],
        "lbmx_checkpoint_config": [C.c_void_p, C.c_int],
        "lbmx_gradient": [C.c_void_p, C.c_int, C.c_double, dp, C.c_double, dp, dp, C.c_size_t],
        "lbmx_get_sensitivity": [C.c_void_p, dp, C.c_size_t],
        "lbmx_get_adjoint": [C.c_void_p, dp, C.c_size_t],
        "lbmx_adjoint_reset": [C.c_void_p, C.c_int],
        "lbmx_adjoint_step": [C.c_void_p, C.c_int],
        "lbmx_adjoint_set_timestep": [C.c_void_p, C.c_int],
        "lbmx_set_design_device": [C.c_void_p, dp, C.c_size_t],
        "lbmx_get_sensitivity_device": [C.c_void_p, P(dp), P(C.c_size_t)],
        "lbmx_get_stats": [C.c_void_p, P(Stats)],
        "lbmx_device_bytes": [C.c_void_p, P(C.c_size_t)],
        "lbmx_time_sweeps": [C.c_void_p, C.c_int, C.c_int, dp],
        "lbmx_stream_copy": [C.c_size_t, C.c_int, dp],
        "lbmx_stream_copy_planes": [C.c_int, C.c_size_t, C.c_int, dp],
        "lbmx_stream_copy_planes16": [C.c_int, C.c_size_t, C.c_int, dp],
        "lbmx_device_count": [P(C.c_int)],
        "lbmx_batch_create": [P(C.c_void_p), C.c_int, P(C.c_void_p)],
        "lbmx_batch_destroy": [C.c_void_p],
        "lbmx_batch_checkpoint_config": [C.c_void_p, C.c_int],
        "lbmx_batch_set_tile_rows": [C.c_void_p, C.c_int],
        "lbmx_batch_set_tile": [C.c_void_p, C.c_int, C.c_int],
        "lbmx_design_change_noalloc": [dp, dp, C.c_longlong, dp, C.c_void_p],
        "lbmx_batch_get_tile": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "lbmx_batch_set_graph": [C.c_void_p, C.c_int],
        "lbmx_batch_autotune": [C.c_void_p, C.c_int, P(C.c_int), dp],
        "lbmx_batch_gradient": [C.c_void_p, C.c_int, C.c_double, dp, C.c_double, dp, P(dp)],
        "lbmx_batch_get_stats": [C.c_void_p, P(Stats)],
        "lbmx_batch_device_bytes": [C.c_void_p, P(C.c_size_t)],
        "lbmx_describe_boundaries": [C.c_void_p, P(C.c_int), C.c_int, P(C.c_int)],
        "lbmx_checkpoint_simulate": [C.c_int, C.c_int, P(C.c_longlong), P(C.c_int)],
        "lbmx_checkpoint_simulate_mode": [C.c_int, C.c_int, C.c_int, P(C.c_longlong), P(C.c_int)],
        "lbmx_checkpoint_mode": [C.c_void_p, C.c_int],
        "lbmx_batch_checkpoint_mode": [C.c_void_p, C.c_int],
        "lbmx_filter_create": [P(FilterDesc), P(C.c_void_p)],
        "lbmx_filter_destroy": [C.c_void_p],
        "lbmx_filter_set_projection": [C.c_void_p, C.c_int, C.c_double, C.c_double],
        "lbmx_filter_increase_beta": [C.c_void_p, C.c_double],
        "lbmx_filter_set_eta_field": [C.c_void_p, C.c_int, dp, C.c_size_t],
        "lbmx_filter_set_background": [C.c_void_p, dp, C.c_size_t],
        "lbmx_filter_design": [C.c_void_p, dp, C.c_size_t, dp, C.c_size_t],
        "lbmx_filter_design_to_solver": [C.c_void_p, dp, C.c_size_t, C.c_void_p],
        "lbmx_filter_sensitivities": [C.c_void_p, dp, C.c_size_t, dp, C.c_size_t],
        "lbmx_filter_solver_sensitivities": [C.c_void_p, C.c_void_p, dp, C.c_size_t],
        "lbmx_filter_design_device": [C.c_void_p, dp, C.c_size_t, dp, C.c_size_t],
        "lbmx_filter_design_to_solver_device": [C.c_void_p, dp, C.c_size_t, C.c_void_p],
        "lbmx_filter_sensitivities_device": [C.c_void_p, dp, C.c_size_t, dp, C.c_size_t],
        "lbmx_filter_solver_sensitivities_device": [C.c_void_p, C.c_void_p, dp, C.c_size_t],
        "lbmx_filter_get_field": [C.c_void_p, C.c_int, dp, C.c_size_t],
        "lbmx_filter_timing": [C.c_void_p, dp],
        "lbmx_vec_write_binary": [C.c_char_p, dp, C.c_size_t],
        "lbmx_vec_load_binary": [C.c_char_p, dp, C.c_size_t],
        "lbmx_write_vts": [C.c_void_p, C.c_char_p],
        "lbmx_dmda_global_to_natural": [P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), C.c_int, dp, dp],
        "lbmx_dmda_natural_to_global": [P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), C.c_int, dp, dp],
        "lbmx_mma_create": [C.c_longlong, C.c_int, P(C.c_void_p)],
        "lbmx_mma_destroy": [C.c_void_p],
        "lbmx_mma_set_robust_constants": [C.c_void_p, dp],
        "lbmx_mma_set_asymptotes": [C.c_void_p, C.c_double, C.c_double, C.c_double],
        "lbmx_mma_set_options": [C.c_void_p, C.c_int, C.c_int],
        "lbmx_mma_update": [C.c_void_p, dp, dp, dp, P(dp), C.c_double],
        "lbmx_mma_design_change": [C.c_void_p, dp, dp],
        "lbmx_mma_update_device": [C.c_void_p, dp, dp, dp, P(dp), C.c_double],
        "lbmx_mma_design_change_device": [C.c_void_p, dp, dp],
        "lbmx_mma_get_state": [C.c_void_p, dp, dp, dp, P(C.c_int), dp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.lbmx_last_error.argtypes = []
    lib.lbmx_last_error.restype = C.c_char_p
    _lib = lib
    return lib


EXPORTED = ["lbmx_create", "lbmx_destroy", "lbmx_set_stream", "lbmx_last_error", "lbmx_add_boundary",
            "lbmx_set_design", "lbmx_get_design", "lbmx_set_objective", "lbmx_set_objective_terms",
            "lbmx_init_equilibrium",
            "lbmx_init_equilibrium_field", "lbmx_forward", "lbmx_get_state", "lbmx_set_state",
            "lbmx_get_macros", "lbmx_get_timestep", "lbmx_checkpoint_config", "lbmx_gradient",
            "lbmx_get_sensitivity", "lbmx_get_adjoint", "lbmx_get_stats", "lbmx_device_bytes",
            "lbmx_describe_boundaries", "lbmx_checkpoint_simulate", "lbmx_time_sweeps", "lbmx_stream_copy", "lbmx_stream_copy_planes", "lbmx_stream_copy_planes16",
            "lbmx_device_count", "lbmx_batch_create", "lbmx_batch_destroy", "lbmx_batch_checkpoint_config",
            "lbmx_batch_set_tile_rows", "lbmx_batch_set_graph", "lbmx_batch_autotune", "lbmx_batch_gradient", "lbmx_batch_get_stats",
            "lbmx_batch_device_bytes", "lbmx_filter_create", "lbmx_filter_destroy", "lbmx_filter_set_projection",
            "lbmx_filter_set_eta_field", "lbmx_filter_increase_beta", "lbmx_filter_set_background", "lbmx_filter_design",
            "lbmx_filter_design_to_solver", "lbmx_filter_sensitivities", "lbmx_filter_solver_sensitivities",
            "lbmx_filter_get_field", "lbmx_filter_timing", "lbmx_vec_write_binary", "lbmx_vec_load_binary",
            "lbmx_write_vts", "lbmx_mma_create", "lbmx_mma_destroy",
            "lbmx_mma_set_robust_constants", "lbmx_mma_set_asymptotes", "lbmx_mma_set_options", "lbmx_mma_update",
            "lbmx_mma_design_change", "lbmx_mma_get_state", "lbmx_adjoint_reset", "lbmx_adjoint_step", "lbmx_adjoint_set_timestep",
            "lbmx_set_design_device", "lbmx_get_sensitivity_device", "lbmx_filter_design_device",
            "lbmx_filter_design_to_solver_device", "lbmx_filter_sensitivities_device",
            "lbmx_filter_solver_sensitivities_device", "lbmx_mma_update_device", "lbmx_mma_design_change_device",
            "lbmx_dmda_global_to_natural", "lbmx_dmda_natural_to_global", "lbmx_checkpoint_mode",
            "lbmx_checkpoint_simulate_mode", "lbmx_batch_checkpoint_mode",
            "lbmx_batch_set_tile", "lbmx_batch_get_tile", "lbmx_add_boundary_table",
            "lbmx_design_change_noalloc"]


def checkpoint_simulate(nslots, nsteps, mode=CHECKPOINT_DYNAMIC):
    """Host replay of the gradient's checkpoint schedule: (recompute advances, max slots used)."""
    adv = C.c_longlong(0)
    mx = C.c_int(0)
    _check(library().lbmx_checkpoint_simulate_mode(nslots, nsteps, mode, C.byref(adv), C.byref(mx)))
    return adv.value, mx.value


def device_count():
    """Number of visible HIP devices."""
    n = C.c_int(0)
    _check(library().lbmx_device_count(C.byref(n)))
    return n.value


def stream_copy_gbs(nbytes=1 << 31, iters=20):
    """Measured device copy bandwidth (GB/s, read + write bytes)."""
    g = C.c_double(0.0)
    _check(library().lbmx_stream_copy(nbytes, iters, C.byref(g)))
    return g.value


def write_petsc_vec(path, v):
    """Write v in PETSc's binary Vec format (writeVectorToBinary)."""
    a = np.ascontiguousarray(v, dtype=np.float64)
    _check(library().lbmx_vec_write_binary(os.fsencode(path), _dptr(a), a.size))


def load_petsc_vec(path, n):
    """Read n values from a PETSc binary Vec file (loadVectorFromBinary)."""
    a = np.empty(n)
    _check(library().lbmx_vec_load_binary(os.fsencode(path), _dptr(a), n))
    return a


def _dmda_call(name, dims, lx, ly, lz, dof, arr):
    dims3 = list(dims) + [1] * (3 - len(dims))
    own = [np.asarray(v, dtype=np.int32) if v is not None else None for v in (lx, ly, lz)]
    procs = [len(v) if v is not None else 1 for v in own]
    a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
    out = np.empty_like(a)
    ip = C.POINTER(C.c_int)
    cptr = [v.ctypes.data_as(ip) if v is not None else None for v in own]
    _check(getattr(library(), name)((C.c_int * 3)(*dims3), (C.c_int * 3)(*procs), cptr[0], cptr[1], cptr[2], dof,
                                    _dptr(a), _dptr(out)))
    return out


def dmda_global_to_natural(arr, dims, lx, ly, lz=None, dof=1):
    """A DMDA global Vec's array (ranks' blocks in PETSc order) in natural order; lx, ly, lz:
    nodes owned per rank along each axis (DMDAGetOwnershipRanges)."""
    return _dmda_call("lbmx_dmda_global_to_natural", dims, lx, ly, lz, dof, arr)


def dmda_natural_to_global(arr, dims, lx, ly, lz=None, dof=1):
    return _dmda_call("lbmx_dmda_natural_to_global", dims, lx, ly, lz, dof, arr)


def plane_copy_gbs(planes, nodes, iters=20):
    """Measured copy bandwidth (GB/s) with a sweep's access pattern: planes FP64 planes in and out."""
    g = C.c_double(0.0)
    _check(library().lbmx_stream_copy_planes(planes, nodes, iters, C.byref(g)))
    return g.value


def plane_copy16_gbs(planes, nodes, iters=20):
    """The same for 16-bit planes, two nodes per thread (the 16-bit forms' forward sweep)."""
    g = C.c_double(0.0)
    _check(library().lbmx_stream_copy_planes16(planes, nodes, iters, C.byref(g)))
    return g.value


class LbmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lbmx error {code}: {msg}")
        self.code = code


def _check(code):
    if code != 0:
        msg = library().lbmx_last_error()
        raise LbmxError(code, msg.decode() if msg else "")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _tptr(t, n=None):
    """Device pointer of a contiguous float64 GPU tensor (the *_device entry points). Work
    queued on torch's current stream is finished first: the engine runs on its own streams."""
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise TypeError("expected a contiguous float64 tensor on the GPU")
    if n is not None and t.numel() != n:
        raise ValueError(f"tensor holds {t.numel()} values, expected {n}")
    torch.cuda.current_stream(t.device).synchronize()
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))


class _DeviceView:
    """A borrowed device array, exposed to torch through __cuda_array_interface__ (torch takes
    no read-only views: writing through it changes the engine's own array)."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f8", "data": (ptr, False), "version": 2}


class Solver:
    """One lattice (forward + adjoint) on the current GPU; mirrors IsothermalLBSolver2d/3d +
    AdjointLBSolver of the reference (src/latticeBoltzmann/isothermalLBSolver2d.hh:20-110,
    src/adjointLatticeBoltzmann/adjointLBSolver.hh:11-46)."""

    def __init__(self, desc):
        self.lib = library()
        self.desc = desc
        self.nodes = desc.nx * desc.ny * desc.nz
        self.Q = Q_OF[desc.lattice]
        h = C.c_void_p()
        _check(self.lib.lbmx_create(C.byref(desc), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.lbmx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        _check(self.lib.lbmx_set_stream(self.h, C.c_void_p(stream_ptr)))

    def add_boundary(self, box, bc):
        _check(self.lib.lbmx_add_boundary(self.h, C.byref(box), C.byref(bc)))

    def add_boundary_table(self, box, bc, profile=None, ncomp=1, time_factors=None):
        """lbmx_add_boundary_table: value * a(i,j,k) * b(t); profile holds ncomp values per node
        of box (natural order), time_factors b(t) for t = 0..len-1 (either may be None)."""
        prof = None if profile is None else np.ascontiguousarray(profile, dtype=np.float64).ravel()
        tf = None if time_factors is None else np.ascontiguousarray(time_factors, dtype=np.float64).ravel()
        _check(self.lib.lbmx_add_boundary_table(self.h, C.byref(box), C.byref(bc),
                                                None if prof is None else _dptr(prof), 0 if prof is None else prof.size,
                                                ncomp, None if tf is None else _dptr(tf), 0 if tf is None else tf.size))

    def set_design(self, gamma):
        g = np.ascontiguousarray(gamma, dtype=np.float64).ravel()
        _check(self.lib.lbmx_set_design(self.h, _dptr(g), g.size))

    def get_design(self):
        g = np.empty(self.nodes)
        _check(self.lib.lbmx_get_design(self.h, _dptr(g), g.size))
        return g

    def set_objective(self, obj):
        _check(self.lib.lbmx_set_objective(self.h, C.byref(obj)))

    def set_objective_terms(self, terms, time_const=1.0):
        """terms: [(kind, weight, box), ...] (lbmx_set_objective_terms)."""
        arr = make_terms(terms)
        _check(self.lib.lbmx_set_objective_terms(self.h, arr, len(terms), time_const))

    def init_equilibrium(self, rho0=1.0, u0=(0.0, 0.0, 0.0), T0=0.0):
        u = np.asarray(u0, dtype=np.float64)
        _check(self.lib.lbmx_init_equilibrium(self.h, rho0, _dptr(u), T0))

    def init_equilibrium_field(self, macros):
        m = np.ascontiguousarray(macros, dtype=np.float64).ravel()
        _check(self.lib.lbmx_init_equilibrium_field(self.h, _dptr(m), m.size))

    def forward(self, nsteps):
        J = C.c_double(0.0)
        _check(self.lib.lbmx_forward(self.h, nsteps, C.byref(J)))
        return J.value

    def get_state(self):
        f = np.empty(self.nodes * self.Q)
        _check(self.lib.lbmx_get_state(self.h, _dptr(f), f.size))
        return f

    def set_state(self, f, timestep):
        a = np.ascontiguousarray(f, dtype=np.float64).ravel()
        _check(self.lib.lbmx_set_state(self.h, _dptr(a), a.size, timestep))

    def get_macros(self):
        m = np.empty(self.nodes * 5)
        _check(self.lib.lbmx_get_macros(self.h, _dptr(m), m.size))
        return m.reshape(self.nodes, 5)

    def timestep(self):
        t = C.c_int(0)
        _check(self.lib.lbmx_get_timestep(self.h, C.byref(t)))
        return t.value

    def checkpoint_config(self, nslots, mode=None):
        _check(self.lib.lbmx_checkpoint_config(self.h, nslots))
        if mode is not None:
            _check(self.lib.lbmx_checkpoint_mode(self.h, mode))

    def gradient(self, nsteps, rho0=1.0, u0=(0.0, 0.0, 0.0), T0=0.0, want_sens=True):
        u = np.asarray(u0, dtype=np.float64)
        J = np.zeros(1)
        sens = np.empty(self.nodes) if want_sens else None
        _check(self.lib.lbmx_gradient(self.h, nsteps, rho0, _dptr(u), T0, _dptr(J),
                                      _dptr(sens) if want_sens else None,
                                      self.nodes if want_sens else 0))
        return J[0], sens

    def get_sensitivity(self):
        s = np.empty(self.nodes)
        _check(self.lib.lbmx_get_sensitivity(self.h, _dptr(s), s.size))
        return s

    def get_adjoint(self):
        a = np.empty(self.nodes * self.Q)
        _check(self.lib.lbmx_get_adjoint(self.h, _dptr(a), a.size))
        return a

    def adjoint_reset(self, timestep):
        """resetAdjoints + setCurrentTimestep: the next adjoint_step reverses step `timestep`."""
        _check(self.lib.lbmx_adjoint_reset(self.h, int(timestep)))

    def adjoint_set_timestep(self, timestep):
        _check(self.lib.lbmx_adjoint_set_timestep(self.h, int(timestep)))

    def adjoint_step(self, with_source=True):
        """One reverse step (adjointCollide, computeCollideSource, adjointStream) on the current
        forward state, which must be at the adjoint's step."""
        _check(self.lib.lbmx_adjoint_step(self.h, int(bool(with_source))))

    def set_design_device(self, gamma):
        """setFieldFromDesignVec from a float64 GPU tensor (natural node order)."""
        _check(self.lib.lbmx_set_design_device(self.h, _tptr(gamma, self.nodes), self.nodes))

    def sensitivity_device(self):
        """dJ/dgamma as a torch view of the solver's own device array (borrowed: valid until the
        next gradient, adjoint reset or close)."""
        import torch
        p = C.POINTER(C.c_double)()
        n = C.c_size_t(0)
        _check(self.lib.lbmx_get_sensitivity_device(self.h, C.byref(p), C.byref(n)))
        return torch.as_tensor(_DeviceView(C.cast(p, C.c_void_p).value, n.value), device="cuda")

    def stats(self):
        st = Stats()
        _check(self.lib.lbmx_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def write_vts(self, path):
        """Macros and design of the current state as a VTK StructuredGrid (outputFields)."""
        _check(self.lib.lbmx_write_vts(self.h, os.fsencode(path)))

    def time_sweeps(self, kind, nsweeps):
        """Device time (ms, HIP events) of nsweeps forward (kind 0) or adjoint (kind 1) sweeps."""
        ms = C.c_double(0.0)
        _check(self.lib.lbmx_time_sweeps(self.h, kind, nsweeps, C.byref(ms)))
        return ms.value

    def boundaries(self):
        cnt = C.c_int(0)
        _check(self.lib.lbmx_describe_boundaries(self.h, None, 0, C.byref(cnt)))
        buf = (C.c_int * (10 * max(1, cnt.value)))()
        _check(self.lib.lbmx_describe_boundaries(self.h, buf, cnt.value, C.byref(cnt)))
        out = []
        for k in range(cnt.value):
            v = buf[10 * k:10 * k + 10]
            out.append((v[0], tuple(v[1:4]), tuple(v[4:10])))
        return out

    def device_bytes(self):
        b = C.c_size_t(0)
        _check(self.lib.lbmx_device_bytes(self.h, C.byref(b)))
        return b.value


class Batch:
    """Batched robust-ensemble form: member Solvers (same lattice/operator/coupling/precision)
    advanced in lockstep, one launch per sweep over all members. Members must outlive it."""

    def __init__(self, members):
        self.lib = library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        self.h = C.c_void_p()
        _check(self.lib.lbmx_batch_create(arr, len(self.members), C.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.lbmx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def checkpoint_config(self, nslots, mode=None):
        _check(self.lib.lbmx_batch_checkpoint_config(self.h, nslots))
        if mode is not None:
            _check(self.lib.lbmx_batch_checkpoint_mode(self.h, mode))

    def set_tile_rows(self, rows):
        _check(self.lib.lbmx_batch_set_tile_rows(self.h, rows))

    def set_tile(self, rows, paired):
        _check(self.lib.lbmx_batch_set_tile(self.h, rows, 1 if paired else 0))

    def tile(self):
        """(rows, paired) of the tile shape in use."""
        r, p = C.c_int(0), C.c_int(0)
        _check(self.lib.lbmx_batch_get_tile(self.h, C.byref(r), C.byref(p)))
        return r.value, bool(p.value)

    def set_graph(self, enable=True):
        _check(self.lib.lbmx_batch_set_graph(self.h, 1 if enable else 0))

    def autotune(self, nsweeps=10):
        """Returns (chosen tile rows, ms per forward+adjoint sweep pair for rows 1, 2, 4)."""
        r = C.c_int(0)
        ms = (C.c_double * 3)()
        _check(self.lib.lbmx_batch_autotune(self.h, nsweeps, C.byref(r), ms))
        return r.value, list(ms)

    def gradient(self, nsteps, rho0=1.0, u0=(0.0, 0.0, 0.0), T0=0.0, want_sens=True):
        """Returns (J per member, list of dJ/dgamma arrays or None)."""
        n = len(self.members)
        J = np.zeros(n)
        u = np.asarray(u0, dtype=np.float64)
        sens = [np.empty(m.nodes) for m in self.members] if want_sens else None
        ptrs = (C.POINTER(C.c_double) * n)(*[_dptr(a) for a in sens]) if want_sens else None
        _check(self.lib.lbmx_batch_gradient(self.h, nsteps, rho0, _dptr(u), T0, _dptr(J), ptrs))
        return J, sens

    def stats(self):
        st = Stats()
        _check(self.lib.lbmx_batch_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in Stats._fields_}

    def device_bytes(self):
        b = C.c_size_t(0)
        _check(self.lib.lbmx_batch_device_bytes(self.h, C.byref(b)))
        return b.value



def make_filter_desc(nx, ny, nz, design_box, filter_box=None, stages=((8, True, 1.0, 0.5),)):
    """lbmx_filter_desc for a chain of (radius, project, beta, eta) stages (1 or 2)."""
    d = FilterDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.design_box = design_box
    d.filter_box = filter_box if filter_box is not None else design_box
    d.nstages = len(stages)
    for k, (r, proj, beta, eta) in enumerate(stages):
        d.radius[k] = r
        d.project[k] = 1 if proj else 0
        d.beta[k] = beta
        d.eta[k] = eta
    return d


class Filter:
    """Design pipeline on the device: design-box mapping, density filter and tanh projection
    stages (DensityFilter2d/3d, ProjectionFilter2d/3d, DoubleProjectionFilter2d) and their
    sensitivity map. Fields are FP64 numpy arrays in natural order."""

    def __init__(self, desc):
        self.lib = library()
        self.desc = desc
        self.h = C.c_void_p()
        _check(self.lib.lbmx_filter_create(C.byref(desc), C.byref(self.h)))
        self.n_full = desc.nx * desc.ny * desc.nz
        b = desc.design_box
        self.n_design = (b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1) * (b.z1 - b.z0 + 1)

    def close(self):
        if self.h:
            self.lib.lbmx_filter_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_projection(self, stage, beta, eta):
        _check(self.lib.lbmx_filter_set_projection(self.h, stage, beta, eta))

    def increase_beta(self, factor):
        _check(self.lib.lbmx_filter_increase_beta(self.h, factor))

    def set_eta_field(self, stage, eta):
        """Per-node projection thresholds of a stage (full grid), or None for its constant eta."""
        if eta is None:
            _check(self.lib.lbmx_filter_set_eta_field(self.h, stage, None, 0))
            return
        a = np.ascontiguousarray(eta, dtype=np.float64)
        _check(self.lib.lbmx_filter_set_eta_field(self.h, stage, _dptr(a), a.size))

    def set_background(self, full):
        a = np.ascontiguousarray(full, dtype=np.float64)
        _check(self.lib.lbmx_filter_set_background(self.h, _dptr(a), a.size))

    def design(self, design):
        """Physical (filtered and projected) full-grid field of a design-box vector."""
        d = np.ascontiguousarray(design, dtype=np.float64)
        out = np.empty(self.n_full)
        _check(self.lib.lbmx_filter_design(self.h, _dptr(d), d.size, _dptr(out), out.size))
        return out

    def design_to_solver(self, design, solver):
        d = np.ascontiguousarray(design, dtype=np.float64)
        _check(self.lib.lbmx_filter_design_to_solver(self.h, _dptr(d), d.size, solver.h))

    def sensitivities(self, full_sens):
        s = np.ascontiguousarray(full_sens, dtype=np.float64)
        out = np.empty(self.n_design)
        _check(self.lib.lbmx_filter_sensitivities(self.h, _dptr(s), s.size, _dptr(out), out.size))
        return out

    def solver_sensitivities(self, solver):
        out = np.empty(self.n_design)
        _check(self.lib.lbmx_filter_solver_sensitivities(self.h, solver.h, _dptr(out), out.size))
        return out

    def design_device(self, design, out=None):
        """design() on float64 GPU tensors; returns the physical field (out, or a new tensor)."""
        import torch
        if out is None:
            out = torch.empty(self.n_full, dtype=torch.float64, device=design.device)
        _check(self.lib.lbmx_filter_design_device(self.h, _tptr(design, self.n_design), self.n_design,
                                                  _tptr(out, self.n_full), self.n_full))
        return out

    def design_to_solver_device(self, design, solver):
        _check(self.lib.lbmx_filter_design_to_solver_device(self.h, _tptr(design, self.n_design), self.n_design,
                                                            solver.h))

    def sensitivities_device(self, full_sens, out=None):
        import torch
        if out is None:
            out = torch.empty(self.n_design, dtype=torch.float64, device=full_sens.device)
        _check(self.lib.lbmx_filter_sensitivities_device(self.h, _tptr(full_sens, self.n_full), self.n_full,
                                                         _tptr(out, self.n_design), self.n_design))
        return out

    def solver_sensitivities_device(self, solver, out=None):
        import torch
        if out is None:
            out = torch.empty(self.n_design, dtype=torch.float64, device="cuda")
        _check(self.lib.lbmx_filter_solver_sensitivities_device(self.h, solver.h, _tptr(out, self.n_design),
                                                                self.n_design))
        return out

    def field(self, which):
        out = np.empty(self.n_full)
        _check(self.lib.lbmx_filter_get_field(self.h, which, _dptr(out), out.size))
        return out

    def timing(self):
        ms = (C.c_double * 2)()
        _check(self.lib.lbmx_filter_timing(self.h, ms))
        return ms[0], ms[1]


class MMA:
    """Method of moving asymptotes on the device (MMA of src/topOpt/MMA.cc): n design
    variables, m constraints (1..8). update() returns the new design."""

    def __init__(self, n, m, a=None):
        self.lib = library()
        self.n, self.m = n, m
        self.h = C.c_void_p()
        _check(self.lib.lbmx_mma_create(n, m, C.byref(self.h)))
        if a is not None:
            self.set_robust_constants(a)

    def close(self):
        if self.h:
            self.lib.lbmx_mma_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_robust_constants(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(self.lib.lbmx_mma_set_robust_constants(self.h, _dptr(a)))

    def set_asymptotes(self, init, dec, inc):
        _check(self.lib.lbmx_mma_set_asymptotes(self.h, init, dec, inc))

    def set_options(self, constraint_modification=False, robust_asymptotes_type=0):
        _check(self.lib.lbmx_mma_set_options(self.h, int(constraint_modification), int(robust_asymptotes_type)))

    def update(self, x, dfdx, gx, dgdx, movlim):
        x = np.array(x, dtype=np.float64)
        dfdx = np.ascontiguousarray(dfdx, dtype=np.float64)
        gx = np.ascontiguousarray(gx, dtype=np.float64)
        rows = [np.ascontiguousarray(r, dtype=np.float64) for r in np.asarray(dgdx, dtype=np.float64).reshape(self.m, self.n)]
        ptrs = (C.POINTER(C.c_double) * self.m)(*[_dptr(r) for r in rows])
        _check(self.lib.lbmx_mma_update(self.h, _dptr(x), _dptr(dfdx), _dptr(gx), ptrs, movlim))
        return x

    def design_change(self, xnew=None):
        ch = C.c_double(0.0)
        if xnew is None:
            _check(self.lib.lbmx_mma_design_change(self.h, None, C.byref(ch)))
        else:
            a = np.ascontiguousarray(xnew, dtype=np.float64)
            _check(self.lib.lbmx_mma_design_change(self.h, _dptr(a), C.byref(ch)))
        return ch.value

    def update_device(self, x, dfdx, gx, dgdx, movlim):
        """update() on float64 GPU tensors: x (updated in place), dfdx and the m rows dgdx; gx
        is a host sequence of m constraint values."""
        gx = np.ascontiguousarray(gx, dtype=np.float64)
        if len(dgdx) != self.m:
            raise ValueError(f"expected {self.m} constraint sensitivity rows")
        ptrs = (C.POINTER(C.c_double) * self.m)(*[_tptr(r, self.n) for r in dgdx])
        _check(self.lib.lbmx_mma_update_device(self.h, _tptr(x, self.n), _tptr(dfdx, self.n), _dptr(gx), ptrs,
                                               movlim))
        return x

    def design_change_device(self, xnew=None):
        ch = C.c_double(0.0)
        _check(self.lib.lbmx_mma_design_change_device(self.h, None if xnew is None else _tptr(xnew, self.n),
                                                      C.byref(ch)))
        return ch.value

    def state(self):
        L, U, lam = np.empty(self.n), np.empty(self.n), np.empty(self.m)
        it, ms = C.c_int(0), C.c_double(0.0)
        _check(self.lib.lbmx_mma_get_state(self.h, _dptr(L), _dptr(U), _dptr(lam), C.byref(it), C.byref(ms)))
        return {"L": L, "U": U, "lam": lam, "newton_iters": it.value, "ms": ms.value}
