"""minimmerflow_amd -- MI355X-native finite-volume time step of minimmerflow.

Python front-end over the C-ABI of lib/libmmf.so (include/mmf.h).  Every
compute call runs the hand-written HIP kernels; there is no CPU fallback: if
the library is missing or no GPU is present the calls raise.

The classes mirror the reference's entry points:
  Solver.compute_rhs   -> euler::computeRHS            (euler.cpp:127-249)
  Solver.advance/run   -> the RK3 loop of computation() (main.cpp:374-506)
  Solver.set_bodies    -> body::isPointFluid + flags   (body.cpp:80-95, main.cpp:222-277)
  Batch                -> many independent 2D runs of computation() advanced together
  Problem              -> problem::*                   (problem.cpp)
  batch_* functions    -> utils::* / euler::eval*      (utils.cpp, euler.cpp)
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmmf.so")

NF = 5
OK, ERR_INVALID, ERR_UNSUPPORTED_ORDER, ERR_HIP, ERR_NO_DEVICE, ERR_OOM, ERR_STATE = range(7)
BC_NONE, BC_FREE_FLOW, BC_REFLECTING, BC_WALL, BC_DIRICHLET, BC_PERIODIC = -1, 0, 1, 2, 3, 4
CELL_FLUID, CELL_SOLVED = 1, 2
KERNEL_AUTO, KERNEL_GATHER, KERNEL_MARCH, KERNEL_TILE = 0, 1, 2, 3
(VORTEX_XY, VORTEX_ZX, VORTEX_YZ, RADSOD, SOD_X, SOD_Y, SOD_Z, FFSTEP, FREESTREAM) = range(9)
PROBLEM_KEYS = {"vortex_xy": VORTEX_XY, "vortex_zx": VORTEX_ZX, "vortex_yz": VORTEX_YZ,
                "radsod": RADSOD, "sod3d_x": SOD_X, "sod3d_y": SOD_Y, "sod3d_z": SOD_Z,
                "ffstep": FFSTEP, "freestream": FREESTREAM}


STATE_U, STATE_W, STATE_RHS = 0, 1, 2


class MMFError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__("mmf status %d (%s)%s" % (status, _status_string(status), (": " + msg) if msg else ""))


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ncells", C.c_int64), ("nfaces", C.c_int64),
                ("cell_volume", C.POINTER(C.c_double)), ("min_cell_size", C.c_double),
                ("face_owner", C.POINTER(C.c_int64)), ("face_neigh", C.POINTER(C.c_int64)),
                ("face_area", C.POINTER(C.c_double)), ("face_normal", C.POINTER(C.c_double)),
                ("face_bc", C.POINTER(C.c_int32)), ("face_info", C.POINTER(C.c_double))]


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("order", C.c_int32), ("n", C.c_int64 * 3),
                ("origin", C.c_double * 3), ("h", C.c_double), ("side_bc", C.c_int32 * 6),
                ("side_info", (C.c_double * 5) * 6)]


class Problem(C.Structure):
    _fields_ = [("type", C.c_int32), ("dim", C.c_int32), ("vortex_beta", C.c_double),
                ("vortex_center", C.c_double * 3), ("vortex_periodic", C.c_int32),
                ("domain_origin", C.c_double * 3), ("domain_length", C.c_double),
                ("sod_r0", C.c_double), ("sod_p_in", C.c_double), ("sod_t_in", C.c_double),
                ("sod_p_out", C.c_double), ("sod_t_out", C.c_double), ("fs_prim", C.c_double * 5)]


_lib = None
# kernel variant applied to every new Solver: (kind, tile_rows, chunk)
_default_kernel = (KERNEL_AUTO, 0, 0)


def set_default_kernel(kind=KERNEL_AUTO, tile_rows=0, chunk=0):
    """Stage-kernel variant for Solvers created afterwards (AUTO restores the default).
    A variant that does not apply to a grid (the marching kernel at order 2)
    leaves that grid on its automatic choice."""
    global _default_kernel
    _default_kernel = (int(kind), int(tile_rows), int(chunk))


def _status_string(s):
    try:
        return lib().mmf_status_string(int(s)).decode()
    except Exception:  # library not loadable: plain code
        return str(s)


def lib():
    """Load lib/libmmf.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("minimmerflow_amd: %s is missing; build it with "
                          "`python -m minimmerflow_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, dp, i64p, u8p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_uint8)
    i32p = C.POINTER(C.c_int32)
    fp = C.POINTER(C.c_float)
    sig = {
        "mmf_create": ([C.POINTER(GridDesc), C.POINTER(C.c_void_p)], C.c_int),
        "mmf_destroy": ([vp], C.c_int),
        "mmf_last_error": ([vp], C.c_char_p),
        "mmf_status_string": ([C.c_int], C.c_char_p),
        "mmf_get_sizes": ([vp, i64p, i64p], C.c_int),
        "mmf_set_cell_flags": ([vp, u8p, u8p], C.c_int),
        "mmf_set_bodies": ([vp, dp, C.c_int32, dp, C.c_int32], C.c_int),
        "mmf_get_cell_flags": ([vp, u8p], C.c_int),
        "mmf_set_state": ([vp, dp], C.c_int),
        "mmf_get_state": ([vp, dp], C.c_int),
        "mmf_set_raw_order": ([vp, i64p, C.c_int64], C.c_int),
        "mmf_compute_rhs": ([vp, dp, dp], C.c_int),
        "mmf_set_time": ([vp, C.c_double, C.c_double, C.c_double], C.c_int),
        "mmf_advance": ([vp, C.c_int64], C.c_int),
        "mmf_set_step_graphs": ([vp, C.c_int32], C.c_int),
        "mmf_run": ([vp, C.c_int64, i64p, dp], C.c_int),
        "mmf_get_time": ([vp, dp, dp, i64p, dp], C.c_int),
        "mmf_get_nonfinite_steps": ([vp, i64p], C.c_int),
        "mmf_compute_rhs_noalloc": ([C.POINTER(GridDesc), vp, vp, C.c_int64, vp, vp, vp, vp], C.c_int),
        "mmf_noalloc_last_error": ([], C.c_char_p),
        "mmf_synchronize": ([vp], C.c_int),
        "mmf_write_vtk": ([vp, C.c_char_p], C.c_int),
        "mmf_write_vtk_async": ([vp, C.c_char_p], C.c_int),
        "mmf_wait_writes": ([vp], C.c_int),
        "mmf_error_norms": ([vp, dp, dp, dp], C.c_int),
        "mmf_get_positivity_counters": ([vp, i64p, i64p], C.c_int),
        "mmf_reset_positivity_counters": ([vp], C.c_int),
        "mmf_enable_timing": ([vp, C.c_int32], C.c_int),
        "mmf_read_timing": ([vp, dp, i64p], C.c_int),
        "mmf_batch_cons2prim": ([dp, C.c_int64, dp], C.c_int),
        "mmf_batch_prim2cons": ([dp, C.c_int64, dp], C.c_int),
        "mmf_batch_eval_fluxes": ([dp, dp, dp, C.c_int64, dp], C.c_int),
        "mmf_batch_eval_splitting": ([dp, dp, dp, C.c_int64, dp, dp], C.c_int),
        "mmf_batch_eval_bc_values": ([C.c_int32, dp, dp, dp, C.c_int64, dp], C.c_int),
        "mmf_problem_defaults": ([C.POINTER(Problem), C.c_int32, C.c_int32], C.c_int),
        "mmf_problem_default_end_time": ([C.POINTER(Problem), dp], C.c_int),
        "mmf_problem_exact": ([C.POINTER(Problem), dp, C.c_int64, C.c_double, dp], C.c_int),
        "mmf_problem_border_bc": ([C.POINTER(Problem), dp, i32p, dp], C.c_int),
        "mmf_problem_side_bcs": ([C.POINTER(Problem), C.POINTER(GridDesc)], C.c_int),
        "mmf_problem_border_bcs": ([C.POINTER(Problem), C.POINTER(GridDesc), i32p, dp], C.c_int),
        "mmf_border_face_count": ([C.POINTER(GridDesc), i64p], C.c_int),
        "mmf_set_border_bcs": ([vp, i32p, dp], C.c_int),
        "mmf_compute_rhs_ex": ([vp, u8p, C.c_int32, dp, dp, dp], C.c_int),
        "mmf_rk_rhs": ([vp, C.c_int32, dp], C.c_int),
        "mmf_rk_dt": ([vp, C.c_double, dp], C.c_int),
        "mmf_rk_update": ([vp, C.c_int32, C.c_double], C.c_int),
        "mmf_pack_cells": ([vp, C.c_int32, i64p, C.c_int64, dp], C.c_int),
        "mmf_unpack_cells": ([vp, C.c_int32, i64p, C.c_int64, dp], C.c_int),
        "mmf_get_device_state": ([vp, C.c_int32, C.POINTER(C.c_void_p), i64p], C.c_int),
        "mmf_set_initial_state": ([vp, C.POINTER(Problem)], C.c_int),
        "mmf_error_vs_exact": ([vp, C.POINTER(Problem), C.c_double, dp, dp], C.c_int),
        "mmf_get_centroids": ([vp, dp], C.c_int),
        "mmf_calibration_copy": ([C.c_int64, C.c_int32, dp], C.c_int),
        "mmf_set_kernel": ([vp, C.c_int32, C.c_int32, C.c_int32], C.c_int),
        "mmf_get_kernel": ([vp, i32p, i32p, i32p], C.c_int),
        "mmf_set_march_shape": ([vp, C.c_int32, C.c_int32, C.c_int32], C.c_int),
        "mmf_set_march_blocks": ([vp, C.c_int32], C.c_int),
        "mmf_batch_create": ([C.POINTER(GridDesc), C.c_int32, C.POINTER(C.c_void_p)], C.c_int),
        "mmf_batch_destroy": ([vp], C.c_int),
        "mmf_batch_last_error": ([vp], C.c_char_p),
        "mmf_batch_sizes": ([vp, i64p, i32p], C.c_int),
        "mmf_batch_set_state": ([vp, C.c_int32, dp], C.c_int),
        "mmf_batch_get_state": ([vp, C.c_int32, dp], C.c_int),
        "mmf_batch_set_initial_states": ([vp, C.POINTER(Problem)], C.c_int),
        "mmf_batch_set_time": ([vp, dp, dp, dp], C.c_int),
        "mmf_batch_set_shape": ([vp, C.c_int32, C.c_int32], C.c_int),
        "mmf_batch_set_tail": ([vp, C.c_int32], C.c_int),
        "mmf_batch_get_tail": ([vp, i32p], C.c_int),
        "mmf_batch_get_shape": ([vp, i32p, i32p], C.c_int),
        "mmf_batch_advance": ([vp, C.c_int64], C.c_int),
        "mmf_batch_run": ([vp, C.c_int64, i64p], C.c_int),
        "mmf_batch_get_times": ([vp, dp, i64p], C.c_int),
        "mmf_batch_error_vs_exact": ([vp, C.POINTER(Problem), dp, dp], C.c_int),
        "mmf_batch_enable_timing": ([vp, C.c_int32], C.c_int),
        "mmf_batch_read_timing": ([vp, dp, i64p], C.c_int),
        "mmf_batch_autotune": ([vp, C.c_int32, i32p, i32p, dp], C.c_int),
        "mmf_learn_create": ([C.c_uint64, C.c_int64, C.POINTER(C.c_void_p)], C.c_int),
        "mmf_learn_destroy": ([vp], C.c_int),
        "mmf_learn_last_error": ([vp], C.c_char_p),
        "mmf_learn_get_params": ([vp, fp], C.c_int),
        "mmf_learn_set_params": ([vp, fp], C.c_int),
        "mmf_learn_collect": ([vp, vp, C.c_int32, i64p, C.c_int32, C.c_double], C.c_int),
        "mmf_learn_num_samples": ([vp, i64p], C.c_int),
        "mmf_learn_get_samples": ([vp, C.c_int64, C.c_int64, fp, fp, fp], C.c_int),
        "mmf_learn_set_samples": ([vp, fp, fp, fp, C.c_int64], C.c_int),
        "mmf_learn_loss_grad": ([vp, fp, fp, fp, C.c_int64, dp, fp], C.c_int),
        "mmf_learn_adam_step": ([vp, fp, C.c_double], C.c_int),
        "mmf_learn_train": ([vp, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_uint64, dp], C.c_int),
        "mmf_learn_forward": ([vp, fp, C.c_int64, C.c_int32, fp], C.c_int),
        "mmf_set_learned_reconstruction": ([vp, vp, C.c_double, C.c_int32], C.c_int),
        "mmf_set_learned_options": ([vp, C.c_int32, C.c_double], C.c_int),
        "mmf_get_learned_counts": ([vp, i64p, i64p], C.c_int),
        "mmf_reset_learned_counts": ([vp], C.c_int),
        "mmf_get_face_states": ([vp, dp], C.c_int),
        "mmf_restrict_state": ([vp, C.c_int32, dp], C.c_int),
        "mmf_mesh_create": ([C.POINTER(MeshDesc), C.POINTER(C.c_void_p)], C.c_int),
        "mmf_mesh_destroy": ([vp], C.c_int),
        "mmf_mesh_last_error": ([vp], C.c_char_p),
        "mmf_mesh_sizes": ([vp, i64p, i64p], C.c_int),
        "mmf_mesh_set_cell_flags": ([vp, u8p, u8p], C.c_int),
        "mmf_mesh_set_state": ([vp, dp], C.c_int),
        "mmf_mesh_get_state": ([vp, dp], C.c_int),
        "mmf_mesh_compute_rhs": ([vp, dp, dp], C.c_int),
        "mmf_mesh_set_time": ([vp, C.c_double, C.c_double, C.c_double], C.c_int),
        "mmf_mesh_advance": ([vp, C.c_int64], C.c_int),
        "mmf_mesh_set_step_graphs": ([vp, C.c_int32], C.c_int),
        "mmf_mesh_run": ([vp, C.c_int64, i64p, dp], C.c_int),
        "mmf_mesh_get_time": ([vp, dp, dp, i64p, dp], C.c_int),
        "mmf_mesh_get_counters": ([vp, i64p, i64p, i64p], C.c_int),
        "mmf_mesh_enable_timing": ([vp, C.c_int32], C.c_int),
        "mmf_mesh_read_timing": ([vp, dp, dp, i64p], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names the header declares (for the ABI test)."""
    import re
    hdr = open(os.path.join(os.path.dirname(HERE), "include", "mmf.h")).read()
    return sorted(set(re.findall(r"\b(mmf_[a-z0-9_]+)\s*\(", hdr)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _check(rc, ctx=None):
    if rc != OK:
        msg = ""
        if ctx:
            m = lib().mmf_last_error(ctx)
            msg = m.decode() if m else ""
        raise MMFError(rc, msg)


# ---- problem catalogue -----------------------------------------------------------

def problem(ptype, dim=2, **overrides):
    """Problem with the reference defaults (problem.cpp) and optional overrides."""
    if isinstance(ptype, str):
        ptype = PROBLEM_KEYS[ptype]
    pb = Problem()
    _check(lib().mmf_problem_defaults(C.byref(pb), int(ptype), int(dim)))
    for k, v in overrides.items():
        if isinstance(v, (list, tuple, np.ndarray)):
            arr = getattr(pb, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(pb, k, v)
    return pb


def default_end_time(pb):
    t = C.c_double()
    _check(lib().mmf_problem_default_end_time(C.byref(pb), C.byref(t)))
    return t.value


def exact_conservatives(pb, points, t):
    pts = _f64(points).reshape(-1, 3)
    out = np.empty((pts.shape[0], NF))
    _check(lib().mmf_problem_exact(C.byref(pb), _dp(pts), pts.shape[0], float(t), _dp(out)))
    return out


# ---- batched point functions ------------------------------------------------------

def batch_cons2prim(cons):
    c = _f64(cons).reshape(-1, NF)
    out = np.empty_like(c)
    _check(lib().mmf_batch_cons2prim(_dp(c), c.shape[0], _dp(out)))
    return out


def batch_prim2cons(prim):
    p = _f64(prim).reshape(-1, NF)
    out = np.empty_like(p)
    _check(lib().mmf_batch_prim2cons(_dp(p), p.shape[0], _dp(out)))
    return out


def batch_eval_fluxes(cons, prim, normals):
    c, p, n = _f64(cons).reshape(-1, NF), _f64(prim).reshape(-1, NF), _f64(normals).reshape(-1, 3)
    out = np.empty_like(c)
    _check(lib().mmf_batch_eval_fluxes(_dp(c), _dp(p), _dp(n), c.shape[0], _dp(out)))
    return out


def batch_eval_splitting(L, R, normals):
    L, R, n = _f64(L).reshape(-1, NF), _f64(R).reshape(-1, NF), _f64(normals).reshape(-1, 3)
    F = np.empty_like(L)
    lam = np.empty(L.shape[0])
    _check(lib().mmf_batch_eval_splitting(_dp(L), _dp(R), _dp(n), L.shape[0], _dp(F), _dp(lam)))
    return F, lam


def batch_eval_bc_values(bc, normals, cons, info=None):
    n, c = _f64(normals).reshape(-1, 3), _f64(cons).reshape(-1, NF)
    i = None if info is None else _f64(info).reshape(-1, NF)
    out = np.empty_like(c)
    _check(lib().mmf_batch_eval_bc_values(int(bc), _dp(n), None if i is None else _dp(i), _dp(c),
                                          c.shape[0], _dp(out)))
    return out


def calibration_copy(nbytes, reps=20):
    ms = C.c_double()
    _check(lib().mmf_calibration_copy(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


# ---- solver ------------------------------------------------------------------------

class Solver:
    """Device-resident solver of one uniform Cartesian grid."""

    def __init__(self, dim, n, origin, h, side_bc, side_info=None, order=1):
        d = GridDesc()
        d.dim = int(dim)
        d.order = int(order)
        n = list(n) + [1] * (3 - len(n))
        for a in range(3):
            d.n[a] = int(n[a])
            d.origin[a] = float(origin[a]) if a < len(origin) else 0.0
        d.h = float(h)
        for s in range(6):
            d.side_bc[s] = int(side_bc[s]) if s < len(side_bc) else BC_FREE_FLOW
            if side_info is not None:
                for k in range(NF):
                    d.side_info[s][k] = float(side_info[s][k])
        self.desc = d
        self.dim = int(dim)
        self.n = tuple(int(x) for x in n)
        self.h = float(h)
        self.order = int(order)
        ctx = C.c_void_p()
        _check(lib().mmf_create(C.byref(d), C.byref(ctx)))
        self.ctx = ctx
        nc, nf = C.c_int64(), C.c_int64()
        _check(lib().mmf_get_sizes(self.ctx, C.byref(nc), C.byref(nf)))
        self.ncells, self.ninterfaces = nc.value, nf.value
        if _default_kernel != (KERNEL_AUTO, 0, 0):
            rc = lib().mmf_set_kernel(self.ctx, *_default_kernel)
            if rc not in (OK, ERR_INVALID):
                _check(rc, self.ctx)

    @classmethod
    def for_problem(cls, pb, ncells, order=1, side_bc=None, origin=None, length=None):
        """Grid as computation() builds it: length/nCells cells (main.cpp:147-149),
        border BCs from the problem (main.cpp:251-260) unless given."""
        dim = pb.dim
        origin = list(pb.domain_origin) if origin is None else list(origin)
        length = pb.domain_length if length is None else float(length)
        n = [int(ncells)] * dim
        h = length / int(ncells)
        d = GridDesc()
        d.dim, d.order, d.h = dim, order, h
        for a in range(3):
            d.n[a] = n[a] if a < dim else 1
            d.origin[a] = origin[a]
        _check(lib().mmf_problem_side_bcs(C.byref(pb), C.byref(d)))
        bcs = [d.side_bc[s] for s in range(6)] if side_bc is None else list(side_bc)
        info = [[d.side_info[s][k] for k in range(NF)] for s in range(6)]
        s = cls(dim, n, origin, h, bcs, info, order)
        if side_bc is None:
            # a side whose faces differ (FFSTEP past x = 3.2) gets them face by face
            fb, fi = problem_border_bcs(pb, d)
            if not _sides_uniform(d, fb, fi):
                s.set_border_bcs(fb, fi)
        return s

    def close(self):
        if getattr(self, "ctx", None):
            lib().mmf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.ctx)

    # state
    def set_state(self, U):
        U = _f64(U).reshape(self.ncells, NF)
        self._c(lib().mmf_set_state(self.ctx, _dp(U)))

    def get_state(self):
        U = np.empty((self.ncells, NF))
        self._c(lib().mmf_get_state(self.ctx, _dp(U)))
        return U

    def set_initial_state(self, pb):
        self._c(lib().mmf_set_initial_state(self.ctx, C.byref(pb)))

    def centroids(self):
        out = np.empty((self.ncells, 3))
        self._c(lib().mmf_get_centroids(self.ctx, _dp(out)))
        return out

    def set_raw_order(self, cell_of_raw):
        if cell_of_raw is None:
            self._c(lib().mmf_set_raw_order(self.ctx, None, 0))
            return
        a = np.ascontiguousarray(cell_of_raw, dtype=np.int64)
        self._raw = a
        self._c(lib().mmf_set_raw_order(self.ctx, a.ctypes.data_as(C.POINTER(C.c_int64)), a.size))

    # stage-kernel variant
    def set_kernel(self, kind=KERNEL_AUTO, tile_rows=0, chunk=0):
        self._c(lib().mmf_set_kernel(self.ctx, int(kind), int(tile_rows), int(chunk)))

    def set_march_shape(self, tile_x=0, tile_y=0, chunk=0):
        """3D marching tile (64 x 4, 64 x 8, 32 x 8, 32 x 16) and chunk; 0 keeps the default."""
        self._c(lib().mmf_set_march_shape(self.ctx, int(tile_x), int(tile_y), int(chunk)))

    def set_march_blocks(self, blocks):
        """3D order 1, 64 x 8 tiles: the planes of all tiles split over this many marching blocks in
        shares of equal cost (0: the default)."""
        self._c(lib().mmf_set_march_blocks(self.ctx, int(blocks)))

    def kernel(self):
        k, ty, ch = C.c_int32(), C.c_int32(), C.c_int32()
        self._c(lib().mmf_get_kernel(self.ctx, C.byref(k), C.byref(ty), C.byref(ch)))
        return k.value, ty.value, ch.value

    # bodies / flags
    def set_cell_flags(self, fluid, solved=None):
        f = np.ascontiguousarray(fluid, dtype=np.uint8)
        s = None if solved is None else np.ascontiguousarray(solved, dtype=np.uint8)
        self._c(lib().mmf_set_cell_flags(self.ctx, f.ctypes.data_as(C.POINTER(C.c_uint8)),
                                         None if s is None else s.ctypes.data_as(C.POINTER(C.c_uint8))))

    def set_bodies(self, boxes=(), spheres=()):
        b = _f64(np.asarray(boxes, dtype=np.float64).reshape(-1, 6))
        s = _f64(np.asarray(spheres, dtype=np.float64).reshape(-1, 4))
        self._c(lib().mmf_set_bodies(self.ctx, _dp(b) if len(b) else None, len(b),
                                     _dp(s) if len(s) else None, len(s)))

    def set_border_bcs(self, bc, info=None):
        """Border condition of every border face (main.cpp:251-260), in the
        layout of border_face_count: sides x-, x+, y-, y+ (, z-, z+), a side's
        faces ordered by its two other axes, the lower one fastest."""
        b = np.ascontiguousarray(bc, dtype=np.int32).ravel()
        if len(b) != border_face_count(self.desc):
            raise ValueError("one border condition per border face: %d" % border_face_count(self.desc))
        inf = None if info is None else _f64(np.asarray(info, dtype=np.float64).reshape(len(b), NF))
        self._c(lib().mmf_set_border_bcs(self.ctx, b.ctypes.data_as(C.POINTER(C.c_int32)),
                                         _dp(inf) if inf is not None else None))

    def cell_flags(self):
        out = np.empty(self.ncells, dtype=np.uint8)
        self._c(lib().mmf_get_cell_flags(self.ctx, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    # euler::computeRHS
    def compute_rhs(self, want_rhs=True):
        R = np.empty((self.ncells, NF)) if want_rhs else None
        me = C.c_double()
        self._c(lib().mmf_compute_rhs(self.ctx, _dp(R) if want_rhs else None, C.byref(me)))
        return R, me.value

    def compute_rhs_ex(self, solved, order, U=None):
        """euler::computeRHS with the reference's per-call arguments: solved
        flags (None keeps the context's) and order (euler.hpp:41-43)."""
        sv = None if solved is None else np.ascontiguousarray(solved, dtype=np.uint8)
        u = None if U is None else _f64(U)
        R = np.empty((self.ncells, NF))
        me = C.c_double()
        self._c(lib().mmf_compute_rhs_ex(self.ctx, None if sv is None else sv.ctypes.data_as(C.POINTER(C.c_uint8)),
                                         int(order), None if u is None else _dp(u), _dp(R), C.byref(me)))
        return R, me.value

    # one RK3 stage at a time
    def rk_rhs(self, state=0):
        me = C.c_double()
        self._c(lib().mmf_rk_rhs(self.ctx, int(state), C.byref(me)))
        return me.value

    def rk_dt(self, max_eig):
        dt = C.c_double()
        self._c(lib().mmf_rk_dt(self.ctx, float(max_eig), C.byref(dt)))
        return dt.value

    def rk_update(self, stage, dt):
        self._c(lib().mmf_rk_update(self.ctx, int(stage), float(dt)))

    def pack_cells(self, state, cells):
        c = np.ascontiguousarray(cells, dtype=np.int64)
        buf = np.empty((len(c), NF))
        self._c(lib().mmf_pack_cells(self.ctx, int(state), c.ctypes.data_as(C.POINTER(C.c_int64)), len(c), _dp(buf)))
        return buf

    def unpack_cells(self, state, cells, buf):
        c = np.ascontiguousarray(cells, dtype=np.int64)
        b = _f64(np.asarray(buf, dtype=np.float64).reshape(len(c), NF))
        self._c(lib().mmf_unpack_cells(self.ctx, int(state), c.ctypes.data_as(C.POINTER(C.c_int64)), len(c), _dp(b)))

    def device_state(self, state):
        """(device address, field stride) of a state's SoA planes."""
        p, fs = C.c_void_p(), C.c_int64()
        self._c(lib().mmf_get_device_state(self.ctx, int(state), C.byref(p), C.byref(fs)))
        return p.value, fs.value

    # time loop
    def set_time(self, t, tmax, cfl):
        self._c(lib().mmf_set_time(self.ctx, float(t), float(tmax), float(cfl)))

    def advance(self, nsteps):
        self._c(lib().mmf_advance(self.ctx, int(nsteps)))

    def set_step_graphs(self, on=True):
        self._c(lib().mmf_set_step_graphs(self.ctx, 1 if on else 0))

    def run(self, max_steps=-1):
        steps, t = C.c_int64(), C.c_double()
        self._c(lib().mmf_run(self.ctx, int(max_steps), C.byref(steps), C.byref(t)))
        return steps.value, t.value

    def time(self):
        t, dt, step = C.c_double(), C.c_double(), C.c_int64()
        eig = np.empty(3)
        self._c(lib().mmf_get_time(self.ctx, C.byref(t), C.byref(dt), C.byref(step), _dp(eig)))
        return t.value, dt.value, step.value, eig

    def synchronize(self):
        self._c(lib().mmf_synchronize(self.ctx))

    def nonfinite_steps(self):
        """Steps that met a NaN or infinite state (their time did not advance)."""
        n = C.c_int64()
        self._c(lib().mmf_get_nonfinite_steps(self.ctx, C.byref(n)))
        return n.value

    # output (solver_writer.cpp:56-119)
    def write_vtk(self, path, wait=True):
        """VTK image data (.vti) of the current state; wait=False returns at once
        and a host thread writes the file (see wait_writes)."""
        fn = lib().mmf_write_vtk if wait else lib().mmf_write_vtk_async
        self._c(fn(self.ctx, os.fsencode(path)))

    def wait_writes(self):
        self._c(lib().mmf_wait_writes(self.ctx))

    # diagnostics
    def error_norms(self, rho_exact):
        r = _f64(rho_exact).reshape(self.ncells)
        l1, l2 = C.c_double(), C.c_double()
        self._c(lib().mmf_error_norms(self.ctx, _dp(r), C.byref(l1), C.byref(l2)))
        return l1.value, l2.value

    def error_vs_exact(self, pb, t):
        l1, l2 = C.c_double(), C.c_double()
        self._c(lib().mmf_error_vs_exact(self.ctx, C.byref(pb), float(t), C.byref(l1), C.byref(l2)))
        return l1.value, l2.value

    def positivity_counters(self):
        a, b = C.c_int64(), C.c_int64()
        self._c(lib().mmf_get_positivity_counters(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def reset_positivity_counters(self):
        self._c(lib().mmf_reset_positivity_counters(self.ctx))

    def enable_timing(self, on=True):
        """on: True or 1, an event pair per stage launch (steps launched one
        by one); 2, one pair per advance call with the steps graph-replayed."""
        self._c(lib().mmf_enable_timing(self.ctx, int(on)))

    def read_timing(self):
        ms, n = C.c_double(), C.c_int64()
        self._c(lib().mmf_read_timing(self.ctx, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # learned face reconstruction (2D, order 2)
    def set_learned(self, net, sensor_threshold=0.05, enable=True):
        """Use the network's face values (net None: plain face values through
        the same face-state kernels); enable False restores the plain kernels."""
        self._net = net
        self._c(lib().mmf_set_learned_reconstruction(self.ctx, None if net is None else net.h,
                                                     float(sensor_threshold), 1 if enable else 0))

    def set_learned_options(self, clip=True, blend=1.0):
        self._c(lib().mmf_set_learned_options(self.ctx, 1 if clip else 0, float(blend)))

    def learned_counts(self):
        a, b = C.c_int64(), C.c_int64()
        self._c(lib().mmf_get_learned_counts(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def reset_learned_counts(self):
        self._c(lib().mmf_reset_learned_counts(self.ctx))

    def face_states(self):
        out = np.empty((self.ncells, 4, NF))
        self._c(lib().mmf_get_face_states(self.ctx, _dp(out)))
        return out

    def restrict_state(self, factor):
        n = self.ncells // (int(factor) ** 2)
        out = np.empty((n, NF))
        self._c(lib().mmf_restrict_state(self.ctx, int(factor), _dp(out)))
        return out


# ---- batched mode: many independent 2D grids -------------------------------------

def grid_desc(dim, n, origin, h, side_bc, side_info=None, order=1):
    d = GridDesc()
    d.dim, d.order, d.h = int(dim), int(order), float(h)
    n = list(n) + [1] * (3 - len(n))
    for a in range(3):
        d.n[a] = int(n[a])
        d.origin[a] = float(origin[a]) if a < len(origin) else 0.0
    for s in range(6):
        d.side_bc[s] = int(side_bc[s]) if s < len(side_bc) else BC_FREE_FLOW
        if side_info is not None:
            for k in range(NF):
                d.side_info[s][k] = float(side_info[s][k])
    return d


def border_face_count(desc):
    """Border faces of a descriptor's grid (the entries of set_border_bcs)."""
    c = C.c_int64()
    _check(lib().mmf_border_face_count(C.byref(desc), C.byref(c)))
    return c.value


def problem_border_bcs(pb, desc):
    """problem.cpp:404-477 at every border face centroid of desc's grid:
    (codes [count], Dirichlet primitives [count][5])."""
    m = border_face_count(desc)
    bc = np.empty(m, dtype=np.int32)
    info = np.empty((m, NF))
    _check(lib().mmf_problem_border_bcs(C.byref(pb), C.byref(desc), bc.ctypes.data_as(C.POINTER(C.c_int32)),
                                        _dp(info)))
    return bc, info


def _sides_uniform(desc, bc, info):
    n = [desc.n[a] for a in range(3)]
    off = 0
    for s in range(2 * desc.dim):
        a = s // 2
        cnt = n[1] * n[2] if a == 0 else (n[0] * n[2] if a == 1 else n[0] * n[1])
        b, i = bc[off:off + cnt], info[off:off + cnt]
        if desc.side_bc[2 * a] != BC_PERIODIC and (np.any(b != b[0]) or np.any(i != i[0])):
            return False
        off += cnt
    return True


def problem_grid_desc(pb, ncells, order=1):
    """Grid of computation() for a problem (main.cpp:147-149, 251-260)."""
    d = GridDesc()
    d.dim, d.order, d.h = pb.dim, int(order), pb.domain_length / int(ncells)
    for a in range(3):
        d.n[a] = int(ncells) if a < pb.dim else 1
        d.origin[a] = pb.domain_origin[a]
    _check(lib().mmf_problem_side_bcs(C.byref(pb), C.byref(d)))
    if not _sides_uniform(d, *problem_border_bcs(pb, d)):
        raise MMFError(ERR_INVALID, "the problem's border conditions change along a side; "
                                    "a descriptor holds one per side (Solver.for_problem applies them per face)")
    return d


class Batch:
    """Many independent 2D order-1 grids advanced together on the device.

    Each case keeps the reference's own time loop (own dt, own end time); one
    step of the whole batch is three persistent stage launches plus the time
    update, replayed from a captured HIP graph.  Per case the results are
    bitwise those of a single-grid Solver."""

    def __init__(self, descs):
        self.ncases = len(descs)
        arr = (GridDesc * self.ncases)(*descs)
        self.descs = list(descs)
        self.ncells = [int(d.n[0] * d.n[1]) for d in descs]
        h = C.c_void_p()
        _check(lib().mmf_batch_create(arr, self.ncases, C.byref(h)))
        self.h = h
        tc, ni = C.c_int64(), C.c_int32()
        self._c(lib().mmf_batch_sizes(self.h, C.byref(tc), C.byref(ni)))
        self.total_cells, self.nitems = tc.value, ni.value

    def close(self):
        if getattr(self, "h", None):
            lib().mmf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        if rc != OK:
            m = lib().mmf_batch_last_error(self.h) if getattr(self, "h", None) else None
            raise MMFError(rc, m.decode() if m else "")

    def set_state(self, i, U):
        U = _f64(U).reshape(self.ncells[i], NF)
        self._c(lib().mmf_batch_set_state(self.h, int(i), _dp(U)))

    def get_state(self, i):
        U = np.empty((self.ncells[i], NF))
        self._c(lib().mmf_batch_get_state(self.h, int(i), _dp(U)))
        return U

    def set_initial_states(self, problems):
        arr = (Problem * self.ncases)(*problems)
        self._c(lib().mmf_batch_set_initial_states(self.h, arr))

    def set_time(self, t0, tmax, cfl):
        t0 = _f64(np.broadcast_to(t0, (self.ncases,)))
        tm = _f64(np.broadcast_to(tmax, (self.ncases,)))
        cf = _f64(np.broadcast_to(cfl, (self.ncases,)))
        self._c(lib().mmf_batch_set_time(self.h, _dp(t0), _dp(tm), _dp(cf)))

    def set_shape(self, chunk=0, blocks=0):
        self._c(lib().mmf_batch_set_shape(self.h, int(chunk), int(blocks)))
        tc, ni = C.c_int64(), C.c_int32()
        self._c(lib().mmf_batch_sizes(self.h, C.byref(tc), C.byref(ni)))
        self.nitems = ni.value

    def shape(self):
        a, b = C.c_int32(), C.c_int32()
        self._c(lib().mmf_batch_get_shape(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_tail(self, rows):
        """Rows of the small items that end each stage (0: none)."""
        self._c(lib().mmf_batch_set_tail(self.h, int(rows)))
        self.set_shape(0, 0)

    def tail(self):
        r = C.c_int32()
        self._c(lib().mmf_batch_get_tail(self.h, C.byref(r)))
        return r.value

    def advance(self, nsteps):
        self._c(lib().mmf_batch_advance(self.h, int(nsteps)))

    def run(self, max_steps=-1):
        n = C.c_int64()
        self._c(lib().mmf_batch_run(self.h, int(max_steps), C.byref(n)))
        return n.value

    def times(self):
        t = np.empty(self.ncases)
        st = np.empty(self.ncases, dtype=np.int64)
        self._c(lib().mmf_batch_get_times(self.h, _dp(t), st.ctypes.data_as(C.POINTER(C.c_int64))))
        return t, st

    def error_vs_exact(self, problems):
        arr = (Problem * self.ncases)(*problems)
        l1, l2 = np.empty(self.ncases), np.empty(self.ncases)
        self._c(lib().mmf_batch_error_vs_exact(self.h, arr, _dp(l1), _dp(l2)))
        return l1, l2

    def enable_timing(self, on=True):
        self._c(lib().mmf_batch_enable_timing(self.h, 1 if on else 0))

    def read_timing(self):
        ms, n = C.c_double(), C.c_int64()
        self._c(lib().mmf_batch_read_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def autotune(self, steps_per_trial=3):
        a, b, ms = C.c_int32(), C.c_int32(), C.c_double()
        self._c(lib().mmf_batch_autotune(self.h, int(steps_per_trial), C.byref(a), C.byref(b), C.byref(ms)))
        self.set_shape(0, 0)
        return a.value, b.value, ms.value


# ---- learned face reconstruction -----------------------------------------------------

LEARN_NPARAMS, LEARN_NIN, LEARN_NOUT = 11664, 100, 16


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


class Learner:
    """The network of the learned face reconstruction, its Adam state and a
    device-resident training set (include/mmf.h, mmf_learn_*)."""

    def __init__(self, seed=0, capacity=0):
        h = C.c_void_p()
        _check(lib().mmf_learn_create(int(seed) & ((1 << 64) - 1), int(capacity), C.byref(h)))
        self.h = h
        self.capacity = int(capacity)

    def close(self):
        if getattr(self, "h", None):
            lib().mmf_learn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        if rc != OK:
            m = lib().mmf_learn_last_error(self.h) if getattr(self, "h", None) else None
            raise MMFError(rc, m.decode() if m else "")

    def params(self):
        p = np.empty(LEARN_NPARAMS, dtype=np.float32)
        self._c(lib().mmf_learn_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = _f32(p, (LEARN_NPARAMS,))
        self._c(lib().mmf_learn_set_params(self.h, _fp(p)))

    def collect(self, fine, factor, cells, sensor_threshold=0.05):
        c = np.ascontiguousarray(cells, dtype=np.int64)
        self._c(lib().mmf_learn_collect(self.h, fine.ctx, int(factor), c.ctypes.data_as(C.POINTER(C.c_int64)),
                                        c.size, float(sensor_threshold)))

    def num_samples(self):
        n = C.c_int64()
        self._c(lib().mmf_learn_num_samples(self.h, C.byref(n)))
        return n.value

    def samples(self, first=0, count=None):
        if count is None:
            count = self.num_samples() - first
        x = np.empty((count, LEARN_NIN), dtype=np.float32)
        t = np.empty((count, LEARN_NOUT), dtype=np.float32)
        w = np.empty(count, dtype=np.float32)
        self._c(lib().mmf_learn_get_samples(self.h, int(first), int(count), _fp(x), _fp(t), _fp(w)))
        return x, t, w

    def set_samples(self, x, t, w):
        x, t, w = _f32(x, (-1, LEARN_NIN)), _f32(t, (-1, LEARN_NOUT)), _f32(w, (-1,))
        self._c(lib().mmf_learn_set_samples(self.h, _fp(x), _fp(t), _fp(w), x.shape[0]))

    def loss_grad(self, x, t, w):
        x, t, w = _f32(x, (-1, LEARN_NIN)), _f32(t, (-1, LEARN_NOUT)), _f32(w, (-1,))
        loss = C.c_double()
        g = np.empty(LEARN_NPARAMS, dtype=np.float32)
        self._c(lib().mmf_learn_loss_grad(self.h, _fp(x), _fp(t), _fp(w), x.shape[0], C.byref(loss), _fp(g)))
        return loss.value, g

    def adam_step(self, grad, lr):
        g = _f32(grad, (LEARN_NPARAMS,))
        self._c(lib().mmf_learn_adam_step(self.h, _fp(g), float(lr)))

    def train(self, steps, batch=8192, lr0=2e-3, lr1=2e-5, seed=1):
        loss = C.c_double()
        self._c(lib().mmf_learn_train(self.h, int(steps), int(batch), float(lr0), float(lr1),
                                      int(seed) & ((1 << 64) - 1), C.byref(loss)))
        return loss.value

    def forward(self, x, mfma=True):
        x = _f32(x, (-1, LEARN_NIN))
        y = np.empty((x.shape[0], LEARN_NOUT), dtype=np.float32)
        self._c(lib().mmf_learn_forward(self.h, _fp(x), x.shape[0], 1 if mfma else 0, _fp(y)))
        return y


# ---- general mesh: an explicit interface list ------------------------------------

class Mesh:
    """Device-resident solver of an explicit face list (the general
    MeshGeometricalInfo); fm is a meshes.FaceMesh or anything with its arrays."""

    def __init__(self, fm):
        d = MeshDesc()
        self._keep = [np.ascontiguousarray(fm.volume, dtype=np.float64),
                      np.ascontiguousarray(fm.owner, dtype=np.int64), np.ascontiguousarray(fm.neigh, dtype=np.int64),
                      np.ascontiguousarray(fm.area, dtype=np.float64),
                      np.ascontiguousarray(fm.normal, dtype=np.float64).reshape(-1, 3),
                      np.ascontiguousarray(np.maximum(fm.face_bc, 0), dtype=np.int32),
                      np.ascontiguousarray(fm.face_info, dtype=np.float64).reshape(-1, NF)]
        vol, own, nei, ar, nrm, fbc, finfo = self._keep
        d.dim = int(fm.dim)
        d.ncells = len(vol)
        d.nfaces = len(own)
        d.cell_volume = _dp(vol)
        d.min_cell_size = float(fm.min_cell_size)
        d.face_owner = own.ctypes.data_as(C.POINTER(C.c_int64))
        d.face_neigh = nei.ctypes.data_as(C.POINTER(C.c_int64))
        d.face_area = _dp(ar)
        d.face_normal = _dp(nrm)
        d.face_bc = fbc.ctypes.data_as(C.POINTER(C.c_int32))
        d.face_info = _dp(finfo)
        h = C.c_void_p()
        _check(lib().mmf_mesh_create(C.byref(d), C.byref(h)))
        self.h = h
        self.ncells, self.nfaces = d.ncells, d.nfaces
        self.dim = d.dim

    def close(self):
        if getattr(self, "h", None):
            lib().mmf_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        if rc != OK:
            m = lib().mmf_mesh_last_error(self.h)
            raise MMFError(rc, m.decode() if m else "")

    def set_cell_flags(self, fluid, solved=None):
        f = np.ascontiguousarray(fluid, dtype=np.uint8)
        s = None if solved is None else np.ascontiguousarray(solved, dtype=np.uint8)
        self._c(lib().mmf_mesh_set_cell_flags(self.h, f.ctypes.data_as(C.POINTER(C.c_uint8)),
                                              None if s is None else s.ctypes.data_as(C.POINTER(C.c_uint8))))

    def set_state(self, U):
        U = _f64(U).reshape(self.ncells, NF)
        self._c(lib().mmf_mesh_set_state(self.h, _dp(U)))

    def get_state(self):
        U = np.empty((self.ncells, NF))
        self._c(lib().mmf_mesh_get_state(self.h, _dp(U)))
        return U

    def compute_rhs(self):
        R = np.empty((self.ncells, NF))
        me = C.c_double()
        self._c(lib().mmf_mesh_compute_rhs(self.h, _dp(R), C.byref(me)))
        return R, me.value

    def set_time(self, t, tmax, cfl):
        self._c(lib().mmf_mesh_set_time(self.h, float(t), float(tmax), float(cfl)))

    def advance(self, nsteps):
        self._c(lib().mmf_mesh_advance(self.h, int(nsteps)))

    def set_step_graphs(self, on=True):
        self._c(lib().mmf_mesh_set_step_graphs(self.h, 1 if on else 0))

    def run(self, max_steps=-1):
        steps, t = C.c_int64(), C.c_double()
        self._c(lib().mmf_mesh_run(self.h, int(max_steps), C.byref(steps), C.byref(t)))
        return steps.value, t.value

    def time(self):
        t, dt, step = C.c_double(), C.c_double(), C.c_int64()
        eig = np.empty(3)
        self._c(lib().mmf_mesh_get_time(self.h, C.byref(t), C.byref(dt), C.byref(step), _dp(eig)))
        return t.value, dt.value, step.value, eig

    def counters(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._c(lib().mmf_mesh_get_counters(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def enable_timing(self, on=True):
        self._c(lib().mmf_mesh_enable_timing(self.h, 1 if on else 0))

    def read_timing(self):
        f, c, n = C.c_double(), C.c_double(), C.c_int64()
        self._c(lib().mmf_mesh_read_timing(self.h, C.byref(f), C.byref(c), C.byref(n)))
        return f.value, c.value, n.value
