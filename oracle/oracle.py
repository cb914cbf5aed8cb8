"""ctypes front-end of the CPU oracle (oracle/mmf_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It builds oracle/_build/liboracle.so with oracle/Makefile when
the library is missing.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

NF = 5
BC_NONE, BC_FREE_FLOW, BC_REFLECTING, BC_WALL, BC_DIRICHLET, BC_PERIODIC = -1, 0, 1, 2, 3, 4
(VORTEX_XY, VORTEX_ZX, VORTEX_YZ, RADSOD, SOD_X, SOD_Y, SOD_Z, FFSTEP) = range(8)
PROBLEM_KEYS = {"vortex_xy": VORTEX_XY, "vortex_zx": VORTEX_ZX, "vortex_yz": VORTEX_YZ,
                "radsod": RADSOD, "sod3d_x": SOD_X, "sod3d_y": SOD_Y, "sod3d_z": SOD_Z,
                "ffstep": FFSTEP}


class Problem(C.Structure):
    _fields_ = [
        ("type", C.c_int), ("dim", C.c_int),
        ("vortex_beta", C.c_double), ("vortex_center", C.c_double * 3),
        ("vortex_periodic", C.c_int),
        ("domain_origin", C.c_double * 3), ("domain_length", C.c_double),
        ("sod_r0", C.c_double), ("sod_p_in", C.c_double), ("sod_t_in", C.c_double),
        ("sod_p_out", C.c_double), ("sod_t_out", C.c_double),
        ("fs_prim", C.c_double * 5),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def build_variant(defines, outdir):
    """Build a copy of the oracle with extra -D flags (used to plant a mistake)."""
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "liboracle_%s.so" % "_".join(d.lower() for d in defines))
    cmd = ["gcc", "-O3", "-std=gnu99", "-fPIC", "-ffp-contract=off", "-fopenmp", "-shared"]
    cmd += ["-D" + d for d in defines]
    cmd += ["-o", out, os.path.join(HERE, "mmf_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True)
    return out


class use_library:
    """Context manager: route this module's calls to another build of the oracle."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = _lib
        _lib = None
        _load(self.path)
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved


def lib():
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    return _load(LIB_PATH)


def _load(path):
    global _lib
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    i64p = C.POINTER(C.c_int64)
    L.orc_problem_defaults.argtypes = [C.POINTER(Problem), C.c_int, C.c_int]
    L.orc_cons2prim.argtypes = [dp, dp]
    L.orc_prim2cons.argtypes = [dp, dp]
    L.orc_normal_velocity.argtypes = [dp, dp]
    L.orc_normal_velocity.restype = C.c_double
    L.orc_eval_fluxes.argtypes = [dp, dp, dp, dp, i64p, i64p]
    L.orc_eval_splitting.argtypes = [dp, dp, dp, dp, dp, i64p, i64p]
    L.orc_eval_bc_values.argtypes = [C.c_int, dp, dp, dp, dp]
    L.orc_exact_conservatives.argtypes = [C.POINTER(Problem), dp, C.c_double, dp]
    L.orc_solver_alloc.restype = C.c_void_p
    L.orc_solver_dealloc.argtypes = [C.c_void_p]
    L.orc_solver_init.argtypes = [C.c_void_p, C.POINTER(Problem), C.c_int, i64p, dp, C.c_double,
                                  C.POINTER(C.c_int), dp, C.POINTER(C.c_uint8)]
    L.orc_solver_init.restype = C.c_int
    L.orc_solver_ncells.argtypes = [C.c_void_p]
    L.orc_solver_ncells.restype = C.c_int64
    L.orc_solver_nfaces.argtypes = [C.c_void_p]
    L.orc_solver_nfaces.restype = C.c_int64
    L.orc_solver_counters.argtypes = [C.c_void_p, i64p, i64p]
    L.orc_solver_cell_centroids.argtypes = [C.c_void_p, dp]
    L.orc_solver_flags.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]
    L.orc_solver_set_solved.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
    L.orc_solver_set_solved.restype = C.c_int
    L.orc_solver_face_bc.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.orc_initial_state.argtypes = [C.c_void_p, dp]
    L.orc_compute_polynomials.argtypes = [C.c_void_p, dp]
    L.orc_compute_rhs.argtypes = [C.c_void_p, dp, dp, dp]
    L.orc_compute_rhs.restype = C.c_int
    L.orc_rk3_step.argtypes = [C.c_void_p, dp, dp, dp, C.c_double, C.c_double, C.c_double, dp]
    L.orc_rk3_step.restype = C.c_double
    L.orc_rk3_step_threads.argtypes = [C.c_void_p, dp, dp, dp, C.c_double, C.c_double, C.c_double, dp, C.c_int]
    L.orc_rk3_step_threads.restype = C.c_double
    L.orc_compute_rhs_threads.argtypes = [C.c_void_p, dp, dp, dp, C.c_int]
    L.orc_compute_rhs_threads.restype = C.c_int
    L.orc_compute_rhs_faces_threads.argtypes = [C.c_void_p, dp, dp, dp, C.c_int]
    L.orc_compute_rhs_faces_threads.restype = C.c_int
    L.orc_run.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_int64, dp]
    L.orc_run.restype = C.c_int64
    L.orc_error.argtypes = [C.c_void_p, dp, C.c_double, dp, dp]
    L.orc_reconstruct.argtypes = [C.c_void_p, C.c_int64, dp, dp, dp]
    L.orc_mark_fluid.argtypes = [C.c_void_p, dp, C.c_int, dp, C.c_int, C.POINTER(C.c_uint8)]
    L.orc_solver_init_faces.argtypes = [C.c_void_p, C.POINTER(Problem), C.c_int, C.c_int64, C.c_int64, dp, dp,
                                        C.c_double, i64p, i64p, dp, dp, C.POINTER(C.c_int), dp,
                                        C.POINTER(C.c_uint8)]
    L.orc_solver_init_faces.restype = C.c_int
    L.orc_solver_face_list.argtypes = [C.c_void_p, i64p, i64p, dp, dp]
    L.orc_solver_cell_volumes.argtypes = [C.c_void_p, dp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def problem(ptype, dim, **kw):
    """Problem with the reference defaults (problem.cpp) and optional overrides."""
    pb = Problem()
    lib().orc_problem_defaults(C.byref(pb), int(ptype), int(dim))
    for k, v in kw.items():
        if isinstance(v, (list, tuple, np.ndarray)):
            arr = getattr(pb, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(pb, k, v)
    return pb


# ---- point functions -------------------------------------------------------

def cons2prim(c):
    c = _f64(c)
    p = np.empty(NF)
    lib().orc_cons2prim(_dp(c), _dp(p))
    return p


def prim2cons(p):
    p = _f64(p)
    c = np.empty(NF)
    lib().orc_prim2cons(_dp(p), _dp(c))
    return c


def eval_fluxes(cons, prim, n):
    cons, prim, n = _f64(cons), _f64(prim), _f64(n)
    f = np.empty(NF)
    lib().orc_eval_fluxes(_dp(cons), _dp(prim), _dp(n), _dp(f), None, None)
    return f


def eval_splitting(L, R, n):
    L, R, n = _f64(L), _f64(R), _f64(n)
    F = np.empty(NF)
    lam = C.c_double()
    lib().orc_eval_splitting(_dp(L), _dp(R), _dp(n), _dp(F), C.byref(lam), None, None)
    return F, lam.value


def eval_bc_values(bc, normal, cons, info=None):
    normal, cons = _f64(normal), _f64(cons)
    info = _f64(np.zeros(NF) if info is None else info)
    out = np.empty(NF)
    lib().orc_eval_bc_values(int(bc), _dp(normal), _dp(info), _dp(cons), _dp(out))
    return out


def exact_conservatives(pb, point, t):
    point = _f64(point)
    out = np.empty(NF)
    lib().orc_exact_conservatives(C.byref(pb), _dp(point), float(t), _dp(out))
    return out


# ---- solver ------------------------------------------------------------------

class Solver:
    """Oracle solver on a uniform Cartesian grid (mesh_info closed forms)."""

    def __init__(self, pb, order, n, origin, h, side_bc=None, side_info=None, fluid=None):
        L = lib()
        self.pb = pb
        self.dim = pb.dim
        n = list(n) + [1] * (3 - len(n))
        self.n = tuple(int(x) for x in n)
        self._n = np.array(self.n, dtype=np.int64)
        self._origin = _f64(origin)
        self._side_bc = None if side_bc is None else np.ascontiguousarray(side_bc, dtype=np.int32)
        self._side_info = None if side_info is None else _f64(side_info)
        self._fluid = None if fluid is None else np.ascontiguousarray(fluid, dtype=np.uint8)
        self.h = float(h)
        self.s = L.orc_solver_alloc()
        rc = L.orc_solver_init(
            self.s, C.byref(pb), int(order), self._n.ctypes.data_as(C.POINTER(C.c_int64)),
            _dp(self._origin), self.h,
            None if self._side_bc is None else self._side_bc.ctypes.data_as(C.POINTER(C.c_int)),
            None if self._side_info is None else _dp(self._side_info),
            None if self._fluid is None else self._fluid.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            L.orc_solver_dealloc(self.s)
            self.s = None
            raise ValueError("oracle solver init failed (%d)" % rc)
        self.ncells = L.orc_solver_ncells(self.s)
        self.nfaces = L.orc_solver_nfaces(self.s)

    @classmethod
    def from_faces(cls, pb, dim, volume, owner, neigh, area, normal, min_cell_size, border_bc=None,
                   border_info=None, fluid=None, centroids=None):
        """Oracle solver on an explicit interface list (order 1): the general
        MeshGeometricalInfo; the face loop runs in the given face order."""
        L = lib()
        self = cls.__new__(cls)
        self.pb = pb
        self.dim = dim
        self._keep = [np.ascontiguousarray(volume, dtype=np.float64),
                      np.ascontiguousarray(owner, dtype=np.int64), np.ascontiguousarray(neigh, dtype=np.int64),
                      np.ascontiguousarray(area, dtype=np.float64),
                      np.ascontiguousarray(normal, dtype=np.float64).reshape(-1, 3),
                      None if border_bc is None else np.ascontiguousarray(border_bc, dtype=np.int32),
                      None if border_info is None else np.ascontiguousarray(border_info, dtype=np.float64),
                      None if fluid is None else np.ascontiguousarray(fluid, dtype=np.uint8),
                      None if centroids is None else np.ascontiguousarray(centroids, dtype=np.float64)]
        vol, own, nei, ar, nrm, bbc, binfo, fl, cen = self._keep
        self.h = float(min_cell_size)
        self.s = L.orc_solver_alloc()
        rc = L.orc_solver_init_faces(
            self.s, C.byref(pb), int(dim), len(vol), len(own), _dp(vol), None if cen is None else _dp(cen),
            self.h, own.ctypes.data_as(C.POINTER(C.c_int64)), nei.ctypes.data_as(C.POINTER(C.c_int64)), _dp(ar),
            _dp(nrm), None if bbc is None else bbc.ctypes.data_as(C.POINTER(C.c_int)),
            None if binfo is None else _dp(binfo), None if fl is None else fl.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            L.orc_solver_dealloc(self.s)
            self.s = None
            raise ValueError("oracle face-list solver init failed (%d)" % rc)
        self.ncells = L.orc_solver_ncells(self.s)
        self.nfaces = L.orc_solver_nfaces(self.s)
        return self

    def face_list(self):
        """(owner, neighbour, normal [nfaces][3], area) of this solver's faces."""
        o = np.empty(self.nfaces, dtype=np.int64)
        n = np.empty(self.nfaces, dtype=np.int64)
        nrm = np.empty((self.nfaces, 3))
        a = np.empty(self.nfaces)
        lib().orc_solver_face_list(self.s, o.ctypes.data_as(C.POINTER(C.c_int64)),
                                   n.ctypes.data_as(C.POINTER(C.c_int64)), _dp(nrm), _dp(a))
        return o, n, nrm, a

    def cell_volumes(self):
        v = np.empty(self.ncells)
        lib().orc_solver_cell_volumes(self.s, _dp(v))
        return v

    def __del__(self):
        if getattr(self, "s", None):
            lib().orc_solver_dealloc(self.s)
            self.s = None

    def initial_state(self):
        U = np.empty((self.ncells, NF))
        lib().orc_initial_state(self.s, _dp(U))
        return U

    def centroids(self):
        out = np.empty((self.ncells, 3))
        lib().orc_solver_cell_centroids(self.s, _dp(out))
        return out

    def flags(self):
        f = np.empty(self.ncells, dtype=np.uint8)
        s = np.empty(self.ncells, dtype=np.uint8)
        lib().orc_solver_flags(self.s, f.ctypes.data_as(C.POINTER(C.c_uint8)),
                               s.ctypes.data_as(C.POINTER(C.c_uint8)))
        return f, s

    def set_solved(self, solved):
        """Cells solved here (main.cpp:231-235: ghosts fluid but unsolved)."""
        sv = np.ascontiguousarray(solved, dtype=np.uint8)
        if lib().orc_solver_set_solved(self.s, sv.ctypes.data_as(C.POINTER(C.c_uint8))) != 0:
            raise ValueError("a solved cell must be fluid")

    def face_bc(self):
        out = np.empty(self.nfaces, dtype=np.int32)
        lib().orc_solver_face_bc(self.s, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def counters(self):
        a, b = C.c_int64(), C.c_int64()
        lib().orc_solver_counters(self.s, C.byref(a), C.byref(b))
        return a.value, b.value

    def compute_rhs(self, U, threads=1):
        """Residual and maximum eigenvalue; threads > 1 runs the threaded face
        loop (bit for bit the serial one)."""
        U = _f64(U)
        R = np.empty((self.ncells, NF))
        me = C.c_double()
        lib().orc_compute_polynomials(self.s, _dp(U))
        if threads > 1 and getattr(self, "_keep", None) is not None:
            lib().orc_compute_rhs_faces_threads(self.s, _dp(U), _dp(R), C.byref(me), int(threads))
        elif threads > 1:
            lib().orc_compute_rhs_threads(self.s, _dp(U), _dp(R), C.byref(me), int(threads))
        else:
            lib().orc_compute_rhs(self.s, _dp(U), _dp(R), C.byref(me))
        return R, me.value

    def face_states(self, U):
        """Order-2 face values of every cell, faces x-, x+, y-, y+ (, z-, z+):
        orc_compute_polynomials + orc_reconstruct per face."""
        U = _f64(U)
        L = lib()
        L.orc_compute_polynomials(self.s, _dp(U))
        nfc = 2 * self.dim
        out = np.empty((self.ncells, nfc, NF))
        v = np.empty(NF)
        for c in range(self.ncells):
            for f in range(nfc):
                n = np.zeros(3)
                n[f // 2] = -1.0 if f % 2 == 0 else 1.0
                L.orc_reconstruct(self.s, c, _dp(n), _dp(U), _dp(v))
                out[c, f] = v
        return out

    def rk3_step(self, U, W, cfl, t, tmax, threads=1):
        """In-place step of U (and work W); returns (dt, stage max eigenvalues).
        threads > 1: the threaded step (bit for bit the serial one)."""
        R = np.empty((self.ncells, NF))
        eig = np.empty(3)
        if threads > 1:
            dt = lib().orc_rk3_step_threads(self.s, _dp(U), _dp(W), _dp(R), float(cfl), float(t),
                                            float(tmax), _dp(eig), int(threads))
        else:
            dt = lib().orc_rk3_step(self.s, _dp(U), _dp(W), _dp(R), float(cfl), float(t),
                                    float(tmax), _dp(eig))
        return dt, eig

    def run(self, U, cfl, t0, tmax, max_steps=-1):
        t = C.c_double()
        steps = lib().orc_run(self.s, _dp(U), float(cfl), float(t0), float(tmax),
                              int(max_steps), C.byref(t))
        return steps, t.value

    def error(self, U, t):
        U = _f64(U)
        l1, l2 = C.c_double(), C.c_double()
        lib().orc_error(self.s, _dp(U), float(t), C.byref(l1), C.byref(l2))
        return l1.value, l2.value


# ---- the reference's own test cases ---------------------------------------------

def reference_case(ptype, dim, ncells, tend, cfl=0.45, order=1):
    """Configuration exactly as computation() builds it (main.cpp:77-155)."""
    pb = problem(ptype, dim)
    origin = list(pb.domain_origin)
    length = pb.domain_length
    h = length / ncells
    n = [ncells] * dim
    return pb, order, n, origin, h, cfl, 0.0, tend


def run_reference_case(ptype, dim, ncells, tend, cfl=0.45):
    pb, order, n, origin, h, cfl, t0, tend = reference_case(ptype, dim, ncells, tend, cfl)
    s = Solver(pb, order, n, origin, h)
    U = s.initial_state()
    steps, t = s.run(U, cfl, t0, tend)
    l1, l2 = s.error(U, tend)
    return l1, l2, steps, s, U
