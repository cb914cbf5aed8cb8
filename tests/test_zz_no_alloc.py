"""mmf_compute_rhs_noalloc: the RHS on caller-owned device buffers with no
allocation in the library, against mmf_compute_rhs on a context.

Calls go through the C-ABI with device buffers from the HIP runtime.  Each
sequence makes calls in a row on one GPU, later ones on smaller grids with a
smaller eigenvalue maximum and no negative states, so anything the shared
scratch kept from an earlier call would show in the later results.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(int(nbytes), 8)) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free_all(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def state(mmf, s, seed, speed, neg=()):
    rng = np.random.default_rng(seed)
    cen = s.centroids()
    r = np.sqrt((cen ** 2).sum(1))
    prim = np.empty((s.ncells, 5))
    prim[:, 0] = np.where(r < 0.45, 1.0, 0.2) * (1 + 0.05 * rng.uniform(-1, 1, s.ncells))
    prim[:, 1:4] = speed * rng.uniform(-1, 1, (s.ncells, 3))
    if s.dim == 2:
        prim[:, 3] = 0.0
    prim[:, 4] = np.where(r < 0.45, 1.0, 0.8) * (1 + 0.05 * rng.uniform(-1, 1, s.ncells))
    U = mmf.batch_prim2cons(prim)
    for c in neg:
        U[c] = [-0.5, 0.0, 0.0, 0.0, -1.0]  # rho < 0 and p = rho T < 0 with T = 0.8 > 0 (finite eigenvalue)
    return U


def reference(mmf, s, U, fluid):
    if fluid is not None:
        s.set_cell_flags(fluid)
    s.set_state(U)
    s.reset_positivity_counters()
    R, me = s.compute_rhs()
    return R, me, s.positivity_counters()


def noalloc(mmf, hip, s, U, fluid, pad):
    N = s.ncells
    fs = (N + 63) // 64 * 64 + pad
    soa = np.zeros((5, fs))
    soa[:, :N] = U.T
    dU = hip.put(soa)
    drhs = hip.put(np.full((5, fs), 7.0))  # stale contents the call must overwrite
    dme = hip.put(np.full(1, 1e300))
    dcnt = hip.put(np.full(2, 12345, dtype=np.int64))
    dfl = None
    if fluid is not None:
        fl = np.where(fluid != 0, mmf.CELL_FLUID | mmf.CELL_SOLVED, 0).astype(np.uint8)
        dfl = hip.put(fl)
    rc = mmf.lib().mmf_compute_rhs_noalloc(C.byref(s.desc), dfl, dU, fs, drhs, dme, dcnt, None)
    assert rc == 0, mmf.lib().mmf_noalloc_last_error()
    hip.sync()
    R = hip.get(drhs, (5, fs), np.float64)[:, :N].T.copy()
    me = float(hip.get(dme, (1,), np.float64)[0])
    cnt = tuple(int(x) for x in hip.get(dcnt, (2,), np.int64))
    return R, me, cnt


def grids(mmf):
    """(solver, fluid mask or None, state) in call order: sizes and maxima shrink."""
    out = []
    # 128 x 128 x 64 = 2^20 cells: the 3D marching kernel; negative densities; fast
    s = mmf.Solver(3, [128, 128, 64], [-1.0, -1.0, -1.0], 2.0 / 128, [0, 0, 3, 1, 4, 4],
                   [[1.0, 0.3, -0.1, 0.05, 1.0]] * 6, 1)
    out.append((s, None, state(mmf, s, 1, 0.9, neg=[5, 70000, 900001])))
    # 3D order 2 with a box body (cell flags: the gather kernel); slower
    s = mmf.Solver(3, [40, 24, 20], [-1.0, -1.0, -1.0], 2.0 / 40, [3, 0, 1, 1, 0, 2],
                   [[1.0, 0.3, -0.1, 0.05, 1.0]] * 6, 2)
    cen = s.centroids()
    box = np.array([-0.4, -0.3, -0.5, 0.1, 0.35, 0.2])
    fluid = (~np.all((cen >= box[:3]) & (cen <= box[3:]), axis=1)).astype(np.uint8)
    out.append((s, fluid, state(mmf, s, 2, 0.5)))
    # 2D order 1, periodic
    s = mmf.Solver(2, [70, 33], [-1.0, -1.0, 0.0], 2.0 / 70, [4, 4, 4, 4], None, 1)
    out.append((s, None, state(mmf, s, 3, 0.2)))
    return out


def test_noalloc_equals_context_rhs(mmf):
    hip = Hip()
    try:
        cases = grids(mmf)
        refs = [reference(mmf, s, U, fl) for s, fl, U in cases]
        assert refs[0][2][1] > 0 and refs[1][2] == (0, 0)
        assert refs[0][1] > refs[1][1] > refs[2][1]
        for rnd in range(2):  # twice through, large then smaller grids each time
            for i, ((s, fl, U), (Rr, mr, cr)) in enumerate(zip(cases, refs)):
                R, me, cnt = noalloc(mmf, hip, s, U, fl, pad=64 * rnd)
                print("noalloc call: grid %d round %d: %d cells, maxEig %.6f" % (i, rnd, s.ncells, me), flush=True)
                assert np.array_equal(R, Rr), (i, rnd)
                assert me == mr, (i, rnd)
                assert cnt == cr, (i, rnd)
            hip.free_all()
    finally:
        hip.free_all()
        for s, _, _ in locals().get("cases", []):
            s.close()


def test_noalloc_rejects_what_it_cannot_take(mmf):
    hip = Hip()
    try:
        s = mmf.Solver(2, [70, 33], [-1.0, -1.0, 0.0], 2.0 / 70, [0] * 4, None, 1)
        dU = hip.alloc(8 * 5 * 2368)
        dR = hip.alloc(8 * 5 * 2368)
        L = mmf.lib()
        assert L.mmf_compute_rhs_noalloc(C.byref(s.desc), None, dU, 2300, dR, None, None, None) == mmf.ERR_INVALID
        assert b"multiple of 64" in L.mmf_noalloc_last_error()
        assert L.mmf_compute_rhs_noalloc(C.byref(s.desc), None, dU, 2240, dR, None, None, None) == mmf.ERR_INVALID
        d = mmf.grid_desc(2, [1 << 31, 2], [0.0, 0.0, 0.0], 1.0, [0] * 6)
        assert L.mmf_compute_rhs_noalloc(C.byref(d), None, dU, 1 << 32, dR, None, None, None) == mmf.ERR_INVALID
        assert L.mmf_compute_rhs_noalloc(None, None, dU, 2368, dR, None, None, None) == mmf.ERR_INVALID
        s.close()
    finally:
        hip.free_all()
