"""Pins of the oracle's immersed-body path, sphere mask and L2 norm (CPU).

The reference's known answers have no body, so the body branch of the
interface loop (main.cpp:251-277 wall faces; euler.cpp:181-247 skipped faces,
the virtual state in the non-fluid slot, the flipped normal for the boundary
evaluation) is pinned here by an identity the mathematics fixes:

  a solid slab of cells at one end of an axis leaves, bit for bit, the
  truncated domain with a REFLECTING border on that side.

Why bit for bit: the wall state is the reflecting state (euler.cpp:347-356);
the mirror u - 2(u.n)n is the same for n and -n; the flux F(virtual, fluid, n)
of a face whose owner is solid is exactly -F(fluid, virtual, -n) (the Rusanov
flux is exactly antisymmetric); x - a*F == x + a*(-F); faces between two
solid cells are skipped, so the eigenvalue maximum and dt see the same faces;
and each fluid cell accumulates its faces in the same order in both grids.
A slab on the low side makes the solid cell the face OWNER (the virtual state
goes to the owner slot and the boundary normal is flipped); one on the high
side makes it the NEIGHBOUR.  Planted mistakes in copies of the oracle show
that each plausible slip in that branch fails the identity.
"""
import numpy as np
import pytest

GAMMA = 1.4


def random_state(orc, rng, ncells, dim, speed=0.6):
    prim = np.empty((ncells, 5))
    prim[:, 0] = rng.uniform(0.4, 2.0, ncells)
    prim[:, 1:4] = rng.uniform(-speed, speed, (ncells, 3))
    if dim == 2:
        prim[:, 3] = 0.0
    prim[:, 4] = rng.uniform(0.5, 2.0, ncells)
    return np.array([orc.prim2cons(p) for p in prim])


def slab_pair(orc, dim, axis, side, nslab, n_fluid, side_bc, seed, order=1):
    """(full grid with a solid slab, truncated grid, fluid mask, fluid index map)."""
    pb = orc.problem(orc.RADSOD, dim)
    h = 0.25
    n_full = list(n_fluid)
    n_full[axis] += nslab
    origin_t = [0.0, 0.0, 0.0]
    origin_f = [0.0, 0.0, 0.0]
    if side == 0:
        origin_f[axis] = -nslab * h
    idx = np.indices(n_full[::-1]).reshape(dim, -1)[::-1]  # idx[a] = coordinate along axis a, cells x-fastest
    q = idx[axis]
    fluid = ((q >= nslab) if side == 0 else (q < n_fluid[axis])).astype(np.uint8)
    bc_t = list(side_bc)
    bc_t[2 * axis + side] = orc.BC_REFLECTING
    info = np.tile([0.9, 0.3, -0.2, 0.1 if dim == 3 else 0.0, 1.1], (6, 1))  # Dirichlet primitives
    full = orc.Solver(pb, order, n_full, origin_f, h, side_bc=side_bc, side_info=info, fluid=fluid)
    trunc = orc.Solver(pb, order, n_fluid, origin_t, h, side_bc=bc_t, side_info=info)
    rng = np.random.default_rng(seed)
    Ut = random_state(orc, rng, trunc.ncells, dim)
    Uf = random_state(orc, rng, full.ncells, dim, speed=3.0)  # solid cells: fast, so a stray face would move dt
    Uf[fluid == 1] = Ut
    return full, trunc, fluid, Uf, Ut


CASES = [
    (2, 0, 0), (2, 0, 1), (2, 1, 0), (2, 1, 1),
    (3, 0, 0), (3, 1, 1), (3, 2, 0), (3, 2, 1),
]


def _bcs(orc, dim):
    bc = [orc.BC_FREE_FLOW, orc.BC_REFLECTING, orc.BC_FREE_FLOW, orc.BC_DIRICHLET, orc.BC_REFLECTING,
          orc.BC_FREE_FLOW]
    return bc[:2 * dim] + [orc.BC_FREE_FLOW] * (6 - 2 * dim)


def _n_fluid(dim):
    return [5, 4] if dim == 2 else [4, 3, 5]


@pytest.mark.parametrize("dim,axis,side", CASES)
@pytest.mark.parametrize("order", [1, 2])
def test_solid_slab_is_a_reflecting_border(orc, dim, axis, side, order):
    nf = _n_fluid(dim)
    full, trunc, fluid, Uf, Ut = slab_pair(orc, dim, axis, side, 2, nf, _bcs(orc, dim), 10 * dim + 3 * axis + side,
                                           order)
    fb = full.face_bc()
    assert int((fb == orc.BC_WALL).sum()) == trunc.ncells // nf[axis]  # one wall face per fluid row
    Rf, mf = full.compute_rhs(Uf)
    Rt, mt = trunc.compute_rhs(Ut)
    assert mf == mt
    assert np.array_equal(Rf[fluid == 1], Rt)
    assert np.all(Rf[fluid == 0] == 0.0)  # euler.cpp:142-148: zeroed, never accumulated
    # three RK3 steps: the fluid part follows the truncated grid bit for bit, solid cells keep their state
    Wf, Wt = Uf.copy(), Ut.copy()
    U0 = Uf.copy()
    t = 0.0
    for _ in range(3):
        dtf, ef = full.rk3_step(Uf, Wf, 0.5, t, 1e30)
        dtt, et = trunc.rk3_step(Ut, Wt, 0.5, t, 1e30)
        assert dtf == dtt and np.array_equal(ef, et)
        t += dtf
    assert np.array_equal(Uf[fluid == 1], Ut)
    assert np.array_equal(Uf[fluid == 0], U0[fluid == 0])


def _slab_identity_holds(orc, dim, axis, side):
    full, trunc, fluid, Uf, Ut = slab_pair(orc, dim, axis, side, 2, _n_fluid(dim), _bcs(orc, dim),
                                           10 * dim + 3 * axis + side)
    Wf, Wt = Uf.copy(), Ut.copy()
    U0 = Uf.copy()
    dtf, _ = full.rk3_step(Uf, Wf, 0.5, 0.0, 1e30)
    dtt, _ = trunc.rk3_step(Ut, Wt, 0.5, 0.0, 1e30)
    return dtf == dtt and np.array_equal(Uf[fluid == 1], Ut) and np.array_equal(Uf[fluid == 0], U0[fluid == 0])


@pytest.mark.parametrize("plant", ["ORC_PLANT_WALL_DROP", "ORC_PLANT_NO_SKIP", "ORC_PLANT_SOLID_UPDATE"])
def test_planted_body_mistakes_are_caught(orc, tmp_path, plant):
    """Mistakes only the body branch makes break the slab identity on every slab."""
    path = orc.build_variant([plant], str(tmp_path))
    with orc.use_library(path):
        held = [_slab_identity_holds(orc, *c) for c in CASES]
    assert not any(held), (plant, held)


def test_planted_slot_mistake_is_caught(orc, tmp_path):
    """The virtual state in the fluid cell's slot is shared by border and wall
    faces, so the slab identity still holds; the reference's reflecting known
    answer (2D radial Sod, test/2dRadSod/CMakeLists.txt:33) catches it, and the
    identity carries that pin over to the wall faces."""
    path = orc.build_variant(["ORC_PLANT_WALL_SLOT"], str(tmp_path))
    with orc.use_library(path):
        l1, _, _, _, _ = orc.run_reference_case(orc.RADSOD, 2, 64, 1.0)
        assert all(_slab_identity_holds(orc, *c) for c in CASES)
    assert not abs(l1 - 1.091296206337) <= 1e-8 * 1.091296206337, l1  # off, or not finite


def test_wall_equals_reflecting_under_either_normal(orc):
    """euler.cpp:347-356 (wall = reflecting) and the mirror's n -> -n invariance,
    which makes the flipped boundary normal (euler.cpp:219-222) exact."""
    rng = np.random.default_rng(7)
    for c in random_state(orc, rng, 50, 3, speed=2.0):
        for a in range(3):
            n = np.zeros(3)
            n[a] = 1.0
            r1 = orc.eval_bc_values(orc.BC_WALL, n, c)
            assert np.array_equal(r1, orc.eval_bc_values(orc.BC_REFLECTING, n, c))
            assert np.array_equal(r1, orc.eval_bc_values(orc.BC_WALL, -n, c))


def test_body_faces_and_flags_of_a_box(orc):
    """main.cpp:222-277: solved = fluid; a face is a wall iff exactly one side is
    fluid; border faces keep the side's code.  Box bounds are inclusive
    (body.cpp:83-89): a centroid on the box face is solid."""
    pb = orc.problem(orc.RADSOD, 2)
    n, h = 8, 0.5
    box = np.array([1.25, 0.75, -1.0, 2.75, 2.25, 1.0])  # x in [1.25, 2.75]: centroids 1.25 .. 2.75 inclusive
    # orc_mark_fluid through the library (boxes and no sphere)
    import ctypes as C
    L = orc.lib()
    mesh_fluid = np.empty(n * n, dtype=np.uint8)
    s_box = orc.Solver(pb, 1, [n, n], [0.0] * 3, h)
    L.orc_mark_fluid(C.c_void_p(s_box.s), box.ctypes.data_as(C.POINTER(C.c_double)), 1, None, 0,
                     mesh_fluid.ctypes.data_as(C.POINTER(C.c_uint8)))
    i = np.arange(n * n) % n
    j = np.arange(n * n) // n
    expect_solid = (i >= 2) & (i <= 5) & (j >= 1) & (j <= 4)  # x = 0.25 + 0.5 i in [1.25, 2.75]
    assert np.array_equal(mesh_fluid == 0, expect_solid)
    s = orc.Solver(pb, 1, [n, n], [0.0] * 3, h, fluid=mesh_fluid)
    f, so = s.flags()
    assert np.array_equal(f, mesh_fluid) and np.array_equal(so, mesh_fluid)
    fb = s.face_bc()
    # 2 x 4 block of solid cells: perimeter 2 * (4 + 4) wall faces
    assert int((fb == orc.BC_WALL).sum()) == 16


def test_sphere_mask_is_the_level_set(orc):
    """Sphere extension: fluid iff |x - c| - r >= 0 at the centroid.  Checked
    against numpy's hypot (another evaluation order) on a grid where no
    centroid lies within 1e-9 of the surface, plus an exact on-surface
    centroid that must stay fluid; the solid count matches the ball's volume."""
    import ctypes as C
    pb = orc.problem(orc.RADSOD, 3)
    n, h = 24, 0.25
    s = orc.Solver(pb, 1, [n] * 3, [-3.0] * 3, h)
    cen = s.centroids()
    L = orc.lib()

    def mark(spheres):
        sp = np.ascontiguousarray(spheres, dtype=np.float64)
        out = np.empty(s.ncells, dtype=np.uint8)
        L.orc_mark_fluid(C.c_void_p(s.s), None, 0, sp.ctypes.data_as(C.POINTER(C.c_double)), len(sp),
                         out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out

    c, r = np.array([0.1, -0.2, 0.3]), 1.3
    fl = mark([[*c, r]])
    d = np.hypot(np.hypot(cen[:, 0] - c[0], cen[:, 1] - c[1]), cen[:, 2] - c[2]) - r
    assert np.min(np.abs(d)) > 1e-9
    assert np.array_equal(fl == 1, d >= 0)
    vol_cells = int((fl == 0).sum()) * h ** 3
    assert abs(vol_cells - 4.0 / 3.0 * np.pi * r ** 3) < 0.05 * 4.0 / 3.0 * np.pi * r ** 3
    # a centroid exactly on the surface (distance 0.5 along x from a centroid centre) is fluid
    c0 = cen[0]
    fl2 = mark([[c0[0], c0[1], c0[2], 0.5]])
    on = np.where((np.abs(cen[:, 0] - c0[0]) == 0.5) & (cen[:, 1] == c0[1]) & (cen[:, 2] == c0[2]))[0]
    assert len(on) == 1 and fl2[on[0]] == 1 and fl2[0] == 0
    # two spheres: solid is the union
    fl3 = mark([[*c, r], [1.5, 1.5, 1.5, 0.6]])
    d2 = np.hypot(np.hypot(cen[:, 0] - 1.5, cen[:, 1] - 1.5), cen[:, 2] - 1.5) - 0.6
    assert np.array_equal(fl3 == 1, (d >= 0) & (d2 >= 0))


def test_error_norms_closed_form(orc):
    """L1 = |delta| |Omega| and L2 = |delta| sqrt(|Omega|) for a constant density
    offset delta; two offsets on halves: L2 = sqrt((d1^2 + d2^2)/2 |Omega|).
    Dyadic data, so every difference and sum is exact."""
    pb = orc.problem(orc.RADSOD, 2)  # exact density 1 or 0.125 at any t
    n, h = 16, 0.5  # |Omega| = 64
    s = orc.Solver(pb, 1, [n, n], [0.0] * 3, h)
    U = s.initial_state()
    delta = 2.0 ** -6
    V = U.copy()
    V[:, 0] += delta
    l1, l2 = s.error(V, 0.7)
    assert l1 == delta * 64.0
    assert l2 == delta * 8.0
    V = U.copy()
    V[: n * n // 2, 0] += 3 * delta
    V[n * n // 2:, 0] -= 4 * delta
    l1, l2 = s.error(V, 0.7)
    assert l1 == 3.5 * delta * 64.0
    assert l2 == pytest.approx(np.sqrt(12.5 * delta ** 2 * 64.0), rel=1e-15)
    assert s.error(U, 0.7) == (0.0, 0.0)
