"""The marching LDS-tile kernel against the per-cell gather kernel.

Both evaluate every face with the same operation sequence and accumulate each
cell's residual in the reference's face order, so they must agree bit for
bit on every grid shape: partial tiles, chunks that do not divide the march
axis, periodic axes, bodies, every border BC.  Each case also checks the
oracle on the RHS.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(8, 32), (4, 7), (8, 128), (8, 1)]
# 32-wide tiles (cols, rows, chunk)
SHAPES32 = [(32, 8, 32), (32, 8, 5), (32, 16, 128), (32, 16, 3)]


def build(mmf, dim, n, side_bc, info=None, body=None, order=1, seed=0):
    h = 2.0 / max(n)
    s = mmf.Solver(dim, n, [-1.0, -1.0, -1.0], h, side_bc, info, order)
    if body is not None:
        s.set_bodies(**body)
    rng = np.random.default_rng(seed)
    cen = s.centroids()
    r = np.sqrt((cen ** 2).sum(1))
    prim = np.empty((s.ncells, 5))
    prim[:, 0] = np.where(r < 0.45, 1.0, 0.1) * (1 + 0.05 * rng.uniform(-1, 1, s.ncells))
    prim[:, 1:4] = 0.2 * rng.uniform(-1, 1, (s.ncells, 3))
    if dim == 2:
        prim[:, 3] = 0.0
    prim[:, 4] = np.where(r < 0.45, 1.0, 0.8) * (1 + 0.05 * rng.uniform(-1, 1, s.ncells))
    U = mmf.batch_prim2cons(prim)
    return s, U


def run_variant(mmf, s, U, kind, rows=0, chunk=0, steps=3, cols=0):
    s.set_kernel(kind, rows, chunk)
    if cols:
        s.set_march_shape(cols, rows, chunk)
    s.set_state(U)
    R, me = s.compute_rhs()
    s.set_time(0.0, 10.0, 0.5)
    s.advance(steps)
    t, dt, step, eig = s.time()
    return R, me, s.get_state(), (t, dt, step, tuple(eig)), s.positivity_counters()


CASES = [
    ("3d cube free", 3, (48, 48, 48), [0] * 6, None),
    ("3d ragged", 3, (70, 13, 37), [0, 1, 2, 0, 1, 0], None),
    ("3d periodic xyz", 3, (66, 10, 12), [4] * 6, None),
    ("3d periodic x only", 3, (64, 9, 11), [4, 4, 1, 1, 0, 0], None),
    ("3d dirichlet+box", 3, (40, 24, 20), [3, 0, 1, 1, 0, 2], {"boxes": [[-0.4, -0.3, -0.5, 0.1, 0.35, 0.2]]}),
    ("3d dirichlet free", 3, (40, 24, 20), [3, 0, 0, 3, 0, 3], None),
    ("3d sphere", 3, (36, 36, 36), [3, 0, 0, 0, 0, 0], {"spheres": [[0.1, 0.0, -0.1, 0.35]]}),
    # a body against the periodic x and y sides: the tiles across the wrap read it
    ("3d periodic body", 3, (130, 20, 24), [4, 4, 4, 4, 0, 3],
     {"boxes": [[0.9, -0.75, -0.9, 1.0, -0.68, -0.8]]}),
    # a body on the periodic z sides: the planes across the wrap need the masked code
    ("3d periodic z body", 3, (70, 18, 40), [0, 3, 1, 0, 4, 4], {"boxes": [[-0.2, -0.9, -1.0, 0.3, -0.7, -0.95]]}),
    # full 64 x 8 tiles with copy / Dirichlet x and y sides: the flat plane
    # pass (border faces from LDS slots, counts without branches)
    ("3d full tiles dirichlet", 3, (128, 16, 20), [3, 0, 0, 3, 0, 3], None),
    ("3d full tiles periodic z", 3, (64, 24, 30), [0, 3, 3, 0, 4, 4], None),
    ("3d full tiles all dirichlet", 3, (64, 16, 18), [3] * 6, None),
    ("2d free", 2, (64, 64), [0] * 4, None),
    ("2d ragged periodic", 2, (100, 37), [4, 4, 4, 4], None),
    ("2d box reflecting", 2, (77, 50), [1, 1, 1, 1], {"boxes": [[-0.2, -0.2, -1, 0.3, 0.1, 1]]}),
]


@pytest.mark.parametrize("name,dim,n,bc,body", CASES, ids=[c[0] for c in CASES])
def test_march_equals_gather(mmf, orc, name, dim, n, bc, body):
    info = np.zeros((6, 5))
    info[:] = [1.0, 0.3, -0.1, 0.05 if dim == 3 else 0.0, 1.0]
    # a different Dirichlet state per side (p, u, v, w, T), so that a side's
    # state read from another side's slot changes the result
    info += np.arange(6)[:, None] * np.array([0.05, 0.02, 0.01, -0.01 if dim == 3 else 0.0, 0.03])
    s, U = build(mmf, dim, list(n), bc, info, body)
    ref = run_variant(mmf, s, U, mmf.KERNEL_GATHER)
    shapes = [(64, r, c) for r, c in SHAPES] + SHAPES32 if dim == 3 else [(0, 1, 16), (0, 1, 5), (0, 1, 1000)]
    for cols, rows, chunk in shapes:
        got = run_variant(mmf, s, U, mmf.KERNEL_MARCH, rows, chunk, cols=cols)
        assert s.kernel()[0] == mmf.KERNEL_MARCH
        assert np.array_equal(got[0], ref[0]), (name, rows, chunk, "rhs")
        assert got[1] == ref[1], (name, rows, chunk, "maxEig")
        assert np.array_equal(got[2], ref[2]), (name, rows, chunk, "state after 3 steps")
        assert got[3] == ref[3], (name, rows, chunk, "time control")
        assert got[4] == ref[4], (name, rows, chunk, "positivity counters")
    if dim == 3:
        # the planes split over a number of blocks in shares of equal cost
        # (segments that start and end at arbitrary planes of a tile, cut
        # where the code changes between masked and unmasked)
        ntiles = -(-n[0] // 64) * -(-n[1] // 8)
        for blocks in sorted({ntiles, ntiles + 3, 2 * ntiles + 1, 7 * ntiles - 2, ntiles * n[2]}):
            s.set_kernel(mmf.KERNEL_MARCH, 8)
            s.set_march_blocks(blocks)
            assert s.kernel()[2] == 0
            s.set_state(U)
            R, me = s.compute_rhs()
            s.set_time(0.0, 10.0, 0.5)
            s.advance(3)
            t, dt, step, eig = s.time()
            got = (R, me, s.get_state(), (t, dt, step, tuple(eig)), s.positivity_counters())
            assert np.array_equal(got[0], ref[0]), (name, blocks, "rhs")
            assert got[1] == ref[1], (name, blocks, "maxEig")
            assert np.array_equal(got[2], ref[2]), (name, blocks, "state after 3 steps")
            assert got[3] == ref[3], (name, blocks, "time control")
            assert got[4] == ref[4], (name, blocks, "positivity counters")
        s.set_march_blocks(0)
    # and the oracle on the RHS
    fl = s.cell_flags() & 1
    po = orc.problem(orc.RADSOD, dim)
    o = orc.Solver(po, 1, list(n), [-1.0, -1.0, -1.0], 2.0 / max(n), side_bc=bc, side_info=info, fluid=fl)
    Ro, mo = o.compute_rhs(U)
    den = np.max(np.abs(Ro))
    assert np.max(np.abs(ref[0] - Ro)) <= 1e-12 * den
    assert ref[1] == mo


# negative densities placed across tile / strip edges: 3D rows 7 | 8 (64 x 8 tiles), 2D columns
# 61 | 62 (62-cell strips); on the full-tile grid (the flat plane pass) also at the x and y
# borders and across planes
NEG_RHO = {(3, 20): [943, 963], (2, 70): [61, 412],
           (3, 64): [2499, 2563, 3392, 4479, 5130, 7114, 7178]}


@pytest.mark.parametrize("dim,n", [(3, [20, 20, 20]), (2, [70, 33]), (3, [64, 16, 12])])
def test_march_positivity_counter(mmf, orc, dim, n):
    """Negative pressures and densities are counted per face side as the oracle
    counts them (the marching kernels sum per thread and add once per wave)."""
    s, U = build(mmf, dim, n, [0] * 6)
    U = U.copy()
    U[[5, 777, 2000], 4] = 0.0  # zero energy -> negative temperature and pressure
    U[NEG_RHO[(dim, n[0])], 0] = -0.5
    h = 2.0 / max(n)
    o = orc.Solver(orc.problem(orc.RADSOD, dim), 1, n, [-1.0] * 3, h, side_bc=[0] * 6)
    o.compute_rhs(U)
    assert o.counters()[0] > 0 and o.counters()[1] > 0
    for kind in (mmf.KERNEL_GATHER, mmf.KERNEL_MARCH):
        s.set_kernel(kind)
        assert s.kernel()[0] == kind
        s.reset_positivity_counters()
        s.set_state(U)
        s.compute_rhs()
        assert s.positivity_counters() == o.counters(), kind


O2_CASES = [
    ("2d o2 free", (64, 64), [0] * 4),
    ("2d o2 periodic ragged", (121, 37), [4, 4, 4, 4]),
    ("2d o2 mixed bcs", (61, 60), [3, 1, 2, 0]),
    ("2d o2 periodic x only", (60, 9), [4, 4, 1, 3]),
    ("2d o2 thin", (3, 70), [4, 4, 0, 1]),
    ("2d o2 one row", (130, 1), [0, 2, 1, 3]),
    ("2d o2 large ragged", (203, 170), [4, 4, 0, 1]),
]


@pytest.mark.parametrize("name,n,bc", O2_CASES, ids=[c[0] for c in O2_CASES])
def test_march_order2_2d_equals_gather(mmf, orc, name, n, bc):
    """The 2D order-2 strip kernel (limited linear, each face once) against the
    per-cell gather kernel: bit-identical RHS, steps, time control and
    positivity counts for every chunk shape; and the oracle on the RHS."""
    info = np.zeros((6, 5))
    info[:] = [1.0, 0.3, -0.1, 0.0, 1.0]
    s, U = build(mmf, 2, list(n), bc, info, None, order=2, seed=5)
    ref = run_variant(mmf, s, U, mmf.KERNEL_GATHER)
    for chunk in (16, 5, 1, 2, 1000):
        got = run_variant(mmf, s, U, mmf.KERNEL_MARCH, 1, chunk)
        assert s.kernel()[0] == mmf.KERNEL_MARCH
        assert np.array_equal(got[0], ref[0]), (name, chunk, "rhs")
        assert got[1] == ref[1], (name, chunk, "maxEig")
        assert np.array_equal(got[2], ref[2]), (name, chunk, "state after 3 steps")
        assert got[3] == ref[3], (name, chunk, "time control")
        assert got[4] == ref[4], (name, chunk, "positivity counters")
    po = orc.problem(orc.RADSOD, 2)
    o = orc.Solver(po, 2, list(n), [-1.0, -1.0, -1.0], 2.0 / max(n), side_bc=bc + [0, 0], side_info=info)
    Ro, mo = o.compute_rhs(U)
    assert np.max(np.abs(ref[0] - Ro)) <= 1e-12 * np.max(np.abs(Ro))
    assert ref[1] == mo


O2_3D_CASES = [
    ("3d o2 cube free", (48, 40, 36), [0] * 6, None),
    ("3d o2 ragged mixed", (70, 13, 37), [0, 1, 2, 0, 1, 3], None),
    ("3d o2 periodic xyz", (66, 10, 12), [4] * 6, None),
    ("3d o2 periodic z thin", (5, 3, 9), [0, 1, 3, 0, 4, 4], None),
    ("3d o2 dirichlet box", (40, 24, 20), [3, 0, 1, 1, 0, 2], {"boxes": [[-0.4, -0.3, -0.5, 0.1, 0.35, 0.2]]}),
    ("3d o2 sphere", (36, 36, 36), [3, 0, 0, 0, 0, 0], {"spheres": [[0.1, 0.0, -0.1, 0.35]]}),
    ("3d o2 periodic body", (130, 20, 24), [4, 4, 4, 4, 0, 3], {"boxes": [[0.9, -0.75, -0.9, 1.0, -0.68, -0.8]]}),
]


@pytest.mark.parametrize("name,n,bc,body", O2_3D_CASES, ids=[c[0] for c in O2_3D_CASES])
def test_march_order2_3d_equals_gather(mmf, orc, name, n, bc, body):
    """The 3D order-2 marching kernel (LDS-staged limited reconstruction,
    each z face once) against the per-cell gather kernel: bit-identical RHS,
    steps, time control and positivity counts for every chunk, bodies and
    periodic axes included; and the oracle on the RHS."""
    info = np.zeros((6, 5))
    info[:] = [1.0, 0.3, -0.1, 0.05, 1.0]
    s, U = build(mmf, 3, list(n), bc, info, body, order=2, seed=7)
    ref = run_variant(mmf, s, U, mmf.KERNEL_GATHER)
    for chunk in (32, 5, 1, 2, 1000):
        got = run_variant(mmf, s, U, mmf.KERNEL_MARCH, 8, chunk)
        assert s.kernel()[0] == mmf.KERNEL_MARCH
        assert np.array_equal(got[0], ref[0]), (name, chunk, "rhs")
        assert got[1] == ref[1], (name, chunk, "maxEig")
        assert np.array_equal(got[2], ref[2]), (name, chunk, "state after 3 steps")
        assert got[3] == ref[3], (name, chunk, "time control")
        assert got[4] == ref[4], (name, chunk, "positivity counters")
    fl = s.cell_flags() & 1
    po = orc.problem(orc.RADSOD, 3)
    o = orc.Solver(po, 2, list(n), [-1.0, -1.0, -1.0], 2.0 / max(n), side_bc=bc, side_info=info, fluid=fl)
    Ro, mo = o.compute_rhs(U)
    assert np.max(np.abs(ref[0] - Ro)) <= 1e-12 * np.max(np.abs(Ro))
    assert ref[1] == mo


@pytest.mark.parametrize("name,n,bc", O2_CASES, ids=[c[0] for c in O2_CASES])
def test_tile_order2_2d_equals_gather(mmf, name, n, bc):
    """The 8x8 LDS tile kernel (ring cells recomputed, each face once, the
    time advance in the stage-3 launch) against the per-cell gather kernel:
    bit-identical RHS, steps, time control and positivity counts, also when
    chosen by AUTO."""
    info = np.zeros((6, 5))
    info[:] = [1.0, 0.3, -0.1, 0.0, 1.0]
    s, U = build(mmf, 2, list(n), bc, info, None, order=2, seed=5)
    ref = run_variant(mmf, s, U, mmf.KERNEL_GATHER)
    assert s.kernel()[0] == mmf.KERNEL_GATHER
    for kind in (mmf.KERNEL_TILE, mmf.KERNEL_AUTO):
        got = run_variant(mmf, s, U, kind)
        assert s.kernel()[0] == mmf.KERNEL_TILE
        assert np.array_equal(got[0], ref[0]), (name, kind, "rhs")
        assert got[1] == ref[1], (name, kind, "maxEig")
        assert np.array_equal(got[2], ref[2]), (name, kind, "state after 3 steps")
        assert got[3] == ref[3], (name, kind, "time control")
        assert got[4] == ref[4], (name, kind, "positivity counters")


def test_tile_kernel_refused_where_it_does_not_apply(mmf):
    s = mmf.Solver(2, [16, 16], [-1.0, -1.0, 0.0], 0.125, [0] * 6, None, 1)
    with pytest.raises(mmf.MMFError):
        s.set_kernel(mmf.KERNEL_TILE)
    s3 = mmf.Solver(3, [8, 8, 8], [-1.0, -1.0, -1.0], 0.25, [0] * 6, None, 2)
    with pytest.raises(mmf.MMFError):
        s3.set_kernel(mmf.KERNEL_TILE)


def test_bodies_changed_between_graph_replayed_steps(mmf):
    """Flags changed on a live context after graph-captured steps: a body moved
    (the tile table rewritten under the same pointer), removed (no flags at all)
    and set again; every phase bit-identical to the gather kernel, with one
    (tile, chunk) per block and with the planes split over 7 blocks (the
    segment table rebuilt from the new flags)."""
    info = np.zeros((6, 5))
    info[:] = [1.0, 0.3, -0.1, 0.05, 1.0]
    bodies = [{"boxes": [[-0.4, -0.3, -0.5, 0.1, 0.35, 0.2]]}, {"spheres": [[0.2, 0.1, -0.1, 0.3]]}, None,
              {"boxes": [[0.5, -0.9, -0.9, 0.8, -0.5, 0.0]]}]
    out = {}
    for variant in ("gather", "march", "march segments"):
        kind = mmf.KERNEL_GATHER if variant == "gather" else mmf.KERNEL_MARCH
        s, U = build(mmf, 3, [64, 24, 20], [3, 0, 0, 3, 0, 3], info)
        s.set_state(U)
        s.set_time(0.0, 10.0, 0.5)
        res = []
        for b in bodies:
            if b is None:
                s.set_cell_flags(np.ones(s.ncells, dtype=np.uint8))
            else:
                s.set_bodies(**b)
            s.set_kernel(kind, 8, 8)
            if variant == "march segments":
                s.set_march_blocks(7)
                assert s.kernel()[2] == 0
            assert s.kernel()[0] == kind
            s.advance(17)  # one 16-step graph and one 1-step graph
            res.append((s.get_state(), s.time()[:3], s.cell_flags()))
        out[variant] = res
        s.close()
    for variant in ("march", "march segments"):
        for (a, ta, fa), (b, tb, fb) in zip(out["gather"], out[variant]):
            assert np.array_equal(fa, fb), variant
            assert ta == tb, variant
            assert np.array_equal(a, b), variant


def test_timing_modes_count_launches_and_keep_results(mmf):
    """Per-launch event pairs (steps launched one by one) and one pair per
    advance call (steps replayed from graphs) count three stage launches per
    step and leave the results as with timing off."""
    s, U = build(mmf, 3, [64, 40, 36], [0] * 6)
    s.set_kernel(mmf.KERNEL_MARCH)
    out = []
    for mode in (0, 1, 2):
        s.set_state(U)
        s.set_time(0.0, 10.0, 0.5)
        s.enable_timing(mode)
        s.advance(5)
        s.advance(2)
        ms, n = s.read_timing()
        s.enable_timing(0)
        if mode:
            assert n == 21 and ms > 0.0, (mode, n, ms)
        out.append((s.get_state(), s.time()))
    for st, tm in out[1:]:
        assert np.array_equal(st, out[0][0]) and tm[:3] == out[0][1][:3]
    with pytest.raises(mmf.MMFError):
        s.enable_timing(3)
    s.close()


@pytest.mark.parametrize("covered", [True, False])
def test_early_dt_dirichlet_side_behind_a_body(mmf, covered):
    """A Dirichlet side whose border cells are all solid has no face in the
    stage maximum (euler.cpp:181-183): the marching kernel's early dt on a grid
    with flags leaves that side's state out, as the gather kernel does; with
    the side open its state (faster than every cell) sets the first dt."""
    info = np.zeros((6, 5))
    info[0] = [1.0, 4.0, 0.0, 0.0, 1.0]  # |u| + c = 4 + sqrt(1.4): above every cell's
    # the first cell column (centroid x = -0.984) solid, or a box away from the side
    box = [-1.0, -1.0, -1.0, -0.98, 1.0, 1.0] if covered else [0.2, -0.9, -0.9, 0.4, -0.7, -0.7]
    h = 2.0 / 64
    dt_inflow = 0.9 * 0.5 * h / (4.0 + np.sqrt(1.4))
    out = []
    for kind in (mmf.KERNEL_GATHER, mmf.KERNEL_MARCH):
        s, U = build(mmf, 3, [64, 16, 16], [3, 0, 0, 0, 0, 0], info, {"boxes": [box]})
        s.set_kernel(kind)
        assert s.kernel()[0] == kind
        s.set_state(U)
        s.set_time(0.0, 10.0, 0.5)
        s.advance(1)
        dt1 = s.time()[1]
        assert abs(dt1 - dt_inflow) <= 1e-12 * dt_inflow if not covered else dt1 > 1.01 * dt_inflow, (kind, dt1)
        s.advance(2)
        out.append((s.get_state(), s.time(), dt1))
        s.close()
    (Ug, tg, d1g), (Um, tm, d1m) = out
    assert d1g == d1m
    assert tg[:3] == tm[:3] and np.array_equal(tg[3], tm[3]), (tg, tm)
    assert np.array_equal(Ug, Um)
