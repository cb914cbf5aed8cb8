"""One RK3 stage at a time (mmf_rk_rhs / mmf_rk_dt / mmf_rk_update, the
reference's loop main.cpp:377-506 with a hook between stages) and the
drop-in computeRHS with per-call solved flags and order, on the GPU:

- the staged step gives the fused step's bits (state, t, dt, step count,
  stage maxima) on the gather, marching (2D strips, 3D with and without the
  early dt) and tile kernels;
- a grid split into two sub-grids ("ranks") with one or two ghost layers
  (fluid, unsolved: main.cpp:231-235), their ghosts exchanged after every
  stage through mmf_pack_cells / mmf_unpack_cells (main.cpp:425-430,
  461-466, 497-502) and the eigenvalue reduced over both (main.cpp:390-395),
  reproduces the whole grid bit for bit;
- mmf_compute_rhs_ex follows solved flags and order changed between calls,
  against the oracle with the same flags.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_state(orc, rng, n, dim, speed=0.6):
    return np.array([orc.prim2cons([rng.uniform(0.5, 2), *rng.uniform(-speed, speed, 2),
                                    rng.uniform(-speed, speed) if dim == 3 else 0.0, rng.uniform(0.5, 2)])
                     for _ in range(n)])


def one_staged_step(mmf, s):
    e0 = s.rk_rhs(mmf.STATE_U)
    dt = s.rk_dt(e0)
    s.rk_update(1, dt)
    s.rk_rhs(mmf.STATE_W)
    s.rk_update(2, dt)
    s.rk_rhs(mmf.STATE_W)
    s.rk_update(3, dt)


CASES = [
    # (label, dim, n, order, side_bc, kernel, body)
    ("3d gather body", 3, [20, 18, 16], 1, [3, 0, 1, 2, 0, 1], None, True),
    ("3d gather order 2 body", 3, [20, 18, 16], 2, [3, 0, 1, 2, 4, 4], None, True),
    ("3d march early dt", 3, [64, 40, 36], 1, [0, 0, 0, 0, 4, 4], "march", False),
    ("3d march reflecting", 3, [64, 40, 36], 1, [1, 0, 0, 1, 0, 0], "march", False),
    ("3d march body", 3, [64, 40, 36], 1, [0] * 6, "march", True),
    ("2d march strips", 2, [96, 70], 1, [1, 0, 3, 0], "march", False),
    ("2d tile order 2", 2, [60, 52], 2, [0, 1, 4, 4], None, False),
]


@pytest.mark.parametrize("label,dim,n,order,side_bc,kernel,body", CASES, ids=[c[0] for c in CASES])
def test_staged_step_equals_the_fused_step(mmf, orc, label, dim, n, order, side_bc, kernel, body):
    info = np.tile([1.0, 0.4, 0.1, 0.0, 1.0 / 1.4], (6, 1))
    rng = np.random.default_rng(len(label))
    U0 = random_state(orc, rng, int(np.prod(n)), dim)
    out = []
    for staged in (False, True):
        s = mmf.Solver(dim, n, [0.0] * 3, 0.125, side_bc, info, order)
        if kernel == "march":
            s.set_kernel(mmf.KERNEL_MARCH)
        if body:
            s.set_bodies(boxes=[[1.0, 1.0, 0.5, 1.6, 1.4, 1.2]])
        used = s.kernel()[0]
        s.set_state(U0)
        s.set_time(0.0, 1e30, 0.45)
        for _ in range(3):
            if staged:
                one_staged_step(mmf, s)
            else:
                s.advance(1)
        out.append((s.get_state(), s.time(), used))
        s.close()
    print("%s: kernel %d" % (label, out[0][2]), flush=True)
    (Uf, tf, _), (Us, ts, _) = out
    assert tf[:3] == ts[:3] and np.array_equal(tf[3], ts[3]), (label, tf, ts)
    assert np.array_equal(Uf, Us), label


def split_case(mmf, full_n, dim, axis, cut, ghosts, origin, h, side_bc, info, order):
    """Two sub-grids of the full grid split at 'cut' along 'axis', each with
    'ghosts' layers of the other's cells (fluid, unsolved)."""
    parts = []
    for lo, hi in ((0, cut + ghosts), (cut - ghosts, full_n[axis])):
        n = list(full_n)
        n[axis] = hi - lo
        org = list(origin)
        org[axis] = origin[axis] + lo * h
        s = mmf.Solver(dim, n, org, h, side_bc, info, order)
        g = np.arange(lo, hi)
        parts.append({"s": s, "lo": lo, "hi": hi, "n": n, "g": g})
    return parts


def cells_of(n, axis, local_range):
    """Cell indices (i + n0 (j + n1 k)) of the planes local_range along axis."""
    idx = np.arange(int(np.prod(n))).reshape(n[2] if len(n) == 3 else 1, n[1], n[0])
    sl = [slice(None)] * 3
    sl[2 - axis] = slice(local_range[0], local_range[1])
    return idx[tuple(sl)].ravel()


@pytest.mark.parametrize("dim,order,kernel", [(2, 1, None), (2, 2, None), (3, 1, "march"), (3, 2, None)])
def test_two_subgrids_with_ghost_exchange_equal_the_whole_grid(mmf, orc, dim, order, kernel):
    full_n = [48, 40] if dim == 2 else [40, 36, 64]
    axis = dim - 1  # split the last axis
    cut = full_n[axis] // 2
    G = order  # ghost layers: the stencil's reach
    h = 0.125
    origin = [0.0, 0.0, 0.0]
    side_bc = [1, 0, 3, 1, 0, 1][:2 * dim] + [0] * (6 - 2 * dim)
    info = np.zeros((6, 5))
    info[2] = [1.0, 0.2, 0.4, 0.0, 1.0 / 1.4]
    rng = np.random.default_rng(dim * 10 + order)
    N = int(np.prod(full_n))
    U0 = random_state(orc, rng, N, dim)
    # a body across the cut (2D centroids sit at z = 0)
    box = [2.0, 1.5, -1.0, 2.9, 2.6, 1.0] if dim == 2 else [2.0, 1.5, 3.5, 2.9, 2.6, 4.6]
    # the whole grid
    w = mmf.Solver(dim, full_n, origin, h, side_bc, info, order)
    w.set_bodies(boxes=[box])
    fluid_full = (w.cell_flags() & 1).astype(np.uint8)
    w.set_state(U0)
    w.set_time(0.0, 1e30, 0.45)
    parts = split_case(mmf, full_n, dim, axis, cut, G, origin, h, side_bc, info, order)
    U0r = U0.reshape([full_n[2] if dim == 3 else 1, full_n[1], full_n[0], 5])
    fl = fluid_full.reshape(U0r.shape[:3])
    for k, p in enumerate(parts):
        sl = [slice(None)] * 3
        sl[2 - axis] = slice(p["lo"], p["hi"])
        sub = U0r[tuple(sl)].reshape(-1, 5)
        fsub = fl[tuple(sl)].ravel().copy()
        solved = fsub.copy()
        ghost_range = (p["n"][axis] - G, p["n"][axis]) if k == 0 else (0, G)
        solved[cells_of(p["n"], axis, ghost_range)] = 0
        p["s"].set_cell_flags(fsub, solved)
        if kernel == "march":
            p["s"].set_kernel(mmf.KERNEL_MARCH)
        p["s"].set_state(sub)
        p["s"].set_time(0.0, 1e30, 0.45)
        p["ghost"] = cells_of(p["n"], axis, ghost_range)
        # the cells the other part holds as ghosts, in the same order
        p["send"] = cells_of(p["n"], axis, (p["n"][axis] - 2 * G, p["n"][axis] - G) if k == 0 else (G, 2 * G))
    a, b = parts

    def exchange(state):
        ba = a["s"].pack_cells(state, a["send"])
        bb = b["s"].pack_cells(state, b["send"])
        b["s"].unpack_cells(state, b["ghost"], ba)
        a["s"].unpack_cells(state, a["ghost"], bb)

    for step in range(4):
        e = max(a["s"].rk_rhs(mmf.STATE_U), b["s"].rk_rhs(mmf.STATE_U))  # MPI_Allreduce(MAX)
        dt = a["s"].rk_dt(e)
        assert dt == b["s"].rk_dt(e)
        for stage in (1, 2, 3):
            for p in parts:
                p["s"].rk_update(stage, dt)
            exchange(mmf.STATE_W if stage < 3 else mmf.STATE_U)
            if stage < 3:
                a["s"].rk_rhs(mmf.STATE_W)
                b["s"].rk_rhs(mmf.STATE_W)
        w.advance(1)
    tw = w.time()
    Uw = w.get_state().reshape(U0r.shape)
    for k, p in enumerate(parts):
        t = p["s"].time()
        assert t[:3] == tw[:3], (k, t, tw)
        own = (0, p["n"][axis] - G) if k == 0 else (G, p["n"][axis])
        Up = p["s"].get_state()[cells_of(p["n"], axis, own)]
        sl = [slice(None)] * 3
        glo = p["lo"] + own[0]
        sl[2 - axis] = slice(glo, glo + own[1] - own[0])
        assert np.array_equal(Up, Uw[tuple(sl)].reshape(-1, 5)), k
    print("%dD order %d split at %d with %d ghost layers: bit-identical" % (dim, order, cut, G), flush=True)
    for x in (w, a["s"], b["s"]):
        x.close()


def test_compute_rhs_ex_follows_solved_flags_and_order(mmf, orc):
    n = [18, 16, 14]
    h = 0.125
    side_bc = [3, 0, 1, 2, 0, 1]
    info = np.zeros((6, 5))
    info[0] = [1.0, 0.4, 0.1, 0.0, 1.0 / 1.4]
    pb = orc.problem(orc.RADSOD, 3)
    s = mmf.Solver(3, n, [0.0] * 3, h, side_bc, info, 1)
    s.set_bodies(boxes=[[0.6, 0.6, 0.6, 1.2, 1.1, 0.9]])
    fluid = (s.cell_flags() & 1).astype(np.uint8)
    rng = np.random.default_rng(4)
    U = random_state(orc, rng, s.ncells, 3)
    ghost = np.zeros(s.ncells, dtype=bool)
    ghost[rng.choice(s.ncells, 300, replace=False)] = True
    for k, (order, unsolved) in enumerate(((1, None), (1, ghost), (2, ghost), (2, None), (1, ghost))):
        solved = fluid.copy()
        if unsolved is not None:
            solved[unsolved] = 0
        o = orc.Solver(pb, order, n, [0.0] * 3, h, side_bc=side_bc, side_info=info, fluid=fluid)
        o.set_solved(solved)
        Ro, mo = o.compute_rhs(U)
        Rg, mg = s.compute_rhs_ex(solved, order, U)
        assert np.array_equal(Rg, Ro) and mg == mo, (k, order)
        assert np.array_equal(s.cell_flags() >> 1 & 1, solved), k
    with pytest.raises(mmf.MMFError) as e:
        s.compute_rhs_ex(None, 3, U)
    assert e.value.status == mmf.ERR_UNSUPPORTED_ORDER
    s.close()
