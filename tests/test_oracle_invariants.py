"""Unit-level pins of the oracle (CPU), chosen so that a plausible mistake fails.

The reference ships no unit tests or vectors; besides its five end-to-end known
answers (test_oracle_known_answers.py) the oracle is pinned here by what the
mathematics fixes: closed forms, exact symmetries of the Rusanov flux,
identities of the boundary states, zero residual of uniform states, discrete
conservation, and a brute-force face sum on tiny grids.  The last tests plant
mistakes in a copy of the oracle and check that the known answers catch them.
"""
import os

import numpy as np
import pytest

GAMMA = 1.4


def rand_prim(rng, n):
    p = np.empty((n, 5))
    p[:, 0] = rng.uniform(0.05, 3.0, n)
    p[:, 1:4] = rng.uniform(-2.0, 2.0, (n, 3))
    p[:, 4] = rng.uniform(0.2, 3.0, n)
    return p


def test_cons_prim_round_trip(orc):
    rng = np.random.default_rng(0)
    for prim in rand_prim(rng, 200):
        c = orc.prim2cons(prim)
        back = orc.cons2prim(c)
        assert np.allclose(back, prim, rtol=4e-15, atol=1e-15)
        # p = rho T with R = 1, E = rho T/(gamma-1) + rho |u|^2 / 2 (utils.cpp:71-83)
        rho = prim[0] / prim[4]
        assert c[0] == pytest.approx(rho, rel=1e-15)
        assert c[4] == pytest.approx(rho * prim[4] / (GAMMA - 1) + 0.5 * rho * np.dot(prim[1:4], prim[1:4]), rel=1e-14)


def test_flux_closed_form(orc):
    """Textbook Euler flux along n (state at rest: pure pressure flux)."""
    rng = np.random.default_rng(1)
    for prim in rand_prim(rng, 100):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        c = orc.prim2cons(prim)
        pr = orc.cons2prim(c)
        f = orc.eval_fluxes(c, pr, n)
        rho, u, p = c[0], pr[1:4], pr[0]
        un = u @ n
        E = p / (GAMMA - 1) + 0.5 * rho * (u @ u)
        assert np.allclose(f, [rho * un, *(rho * un * u + p * n), un * (E + p)], rtol=1e-13, atol=1e-13)
    rest = orc.prim2cons([2.0, 0.0, 0.0, 0.0, 1.5])
    f = orc.eval_fluxes(rest, orc.cons2prim(rest), [0.0, 1.0, 0.0])
    assert f[0] == 0.0 and f[1] == 0.0 and f[3] == 0.0 and f[4] == 0.0 and f[2] == 2.0


def test_splitting_consistency_and_antisymmetry(orc):
    rng = np.random.default_rng(2)
    P = rand_prim(rng, 300)
    C = np.array([orc.prim2cons(p) for p in P])
    for i in range(150):
        L, R = C[i], C[150 + i]
        n = np.zeros(3)
        n[rng.integers(3)] = rng.choice([-1.0, 1.0])
        F, lam = orc.eval_splitting(L, L, n)
        assert np.array_equal(F, orc.eval_fluxes(L, orc.cons2prim(L), n))  # consistency, exact
        F1, l1 = orc.eval_splitting(L, R, n)
        F2, l2 = orc.eval_splitting(R, L, -n)
        assert np.array_equal(F1, -F2) and l1 == l2  # exact antisymmetry
        pl, pr = orc.cons2prim(L), orc.cons2prim(R)
        lam_exp = max(abs(pl[1:4] @ n) + np.sqrt(GAMMA * pl[4]), abs(pr[1:4] @ n) + np.sqrt(GAMMA * pr[4]))
        assert l1 == pytest.approx(lam_exp, rel=1e-15)


def test_bc_states(orc):
    rng = np.random.default_rng(3)
    for prim in rand_prim(rng, 100):
        c = orc.prim2cons(prim)
        a = rng.integers(3)
        n = np.zeros(3)
        n[a] = rng.choice([-1.0, 1.0])
        assert np.array_equal(orc.eval_bc_values(orc.BC_FREE_FLOW, n, c), c)
        r = orc.eval_bc_values(orc.BC_REFLECTING, n, c)
        assert np.array_equal(r, orc.eval_bc_values(orc.BC_WALL, n, c))
        pr = orc.cons2prim(r)
        assert pr[a + 1] == -prim[a + 1] or abs(pr[a + 1] + prim[a + 1]) < 1e-14  # normal velocity negated
        for b in range(3):
            if b != a:
                assert pr[b + 1] == pytest.approx(prim[b + 1], rel=1e-14, abs=1e-15)
        assert pr[0] == pytest.approx(prim[0], rel=1e-14) and pr[4] == pytest.approx(prim[4], rel=1e-14)
        info = rand_prim(rng, 1)[0]
        assert np.array_equal(orc.eval_bc_values(orc.BC_DIRICHLET, n, c, info), orc.prim2cons(info))


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_state_zero_residual(orc, dim):
    pb = orc.problem(orc.RADSOD, dim)
    n = [6] * dim
    for bc in ([orc.BC_FREE_FLOW] * 6, [orc.BC_PERIODIC] * 6):
        s = orc.Solver(pb, 1, n, [0.0, 0.0, 0.0], 0.5, side_bc=bc)
        U = np.tile(orc.prim2cons([0.7, 0.3, -0.2, 0.1 if dim == 3 else 0.0, 1.3]), (s.ncells, 1))
        R, me = s.compute_rhs(U)
        assert np.all(R == 0.0)
        assert me == pytest.approx(max(0.3, 0.2, 0.1) + np.sqrt(GAMMA * 1.3), rel=1e-15)


@pytest.mark.parametrize("order", [1, 2])
def test_conservation(orc, order):
    """Periodic: the residuals sum to zero; reflecting walls: no mass crosses."""
    pb = orc.problem(orc.VORTEX_XY, 2)
    for bc, fields in (([orc.BC_PERIODIC] * 4, range(5)), ([orc.BC_REFLECTING] * 4, [0, 4])):
        s = orc.Solver(pb, order, [24, 24], [-5.0, -5.0, -5.0], 10 / 24, side_bc=bc)
        U = s.initial_state()
        R, _ = s.compute_rhs(U)
        for k in fields:
            assert abs(R[:, k].sum()) <= 1e-13 * np.abs(R[:, k]).sum()


def test_brute_force_face_sum(orc):
    """Residual of every cell of a 3x3 grid = sum over its four faces of the
    point splitting with the outward normal (neighbour or boundary state)."""
    pb = orc.problem(orc.RADSOD, 2)
    bc = [orc.BC_FREE_FLOW, orc.BC_REFLECTING, orc.BC_REFLECTING, orc.BC_FREE_FLOW]
    n, h = 3, 0.5
    s = orc.Solver(pb, 1, [n, n], [0.0, 0.0, 0.0], h, side_bc=bc)
    rng = np.random.default_rng(4)
    U = np.array([orc.prim2cons(p) for p in rand_prim(rng, n * n)])
    U[:, 3] = 0.0
    U[:, 4] = np.array([orc.prim2cons([p[0], p[1], p[2], 0.0, p[4]])[4] for p in (orc.cons2prim(u) for u in U)])
    R, _ = s.compute_rhs(U)
    for j in range(n):
        for i in range(n):
            c = i + n * j
            acc = np.zeros(5)
            for a, d in ((0, -1), (0, 1), (1, -1), (1, 1)):
                q = [i, j]
                q[a] += d
                nrm = np.zeros(3)
                nrm[a] = d
                if 0 <= q[a] < n:
                    other = U[q[0] + n * q[1]]
                else:
                    side = 2 * a + (d > 0)
                    other = orc.eval_bc_values(bc[side], nrm, U[c])
                F, _ = orc.eval_splitting(U[c], other, nrm)
                acc -= h * F
            assert np.allclose(R[c], acc, rtol=1e-13, atol=1e-14), (i, j)


def test_axis_permutation_symmetry(orc):
    """vortex_xy, vortex_yz and vortex_zx are the same flow with permuted axes."""
    n = 10
    out = {}
    for key, perm in (("vortex_xy", (0, 1, 2)), ("vortex_yz", (1, 2, 0)), ("vortex_zx", (2, 0, 1))):
        pb = orc.problem(orc.PROBLEM_KEYS[key], 3)
        s = orc.Solver(pb, 1, [n] * 3, [-5.0] * 3, 1.0)
        U = s.initial_state()
        R, me = s.compute_rhs(U)
        out[key] = (R.reshape(n, n, n, 5), me, perm)
    Rxy, mxy, _ = out["vortex_xy"]
    # xy: (u, v) in (x, y), uniform along z -> yz: (v, w) in (y, z) uniform along x
    Ryz, myz, _ = out["vortex_yz"]
    a = Rxy[..., [0, 1, 2, 3, 4]]                      # index order [k][j][i]
    b = np.transpose(Ryz, (2, 0, 1, 3))[..., [0, 2, 3, 1, 4]]
    assert mxy == myz
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_initial_dt(orc):
    """maxEig of the 64^2 vortex initial state in closed form (dt rule main.cpp:398)."""
    pb = orc.problem(orc.VORTEX_XY, 2)
    s = orc.Solver(pb, 1, [64, 64], [-5.0] * 3, 10 / 64)
    U = s.initial_state()
    P = np.array([orc.cons2prim(u) for u in U])
    lam = np.max(np.maximum(np.abs(P[:, 1]), np.abs(P[:, 2])) + np.sqrt(GAMMA * P[:, 4]))
    _, me = s.compute_rhs(U)
    assert me == lam
    W = U.copy()
    dt, eig = s.rk3_step(U, W, 0.45, 0.0, 2.0)
    assert dt == 0.9 * 0.45 * (10 / 64) / lam
    assert abs(me - 2.923654) < 1e-6


def test_order2_reduces_to_order1_under_flat_limiter(orc):
    """Where every one-sided difference changes sign the limiter is zero and
    the limited-linear reconstruction is the first-order one, bit for bit."""
    pb = orc.problem(orc.RADSOD, 2)
    n = 8
    s1 = orc.Solver(pb, 1, [n, n], [0.0] * 3, 0.25, side_bc=[orc.BC_PERIODIC] * 4)
    s2 = orc.Solver(pb, 2, [n, n], [0.0] * 3, 0.25, side_bc=[orc.BC_PERIODIC] * 4)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    chk = ((i + j) % 2).ravel()
    prim = np.zeros((n * n, 5))
    prim[:, 0] = 1.0 + 0.1 * chk
    prim[:, 1] = 0.2 - 0.1 * chk
    prim[:, 2] = -0.1 + 0.05 * chk
    prim[:, 4] = 1.0 + 0.2 * chk
    U = np.array([orc.prim2cons(p) for p in prim])
    R1, m1 = s1.compute_rhs(U)
    R2, m2 = s2.compute_rhs(U)
    assert np.array_equal(R1, R2) and m1 == m2


def test_order2_exact_on_linear_data(orc):
    """For linear data the limited slope is the exact slope: interior residuals
    equal those of the exact face values (uniform velocity and pressure, density
    linear in x: the mass residual is -u * d(rho)/dx * V exactly in the interior)."""
    pb = orc.problem(orc.RADSOD, 2)
    n, h = 12, 0.5
    s = orc.Solver(pb, 2, [n, n], [0.0] * 3, h, side_bc=[orc.BC_FREE_FLOW] * 4)
    cen = s.centroids()
    u, p, g = 0.5, 1.0, 0.125
    prim = np.zeros((n * n, 5))
    rho = 1.0 + g * cen[:, 0]
    prim[:, 0] = p
    prim[:, 1] = u
    prim[:, 4] = p / rho
    U = np.array([orc.prim2cons(q) for q in prim])
    R, _ = s.compute_rhs(U)
    ii = np.arange(n * n) % n
    inner = (ii >= 2) & (ii <= n - 3)
    assert np.allclose(R[inner, 0], -u * g * h * h, rtol=1e-10)


def test_order2_converges_faster(orc):
    """Smooth vortex, t=0.5: order 1 converges at ~first order (ratio -> 2 per
    refinement); the minmod-limited order 2 is 3x more accurate at 128^2 and
    converges faster (ratio 2.55 at 64 -> 128, clipped extrema keep it below 4)."""
    errs = {}
    for order in (1, 2):
        for nc in (64, 128):
            pb = orc.problem(orc.VORTEX_XY, 2, vortex_periodic=1)
            s = orc.Solver(pb, order, [nc, nc], [-5.0] * 3, 10 / nc, side_bc=[orc.BC_PERIODIC] * 4)
            U = s.initial_state()
            s.run(U, 0.5, 0.0, 0.5)
            errs[(order, nc)] = s.error(U, 0.5)[0]
    assert errs[(1, 128)] / errs[(2, 128)] > 3.0
    assert 1.7 < errs[(1, 64)] / errs[(1, 128)] < 2.1
    assert errs[(2, 64)] / errs[(2, 128)] > 2.4


@pytest.mark.parametrize("plant,case", [
    ("ORC_PLANT_REFLECT", ("radsod", 2, 64, 1.0, "1.091296206337e+00")),
    ("ORC_PLANT_DT", ("vortex_xy", 2, 64, 2.0, "1.388164788184e+00")),
    ("ORC_PLANT_RK", ("vortex_xy", 2, 64, 2.0, "1.388164788184e+00")),
])
def test_planted_mistakes_are_caught(orc, tmp_path, plant, case):
    path = orc.build_variant([plant], str(tmp_path))
    key, dim, nc, tend, expected = case
    with orc.use_library(path):
        l1, _, _, _, _ = orc.run_reference_case(orc.PROBLEM_KEYS[key], dim, nc, tend)
    # the known-answer test accepts 1e-11 relative; every plant lands far outside
    assert abs(l1 - float(expected)) / float(expected) > 1e-8, (plant, l1)
    l1, _, _, _, _ = orc.run_reference_case(orc.PROBLEM_KEYS[key], dim, nc, tend)
    assert "%.12e" % l1 == expected


THREAD_CASES = [
    ("3d free", 3, [21, 13, 17], [0] * 6, None, 1),
    ("3d periodic box", 3, [70, 9, 12], [4, 4, 1, 2, 4, 4], [[0.5, 0.5, 0.5, 1.6, 1.2, 2.0]], 1),
    ("3d order 2 dirichlet", 3, [19, 22, 8], [3, 0, 1, 0, 0, 3], [[1.0, 1.0, -1.0, 2.0, 3.0, 9.0]], 2),
    ("2d ragged", 2, [131, 7], [4, 4, 0, 1], None, 1),
    ("2d order 2 body", 2, [66, 40], [1, 0, 3, 4][:2] + [0, 1], [[2.0, 2.0, -1.0, 4.0, 3.0, 1.0]], 2),
]


@pytest.mark.parametrize("name,dim,n,bc,boxes,order", THREAD_CASES, ids=[c[0] for c in THREAD_CASES])
def test_threaded_steps_equal_serial(orc, name, dim, n, bc, boxes, order):
    """The threaded face loop and RK update (the all-cores CPU baseline) give
    the serial oracle's residual, eigenvalue maximum, counters and steps bit
    for bit, for 1, 3 and 8 threads."""
    import ctypes as C
    pb = orc.problem(orc.RADSOD, dim)
    h = 0.25
    info = np.tile([0.9, 0.3, -0.2, 0.1 if dim == 3 else 0.0, 1.1], (6, 1))
    probe = orc.Solver(pb, order, n, [0.0] * 3, h, side_bc=bc, side_info=info)
    fluid = None
    if boxes:
        fluid = np.empty(probe.ncells, dtype=np.uint8)
        b = np.ascontiguousarray(boxes, dtype=np.float64)
        orc.lib().orc_mark_fluid(C.c_void_p(probe.s), b.ctypes.data_as(C.POINTER(C.c_double)), len(b), None, 0,
                                 fluid.ctypes.data_as(C.POINTER(C.c_uint8)))
        assert 0 < fluid.sum() < probe.ncells
    rng = np.random.default_rng(len(name))
    U0 = np.array([orc.prim2cons([rng.uniform(0.5, 2), *rng.uniform(-0.7, 0.7, 2),
                                  rng.uniform(-0.7, 0.7) if dim == 3 else 0.0, rng.uniform(0.5, 2)])
                   for _ in range(probe.ncells)])
    Uneg = U0.copy()
    Uneg[3] = [-0.5, 0.0, 0.0, 0.0, -1.0]  # counted negative density and pressure
    out = {}
    for threads in (1, 3, 8):
        s = orc.Solver(pb, order, n, [0.0] * 3, h, side_bc=bc, side_info=info, fluid=fluid)
        R, me = s.compute_rhs(Uneg, threads=threads)
        # steps from a state without it (a NaN eigenvalue makes the serial maximum depend on the face order)
        U, W = U0.copy(), U0.copy()
        t, steps = 0.0, []
        for _ in range(2):
            dt, eig = s.rk3_step(U, W, 0.5, t, 1e30, threads=threads)
            t += dt
            steps.append((dt, tuple(eig)))
        out[threads] = (R, me, s.counters(), steps, U)
    ref = out[1]
    assert ref[2][1] > 0
    for threads in (3, 8):
        got = out[threads]
        assert np.array_equal(got[0], ref[0]) and got[1] == ref[1] and got[2] == ref[2]
        assert got[3] == ref[3] and np.array_equal(got[4], ref[4])


def _partitioned_steps(orc, order, G, nsteps=3):
    """Two sub-grids of a 20 x 16 grid split at row 8, each holding G rows of
    the other's cells as ghosts, stepped stage by stage with the ghosts
    exchanged after every stage and the eigenvalue reduced over both; returns
    the sub-grids' own rows joined and the whole grid's state."""
    n, cut, h, cfl = [20, 16], 8, 0.25, 0.45
    side_bc = [1, 0, 3, 1, 0, 0]
    info = np.tile([1.0, 0.2, 0.4, 0.0, 1.0 / 1.4], (6, 1))
    pb = orc.problem(orc.RADSOD, 2)
    rng = np.random.default_rng(order)
    N = n[0] * n[1]
    U0 = np.array([orc.prim2cons([rng.uniform(0.5, 2), *rng.uniform(-0.6, 0.6, 2), 0.0, rng.uniform(0.5, 2)])
                   for _ in range(N)])
    cx, cy = (np.arange(N) % n[0] + 0.5) * h, (np.arange(N) // n[0] + 0.5) * h
    fluid = (~((cx > 1.5) & (cx < 2.6) & (cy > 1.4) & (cy < 2.6))).astype(np.uint8)  # a body across the cut
    whole = orc.Solver(pb, order, n, [0.0] * 3, h, side_bc=side_bc, side_info=info, fluid=fluid)
    parts = []
    for k, (lo, hi) in enumerate([(0, cut + G), (cut - G, n[1])]):
        sl = slice(lo * n[0], hi * n[0])
        s = orc.Solver(pb, order, [n[0], hi - lo], [0.0, lo * h, 0.0], h, side_bc=side_bc, side_info=info,
                       fluid=fluid[sl])
        rows = hi - lo
        ghost = np.arange((rows - G) * n[0], rows * n[0]) if k == 0 else np.arange(0, G * n[0])
        send = np.arange((rows - 2 * G) * n[0], (rows - G) * n[0]) if k == 0 else np.arange(G * n[0], 2 * G * n[0])
        solved = fluid[sl].copy()
        solved[ghost] = 0
        s.set_solved(solved)
        parts.append({"s": s, "U": U0[sl].copy(), "W": U0[sl].copy(), "solved": solved.astype(bool),
                      "ghost": ghost, "send": send})
    a, b = parts

    def exchange(key):
        va, vb = a[key][a["send"]].copy(), b[key][b["send"]].copy()
        b[key][b["ghost"]] = va
        a[key][a["ghost"]] = vb

    Uw, Ww = U0.copy(), U0.copy()
    t = 0.0
    V = h * h
    dts = []
    for _ in range(nsteps):
        R = [p["s"].compute_rhs(p["U"]) for p in parts]
        dt = 0.9 * cfl * h / max(R[0][1], R[1][1])  # main.cpp:390-402
        for p, (Rp, _) in zip(parts, R):
            m = p["solved"]
            p["W"][m] = p["U"][m] + dt * Rp[m] / V  # main.cpp:409-422
        exchange("W")
        for stage in (2, 3):
            R = [p["s"].compute_rhs(p["W"]) for p in parts]
            for p, (Rp, _) in zip(parts, R):
                m = p["solved"]
                if stage == 2:  # main.cpp:445-458
                    p["W"][m] = 0.75 * p["U"][m] + 0.25 * (p["W"][m] + dt * Rp[m] / V)
                else:  # main.cpp:481-494
                    p["U"][m] = (1. / 3) * p["U"][m] + (2. / 3) * (p["W"][m] + dt * Rp[m] / V)
            exchange("W" if stage == 2 else "U")
        dtw, _ = whole.rk3_step(Uw, Ww, cfl, t, 1e30)
        dts.append((dt, dtw))
        t += dtw
    return np.concatenate([a["U"][:cut * n[0]], b["U"][G * n[0]:]]), Uw, dts


@pytest.mark.parametrize("order", [1, 2])
def test_partitioned_oracle_equals_the_whole_grid(orc, order):
    """Domain decomposition (main.cpp:222-237, 377-506 under MPI): ghosts as
    deep as the stencil give the whole grid's steps bit for bit.  Pins the
    oracle's unsolved fluid cells: faces between two of them skipped, the
    cells not updated, their values still read by the reconstruction.  One
    ghost row under the order-2 stencil (which reaches two cells) is seen."""
    parts, whole, dts = _partitioned_steps(orc, order, G=order)
    assert all(d == w for d, w in dts)
    assert np.array_equal(parts, whole)
    if order == 2:
        thin, whole, _ = _partitioned_steps(orc, 2, G=1)
        assert not np.array_equal(thin, whole)
