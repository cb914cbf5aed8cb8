"""HIP path against the oracle, through the C-ABI (include/mmf.h).

Bar: the RHS must match the reference CPU path within 1e-12 normwise relative
L-infinity (max|dR| / max|R|); both sides are built without FMA contraction
and evaluate the reference's expressions in the same order, so in practice
the results are bit-identical and the tests also report that.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))["cases"]
TOL = 1e-12


def normwise(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


def pair(mmf, orc, key, dim, ncells, order=1, side_bc=None, side_info=None, fluid=None, length=None,
         origin=None, **pbkw):
    """Same grid in the oracle and on the GPU."""
    po = orc.problem(orc.PROBLEM_KEYS[key] if key in orc.PROBLEM_KEYS else 8, dim, **pbkw)
    pm = mmf.problem(key, dim, **pbkw)
    dim = pm.dim
    org = list(pm.domain_origin) if origin is None else list(origin)
    L = pm.domain_length if length is None else length
    h = L / ncells
    n = [ncells] * dim
    if side_info is not None:
        gs = mmf.Solver(dim, n, org, h, side_bc, side_info, order)
    else:
        gs = mmf.Solver.for_problem(pm, ncells, order=order, side_bc=side_bc, origin=org, length=L)
    obc = None if side_bc is None else list(side_bc)
    os_ = orc.Solver(po, order, n, org, h, side_bc=obc, side_info=side_info, fluid=fluid)
    if fluid is not None:
        gs.set_cell_flags(fluid)
    return po, pm, os_, gs


def perturbed(U, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    V = U.copy()
    V[:, 0] *= 1.0 + amp * rng.uniform(-1, 1, U.shape[0])
    V[:, 1:4] += amp * rng.uniform(-1, 1, (U.shape[0], 3))
    V[:, 4] *= 1.0 + amp * rng.uniform(-1, 1, U.shape[0])
    return V


def check_rhs(os_, gs, U, label):
    Ro, mo = os_.compute_rhs(U)
    gs.set_state(U)
    Rg, mg = gs.compute_rhs()
    err = normwise(Rg, Ro)
    exact = float(np.mean(Rg == Ro))
    print("%s: normwise %.3e, bit-identical fraction %.6f, maxEig %r vs %r" % (label, err, exact, mg, mo), flush=True)
    assert err <= TOL, label
    assert abs(mg - mo) <= 1e-15 * mo, label
    return err, exact


# ---- point functions ----------------------------------------------------------

def random_states(rng, n):
    prim = np.empty((n, 5))
    prim[:, 0] = rng.uniform(0.05, 3.0, n)
    prim[:, 1:4] = rng.uniform(-2.0, 2.0, (n, 3))
    prim[:, 4] = rng.uniform(0.2, 3.0, n)
    return prim


def test_point_functions_bitwise(mmf, orc):
    rng = np.random.default_rng(1)
    n = 512
    prim = random_states(rng, n)
    cons = np.array([orc.prim2cons(p) for p in prim])
    assert np.array_equal(mmf.batch_prim2cons(prim), cons)
    assert np.array_equal(mmf.batch_cons2prim(cons), np.array([orc.cons2prim(c) for c in cons]))
    normals = rng.normal(size=(n, 3))
    normals /= np.linalg.norm(normals, axis=1)[:, None]
    axis = np.zeros((n, 3))
    axis[np.arange(n), rng.integers(0, 3, n)] = rng.choice([-1.0, 1.0], n)
    for nrm in (normals, axis):
        pr = np.array([orc.cons2prim(c) for c in cons])
        f = mmf.batch_eval_fluxes(cons, pr, nrm)
        assert np.array_equal(f, np.array([orc.eval_fluxes(c, p, m) for c, p, m in zip(cons, pr, nrm)]))
        R = cons[rng.permutation(n)]
        F, lam = mmf.batch_eval_splitting(cons, R, nrm)
        exp = [orc.eval_splitting(a, b, m) for a, b, m in zip(cons, R, nrm)]
        assert np.array_equal(F, np.array([e[0] for e in exp]))
        assert np.array_equal(lam, np.array([e[1] for e in exp]))
        for bc in (orc.BC_FREE_FLOW, orc.BC_REFLECTING, orc.BC_WALL, orc.BC_DIRICHLET):
            info = random_states(rng, n)
            got = mmf.batch_eval_bc_values(bc, nrm, cons, info)
            exp = np.array([orc.eval_bc_values(bc, m, c, i) for m, c, i in zip(nrm, cons, info)])
            assert np.array_equal(got, exp), bc


# ---- euler::computeRHS -------------------------------------------------------------

REF_GRIDS = [("vortex_xy", 2, 64), ("radsod", 2, 64), ("vortex_xy", 3, 32), ("vortex_yz", 3, 32),
             ("radsod", 3, 32)]


@pytest.mark.parametrize("key,dim,nc", REF_GRIDS)
def test_rhs_reference_grids(mmf, orc, key, dim, nc):
    po, pm, os_, gs = pair(mmf, orc, key, dim, nc)
    U0 = os_.initial_state()
    gs.set_initial_state(pm)
    assert np.array_equal(gs.get_state(), U0), "initial state differs"
    check_rhs(os_, gs, U0, "%s %dD %d IC" % (key, dim, nc))
    check_rhs(os_, gs, perturbed(U0, 3), "%s %dD %d perturbed" % (key, dim, nc))


@pytest.mark.parametrize("order", [1, 2])
def test_rhs_bodies_and_all_bcs(mmf, orc, order):
    """Box body (inclusive bounds), wall faces, every border BC on one grid."""
    side_bc = [orc.BC_DIRICHLET, orc.BC_FREE_FLOW, orc.BC_REFLECTING, orc.BC_WALL, orc.BC_FREE_FLOW,
               orc.BC_REFLECTING]
    info = np.zeros((6, 5))
    info[0] = [1.0, 0.4, 0.1, 0.0, 1.0 / 1.4]
    po, pm, os_, gs = pair(mmf, orc, "radsod", 3, 24, order=order, side_bc=side_bc, side_info=info,
                           origin=[-1, -1, -1], length=2.0, sod_r0=0.5)
    cen = os_.centroids()
    box = np.array([-0.3, -0.2, -0.25, 0.1, 0.25, 0.3])
    inside = np.all((cen >= box[:3]) & (cen <= box[3:]), axis=1)
    fluid = (~inside).astype(np.uint8)
    os2 = orc.Solver(po, order, [24] * 3, [-1, -1, -1], 2.0 / 24, side_bc=side_bc, side_info=info, fluid=fluid)
    gs.set_bodies(boxes=[box])
    assert np.array_equal(gs.cell_flags() & 1, fluid)
    U0 = os2.initial_state()
    check_rhs(os2, gs, perturbed(U0, 5), "bodies+bcs order %d" % order)


def test_rhs_sphere_freestream(mmf, orc):
    side_bc = [orc.BC_DIRICHLET] + [orc.BC_FREE_FLOW] * 5
    fs = [1.0 / 1.4, 0.3, 0.0, 0.0, 1.0 / 1.4]
    info = np.array([fs] * 6)
    n = 32
    po, pm, os_, gs = pair(mmf, orc, "freestream", 3, n, side_bc=side_bc, side_info=info,
                           origin=[-2, -2, -2], length=4.0, fs_prim=fs)
    gs.set_bodies(spheres=[[0.0, 0.0, 0.0, 0.5]])
    fl = gs.cell_flags() & 1
    cen = gs.centroids()
    phi = np.hypot(np.hypot(cen[:, 0], cen[:, 1]), cen[:, 2]) - 0.5
    assert np.min(np.abs(phi)) > 1e-12
    assert np.array_equal(fl == 1, phi >= 0.0)  # the level set at the centroids
    assert 0 < int((fl == 0).sum()) < n ** 3
    os2 = orc.Solver(po, 1, [n] * 3, [-2, -2, -2], 4.0 / n, side_bc=side_bc, side_info=info, fluid=fl)
    U0 = os2.initial_state()
    check_rhs(os2, gs, perturbed(U0, 11, 1e-2), "sphere freestream")


def test_rhs_periodic_order2(mmf, orc):
    per = [orc.BC_PERIODIC] * 4
    po, pm, os_, gs = pair(mmf, orc, "vortex_xy", 2, 64, order=2, side_bc=per, vortex_periodic=1)
    U0 = os_.initial_state()
    check_rhs(os_, gs, U0, "periodic vortex order 2")
    check_rhs(os_, gs, perturbed(U0, 9, 1e-2), "periodic vortex order 2 perturbed")


def test_unsupported_order(mmf):
    with pytest.raises(mmf.MMFError) as e:
        mmf.Solver(2, [8, 8], [0, 0, 0], 1.0, [0] * 6, order=3)
    assert e.value.status == mmf.ERR_UNSUPPORTED_ORDER


def test_raw_order_permutation(mmf, orc):
    po, pm, os_, gs = pair(mmf, orc, "vortex_xy", 2, 32)
    U = perturbed(os_.initial_state(), 2)
    Ro, _ = os_.compute_rhs(U)
    perm = np.random.default_rng(4).permutation(gs.ncells)  # raw r holds structured cell perm[r]
    gs.set_raw_order(perm)
    gs.set_state(U[perm])
    Rg, _ = gs.compute_rhs()
    assert np.array_equal(Rg, Ro[perm])
    assert np.array_equal(gs.get_state(), U[perm])  # the transposes run on the device
    gs.set_raw_order(None)
    assert np.array_equal(gs.get_state(), U)
    Rg, _ = gs.compute_rhs()
    assert np.array_equal(Rg, Ro)


# ---- time loop -----------------------------------------------------------------------

@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_known_answers_on_gpu(mmf, orc, case):
    """The reference's five tests run end-to-end on the GPU."""
    pm = mmf.problem(case["problem"], case["dimensions"])
    tend = case["end"] if case["end"] is not None else mmf.default_end_time(pm)
    gs = mmf.Solver.for_problem(pm, case["nCells"])
    gs.set_initial_state(pm)
    gs.set_time(0.0, tend, case["CFL"])
    steps, t = gs.run()
    l1, l2 = gs.error_vs_exact(pm, tend)
    exp = float(case["expected"])
    print("%s: steps %d  L1 %.12e  expected %s" % (case["name"], steps, l1, case["expected"]), flush=True)
    assert t == tend
    assert abs(l1 - exp) / exp < 1e-11
    # and state parity with the oracle
    l1o, _, steps_o, so, Uo = orc.run_reference_case(orc.PROBLEM_KEYS[case["problem"]], case["dimensions"],
                                                     case["nCells"], tend, case["CFL"])
    assert steps == steps_o
    Ug = gs.get_state()
    err = normwise(Ug, Uo)
    print("   state normwise %.3e bit-identical %.6f" % (err, float(np.mean(Ug == Uo))), flush=True)
    assert err <= TOL


def test_stepwise_parity_order2_bodies(mmf, orc):
    """Five RK3 steps with limited reconstruction, a body and mixed BCs."""
    side_bc = [orc.BC_DIRICHLET, orc.BC_FREE_FLOW, orc.BC_REFLECTING, orc.BC_REFLECTING,
               orc.BC_FREE_FLOW, orc.BC_FREE_FLOW]
    info = np.zeros((6, 5))
    info[0] = [1.0, 0.5, 0.0, 0.0, 1.0]
    n = 20
    po = orc.problem(orc.RADSOD, 3, sod_r0=0.5)
    pm = mmf.problem("radsod", 3, sod_r0=0.5)
    gs = mmf.Solver(3, [n] * 3, [-1, -1, -1], 2.0 / n, side_bc, info, order=2)
    gs.set_bodies(boxes=[[0.2, -0.3, -0.3, 0.5, 0.3, 0.3]])
    fl = gs.cell_flags() & 1
    os_ = orc.Solver(po, 2, [n] * 3, [-1, -1, -1], 2.0 / n, side_bc=side_bc, side_info=info, fluid=fl)
    U = os_.initial_state()
    W = U.copy()
    gs.set_state(U)
    gs.set_time(0.0, 10.0, 0.5)
    t = 0.0
    for s in range(5):
        dt, eig = os_.rk3_step(U, W, 0.5, t, 10.0)
        t += dt
        gs.advance(1)
        tg, dtg, stepg, eigg = gs.time()
        assert dtg == dt and tg == t and stepg == s + 1
        assert np.array_equal(eigg, eig)
        Ug = gs.get_state()
        err = normwise(Ug, U)
        print("step %d: dt %.6e normwise %.3e bit-identical %.6f" % (s, dt, err, float(np.mean(Ug == U))), flush=True)
        assert err <= TOL


def test_periodic_vortex_order2_run(mmf, orc):
    """The periodic 2D vortex with limited reconstruction at CFL 0.5 (t=1)."""
    per = [orc.BC_PERIODIC] * 4
    po, pm, os_, gs = pair(mmf, orc, "vortex_xy", 2, 64, order=2, side_bc=per, vortex_periodic=1)
    U = os_.initial_state()
    steps_o, t_o = os_.run(U, 0.5, 0.0, 1.0)
    gs.set_initial_state(pm)
    gs.set_time(0.0, 1.0, 0.5)
    steps, t = gs.run()
    assert steps == steps_o and t == t_o
    Ug = gs.get_state()
    err = normwise(Ug, U)
    l1g, l2g = gs.error_vs_exact(pm, 1.0)
    l1o, l2o = os_.error(U, 1.0)
    print("periodic order 2: steps %d normwise %.3e L1 %.6e vs %.6e" % (steps, err, l1g, l1o), flush=True)
    assert err <= TOL
    assert abs(l1g - l1o) <= 1e-12 * l1o and abs(l2g - l2o) <= 1e-12 * l2o


def test_determinism(mmf):
    pm = mmf.problem("radsod", 3, sod_r0=0.4, domain_origin=[-1, -1, -1], domain_length=2.0)
    outs = []
    for _ in range(2):
        gs = mmf.Solver.for_problem(pm, 48, side_bc=[0] * 6)
        gs.set_initial_state(pm)
        gs.set_time(0.0, 1.0, 0.5)
        gs.advance(4)
        outs.append(gs.get_state())
        gs.close()
    assert np.array_equal(outs[0], outs[1])


def test_sphere_freestream_steps(mmf, orc):
    """The immersed-sphere set-up (Mach 0.3, Dirichlet inflow, free-flow outflow)
    on a 40^3 grid: eight RK3 steps against the oracle, step by step."""
    n = 40
    fs = [1.0 / 1.4, 0.3, 0.0, 0.0, 1.0 / 1.4]
    side_bc = [orc.BC_DIRICHLET] + [orc.BC_FREE_FLOW] * 5
    info = np.array([fs] * 6)
    pm = mmf.problem("freestream", 3, fs_prim=fs)
    gs = mmf.Solver(3, [n] * 3, [-2.0] * 3, 4.0 / n, side_bc, info, 1)
    gs.set_bodies(spheres=[[0.0, 0.0, 0.0, 0.5]])
    fl = gs.cell_flags() & 1
    po = orc.problem(8, 3, fs_prim=fs)
    os_ = orc.Solver(po, 1, [n] * 3, [-2.0] * 3, 4.0 / n, side_bc=side_bc, side_info=info, fluid=fl)
    U = os_.initial_state()
    W = U.copy()
    gs.set_state(U)
    gs.set_time(0.0, 100.0, 0.5)
    t = 0.0
    for s in range(8):
        dt, eig = os_.rk3_step(U, W, 0.5, t, 100.0)
        t += dt
    gs.advance(8)
    tg, dtg, stepg, eigg = gs.time()
    Ug = gs.get_state()
    err = normwise(Ug, U)
    print("sphere 8 steps: normwise %.3e bit-identical %.6f" % (err, float(np.mean(Ug == U))), flush=True)
    assert tg == t and stepg == 8
    assert err <= TOL


@pytest.mark.parametrize("key,dim,nc,order,march", [("vortex_xy", 2, 40, 2, False), ("radsod", 3, 24, 1, False),
                                                     ("radsod", 2, 64, 1, False), ("radsod", 3, 32, 1, True)])
def test_graph_replayed_steps_equal_launched_steps(mmf, key, dim, nc, order, march):
    """Steps replayed from captured graphs (16-step and 1-step graphs, with
    host reads, a state upload and a kernel change between replays) equal
    steps launched one by one.  With march, the 3D marching kernel (whose
    stage-3 launch advances the time) runs first."""
    pm = mmf.problem(key, dim)
    runs = []
    for graphs in (False, True):
        s = mmf.Solver.for_problem(pm, nc, order=order)
        if march:
            s.set_kernel(mmf.KERNEL_MARCH, 8, 8)
            assert s.kernel()[0] == mmf.KERNEL_MARCH
        s.set_step_graphs(graphs)
        s.set_initial_state(pm)
        s.set_time(0.0, 100.0, 0.45)
        states = []
        for n in (17, 1, 3, 16):
            s.advance(n)
            states.append(s.get_state())
            s.set_state(states[-1])
            t, dt, step, _ = s.time()
            states.append(np.array([t, dt, step]))
        s.set_kernel(mmf.KERNEL_GATHER)
        s.advance(5)
        states.append(s.get_state())
        runs.append(states)
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_vortex_64_periodic_order2_to_t10(mmf, orc):
    """The reference CPU configuration end to end: 2D isentropic vortex, 64^2 on
    [-5,5]^2, periodic, beta 5, limited linear + Rusanov, SSP-RK3 at CFL 0.5 to
    t = 10 (one full period of the free stream).  GPU and oracle take the same
    number of steps to the same time and end on the same state."""
    pm = mmf.problem("vortex_xy", 2, vortex_periodic=1)
    po = orc.problem(orc.VORTEX_XY, 2, vortex_periodic=1)
    assert pm.vortex_beta == 5.0
    s = mmf.Solver.for_problem(pm, 64, order=2, side_bc=[mmf.BC_PERIODIC] * 4)
    s.set_initial_state(pm)
    s.set_time(0.0, 10.0, 0.5)
    steps, t = s.run()
    o = orc.Solver(po, 2, [64, 64], list(po.domain_origin), po.domain_length / 64, side_bc=[4] * 4 + [0, 0])
    U = o.initial_state()
    so, to = o.run(U, 0.5, 0.0, 10.0)
    G = s.get_state()
    err = np.max(np.abs(G - U)) / np.max(np.abs(U))
    l1g, _ = s.error_vs_exact(pm, 10.0)
    l1o, _ = o.error(U, 10.0)
    print("steps %d/%d t %.15f/%.15f normwise %.3e identical %.4f L1 %.12e / %.12e"
          % (steps, so, t, to, err, np.mean(G == U), l1g, l1o), flush=True)
    assert steps == so and t == to
    assert err <= 1e-12
    assert abs(l1g - l1o) <= 1e-12 * l1o


def test_error_norms_closed_form(mmf):
    """mmf_error_norms: a constant density offset delta gives L1 = |delta| |Omega|
    and L2 = |delta| sqrt(|Omega|) exactly (dyadic data, every sum exact); two
    offsets on halves give L2 = sqrt((d1^2 + d2^2)/2 |Omega|)."""
    n, h = 32, 0.25  # |Omega| = 64
    gs = mmf.Solver(2, [n, n], [0.0, 0.0, 0.0], h, [mmf.BC_FREE_FLOW] * 4)
    rng = np.random.default_rng(3)
    U = np.zeros((n * n, 5))
    U[:, 0] = rng.choice([0.5, 1.0, 2.0], n * n)
    U[:, 4] = 2.5 * U[:, 0]
    gs.set_state(U)
    d = 2.0 ** -6
    l1, l2 = gs.error_norms(U[:, 0] - d)
    assert l1 == d * 64.0 and l2 == d * 8.0
    ex = U[:, 0].copy()
    ex[: n * n // 2] += 3 * d
    ex[n * n // 2:] -= 4 * d
    l1, l2 = gs.error_norms(ex)
    assert l1 == 3.5 * d * 64.0
    assert l2 == pytest.approx(np.sqrt(12.5 * d * d * 64.0), rel=1e-15)
    assert gs.error_norms(U[:, 0]) == (0.0, 0.0)
    gs.close()
