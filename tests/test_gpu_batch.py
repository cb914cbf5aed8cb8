"""Batched mode: many independent 2D grids advanced by one persistent kernel
per RK stage, the step replayed from a captured HIP graph.

Every case must be bit-identical to the same grid run alone through the
single-grid path (gather and marching kernels), whatever the work-item shape,
and must match the oracle within the 1e-12 normwise bar.  Cases have uneven
sizes, their own CFL and end times (some stop while others continue), and
every border BC kind.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
INFO = [1.0, 0.3, -0.1, 0.0, 1.0]


def cases(mmf):
    """(desc, cfl, tmax) per case; grids of uneven shapes and BCs."""
    out = []
    specs = [
        ((48, 48), [4, 4, 4, 4], 0.5, 10.0),
        ((100, 37), [0, 0, 0, 0], 0.45, 10.0),
        ((63, 64), [1, 2, 3, 0], 0.3, 10.0),
        ((1, 9), [0, 0, 1, 1], 0.5, 10.0),
        ((130, 3), [4, 4, 0, 2], 0.4, 10.0),
        ((62, 62), [3, 3, 4, 4], 0.5, 0.004),  # reaches its end time after a few steps
        ((125, 17), [2, 1, 0, 3], 0.6, 0.0),   # already at its end time: never moves
        ((64, 1), [4, 4, 0, 0], 0.5, 10.0),
    ]
    for i, (n, bc, cfl, tmax) in enumerate(specs):
        h = 2.0 / max(n)
        info = [INFO] * 6
        d = mmf.grid_desc(2, list(n), [-1.0, -1.0, 0.0], h, bc + [0, 0], info)
        out.append((d, bc, cfl, tmax))
    return out


def random_state(mmf, n, h, seed):
    rng = np.random.default_rng(seed)
    nx, ny = n
    x = -1.0 + (np.arange(nx) + 0.5) * h
    y = -1.0 + (np.arange(ny) + 0.5) * h
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2).ravel()
    N = nx * ny
    prim = np.empty((N, 5))
    prim[:, 0] = np.where(r < 0.45, 1.0, 0.125) * (1 + 0.05 * rng.uniform(-1, 1, N))
    prim[:, 1:3] = 0.3 * rng.uniform(-1, 1, (N, 2))
    prim[:, 3] = 0.0
    prim[:, 4] = np.where(r < 0.45, 1.0, 0.1) * (1 + 0.05 * rng.uniform(-1, 1, N))
    return mmf.batch_prim2cons(prim)


def setup(mmf):
    cs = cases(mmf)
    states = [random_state(mmf, (int(d.n[0]), int(d.n[1])), d.h, 10 + i) for i, (d, _, _, _) in enumerate(cs)]
    return cs, states


def make_batch(mmf, cs, states, chunk=0, blocks=0):
    b = mmf.Batch([c[0] for c in cs])
    if chunk or blocks:
        b.set_shape(chunk, blocks)
    for i, U in enumerate(states):
        b.set_state(i, U)
    b.set_time(0.0, [c[3] for c in cs], [c[2] for c in cs])
    return b


def single(mmf, d, bc, cfl, tmax, U, steps, kind):
    s = mmf.Solver(2, [int(d.n[0]), int(d.n[1])], [-1.0, -1.0, 0.0], d.h, bc + [0, 0], [INFO] * 6, 1)
    s.set_kernel(kind)
    s.set_state(U)
    s.set_time(0.0, tmax, cfl)
    s.advance(steps)
    t, dt, step, _ = s.time()
    return s.get_state(), t, step


STEPS = 6


def test_batch_equals_single_grid(mmf):
    cs, states = setup(mmf)
    b = make_batch(mmf, cs, states)
    print("cases %d, cells %d, items %d, shape %s" % (b.ncases, b.total_cells, b.nitems, b.shape()), flush=True)
    b.advance(STEPS)
    t, steps = b.times()
    for i, (d, bc, cfl, tmax) in enumerate(cs):
        got = b.get_state(i)
        for kind in (mmf.KERNEL_GATHER, mmf.KERNEL_MARCH):
            U, ts, ss = single(mmf, d, bc, cfl, tmax, states[i], STEPS, kind)
            assert np.array_equal(got, U), (i, kind)
            assert t[i] == ts and steps[i] == ss, (i, kind, t[i], ts, steps[i], ss)
    assert steps[5] < STEPS and t[5] == cs[5][3]  # stopped at its end time
    assert steps[6] == 0 and t[6] == 0.0
    assert np.array_equal(b.get_state(6), states[6])


@pytest.mark.parametrize("chunk,blocks,tail", [(1, 1, 0), (3, 7, 0), (64, 2048, 0), (1000, 300, 0), (16, 500, 4),
                                               (64, 64, 5), (8, 3, 1)])
def test_batch_shape_independent(mmf, chunk, blocks, tail):
    cs, states = setup(mmf)
    ref = make_batch(mmf, cs, states)
    ref.advance(STEPS)
    b = make_batch(mmf, cs, states, chunk, blocks)
    if tail:
        b.set_tail(tail)
        assert b.tail() == tail
    assert b.shape() == (chunk, blocks)
    b.advance(STEPS)
    assert np.array_equal(b.times()[0], ref.times()[0])
    for i in range(len(cs)):
        assert np.array_equal(b.get_state(i), ref.get_state(i)), (chunk, blocks, i)


def test_batch_graph_replays_with_eager_work_between(mmf):
    """Each replay of the captured step matches a fresh batch advanced the same
    number of steps, with host reads and state uploads between the replays."""
    cs, states = setup(mmf)
    b = make_batch(mmf, cs, states)
    for k in range(1, 4):
        b.advance(1)
        t, _ = b.times()
        ref = make_batch(mmf, cs, states)
        ref.advance(k)
        assert np.array_equal(t, ref.times()[0]), k
        for i in range(len(cs)):
            assert np.array_equal(b.get_state(i), ref.get_state(i)), (k, i)
        b.set_state(0, b.get_state(0))  # host round trip of one case between replays


def test_batch_against_oracle(mmf, orc):
    cs, states = setup(mmf)
    b = make_batch(mmf, cs, states)
    b.advance(3)
    t, steps = b.times()
    po = orc.problem(orc.RADSOD, 2)
    for i, (d, bc, cfl, tmax) in enumerate(cs):
        n = [int(d.n[0]), int(d.n[1])]
        o = orc.Solver(po, 1, n, [-1.0, -1.0, 0.0], d.h, side_bc=bc + [0, 0], side_info=[INFO] * 6)
        U = states[i].copy()
        W = np.empty_like(U)
        to = 0.0
        for _ in range(3):
            if not (to < tmax):
                break
            dt, _ = o.rk3_step(U, W, cfl, to, tmax)
            to += dt
        got = b.get_state(i)
        den = np.max(np.abs(U))
        err = np.max(np.abs(got - U)) / den
        print("case %d n=%s: normwise %.3e, identical %.6f" % (i, n, err, np.mean(got == U)), flush=True)
        assert err <= TOL, i
        assert abs(t[i] - to) <= 1e-15 * max(1.0, to), i


def test_batch_autotune_restores_state(mmf):
    cs, states = setup(mmf)
    ref = make_batch(mmf, cs, states)
    ref.advance(STEPS)
    b = make_batch(mmf, cs, states)
    chunk, blocks, ms = b.autotune(steps_per_trial=1)
    print("autotune: chunk %d blocks %d %.4f ms/step" % (chunk, blocks, ms), flush=True)
    assert b.shape() == (chunk, blocks) and ms > 0
    t, steps = b.times()
    assert np.all(t == 0.0) and np.all(steps == 0)
    b.advance(STEPS)
    assert np.array_equal(b.times()[0], ref.times()[0])
    for i in range(len(cs)):
        assert np.array_equal(b.get_state(i), ref.get_state(i)), i


def test_batch_problem_setup_and_error(mmf, orc):
    """Initial states and the error from the problem catalogue, per case."""
    pv = mmf.problem("vortex_xy", 2)
    pr = mmf.problem("radsod", 2)
    descs = [mmf.problem_grid_desc(pv, 40), mmf.problem_grid_desc(pr, 56), mmf.problem_grid_desc(pv, 24)]
    probs = [pv, pr, pv]
    b = mmf.Batch(descs)
    b.set_initial_states(probs)
    tend = [0.05, 0.02, 0.05]
    b.set_time(0.0, tend, 0.45)
    b.run()
    t, steps = b.times()
    assert np.allclose(t, tend, rtol=1e-14, atol=0)
    l1, _ = b.error_vs_exact(probs)
    for i, (key, nc) in enumerate([(orc.VORTEX_XY, 40), (orc.RADSOD, 56), (orc.VORTEX_XY, 24)]):
        pb = orc.problem(key, 2)
        o = orc.Solver(pb, 1, [nc, nc], list(pb.domain_origin), pb.domain_length / nc)
        U = o.initial_state()
        so, to = o.run(U, 0.45, 0.0, tend[i])
        eo, _ = o.error(U, tend[i])
        assert so == steps[i] and to == t[i], i
        assert abs(l1[i] - eo) <= 1e-12 * eo, (i, l1[i], eo)
