"""The measured workloads at their full sizes, against the oracle.

The 256^3 radial Sod grid runs the marching kernel at its launch shape
(64 x 8 tiles, 128 planes per block, the step's dt gathered by the previous
step's stage 3) and the 192^3 immersed sphere the masked marching kernel with
the wall reflection; a few SSP-RK3 steps of each are compared bit for bit with
the oracle's steps (plain C, about 4 s per 256^3 step) from the same initial
state, and so are two steps of each of the 512 cases of the batched
workload.  Longer 256^3 runs are checked through properties that hold at any
size: the marching and gather kernels agree bit for bit, total mass is
conserved while the waves are inside the domain (free-flow borders at rest
carry no mass flux), and the solution keeps the problem's mirror and
axis-exchange symmetries to round-off.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_SOD = 256
N_SPHERE = 192
FS = [1.0 / 1.4, 0.3, 0.0, 0.0, 1.0 / 1.4]


def radsod_pair(mmf, orc):
    pm = mmf.problem("radsod", 3, sod_r0=0.4, domain_origin=[-1.0] * 3, domain_length=2.0)
    gs = mmf.Solver.for_problem(pm, N_SOD, order=1, side_bc=[mmf.BC_FREE_FLOW] * 6)
    po = orc.problem(orc.RADSOD, 3, sod_r0=0.4)
    os_ = orc.Solver(po, 1, [N_SOD] * 3, [-1.0] * 3, 2.0 / N_SOD, side_bc=[orc.BC_FREE_FLOW] * 6)
    return pm, gs, os_


def radsod_gpu(mmf, kind=None):
    pm = mmf.problem("radsod", 3, sod_r0=0.4, domain_origin=[-1.0] * 3, domain_length=2.0)
    gs = mmf.Solver.for_problem(pm, N_SOD, order=1, side_bc=[mmf.BC_FREE_FLOW] * 6)
    if kind is not None:
        gs.set_kernel(kind)
    gs.set_initial_state(pm)
    gs.set_time(0.0, 1e30, 0.5)
    return gs


def test_radsod_256_steps_bitwise(mmf, orc):
    pm, gs, os_ = radsod_pair(mmf, orc)
    kind, rows, chunk = gs.kernel()
    print("256^3 radial Sod: kernel %d, tile rows %d, chunk %d" % (kind, rows, chunk), flush=True)
    assert (kind, rows) == (mmf.KERNEL_MARCH, 8)  # 128 planes per block on 256 CUs
    U = os_.initial_state()
    gs.set_initial_state(pm)
    assert np.array_equal(gs.get_state(), U)
    gs.set_time(0.0, 1e30, 0.5)
    W = U.copy()
    t = 0.0
    for _ in range(2):
        dt, _ = os_.rk3_step(U, W, 0.5, t, 1e30)
        t += dt
    gs.advance(2)
    tg, dtg, step, _ = gs.time()
    Ug = gs.get_state()
    same = float(np.mean(Ug == U))
    print("256^3 radial Sod, 2 steps: t %r vs %r, bit-identical fraction %.8f" % (tg, t, same), flush=True)
    assert step == 2 and tg == t and dtg == dt
    assert np.array_equal(Ug, U)
    assert gs.positivity_counters() == (0, 0)
    gs.close()


def test_sphere_192_steps_bitwise(mmf, orc):
    side_bc = [mmf.BC_DIRICHLET] + [mmf.BC_FREE_FLOW] * 5
    info = np.array([FS] * 6)
    h = 4.0 / N_SPHERE
    gs = mmf.Solver(3, [N_SPHERE] * 3, [-2.0] * 3, h, side_bc, info, 1)
    gs.set_bodies(spheres=[[0.0, 0.0, 0.0, 0.25]])
    fl = gs.cell_flags() & 1
    # the level set at the centroids (numpy's hypot; no centroid within 1e-12 of the surface)
    cen = gs.centroids()
    phi = np.hypot(np.hypot(cen[:, 0], cen[:, 1]), cen[:, 2]) - 0.25
    assert np.min(np.abs(phi)) > 1e-12
    assert np.array_equal(fl == 1, phi >= 0.0)
    assert 0 < int((fl == 0).sum()) < N_SPHERE ** 3
    kind, rows, chunk = gs.kernel()
    print("192^3 sphere: kernel %d, tile rows %d, chunk %d, %d solid cells" % (kind, rows, chunk, int((fl == 0).sum())),
          flush=True)
    assert kind == mmf.KERNEL_MARCH
    po = orc.problem(orc.FFSTEP + 1, 3, fs_prim=FS)  # the uniform free stream
    os_ = orc.Solver(po, 1, [N_SPHERE] * 3, [-2.0] * 3, h, side_bc=side_bc, side_info=info, fluid=fl)
    U = os_.initial_state()
    gs.set_state(U)
    gs.set_time(0.0, 1e30, 0.5)
    W = U.copy()
    t = 0.0
    for _ in range(3):
        dt, _ = os_.rk3_step(U, W, 0.5, t, 1e30)
        t += dt
    gs.advance(3)
    tg, dtg, step, _ = gs.time()
    Ug = gs.get_state()
    print("192^3 sphere, 3 steps: bit-identical fraction %.8f" % float(np.mean(Ug == U)), flush=True)
    assert step == 3 and tg == t and dtg == dt
    assert np.array_equal(Ug, U)
    gs.close()


def test_radsod_256_march_equals_gather(mmf):
    states = []
    for kind in (mmf.KERNEL_MARCH, mmf.KERNEL_GATHER):
        gs = radsod_gpu(mmf, kind)
        assert gs.kernel()[0] == kind
        gs.advance(6)
        states.append((gs.get_state(), gs.time()[:3]))
        gs.close()
    (a, ta), (b, tb) = states
    print("256^3 radial Sod, 6 steps: march vs gather bit-identical fraction %.8f" % float(np.mean(a == b)), flush=True)
    assert ta == tb
    assert np.array_equal(a, b)


def test_radsod_256_mass_and_symmetry(mmf):
    gs = radsod_gpu(mmf)
    n = N_SOD
    U0 = gs.get_state()
    m0 = float(np.sum(U0[:, 0]))
    gs.advance(40)
    t, _, step, _ = gs.time()
    U = gs.get_state()
    gs.close()
    assert step == 40
    # the initial border states are untouched while the waves are inside
    # (the outer shock runs at about 1.75 from r = 0.4; the nearest border is at 1)
    assert 0.4 + 2.0 * t < 1.0 - 2.0 / n
    m = float(np.sum(U[:, 0]))
    print("256^3 radial Sod, 40 steps to t = %.5f: mass %.17g -> %.17g (rel %.2e)" % (t, m0, m, abs(m - m0) / m0),
          flush=True)
    assert abs(m - m0) <= 1e-13 * m0
    F = U.reshape(n, n, n, 5)  # [z][y][x][component]
    scale = np.max(np.abs(F), axis=(0, 1, 2))
    # mirror x -> -x: rho, rho v, rho w, rho E even, rho u odd
    sign = np.array([1.0, -1.0, 1.0, 1.0, 1.0])
    mir = np.max(np.abs(F - sign * F[:, :, ::-1, :]), axis=(0, 1, 2)) / scale
    # exchange of x and y: rho u <-> rho v
    sw = F.transpose(0, 2, 1, 3)[..., [0, 2, 1, 3, 4]]
    exch = np.max(np.abs(F - sw), axis=(0, 1, 2)) / scale
    print("mirror-x deviation per component %s, x<->y %s" % (mir, exch), flush=True)
    assert np.all(mir <= 1e-12) and np.all(exch <= 1e-12)


def test_batch_512_mixed_set_steps_bitwise(mmf, orc):
    """The 512-case mixed set (periodic vortices and transmissive radial Sod,
    sides 32..256) at the autotuned launch shape: two steps of every case
    equal the oracle's bit for bit."""
    from minimmerflow_amd import workloads as WL
    descs, problems, _ = WL.mixed_2d_cases(mmf, 512, WL.DEFAULT_SEED)
    b = mmf.Batch(descs)
    b.set_initial_states(problems)
    b.set_time(0.0, 1e30, 0.5)
    chunk, blocks, _ = b.autotune(steps_per_trial=3)
    print("512 cases, %d cells: %d rows per item, %d blocks" % (b.total_cells, chunk, blocks), flush=True)
    b.advance(2)
    t, steps = b.times()
    keys = {0: orc.VORTEX_XY, 3: orc.RADSOD}
    same = 0
    for i, (d, pm) in enumerate(zip(descs, problems)):
        po = orc.problem(keys[pm.type], 2, vortex_beta=pm.vortex_beta, vortex_center=list(pm.vortex_center),
                         vortex_periodic=pm.vortex_periodic, domain_origin=list(pm.domain_origin),
                         domain_length=pm.domain_length, sod_r0=pm.sod_r0, sod_p_in=pm.sod_p_in,
                         sod_t_in=pm.sod_t_in, sod_p_out=pm.sod_p_out, sod_t_out=pm.sod_t_out)
        n = [int(d.n[0]), int(d.n[1])]
        o = orc.Solver(po, 1, n, list(d.origin), d.h, side_bc=[d.side_bc[k] for k in range(6)])
        U = o.initial_state()
        W = U.copy()
        to = 0.0
        for _ in range(2):
            dt, _ = o.rk3_step(U, W, 0.5, to, 1e30)
            to += dt
        got = b.get_state(i)
        assert steps[i] == 2 and t[i] == to, (i, n)
        assert np.array_equal(got, U), (i, n)
        same += 1
    print("512 cases, 2 steps: %d bit-identical to the oracle" % same, flush=True)
