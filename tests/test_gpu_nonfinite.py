"""Non-finite states: a NaN injected into the state stops the time loop.

The reference has no such check (it logs negative p / rho, euler.cpp:94-101,
and its dt and t would turn NaN, main.cpp:398-402).  Here a step whose stage
maximum is NaN or infinite leaves t, dt and the step count unchanged and is
counted; mmf_run, mmf_get_time and the batch's run / get_times then return
MMF_ERR_STATE.  The running maximum of the face eigenvalues keeps a NaN, so a
NaN cell is seen in its first step wherever it is the neighbour side of a face;
a corner cell whose lower faces are all Dirichlet borders (the cell only ever
the owner side) is seen one stage later, through its neighbours' fluxes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INFO = [0.8, 0.2, -0.1, 0.05, 1.1]


def grid(mmf, dim, n, order, side_bc):
    h = 2.0 / n[0]
    s = mmf.Solver(dim, n, [-1.0, -1.0, -1.0], h, side_bc, [INFO] * 6, order)
    pb = mmf.problem("radsod", dim, sod_r0=0.4, domain_origin=[-1.0, -1.0, -1.0], domain_length=2.0)
    s.set_initial_state(pb)
    return s


VARIANTS = [
    # dim, cells, order, kernel
    (2, [48, 40], 1, "GATHER"),
    (2, [64, 40], 1, "MARCH"),
    (2, [48, 48], 2, "TILE"),
    (2, [64, 48], 2, "MARCH"),
    (3, [16, 12, 10], 1, "GATHER"),
    (3, [64, 16, 16], 1, "MARCH"),
    (3, [16, 12, 10], 2, "GATHER"),
]


@pytest.mark.parametrize("dim,n,order,kname", VARIANTS)
@pytest.mark.parametrize("where,bcs", [("interior", "free"), ("corner", "dirichlet")])
def test_nan_state_stops_the_run(mmf, dim, n, order, kname, where, bcs):
    side_bc = ([mmf.BC_FREE_FLOW] * 6 if bcs == "free"
               else [mmf.BC_DIRICHLET, mmf.BC_FREE_FLOW] * 3)
    s = grid(mmf, dim, n, order, side_bc)
    s.set_kernel(getattr(mmf, "KERNEL_" + kname))
    assert s.kernel()[0] == getattr(mmf, "KERNEL_" + kname)
    s.set_time(0.0, 1.0, 0.5)
    s.advance(2)
    t2, _, step2, _ = s.time()
    U = s.get_state()
    cell = 0 if where == "corner" else (n[0] // 2) + n[0] * (n[1] // 3)
    U[cell, 4] = np.nan
    s.set_state(U)
    print("nan at cell %d, %dD %s order %d %s" % (cell, dim, kname, order, bcs), flush=True)
    with pytest.raises(mmf.MMFError) as e:
        s.run(max_steps=6)
    assert e.value.status == mmf.ERR_STATE
    bad = s.nonfinite_steps()
    assert 1 <= bad <= 6
    with pytest.raises(mmf.MMFError) as e2:
        s.time()
    assert e2.value.status == mmf.ERR_STATE
    # the time advanced by at most one step past the injection (the corner case), and stays finite
    s.reset_positivity_counters()
    t, dt, step, _ = s.time()
    assert np.isfinite(t) and np.isfinite(dt) and t2 <= t and step2 <= step <= step2 + 1
    s.close()


def test_finite_runs_report_no_nonfinite_steps(mmf):
    s = grid(mmf, 3, [64, 16, 16], 1, [mmf.BC_FREE_FLOW] * 6)
    s.set_kernel(mmf.KERNEL_MARCH)
    s.set_time(0.0, 1.0, 0.5)
    s.run(max_steps=8)
    assert s.nonfinite_steps() == 0
    s.close()


def test_nan_in_one_batch_case(mmf):
    descs = [mmf.grid_desc(2, [40, 36], [-1.0, -1.0, 0.0], 2.0 / 40, [0] * 6, [INFO] * 6),
             mmf.grid_desc(2, [64, 64], [-1.0, -1.0, 0.0], 2.0 / 64, [0] * 6, [INFO] * 6)]
    b = mmf.Batch(descs)
    pbs = [mmf.problem("radsod", 2, sod_r0=0.4, domain_origin=[-1.0, -1.0, -1.0], domain_length=2.0)] * 2
    b.set_initial_states(pbs)
    b.set_time(0.0, 1.0, 0.5)
    U = b.get_state(1)
    U[100, 0] = np.nan
    b.set_state(1, U)
    with pytest.raises(mmf.MMFError) as e:
        b.run(max_steps=4)
    assert e.value.status == mmf.ERR_STATE
    b.close()
