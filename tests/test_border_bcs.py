"""Border conditions per border face (main.cpp:251-260 keeps one per border
interface, problem::getBorderBCType at the face centroid), on the CPU:

(1) the product's problem rule at every border face (mmf_problem_border_bcs,
    host code) equals the oracle's face list face for face, FFSTEP included,
    whose y and z sides change condition at x = 3.2 (problem.cpp:422-432);
(2) the oracle's rule is pinned by a closed form: a uniform state is a fixed
    point of every interior and free-flow face (equal fluxes in and out, so a
    residual of exactly zero) but not of a reflecting face when the normal
    velocity is not zero, nor of the Dirichlet inflow; so on FFSTEP [0,4]^d the
    cells with a non-zero residual are exactly those on the x = 0 side and
    the y / z border cells left of x = 3.2;
(3) the layout's entry count and the refusal of per-side descriptors for a
    problem whose sides are split.
"""
import numpy as np
import pytest

def ffstep(mod, dim, length, origin=0.0):
    return mod.problem(mod.FFSTEP, dim, domain_length=length, domain_origin=[origin] * 3)


CASES = [(2, 20, 4.0, 0.0), (2, 16, 3.2, 0.0), (2, 18, 4.0, -0.4), (3, 10, 4.0, 0.0), (3, 9, 4.5, -0.5)]


@pytest.mark.parametrize("dim,n,length,origin", CASES)
def test_problem_border_bcs_equal_the_oracle_face_rule(mmf, orc, border_index, dim, n, length, origin):
    pm = ffstep(mmf, dim, length, origin)
    po = ffstep(orc, dim, length, origin)
    h = length / n
    nn = [n] * dim
    d = mmf.grid_desc(dim, nn, [origin] * 3, h, [0] * 6)
    bc, info = mmf.problem_border_bcs(pm, d)
    assert len(bc) == mmf.border_face_count(d)
    so = orc.Solver(po, 1, nn, [origin] * 3, h)  # side_bc None: the per-face rule
    owner, neigh, nrm, _ = so.face_list()
    fb = so.face_bc()
    border = neigh < 0
    idx = border_index(nn, dim, owner[border], nrm[border])
    assert sorted(idx) == list(range(len(bc)))
    assert np.array_equal(bc[idx], fb[border])
    dir_ = bc == mmf.BC_DIRICHLET
    assert np.all(info[dir_] == np.array(pm.fs_prim)) and np.all(info[~dir_] == 0.0)
    # the y / z sides are split where their first column lies left of x = 0
    # or their last column right of x = 3.2
    split = origin + 0.5 * h < 1e-10 or origin + length - 0.5 * h > 3.2 - 1e-10
    assert mmf._sides_uniform(d, bc, info) == (not split)


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_ffstep_rule_closed_form(orc, dim):
    n = 20 if dim == 2 else 10
    L = 4.0
    h = L / n
    po = ffstep(orc, dim, L)
    so = orc.Solver(po, 1, [n] * dim, [0.0] * 3, h)
    prim = [0.9, 0.4, 0.3, 0.2 if dim == 3 else 0.0, 1.2]
    U = np.tile(orc.prim2cons(prim), (so.ncells, 1))
    R, _ = so.compute_rhs(U)
    c = so.centroids()
    q = np.stack([np.arange(so.ncells) % n, (np.arange(so.ncells) // n) % n, np.arange(so.ncells) // (n * n)], 1)
    on_side = np.zeros(so.ncells, dtype=bool)
    for a in range(1, dim):
        on_side |= (q[:, a] == 0) | (q[:, a] == n - 1)
    expect_nonzero = (q[:, 0] == 0) | (on_side & (c[:, 0] < 3.2))
    nonzero = np.any(R != 0.0, axis=1)
    assert np.array_equal(nonzero, expect_nonzero)
    assert on_side[c[:, 0] > 3.2].any() and expect_nonzero.sum() < so.ncells


def test_border_face_count_and_split_descriptors(mmf):
    assert mmf.border_face_count(mmf.grid_desc(3, [5, 4, 3], [0.0] * 3, 0.1, [0] * 6)) == 2 * (12 + 15 + 20)
    assert mmf.border_face_count(mmf.grid_desc(2, [5, 4], [0.0] * 3, 0.1, [0] * 4)) == 2 * (4 + 5)
    # FFSTEP on [0, 3.2]: one condition per side; on [0, 4]: split sides
    d = mmf.problem_grid_desc(ffstep(mmf, 2, 3.2), 16)
    assert [d.side_bc[s] for s in range(4)] == [mmf.BC_DIRICHLET, mmf.BC_FREE_FLOW, mmf.BC_REFLECTING,
                                                  mmf.BC_REFLECTING]
    with pytest.raises(mmf.MMFError) as e:
        mmf.problem_grid_desc(ffstep(mmf, 2, 4.0), 20)
    assert e.value.status == mmf.ERR_INVALID
    for key in ("vortex_xy", "radsod", "sod3d_y"):
        mmf.problem_grid_desc(mmf.problem(key, 2), 16)
