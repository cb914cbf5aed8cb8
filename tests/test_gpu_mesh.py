"""The explicit-interface-list path on the GPU (mmf_mesh_*: face kernel +
segmented cell gather) against the oracle's interface loop over the same face
list (orc_solver_init_faces), bit for bit: uniform grids written as face
lists (every border BC, bodies, periodic axes), 2:1 refined meshes with
hanging faces, and a Morton-permuted face order.  Also the structured GPU
path on the same grid, the reference's 2D vortex known answer, graph replay
and non-finite states.
"""
import numpy as np
import pytest

from minimmerflow_amd import meshes

pytestmark = pytest.mark.gpu

INFO = np.tile([0.9, 0.3, -0.2, 0.1, 1.1], (6, 1))


def face_solver(orc, pb, fm, fluid=None):
    return orc.Solver.from_faces(pb, fm.dim, fm.volume, fm.owner, fm.neigh, fm.area, fm.normal, fm.min_cell_size,
                                 border_bc=fm.face_bc, border_info=fm.face_info, fluid=fluid, centroids=fm.centroid)


def random_state(orc, rng, n, dim, speed=0.6):
    return np.array([orc.prim2cons([rng.uniform(0.5, 2), *rng.uniform(-speed, speed, 2),
                                    rng.uniform(-speed, speed) if dim == 3 else 0.0, rng.uniform(0.5, 2)])
                     for _ in range(n)])


def mesh_cases():
    return [
        ("2d cartesian mixed", meshes.cartesian(2, [37, 29], [0.0] * 3, 0.25, [3, 1, 0, 2], INFO), 0.1),
        ("2d cartesian periodic", meshes.cartesian(2, [40, 33], [0.0] * 3, 0.25, [4, 4, 4, 4], INFO), 0.0),
        ("3d cartesian mixed", meshes.cartesian(3, [14, 11, 9], [0.0] * 3, 0.25, [3, 0, 1, 2, 4, 4], INFO), 0.1),
        ("2d refined", meshes.refined(2, [24, 20], [0.0] * 3, 0.5, ([3.0, 2.0, 0], [7.5, 6.0, 0]),
                                      [0, 1, 3, 0], INFO), 0.05),
        ("3d refined", meshes.refined(3, [10, 9, 8], [0.0] * 3, 0.5, ([1.0, 1.0, 1.0], [3.0, 2.5, 2.6]),
                                      [0, 1, 3, 0, 2, 0], INFO), 0.05),
        ("2d general normals and areas", general_geometry(
            meshes.cartesian(2, [30, 21], [0.0] * 3, 0.25, [3, 1, 2, 0], INFO), 1), 0.1),
        ("3d general normals and areas", general_geometry(
            meshes.refined(3, [8, 7, 6], [0.0] * 3, 0.5, ([1.0, 1.0, 1.0], [2.5, 2.0, 2.0]), [1, 3, 0, 2, 0, 1],
                           INFO), 2), 0.05),
    ]


def general_geometry(fm, seed):
    """The same connectivity with every normal rotated off the axes and areas
    jittered face by face: the normal and area arrays are read for every face
    (no axis code, more than 255 distinct areas).  Parity only: the cells no
    longer close."""
    rng = np.random.default_rng(seed)
    c, s_ = np.cos(0.5), np.sin(0.5)
    rot = np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])
    nrm = fm.normal @ rot.T
    area = fm.area * rng.uniform(0.9, 1.1, fm.nfaces)
    return meshes.FaceMesh(fm.dim, fm.volume, fm.size, fm.centroid, fm.owner, fm.neigh, area, nrm, fm.face_centroid,
                           fm.face_bc, fm.face_info)


def fluid_mask(rng, fm, frac):
    if frac == 0.0:
        return None
    return (rng.uniform(size=fm.ncells) > frac).astype(np.uint8)


@pytest.mark.parametrize("permute", [False, True])
def test_mesh_path_equals_oracle(mmf, orc, permute):
    for name, fm, frac in mesh_cases():
        if permute:
            fm = fm.permuted(meshes.morton_order(fm))
        rng = np.random.default_rng(len(name))
        fluid = fluid_mask(rng, fm, frac)
        if fluid is not None:
            fluid[1] = 1
        pb = orc.problem(orc.RADSOD, fm.dim)
        o = face_solver(orc, pb, fm, fluid)
        g = mmf.Mesh(fm)
        if fluid is not None:
            g.set_cell_flags(fluid)
        U0 = random_state(orc, rng, fm.ncells, fm.dim)
        U0[1] = [-0.5, 0.0, 0.0, 0.0, -1.0]  # a counted negative density and pressure (finite eigenvalue)
        g.set_state(U0)
        Rg, mg = g.compute_rhs()
        Ro, mo = o.compute_rhs(U0)
        print("%s%s: %d cells %d faces" % (name, " (Morton order)" if permute else "", fm.ncells, fm.nfaces),
              flush=True)
        assert np.array_equal(Rg, Ro), name
        assert mg == mo, name
        assert g.counters()[:2] == o.counters(), name
        # three RK3 steps from a state without negative values
        U0 = random_state(orc, rng, fm.ncells, fm.dim)
        g.set_state(U0)
        g.set_time(0.0, 1e30, 0.5)
        Uo, Wo = U0.copy(), U0.copy()
        t = 0.0
        for k in range(3):
            g.advance(1)
            dto, eo = o.rk3_step(Uo, Wo, 0.5, t, 1e30)
            t += dto
            tg, dtg, step, eg = g.time()
            assert (tg, dtg, step) == (t, dto, k + 1) and np.array_equal(eg, eo), (name, k)
        assert np.array_equal(g.get_state(), Uo), name
        g.close()


def test_mesh_path_equals_structured_path(mmf):
    """A uniform grid through both GPU paths: the structured kernels and the
    face list give the same bits (state, time control, stage maxima)."""
    side_bc = [3, 0, 1, 2, 0, 1]
    n = [20, 18, 16]
    fm = meshes.cartesian(3, n, [-1.0] * 3, 0.125, side_bc, INFO)
    g = mmf.Mesh(fm)
    s = mmf.Solver(3, n, [-1.0] * 3, 0.125, side_bc, INFO, 1)
    pb = mmf.problem("radsod", 3, sod_r0=0.6, domain_origin=[-1.0] * 3, domain_length=2.5)
    s.set_initial_state(pb)
    U0 = s.get_state()
    g.set_state(U0)
    for x in (s, g):
        x.set_time(0.0, 1e30, 0.45)
        x.advance(20)
    assert s.time()[:3] == g.time()[:3]
    assert np.array_equal(s.get_state(), g.get_state())
    g.close()
    s.close()


def test_mesh_path_reproduces_the_vortex_known_answer(mmf, orc):
    """test/2dIsentropicVortex (CMakeLists.txt:33): 64^2 on [-5,5]^2, free flow,
    CFL 0.45, t = 2; L1 of the GPU state with the oracle's error sum."""
    fm = meshes.cartesian(2, [64, 64], [-5.0, -5.0, -5.0], 10.0 / 64, [0, 0, 0, 0])
    po = orc.problem(orc.VORTEX_XY, 2)
    o = face_solver(orc, po, fm)
    g = mmf.Mesh(fm)
    g.set_state(o.initial_state())
    g.set_time(0.0, 2.0, 0.45)
    steps, t = g.run()
    l1, _ = o.error(g.get_state(), 2.0)
    print("mesh path 2D vortex: %d steps, t=%r, L1 %.12e" % (steps, t, l1), flush=True)
    assert t == 2.0 and "%.12e" % l1 == "1.388164788184e+00"
    g.close()


def test_mesh_graph_replay_equals_launches(mmf, orc):
    fm = meshes.refined(2, [30, 30], [0.0] * 3, 0.5, ([4.0, 4.0, 0], [9.0, 11.0, 0]), [0, 0, 1, 1])
    rng = np.random.default_rng(9)
    U0 = random_state(orc, rng, fm.ncells, 2)
    out = []
    for graphs in (True, False):
        g = mmf.Mesh(fm)
        g.set_step_graphs(graphs)
        g.set_state(U0)
        g.set_time(0.0, 1e30, 0.5)
        g.advance(17)
        out.append((g.get_state(), g.time()[:3]))
        g.close()
    assert out[0][1] == out[1][1] and np.array_equal(out[0][0], out[1][0])


def test_mesh_nan_stops_the_run(mmf, orc):
    fm = meshes.refined(2, [16, 16], [0.0] * 3, 0.5, ([2.0, 2.0, 0], [5.0, 5.0, 0]), [0] * 4)
    g = mmf.Mesh(fm)
    rng = np.random.default_rng(2)
    U0 = random_state(orc, rng, fm.ncells, 2)
    U0[fm.ncells // 2, 4] = np.nan
    g.set_state(U0)
    g.set_time(0.0, 1.0, 0.5)
    with pytest.raises(mmf.MMFError) as e:
        g.run(max_steps=4)
    assert e.value.status == mmf.ERR_STATE
    assert g.counters()[2] >= 1
    g.close()
