"""VTK output (solver_writer.cpp:56-119, main.cpp:284-298): the fields the
reference writes, with its names and order, converted on the GPU.

Each file is read back with a small VTK XML image-data reader and compared
bit for bit with the oracle: velocity, pressure and temperature with the
oracle's utils::conservative2primitive of the same state, density with the
state, the residual fields with the RHS of that state, solved with the cell
flags.  Asynchronous writes must hold the state of the moment they were
issued while later steps run.
"""
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ["solved", "velocity", "pressure", "temperature", "density", "residualC", "residualMX", "residualMY",
          "residualMZ", "residualE"]


def read_vti(path):
    raw = open(path, "rb").read()
    head_end = raw.index(b'<AppendedData encoding="raw">')
    head = raw[:head_end].decode()
    base = raw.index(b"_", head_end) + 1
    m = re.search(r'WholeExtent="([^"]+)" Origin="([^"]+)" Spacing="([^"]+)"', head)
    ext = [int(v) for v in m.group(1).split()]
    meta = {"extent": ext, "origin": [float(v) for v in m.group(2).split()],
            "spacing": [float(v) for v in m.group(3).split()], "names": []}
    arrays = {}
    for a in re.finditer(r'<DataArray type="(\w+)" Name="(\w+)" NumberOfComponents="(\d)" format="appended" '
                         r'offset="(\d+)"/>', head):
        typ, name, comps, off = a.group(1), a.group(2), int(a.group(3)), int(a.group(4))
        meta["names"].append(name)
        nbytes = int(np.frombuffer(raw, dtype="<u8", count=1, offset=base + off)[0])
        dt = {"Int32": "<i4", "Float64": "<f8"}[typ]
        arr = np.frombuffer(raw, dtype=dt, count=nbytes // np.dtype(dt).itemsize, offset=base + off + 8)
        arrays[name] = arr.reshape(-1, comps) if comps > 1 else arr
    return meta, arrays


def expected_fields(orc, U, R, flags):
    prim = np.array([orc.cons2prim(c) for c in U])  # FID_P, FID_U, FID_V, FID_W, FID_T
    return {"solved": ((flags & 2) != 0).astype(np.int32) if flags is not None else np.ones(len(U), np.int32),
            "velocity": prim[:, 1:4], "pressure": prim[:, 0], "temperature": prim[:, 4], "density": U[:, 0],
            "residualC": R[:, 0], "residualMX": R[:, 1], "residualMY": R[:, 2], "residualMZ": R[:, 3],
            "residualE": R[:, 4]}


def check_file(path, orc, s, U, R, flags=None):
    meta, arrays = read_vti(path)
    assert meta["names"] == FIELDS
    n = list(s.n) + [1] * (3 - len(s.n))
    assert meta["extent"] == [0, n[0], 0, n[1], 0, n[2]]
    exp = expected_fields(orc, U, R, flags)
    for name in FIELDS:
        assert np.array_equal(arrays[name], exp[name]), name


def test_vtk_fields_2d_order2(mmf, orc, tmp_path):
    pm = mmf.problem("vortex_xy", 2, vortex_periodic=1)
    s = mmf.Solver.for_problem(pm, 40, order=2, side_bc=[mmf.BC_PERIODIC] * 4)
    s.set_initial_state(pm)
    s.set_time(0.0, 10.0, 0.5)
    s.advance(7)
    U = s.get_state()
    p = str(tmp_path / "vortex.vti")
    s.write_vtk(p)
    R, _ = s.compute_rhs()
    meta, _ = read_vti(p)
    assert meta["origin"][:2] == list(pm.domain_origin[:2]) and meta["spacing"][0] == pm.domain_length / 40
    check_file(p, orc, s, U, R)


def test_vtk_fields_3d_with_body(mmf, orc, tmp_path):
    s = mmf.Solver(3, [20, 12, 9], [-1.0, -1.0, -1.0], 0.1, [3, 0, 1, 1, 0, 2], [[1.0, 0.3, 0.0, 0.1, 1.0]] * 6, 1)
    s.set_bodies(boxes=[[-0.5, -0.4, -0.5, 0.0, 0.1, 0.2]])
    pm = mmf.problem("radsod", 3)
    s.set_initial_state(pm)
    s.set_time(0.0, 10.0, 0.45)
    s.advance(3)
    U = s.get_state()
    flags = s.cell_flags()
    assert (flags & 2).min() == 0  # some cells are not solved
    p = str(tmp_path / "body.vti")
    s.write_vtk(p)
    R, _ = s.compute_rhs()
    check_file(p, orc, s, U, R, flags)


def test_vtk_async_writes_keep_their_state(mmf, orc, tmp_path):
    """Three writes in flight with steps between them: each file holds the
    state of the moment it was issued (two pinned buffers, the third call
    waits for the first)."""
    pm = mmf.problem("radsod", 2)
    s = mmf.Solver.for_problem(pm, 48, order=1)
    s.set_initial_state(pm)
    s.set_time(0.0, 10.0, 0.45)
    states = []
    for i in range(3):
        s.advance(2)
        U = s.get_state()
        R, _ = s.compute_rhs()
        states.append((U, R))
        s.write_vtk(str(tmp_path / ("s%d.vti" % i)), wait=False)
    s.advance(2)
    s.wait_writes()
    for i, (U, R) in enumerate(states):
        check_file(str(tmp_path / ("s%d.vti" % i)), orc, s, U, R)


def test_vtk_write_failure_is_reported(mmf, tmp_path):
    s = mmf.Solver.for_problem(mmf.problem("radsod", 2), 16, order=1)
    s.set_initial_state(mmf.problem("radsod", 2))
    s.write_vtk(str(tmp_path / "missing_dir" / "x.vti"), wait=False)
    with pytest.raises(mmf.MMFError):
        s.wait_writes()
    s.write_vtk(str(tmp_path / "ok.vti"))  # the writer carries on
    assert os.path.getsize(str(tmp_path / "ok.vti")) > 16 * 16 * 92
