"""The C-ABI library loads and exports every symbol include/mmf.h declares.

No compute call here: this runs without a GPU.
"""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_declares_entry_points():
    hdr = open(os.path.join(ROOT, "include", "mmf.h")).read()
    names = set(re.findall(r"\b(mmf_[a-z0-9_]+)\s*\(", hdr))
    for must in ("mmf_create", "mmf_destroy", "mmf_compute_rhs", "mmf_advance", "mmf_run",
                 "mmf_set_state", "mmf_get_state", "mmf_error_norms", "mmf_set_bodies"):
        assert must in names


def test_library_exports_all_declared_symbols(mmf):
    L = ctypes.CDLL(mmf.LIB_PATH)
    missing = [n for n in mmf.exported_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings(mmf):
    L = mmf.lib()
    assert L.mmf_status_string(0) == b"ok"
    assert b"order" in L.mmf_status_string(mmf.ERR_UNSUPPORTED_ORDER)


def test_problem_catalogue_host_side(mmf, orc):
    """problem.cpp restated in the product (host C++) equals the oracle's, bit for bit."""
    import numpy as np
    rng = np.random.default_rng(7)
    for key in ("vortex_xy", "vortex_yz", "vortex_zx", "radsod", "sod3d_x", "sod3d_z", "ffstep"):
        for dim in (2, 3):
            pm = mmf.problem(key, dim)
            po = orc.problem(orc.PROBLEM_KEYS[key], pm.dim)
            pts = rng.uniform(-5, 5, size=(64, 3))
            for t in (0.0, 0.7):
                got = mmf.exact_conservatives(pm, pts, t)
                exp = np.array([orc.exact_conservatives(po, p, t) for p in pts])
                assert np.array_equal(got, exp), (key, dim, t)


def test_batch_create_rejects_bad_cases_without_touching_a_device(mmf):
    """Case validation runs on the host before any device call."""
    import pytest
    ok = mmf.grid_desc(2, [8, 8], [0.0, 0.0, 0.0], 0.125, [0, 0, 0, 0])
    bad_dim = mmf.grid_desc(3, [8, 8, 8], [0.0, 0.0, 0.0], 0.125, [0] * 6)
    bad_order = mmf.grid_desc(2, [8, 8], [0.0, 0.0, 0.0], 0.125, [0] * 4, order=2)
    half_periodic = mmf.grid_desc(2, [8, 8], [0.0, 0.0, 0.0], 0.125, [4, 0, 0, 0])
    bad_h = mmf.grid_desc(2, [8, 8], [0.0, 0.0, 0.0], 0.0, [0] * 4)
    for descs, code in (([ok, bad_dim], mmf.ERR_INVALID), ([bad_order], mmf.ERR_UNSUPPORTED_ORDER),
                        ([ok, half_periodic], mmf.ERR_INVALID), ([bad_h], mmf.ERR_INVALID), ([], mmf.ERR_INVALID)):
        with pytest.raises(mmf.MMFError) as e:
            mmf.Batch(descs)
        assert e.value.status == code


def test_create_rejects_oversized_grids_without_touching_a_device(mmf):
    """Grid sizes beyond 2^30 cells per axis or 2^36 cells in all are refused
    on the host, before the device count is read (which would give
    ERR_NO_DEVICE here)."""
    L = mmf.lib()
    for dim, n in ((2, [1 << 31, 4]), (3, [1 << 13, 1 << 13, 1 << 11]), (3, [1 << 30, 1 << 30, 1 << 30]),
                   (2, [0, 8])):
        d = mmf.grid_desc(dim, n, [0.0, 0.0, 0.0], 0.125, [0] * (2 * dim))
        ctx = ctypes.c_void_p()
        assert L.mmf_create(ctypes.byref(d), ctypes.byref(ctx)) == mmf.ERR_INVALID, (dim, n)
        assert not ctx.value
