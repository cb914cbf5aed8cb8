"""The settings.xml front-end (minimmerflow_amd/lib/minimmerflow).

GPU tests run the reference's five test directories unchanged and compare the
printed error the way test/test_driver.py:42-63 does: the text after
"Final error:" must equal the expected string.  CPU tests cover the
configuration errors the reference raises (problem.cpp:57, body.cpp:69).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(ROOT, "minimmerflow_amd", "lib", "minimmerflow")
CASES = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))["cases"]


def run_driver(workdir, args=(), timeout=120):
    return subprocess.run([DRIVER, *args], cwd=workdir, capture_output=True, text=True, timeout=timeout)


def final_error(stdout):
    for line in stdout.splitlines():
        line = line.strip()
        if "Final error:" in line:  # test_driver.py:44-49
            return line.split("Final error:")[1].strip()
    return None


SETTINGS = """<?xml version="1.0"?>
<minimmerflow version="1">
    <problem>
        <type>{ptype}</type>
    </problem>
    <discretization>
        <space><order>1</order><nCells>8</nCells></space>
        <time><CFL>0.45</CFL></time>
    </discretization>
    {bodies}
</minimmerflow>
"""


@pytest.mark.parametrize("ptype,bodies,msg", [
    ("bogus", "", "Problem bogus is not supported."),
    ("radsod", "<bodies><b1><type>sphere</type></b1></bodies>", 'Body shape "sphere" is not supported.'),
    ("ffstep", "", "The lenght of the domain is manditory."),
])
def test_configuration_errors(tmp_path, ptype, bodies, msg):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    (tmp_path / "settings.xml").write_text(SETTINGS.format(ptype=ptype, bodies=bodies))
    r = run_driver(str(tmp_path))
    assert r.returncode == 1
    assert msg in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_reference_test_directories(tmp_path, case):
    """The reference's settings unchanged; argv = nCells of the settings and
    nSaves = 0 (the reference's own command line, main.cpp:103-121), so only
    the initial, first-step and final states are written instead of one file
    per step."""
    shutil.copy(os.path.join(HERE, "golden", "reference_settings", case["name"], "settings.xml"), tmp_path)
    r = run_driver(str(tmp_path), [str(case["nCells"]), "0"])
    assert r.returncode == 0, r.stderr
    got = final_error(r.stdout)
    print(case["name"], got, case["expected"], flush=True)
    assert got == case["expected"]
    n = case["nCells"]
    names = sorted(os.listdir(str(tmp_path)))
    assert names == ["background_%d.0000.vti" % n, "background_%d.0001.vti" % n, "final_background_%d.vti" % n,
                     "settings.xml"], names


@pytest.mark.gpu
def test_driver_save_schedule(tmp_path):
    """nSaves = 4 over t in [0, 2] (main.cpp:376, 509-523): a save after the
    first step whose time passes 0, 0.5, 1 and 1.5, besides the initial and
    final states; every file holds the reference's ten fields."""
    from test_gpu_output import FIELDS, read_vti
    shutil.copy(os.path.join(HERE, "golden", "reference_settings", "2dIsentropicVortex", "settings.xml"), tmp_path)
    r = run_driver(str(tmp_path), ["32", "4"])
    assert r.returncode == 0, r.stderr
    names = sorted(n for n in os.listdir(str(tmp_path)) if n.endswith(".vti"))
    assert names == ["background_32.%04d.vti" % i for i in range(5)] + ["final_background_32.vti"], names
    for n in names:
        meta, arrays = read_vti(str(tmp_path / n))
        assert meta["names"] == FIELDS and meta["extent"] == [0, 32, 0, 32, 0, 1]
        assert np.all(arrays["density"] > 0) and np.all(arrays["solved"] == 1)


@pytest.mark.gpu
def test_ncells_override_and_body(tmp_path, orc):
    """argv[1] overrides nCells (main.cpp:103-104); a box body gives wall faces."""
    xml = SETTINGS.format(ptype="radsod", bodies="""<bodies><step><type>box</type>
        <xMin>3</xMin><xMax>5</xMax><yMin>0</yMin><yMax>2</yMax><zMin>-1</zMin><zMax>1</zMax></step></bodies>""")
    xml = xml.replace("</problem>", "<time><end>0.5</end></time></problem>")
    (tmp_path / "settings.xml").write_text(xml)
    r = run_driver(str(tmp_path), ["24"])
    assert r.returncode == 0, r.stderr
    got = float(final_error(r.stdout))
    pb = orc.problem(orc.RADSOD, 2)
    s0 = orc.Solver(pb, 1, [24, 24], [0.0] * 3, 8.0 / 24)
    cen = s0.centroids()
    box = [3, 0, -1, 5, 2, 1]
    fluid = [0 if (box[0] <= x <= box[3] and box[1] <= y <= box[4] and box[2] <= z <= box[5]) else 1 for x, y, z in cen]
    s = orc.Solver(pb, 1, [24, 24], [0.0] * 3, 8.0 / 24, fluid=fluid)
    U = s.initial_state()
    s.run(U, 0.45, 0.0, 0.5)
    l1, _ = s.error(U, 0.5)
    assert abs(got - l1) <= 1e-12 * l1
