"""bench.py's host-side contract, without a GPU: the one-GPU refusal and the
shape of the CPU-baseline records (the oracle on small samples)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEYS = {"value", "unit", "cores", "kind", "sample", "single_thread"}


def run_bench(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, cwd=ROOT, env=e,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    for args, env in ((["--gpus", "2"], None), (["--gpus", "8"], None), ([], {"WORLD_SIZE": "2"})):
        r = run_bench(args, env)
        assert r.returncode == 2, (args, env, r.stdout, r.stderr)
        assert "one GPU only" in r.stdout


def test_cpu_baseline_records(orc):
    sys.path.insert(0, ROOT)
    import bench
    rec = bench.cpu_baseline(16, 1, budget_mt=0.2)
    assert set(rec) == KEYS and rec["kind"] == "port" and rec["cores"] == bench.cpu_threads() and rec["value"] > 0
    assert rec["unit"] == "cell-updates/s" and "16^3" in rec["sample"] and "threads" in rec["sample"]
    assert rec["single_thread"]["cores"] == 1 and rec["single_thread"]["value"] > 0
    # OMP_NUM_THREADS, when set, is the thread count used
    os.environ["OMP_NUM_THREADS"] = "3"
    try:
        assert bench.cpu_threads() == 3
    finally:
        del os.environ["OMP_NUM_THREADS"]

    class Flags:  # the sphere's fluid flags at 24^3 (centroid outside r = 0.25)
        def cell_flags(self):
            h = 4.0 / 24
            c = -2.0 + h * (np.arange(24) + 0.5)
            z, y, x = np.meshgrid(c, c, c, indexing="ij")
            return (np.sqrt(x * x + y * y + z * z).reshape(-1) > 0.25).astype(np.uint8)

    rec = bench.cpu_baseline_sphere(Flags(), 24, budget_s=0.2, budget_mt=0.2)
    assert set(rec) == KEYS and rec["kind"] == "port" and rec["value"] > 0
    assert "24^3" in rec["sample"]
