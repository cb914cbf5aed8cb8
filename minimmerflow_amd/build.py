"""Build recipe of the HIP library minimmerflow_amd/lib/libmmf.so (gfx950).

Run as ``python -m minimmerflow_amd.build`` or through __graft_entry__.build().
The library is built in-tree so it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmmf.so")

SOURCES = ["mmf_kernels.hip", "mmf_march.hip", "mmf_march2d.hip", "mmf_learn.hip", "mmf_api.cpp", "mmf_batch.cpp",
           "mmf_learn_api.cpp", "mmf_problem.cpp", "mmf_output.cpp", "mmf_noalloc.hip",
           "mmf_mesh.hip", "mmf_mesh_api.cpp", "mmf_march_o2.hip"]

# -ffp-contract=off: no fused multiply-add, so FP64 results round exactly as
# the reference's expressions are written (bitwise parity with the oracle).
# -packed-fp32-ops: keep the compiler from emitting packed FP32 arithmetic.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


OBJDIR = os.path.join(HERE, "_build")


def _compile(src, verbose):
    """One translation unit to an object (each keeps its own device code)."""
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    flags = [f for f in FLAGS if f != "-shared"]
    cmd = [hipcc()] + flags + ["-c", "-I", INCLUDE, "-I", CSRC, src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    return obj, res


def _report(res, verbose):
    if verbose and res.stderr.strip():
        # the host pass does not know the packed-fp32 feature; that is harmless
        lines = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


def build(verbose=False, jobs=None):
    """Compile every source (in parallel, one hipcc per file) and link libmmf.so."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, verbose), srcs))
    for obj, res in results:
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed building %s" % os.path.basename(obj))
        _report(res, verbose)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + [o for o, _ in results] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libmmf.so")
    _report(res, verbose)
    os.replace(tmp, LIB)
    build_driver(verbose)
    return LIB


DRIVER = os.path.join(LIBDIR, "minimmerflow")


def build_driver(verbose=False):
    """The settings.xml front-end (host C++ only), linked against libmmf.so."""
    src = os.path.join(CSRC, "driver", "main.cpp")
    tmp = DRIVER + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", os.path.join(CSRC, "driver"), src,
           "-L", LIBDIR, "-lmmf", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building the minimmerflow driver")
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
