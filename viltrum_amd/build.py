"""Build the HIP library libviltrum_amd.so in-tree for gfx950 (hipcc cross-compiles; no
GPU needed).  `python -m viltrum_amd.build` or viltrum_amd.build.build().

The engine is several translation units (the C ABI, and the batch pipeline instantiated
per integrand family), compiled in parallel and linked into one shared library."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libviltrum_amd.so")
CSRC = os.path.join(HERE, "csrc")
SOURCES = [os.path.join(CSRC, s) for s in (
    "engine.hip", "batch_poly_f32_a.hip", "batch_poly_f32_b.hip", "batch_poly_f64_a.hip", "batch_poly_f64_b.hip",
    "batch_poly_6.hip", "batch_blobs.hip", "batch_lobes5.hip",
    "batch_fubini.hip", "batch_fubini_inf.hip")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-exact FP64 parity with the reference's uncontracted host arithmetic
    "-ffp-contract=off",
    # no packed FP32 VALU ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed (%d): %s" % (r.returncode, cmd[-1]))


def build(verbose=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(s).replace(".hip", ".o")) for s in SOURCES]
    inc = "-I" + os.path.join(ROOT, "include")
    cmds = [[hipcc] + HIPCC_FLAGS + [inc, "-c", "-o", o, s] for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
