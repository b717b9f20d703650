"""Builds the C++ test programs against the drop-in headers, linked to the in-tree
libviltrum_amd.so: tests/cpp/dropin (a reference-style program with the built-in
integrand families, g++) and tests/cpp/custom (user-defined device integrands, hipcc for
gfx950 with the library's own compile flags)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(HERE, "dropin")
CUSTOM = os.path.join(HERE, "custom")


def build():
    lib = os.path.join(ROOT, "viltrum_amd")
    inc = "-I" + os.path.join(ROOT, "include")
    rpath = "-Wl,-rpath,$ORIGIN/../../viltrum_amd"
    subprocess.run(["g++", "-std=c++17", "-O2", inc, "-o", EXE, os.path.join(HERE, "dropin.cpp"), "-L" + lib,
                    "-lviltrum_amd", rpath], check=True)
    sys.path.insert(0, lib)
    from build import HIPCC_FLAGS  # the library's own flags: the same arithmetic in both
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    r = subprocess.run([hipcc] + HIPCC_FLAGS + [inc, "-o", CUSTOM, os.path.join(HERE, "custom.hip"), "-L" + lib,
                                                "-lviltrum_amd", rpath], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed (%d): %s" % (r.returncode, r.stderr[-4000:]))
    return EXE, CUSTOM


if __name__ == "__main__":
    print(build())
