"""Builds tests/cpp/dropin (a reference-style C++ program against the drop-in header,
linked to the in-tree libviltrum_amd.so)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(HERE, "dropin")


def build():
    lib = os.path.join(ROOT, "viltrum_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), "-o", EXE,
                    os.path.join(HERE, "dropin.cpp"), "-L" + lib, "-lviltrum_amd", "-Wl,-rpath,$ORIGIN/../../viltrum_amd"],
                   check=True)
    return EXE


if __name__ == "__main__":
    print(build())
