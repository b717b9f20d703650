"""Builds tests/native/libharness.so (host build of the product's device arithmetic)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libharness.so")


def build():
    src = os.path.join(HERE, "harness.cpp")
    deps = [src] + [os.path.join(HERE, "..", "..", "viltrum_amd", "csrc", n) for n in ("region_math.h", "philox.h", "features.h", "importance.h")]
    if os.path.exists(LIB) and all(os.path.getmtime(LIB) >= os.path.getmtime(d) for d in deps):
        return LIB
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", LIB, src], check=True)
    return LIB


if __name__ == "__main__":
    print(build())
