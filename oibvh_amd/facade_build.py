"""Compile the C++ facade example (tests/cpp/facade_test.cpp) against
include/oibvh.hpp and liboibvh_hip.so. It is a host-only C++17 program."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_test.cpp")
OUT = os.path.join(ROOT, "tests", "cpp", "facade_test")


def build() -> str:
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", OUT,
           f"-L{os.path.join(ROOT, 'oibvh_amd')}", "-loibvh_hip", "-Wl,-rpath,$ORIGIN/../../oibvh_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return OUT
