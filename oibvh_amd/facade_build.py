"""Compile the host-only C++ programs that use liboibvh_hip.so: the facade
example (tests/cpp/facade_test.cpp, against include/oibvh.hpp) and the
bench's plain-form driver (bench_plain.cpp -> libbench_plain.so, against
include/oibvh.h)."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_test.cpp")
OUT = os.path.join(ROOT, "tests", "cpp", "facade_test")
PLAIN_SRC = os.path.join(ROOT, "bench_plain.cpp")
PLAIN_OUT = os.path.join(ROOT, "libbench_plain.so")


def build() -> str:
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", OUT,
           f"-L{os.path.join(ROOT, 'oibvh_amd')}", "-loibvh_hip", "-Wl,-rpath,$ORIGIN/../../oibvh_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-shared", "-fPIC", f"-I{os.path.join(ROOT, 'include')}", PLAIN_SRC,
           "-o", PLAIN_OUT, f"-L{os.path.join(ROOT, 'oibvh_amd')}", "-loibvh_hip", "-Wl,-rpath,$ORIGIN/oibvh_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"plain driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return OUT
