"""Builds tests/cpp/_build/mirror_driver, the test driver of the C++ mirror
(include/scirs2_linalg.hpp), linked against scirs_amd/lib/libscirs_gpu.so.
Host code only (g++); the library must be built first."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "mirror_driver")


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    libdir = os.path.join(ROOT, "scirs_amd", "lib")
    rel = os.path.relpath(libdir, os.path.dirname(OUT))
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "mirror_driver.cpp"), "-o", OUT + ".tmp",
           "-L", libdir, "-lscirs_gpu", f"-Wl,-rpath,$ORIGIN/{rel}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
