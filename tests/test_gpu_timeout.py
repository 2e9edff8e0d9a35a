"""Bounded waits.  The multi-workgroup LU panel and the chained triangular
solve (one and two block rows per link) hand data between workgroups through
polled, sequence-tagged words;
every poll gives up after SCIRS_GPU_SPIN_LIMIT tries.  With a limit of 1 the
polls give up at once, so the timeout path runs while every kernel still exits
normally: the call must return SCIRS_ERR_TIMEOUT (-4) with the place it timed
out, every kernel queued after the timed-out one must stop (the flag word they
all test is raised), and the next call in the same process must succeed.

The limit is read when the library initialises, so the check runs in a fresh
process.  SCIRS_GPU_TRSV selects the triangular solver per call: the launch-
per-block solver ("big") never polls, the chained one ("chain") does.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r'''
import os, sys
sys.path.insert(0, ROOT)
import numpy as np
import torch  # noqa: F401  (one HIP runtime per process, loaded first)
from scirs_amd import linalg as L, _native as N
rng = np.random.default_rng(1)
print("case: FP64 LU n=2048, spin limit 1", flush=True)
a = rng.standard_normal((2048, 2048)); b = rng.standard_normal(2048)
try:
    L.gesv(a, b)
    print("NO_TIMEOUT panel", flush=True)
except L.GpuSolveError as e:
    print("RC_PANEL", e.rc, "|", N.last_error(), flush=True)
print("case: chained triangular solve n=1000, spin limit 1", flush=True)
os.environ["SCIRS_GPU_TRSV"] = "chain"
t = np.tril(rng.standard_normal((1000, 1000))) + 40.0 * np.eye(1000)
try:
    L.trsv(t, rng.standard_normal(1000), True, False)
    print("NO_TIMEOUT chain", flush=True)
except L.GpuSolveError as e:
    print("RC_CHAIN", e.rc, "|", N.last_error(), flush=True)
print("case: mixed solve n=1000 (two-block refinement links), spin limit 1", flush=True)
a3 = rng.standard_normal((1000, 1000)) + 40.0 * np.eye(1000); b3 = rng.standard_normal(1000)
try:
    x3, info3 = L.gesv_ir(a3, b3)
    r3 = b3 - a3 @ x3
    eta3 = np.abs(r3).max() / (np.abs(a3).sum(axis=1).max() * np.abs(x3).max() + np.abs(b3).max())
    print("PAIR_PATH", info3.path, eta3, flush=True)
except L.GpuSolveError as e:
    print("RC_PAIR", e.rc, "|", N.last_error(), flush=True)
print("case: FP64 LU n=300 after the timeouts", flush=True)
os.environ["SCIRS_GPU_TRSV"] = "big"
a2 = rng.standard_normal((300, 300)); b2 = rng.standard_normal(300)
x, info = L.gesv(a2, b2)
r = b2 - a2 @ x
eta = np.abs(r).max() / (np.abs(a2).sum(axis=1).max() * np.abs(x).max() + np.abs(b2).max())
print("SMALL_ETA", eta, flush=True)
'''


def test_timeouts_report_and_recover():
    env = dict(os.environ, SCIRS_GPU_SPIN_LIMIT="1", SCIRS_GPU_TRSV="big")
    out = subprocess.run([sys.executable, "-c", SCRIPT.replace("ROOT", repr(ROOT))], env=env,
                         capture_output=True, text=True, timeout=240)
    print(out.stdout[-3000:], out.stderr[-2000:])
    assert out.returncode == 0
    lines = out.stdout.splitlines()
    panel = [l for l in lines if l.startswith("RC_PANEL")]
    chain = [l for l in lines if l.startswith("RC_CHAIN")]
    assert panel and panel[0].split()[1] == "-4" and "timed out at column" in panel[0], lines
    assert chain and chain[0].split()[1] == "-4" and "triangular solve timed out" in chain[0], lines
    # the mixed solve's refinement links give up too (and stop the kernels
    # queued after them); the refinement steps then stop halving eta, and
    # GMRES-IR (its preconditioner's solves stopped as well: no
    # preconditioning) or the FP64 LU (single-workgroup panels at n = 1000, no
    # exchange) takes over: the call must return, with an FP64-accurate
    # solution by the solver's own stopping test
    pair = [l.split() for l in lines if l.startswith("PAIR_PATH")]
    assert pair and pair[0][1] in ("3", "4") and float(pair[0][2]) <= 8 * 1000 * 2.0 ** -53, lines
    eta = [float(l.split()[1]) for l in lines if l.startswith("SMALL_ETA")]
    assert eta and eta[0] <= 8 * 300 * 2.0 ** -53, lines
