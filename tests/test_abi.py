"""CPU-side checks of the C ABI: the library builds for gfx950, loads, and
exports every entry point that include/scirs_gpu.h declares.  No compute call
is made (there is no GPU here)."""
import os
import re
import subprocess

import pytest

from scirs_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_declares_entry_points():
    syms = _native.header_symbols()
    for s in ("scirs_gpu_dgesv", "scirs_gpu_dgetrf", "scirs_gpu_dpotrf", "scirs_gpu_dposv",
              "scirs_gpu_dgesv_ir", "scirs_gpu_dgesv_dev", "scirs_gpu_dtrsv"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("library not built")
    lib = _native.load()
    missing = [s for s in _native.header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (scirs_gpu_\w+)", out))
    assert set(_native.header_symbols()) <= exported


def test_python_signatures_cover_header():
    assert set(_native.header_symbols()) <= set(_native.SIGNATURES)


def test_build_targets_gfx950_only():
    from scirs_amd import build
    arch = [f for f in build.CFLAGS if f.startswith("--offload-arch")]
    assert arch == ["--offload-arch=gfx950"]
    assert "-packed-fp32-ops" in build.CFLAGS


def test_no_cpu_fallback_in_product_path():
    # the product package never imports the oracle
    for dirpath, _, files in os.walk(os.path.join(ROOT, "scirs_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in text.lower() or f == "build.py", f


def test_exchange_tag_wraparound():
    """The 32-bit tag count of the exchange buffers: consecutive launches get
    disjoint tag ranges, tag 0 (a cleared word) is never handed out, and a range
    that would run past 2^32 - 1 restarts the count at 1 (the library clears every
    tagged buffer first, so no word written before the restart can match)."""
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("library not built")
    from scirs_amd import linalg
    top = 0xFFFFFFFF
    assert linalg.seq_plan(1, 512) == (0, 1, 513)
    assert linalg.seq_plan(513, 1) == (0, 513, 514)
    assert linalg.seq_plan(0, 512) == (1, 1, 513)           # never tag 0
    assert linalg.seq_plan(top - 512, 512) == (0, top - 512, top)   # last full range
    assert linalg.seq_plan(top - 511, 512) == (1, 1, 513)   # would wrap: restart
    assert linalg.seq_plan(top, 1) == (1, 1, 2)
    # a long run of launches never reuses a tag between two restarts
    nxt, seen = top - 5000, set()
    for span in [512, 1, 512, 7] * 40:
        reset, base, nxt = linalg.seq_plan(nxt, span)
        if reset:
            seen.clear()
        rng = set(range(base, base + span))
        assert 0 not in rng and base + span - 1 <= top and not (rng & seen)
        seen |= rng


def test_noalloc_entry_point_needs_init():
    """The allocation-free backward error never initialises the library itself
    (that allocates): before scirs_gpu_init it returns SCIRS_ERR_UNSUPPORTED
    without touching a device (so this runs without a GPU), and it refuses
    sizes beyond its fixed scratch before anything else."""
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("library not built")
    import ctypes
    lib = _native.load()
    out = ctypes.c_double(0.0)
    p = ctypes.c_void_p(16)   # never dereferenced on these paths
    rc = lib.scirs_gpu_backward_error_dev_noalloc(10, 1, p, 10, p, 1, p, 1, None, ctypes.byref(out))
    assert rc == _native.SCIRS_ERR_UNSUPPORTED and "scirs_gpu_init" in _native.last_error()
    rc = lib.scirs_gpu_backward_error_dev_noalloc(1024, 1024, p, 1024, p, 1024, p, 1024, None, ctypes.byref(out))
    assert rc == _native.SCIRS_ERR_ARG and "fixed scratch" in _native.last_error()
