"""The C ABI library loads on any host and exports every symbol the header
declares; argument validation that needs no device.  No GPU."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def header_functions():
    src = open(os.path.join(ROOT, "include", "mocassin_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(mc_\w+)\s*\(", src, re.M)))


def test_header_declares_the_boundary():
    names = header_functions()
    for want in ("mc_ctx_create", "mc_upload_tables", "mc_transport_run", "mc_walk_replay",
                 "mc_zero_estimators", "mc_fetch_estimators", "mc_ctx_last_error",
                 "mc_ctx_destroy"):
        assert want in names


def test_library_exports_every_declared_symbol():
    from mocassin_amd import _capi
    L = _capi.load()
    for name in header_functions():
        assert hasattr(L, name), name
    assert set(_capi.EXPORTS) <= set(header_functions())


def test_create_rejects_bad_descriptors_without_device():
    from mocassin_amd import _capi
    L = _capi.load()
    h = C.c_void_p()
    assert L.mc_ctx_create(None, C.byref(h)) == 1 and not h.value
    d = _capi.McDesc()
    d.abi_version = 99
    assert L.mc_ctx_create(C.byref(d), C.byref(h)) == 1
    assert b"ABI" in L.mc_ctx_last_error(None)
    d.abi_version = _capi.MC_ABI_VERSION
    assert L.mc_ctx_create(C.byref(d), C.byref(h)) == 1       # empty sizes
    assert b"empty" in L.mc_ctx_last_error(None)


@pytest.mark.parametrize("dE", [float("inf"), float("nan"), 0.0, -1.0])
def test_create_rejects_bad_packet_energy_without_device(dE):
    """deltaE = L/N must be finite and positive; the check comes before any
    device call, so it runs here."""
    import numpy as np
    from mocassin_amd import _capi
    L = _capi.load()
    ax = np.array([0.0, 1.0, 2.0], np.float32)
    act = np.ones(27, np.int32)
    nu = np.array([0.5, 1.0, 2.0], np.float32)
    pos = np.zeros(3, np.float32)
    idx = np.ones(3, np.int32)
    de = np.array([dE], np.float32)
    d = _capi.McDesc()
    d.abi_version = _capi.MC_ABI_VERSION
    d.nx = d.ny = d.nz = 3
    d.xAxis = d.yAxis = d.zAxis = ax.ctypes.data
    d.active = act.ctypes.data
    d.nCells, d.nbins, d.nStars = 27, 3, 1
    d.nuArray, d.starPos, d.starIdx, d.deltaE = nu.ctypes.data, pos.ctypes.data, idx.ctypes.data, de.ctypes.data
    h = C.c_void_p()
    assert L.mc_ctx_create(C.byref(d), C.byref(h)) == 1 and not h.value
    assert b"deltaE" in L.mc_ctx_last_error(None)


def test_calls_on_null_context_fail_cleanly():
    from mocassin_amd import _capi
    L = _capi.load()
    assert L.mc_transport_run(None, 1, 10, 0, None, None, None) == 1
    assert L.mc_upload_tables(None, None, None, None, None, None, None) == 1
    nc, nb = C.c_int32(-1), C.c_int32(-1)
    assert L.mc_ctx_shape(None, C.byref(nc), C.byref(nb)) == 1 and nc.value == nb.value == 0
    assert L.mc_esc_planes(None) == 0
    L.mc_ctx_destroy(None)


def test_copies_and_fills_stay_on_the_context_stream():
    """A context works on a non-blocking stream, which a blocking hipMemcpy
    or hipMemset on the null stream is not ordered with: a fill of a fresh
    re-emission buffer once landed after the kernel that wrote it (seen only
    when the profiler serialised the queues).  The library's copies and
    fills therefore all name a stream."""
    csrc = os.path.join(ROOT, "mocassin_amd", "csrc")
    bad = []
    for name in sorted(os.listdir(csrc)):
        if not name.endswith((".hip", ".h")):
            continue
        for i, line in enumerate(open(os.path.join(csrc, name)), 1):
            code = line.split("//")[0]
            if re.search(r"\bhipMem(cpy|set)(2D)?\s*\(", code):
                bad.append("%s:%d: %s" % (name, i, line.strip()))
    assert not bad, "\n".join(bad)
