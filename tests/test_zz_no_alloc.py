"""The allocation-free batched launch (mc_transport_batch_noalloc, called
through the C ABI) gives what mc_transport_batch gives: the same counters for
every model and its estimators to the FP64 atomics' summation order, on two
calls in a row, the second with fewer models and packets (the device
scratch keeps the first call's contents and must be cleared)."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def _models(specs):
    from mocassin_amd import Transport, synth
    out = []
    for i, (n, npk, teff, nh) in enumerate(specs):
        p = synth.build_problem(n, 2.0e18, 2.0e18, nh, 1.e48, teff, 120, synth.H_HE, n_packets=max(npk, 1))
        T = Transport(p["xAxis"], p["yAxis"], p["zAxis"], p["active"], p["nCells"], p["nbins"],
                      p["nuArray"], p["R_out"], p["geoCorr"], p["ionEdge1"], [p["starPos"]],
                      [p["starIdx"]], [p["deltaE"]], symmetric=True, seed=300 + i, device=0)
        T.upload_tables(p["opacity"], p["specCDF"], p["recPDF"], p["totalLines"])
        out.append(T)
    return out


def _launch(fn, models, npk, offs, tuning):
    from mocassin_amd import _capi
    n = len(models)
    hs = (C.c_void_p * n)(*[m._h for m in models])
    ist = np.ones(n, np.int32)
    pk = np.ascontiguousarray(npk, np.int64)
    of = np.ascontiguousarray(offs, np.uint64)
    tu = _capi.McBatchTuning()
    for k, v in tuning.items():
        setattr(tu, k, int(v))
    cnt = (_capi.McCounters * n)()
    ms = C.c_double()
    for T in models:
        T.zero_estimators()
    rc = fn(hs, n, ist.ctypes.data, pk.ctypes.data, of.ctypes.data, C.byref(tu), cnt, C.byref(ms))
    if rc:
        raise _capi.McError(rc, _capi.load().mc_ctx_last_error(models[0]._h).decode())
    return [c.as_dict() for c in cnt], [T.fetch() for T in models]


SPECS = [(9, 30000, 3.e4, 100.), (15, 0, 4.e4, 1.e3), (21, 120000, 6.e4, 10.), (11, 777, 5.e4, 1.e4),
         (17, 65000, 3.5e4, 300.)]


def test_no_alloc_batch_equals_batch_twice():
    from mocassin_amd import _capi
    L = _capi.load()
    f = L.mc_transport_batch_noalloc
    f.restype = C.c_int
    f.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.POINTER(_capi.McBatchTuning), C.POINTER(_capi.McCounters), C.c_void_p]
    tuning = {"waves_per_cu": 8, "emit_min": 32, "walk_steps": 8, "chunk": 1024}
    models = _models(SPECS)
    try:
        calls = [(models, [s[1] for s in SPECS], [0, 0, 12345, 7, 0]),
                 (models[2:4], [50000, 300], [99, 5])]           # second call: fewer models and packets
        for k, (ms_, npk, offs) in enumerate(calls):
            print("no-alloc batch call %d: %d models" % (k, len(ms_)), flush=True)
            cn, rn = _launch(f, ms_, npk, offs, tuning)
            cb, rb = _launch(L.mc_transport_batch, ms_, npk, offs, tuning)
            for i in range(len(ms_)):
                for key in ("n_seg", "n_abs", "n_esc", "n_line", "n_flights", "n_lost", "n_trapped"):
                    assert cn[i][key] == cb[i][key], (k, i, key)
                assert abs(cn[i]["qphot"] - cb[i]["qphot"]) <= 1e-12 * max(1.0, abs(cb[i]["qphot"]))
                assert np.allclose(rn[i][0], rb[i][0], rtol=1e-12, atol=0)
                assert np.allclose(rn[i][1], rb[i][1], rtol=1e-12, atol=0)
            assert cn[0]["n_seg"] > 0
        # a launch beyond the scratch is refused before anything runs
        with pytest.raises(_capi.McError, match="MC_EINVAL"):
            _launch(f, models[:1], [400_000_000], [0], dict(tuning, chunk=64))
    finally:
        for T in models:
            T.close()
