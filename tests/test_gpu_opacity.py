"""GPU opacity tables (mc_opacity_build) against the FP32 restatement
(oracle/mc_opacity.c, Gaunt factor at each cell's Te) and the unmodified
reference's tables (tests/golden/opacity_*.npz): every photoionisation bin
bit for bit; bin 1 (free-free, never sampled) to the device libm's last
bits."""
import numpy as np
import pytest

import mc_oracle as O
from conftest import balance_fixture, golden, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def _segments(g):
    nb = int(g["ints"][4])
    m = g["seg_meta"]
    last = np.maximum(m[:, 3], np.minimum(m[:, 4], nb))
    return np.stack([m[:, 0], m[:, 1], m[:, 3], last], axis=1).astype(np.int32), g["seg_sigma"], float(g["brems1"][0])


def _ctx(name, permute=False):
    from test_gpu_balance import _context
    g = golden("opacity_" + name)
    b = balance_fixture(name)[0]
    nc, nb = int(g["ints"][3]), int(g["ints"][4])
    assert np.array_equal(np.asarray(b.elem_on, np.int32), g["elem_on"])
    T = _context(nc, nb, b.nuArray, permute=permute)
    T.balance_setup(b)
    T.opacity_setup(*_segments(g))
    return T, g, b


def _state(g, b):
    nc = int(g["ints"][3])
    ion = g["ionDen"].reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0)
    return g["Hden"], g["Te"], g["Ne"], ion


@pytest.mark.parametrize("name,permute", [("c2mini", False), ("c2mini", True), ("dust_pe", False)])
def test_gas_opacity_matches_oracle_and_reference(name, permute):
    print("opacity %s permute=%s" % (name, permute), flush=True)
    T, g, b = _ctx(name, permute)
    try:
        op, sca, ms = T.opacity_build(*_state(g, b))
    finally:
        T.close()
    ex = O.opacity_tables(g, gaunt_cache=False)
    assert np.array_equal(op[:, 1:].view(np.uint32), ex[:, 1:].view(np.uint32))
    assert np.max(np.abs(op[1:, 0] - ex[1:, 0]) / ex[1:, 0]) < 1e-5
    assert np.all(op[0] == 0)
    cells = g["cells"]
    assert np.array_equal(op[cells, 1:].view(np.uint32), g["op_gas"][:, 1:].view(np.uint32))
    assert np.max(np.abs(op[cells, 0] - g["op_gas"][:, 0]) / g["op_gas"][:, 0]) < 0.05
    print("opacity %s: %.3f ms" % (name, ms))


def test_dust_opacity_matches_reference():
    print("opacity with grains", flush=True)
    T, g, b = _ctx("dust_pe")
    bg = golden("balance_dust_pe")
    nS, nSz, nTm, nsm, nSp = (int(v) for v in bg["dust_ints"][:5])
    try:
        T.dust_setup(absXsec=g["dust_abs"].reshape(nS, nSz, -1), grainWeight=g["dust_weight"],
                     grainAbun=g["dust_abun"], emIntegral=bg["dust_emint"].reshape(nTm, nSz, nSp).transpose(2, 1, 0),
                     nTemps=nTm, nSpeciesMax=nsm, scaXsec=g["dust_sca"].reshape(nS, nSz, -1),
                     TdustSublime=g["dust_tsub"], Ndust=g["dust_ndust"])
        op, sca, ms = T.opacity_build(*_state(g, b), Tdust=g["Tdust"])
    finally:
        T.close()
    cells = g["cells"]
    assert np.array_equal(op[cells, 1:].view(np.uint32), g["op_total"][:, 1:].view(np.uint32))
    assert np.array_equal(sca[cells].view(np.uint32), g["op_sca"].view(np.uint32))
    ex = O.opacity_tables(g, gaunt_cache=False)
    tot, sca_o = O.dust_opacity(g, ex)
    assert np.array_equal(op[:, 1:], tot[:, 1:]) and np.array_equal(sca, sca_o)


def test_device_state_after_the_balance():
    """The tables from the state mc_update_cells leaves on the device (FP64,
    rounded to FP32 as the host stores it) equal those from that state
    passed through host memory; the transport then runs on them."""
    print("opacity from the device state", flush=True)
    T, g, b = _ctx("c2mini")
    bb, J, Hden, Te0, Ne0, ion0, ref = balance_fixture("c2mini")
    try:
        T.set_estimators(J.astype(np.float64))
        Te, Ne, ion, fl, nI, _ = T.update_cells(Hden, Te0, Ne0, ion0)
        op_dev, _, _ = T.opacity_build()
        op_host, _, _ = T.opacity_build(Hden, Te.astype(np.float32), Ne.astype(np.float32), ion.astype(np.float32))
    finally:
        T.close()
    assert np.array_equal(op_dev, op_host)
    assert np.all(op_dev[1:, 1:].sum(axis=1) > 0)


def test_transport_runs_on_the_device_tables():
    """A sweep on the tables mc_opacity_build wrote (opacity NULL in
    mc_upload_tables) equals a sweep on the same tables passed through host
    memory: same counters, J to the FP64 atomics' order."""
    print("transport on device-built opacity", flush=True)
    T, g, b = _ctx("c2mini")
    nc, nb = int(g["ints"][3]), int(g["ints"][4])
    spec = np.linspace(0.0, 1.0, nb, dtype=np.float32)
    rec = np.tile(spec, (nc + 1, 1)).astype(np.float32)
    rec[0] = 0.0
    tl = np.zeros(nc + 1, np.float32)
    try:
        op, _, _ = T.opacity_build(*_state(g, b))
        out = []
        for host in (False, True):
            T.upload_tables(op if host else None, spec, rec, tl)
            T.zero_estimators()
            c = T.energy_packet_driver(1, 200000, 0)
            J, E = T.fetch()
            out.append((c, J))
    finally:
        T.close()
    (c1, J1), (c2, J2) = out
    for k in ("n_seg", "n_abs", "n_esc", "n_flights"):
        assert c1[k] == c2[k], k
    assert c1["n_seg"] > 1e5
    assert np.allclose(J1, J2, rtol=1e-12, atol=0)
