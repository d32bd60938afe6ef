"""GPU re-emission tables (mc_emission_setup / mc_emission_build) against the
unmodified reference's recPDF and totalLines of the state after one balance
pass (tests/golden/opacity_*.npz: em_recPDF, em_totalLines) and the
restatement with FP64 level populations (oracle mco_emission_cell_f64); the
tables then serve the transport without a host upload."""
import numpy as np
import pytest

import mc_oracle as O
from conftest import balance_fixture, golden, gpu_available, grid_from_fixture
from mocassin_amd.atomic import emission_tables

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def _ctx(name):
    from mocassin_amd.transport import Transport
    g = golden("opacity_" + name)
    b = balance_fixture(name)[0]
    axes, active, nc, nb, fl, star_idx = grid_from_fixture(g)
    T = Transport(*axes, active, nc, nb, b.nuArray, R_out=float(axes[0][-1]), geoCorr=(0, 0, 0),
                  ionEdge1=1.0, starPos=(0, 0, 0), starIdx=star_idx, deltaE=1.0,
                  symmetric=fl["symmetric"], dust=fl["dust"], gas=True)
    T.balance_setup(b)
    dust = None
    if fl["dust"]:
        bg = golden("balance_" + name)
        nS, nSz, nTm, nsm, nSp = (int(v) for v in bg["dust_ints"][:5])
        T.dust_setup(absXsec=g["dust_abs"].reshape(nS, nSz, -1), grainWeight=g["dust_weight"],
                     grainAbun=g["dust_abun"], emIntegral=bg["dust_emint"].reshape(nTm, nSz, nSp).transpose(2, 1, 0),
                     nTemps=nTm, nSpeciesMax=nsm, scaXsec=g["dust_sca"].reshape(nS, nSz, -1),
                     TdustSublime=g["dust_tsub"], Ndust=g["dust_ndust"])
        dust = dict(absx=g["dust_abs"].reshape(nS, nSz, nb), weight=g["dust_weight"], gabun=g["dust_abun"],
                    tsub=g["dust_tsub"], Tdust=g["Tdust"].reshape(nc + 1, nSz + 1, nsm + 1), Ndust=g["dust_ndust"],
                    Hden=g["Hden"])
    T.emission_setup(**emission_tables(g))
    ion = g["ionDen"].reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0)
    return T, g, b, ion, dust


@pytest.mark.parametrize("name", ["c2mini", "dust_pe"])
def test_host_state_matches_reference_and_oracle(name):
    print("emission %s" % name, flush=True)
    T, g, b, ion, dust = _ctx(name)
    try:
        rec, tl, ms = T.emission_build(g["Hden"], g["Te"], g["Ne"], ion,
                                       Tdust=None if dust is None else g["Tdust"])
        rec2, tl2, _ = T.emission_build(g["Hden"], g["Te"], g["Ne"], ion,
                                        Tdust=None if dust is None else g["Tdust"])
    finally:
        T.close()
    assert np.array_equal(rec, rec2) and np.array_equal(tl, tl2)
    cells = g["cells"]
    ref, tref = g["em_recPDF"], g["em_totalLines"]
    exact = np.mean(rec[cells] == ref)
    d = np.max(np.abs(rec[cells] - ref))
    print("emission %s: %.3f ms, recPDF bit-equal %.4f, max |diff| %.2e, totalLines max rel %.2e"
          % (name, ms, exact, d, np.max(np.abs(tl[1:] / tref[1:] - 1))))
    # every row a CDF ending at 1; equal to the reference's to the last bits
    # of the device's exp/pow against the C library's (most entries bit for
    # bit, the rest within a few FP32 ulps of 1)
    assert np.all(np.diff(rec[1:], axis=1) >= 0) and np.all(rec[1:, -1] == 1.0)
    assert d < 5e-7 and exact > 0.5
    assert np.max(np.abs(tl[1:] / tref[1:] - 1)) < (2e-5 if dust is not None else 2e-6)
    ex, tex = O.Balance(b).emission(O.emission_tables(g)[0], g["Te"], g["Ne"], ion, cells, "f64", dust=dust)
    assert np.max(np.abs(rec[cells] - ex)) < 5e-7


def test_device_tables_feed_the_transport():
    """After a build the sweep runs on the device tables (no host recPDF or
    totalLines); without one the upload still requires them."""
    print("emission tables in the transport", flush=True)
    from mocassin_amd._capi import McError
    T, g, b, ion, dust = _ctx("c2mini")
    nc, nb = T.nCells, T.nbins
    from test_gpu_opacity import _segments
    try:
        T.opacity_setup(*_segments(g))
        T.opacity_build(g["Hden"], g["Te"], g["Ne"], ion, host=False)
        spec = np.linspace(0, 1, nb).astype(np.float32)
        with pytest.raises(McError):
            T.upload_tables(None, spec)
        rec, tl, _ = T.emission_build(g["Hden"], g["Te"], g["Ne"], ion)
        T.upload_tables(None, spec)
        T.zero_estimators()
        c = T.energy_packet_driver(1, 20000, 0)
        assert c["n_abs"] > 0
    finally:
        T.close()


def test_device_state_after_balance():
    """A GPU balance pass from the reference's starting state and estimator,
    then the tables of the device state: within 1e-4 of the reference's."""
    print("emission from the device balance state", flush=True)
    T, g, b, ion, dust = _ctx("c2mini")
    nc, nb = T.nCells, T.nbins
    J = np.zeros((nc + 1) * nb, np.float32)
    J[g["J_idx"]] = g["J_val"]
    ion0 = g["ionDen0"].reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0)
    try:
        with pytest.raises(RuntimeError):
            T.emission_build()
        T.set_estimators(J.reshape(nb, nc + 1).T.astype(np.float64))
        T.update_cells(g["Hden"], g["Te0"], g["Ne0"], ion0)
        rec, tl, ms = T.emission_build()
    finally:
        T.close()
    cells = g["cells"]
    d = np.max(np.abs(rec[cells] - g["em_recPDF"]))
    r = np.max(np.abs(tl[1:] / g["em_totalLines"][1:] - 1))
    print("emission from device state: max |recPDF diff| %.2e, totalLines max rel %.2e, %.3f ms" % (d, r, ms))
    assert d < 1e-4 and r < 1e-4


def test_cell_at_the_1K_floor():
    """A cell the reference's schedule drove to its 1 K floor emits nothing
    (the reference divides 0 by 0): its absorbed packets leave as lines
    (totalLines 1) and its never-sampled row is 0 .. 0 1; a state that is
    not finite is refused."""
    print("emission of a cell at 1 K", flush=True)
    from mocassin_amd._capi import McError
    T, g, b, ion, dust = _ctx("c2mini")
    Te, Ne = g["Te"].copy(), g["Ne"].copy()
    Te[5], Ne[5] = 1.0, 1.e-5
    ion = ion.copy()
    ion[5, :, :] = 0.0
    ion[5, :, 0] = 1.0
    try:
        rec, tl, _ = T.emission_build(g["Hden"], Te, Ne, ion)
        assert tl[5] == 1.0 and np.all(rec[5, :-1] == 0) and rec[5, -1] == 1.0
        assert np.all(tl[6:] < 1)
        Te[7] = np.nan
        with pytest.raises(McError):
            T.emission_build(g["Hden"], Te, Ne, ion)
    finally:
        T.close()
