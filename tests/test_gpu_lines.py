"""GPU line luminosities (mc_line_luminosities) against the FP64
restatement of outputGas's forbVol/HbetaVol sums (oracle mco_line_lum_f64)
on the same state, against the unmodified reference (lines_L, lines_hb of
tests/golden/opacity_*.npz), and, from the GPU's own balance of the
reference's transport estimator, the reference's line ratios within 2 %."""
import numpy as np
import pytest

import mc_oracle as O
from conftest import balance_fixture, golden, gpu_available, grid_from_fixture

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
                  symmetric=fl["symmetric"], gas=True)
    T.balance_setup(b)
    dV, cells = O.cell_volumes(*axes, active, fl["symmetric"])
    ion = g["ionDen"].reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0)
    return T, g, b, dV, cells, ion


@pytest.mark.parametrize("name", ["c2mini", "dust_pe"])
def test_host_state_matches_oracle_and_reference(name):
    print("lines %s" % name, flush=True)
    T, g, b, dV, cells, ion = _ctx(name)
    try:
        forb, hb, ms = T.line_luminosities(g["Hden"], g["Te"], g["Ne"], ion)
        forb2, hb2, _ = T.line_luminosities(g["Hden"], g["Te"], g["Ne"], ion)
    finally:
        T.close()
    # a fixed summation order: the same bits on every call
    assert np.array_equal(forb, forb2) and hb == hb2
    ex, hbx = O.Balance(b).line_lum(g["Hden"], g["Te"], g["Ne"], ion, dV, cells, "f64")
    assert abs(hb / hbx - 1) < 1e-6
    ref = g["lines_L"].reshape(-1, 17, 17).astype(np.float64)
    assert abs(hb / float(g["lines_hb"][0]) - 1) < 1e-6
    if b.ncel == 0:
        assert forb.shape[0] == 0
        return
    big = np.abs(ex) > 1e-30
    assert np.max(np.abs(forb[big] / ex[big] - 1)) < 1e-7
    assert np.all(np.abs(forb[~big]) < 1e-29)
    # zero wherever the level pair is not a radiative transition
    assert np.all(forb[ex == 0] == 0)
    rb = np.abs(ref) > 1e-30
    assert np.max(np.abs(forb[rb] / ref[rb] - 1)) < 1e-5
    print("lines %s: %.3f ms" % (name, ms))


def test_black_cells_are_skipped():
    print("lines with black cells", flush=True)
    T, g, b, dV, cells, ion = _ctx("c2mini")
    nc = T.nCells
    black = np.zeros(nc + 1, np.int32)
    black[1::3] = 1
    try:
        forb, hb, _ = T.line_luminosities(g["Hden"], g["Te"], g["Ne"], ion, lgBlack=black)
    finally:
        T.close()
    keep = cells[black[cells] == 0]
    ex, hbx = O.Balance(b).line_lum(g["Hden"], g["Te"], g["Ne"], ion, dV, keep, "f64")
    big = np.abs(ex) > 1e-30
    assert np.max(np.abs(forb[big] / ex[big] - 1)) < 1e-7
    # each cell's Hbeta is REAL arithmetic: the device powf and the C
    # library's may differ in the last bit
    assert abs(hb / hbx - 1) < 1e-6


def test_device_state_after_balance_gives_reference_ratios():
    """One GPU balance pass from the reference's starting state and
    estimator, then the line luminosities of the device state: every line
    above 1e-3 Hbeta within 2 % of the reference's ratio to Hbeta."""
    print("lines from the device balance state", flush=True)
    T, g, b, dV, cells, ion = _ctx("c2mini")
    nc, nb = T.nCells, T.nbins
    J = np.zeros((nc + 1) * nb, np.float32)
    J[g["J_idx"]] = g["J_val"]
    ion0 = g["ionDen0"].reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0)
    try:
        with pytest.raises(RuntimeError):
            T.line_luminosities()              # no device state yet
        T.set_estimators(J.reshape(nb, nc + 1).T.astype(np.float64))
        Te, Ne, ionD, flags, nIt, _ = T.update_cells(g["Hden"], g["Te0"], g["Ne0"], ion0)
        forb, hb, ms = T.line_luminosities()
    finally:
        T.close()
    assert np.max(np.abs(Te[1:] / g["Te"][1:] - 1)) < 0.01
    ref = g["lines_L"].reshape(-1, 17, 17).astype(np.float64)
    hbr = float(g["lines_hb"][0])
    assert abs(hb / hbr - 1) < 0.02
    strong = ref > 1e-3 * hbr
    assert strong.sum() > 30
    rel = np.abs((forb[strong] / hb) / (ref[strong] / hbr) - 1)
    print("lines from device state: %d lines, max ratio error %.2e, %.3f ms" % (strong.sum(), rel.max(), ms))
    assert rel.max() < 0.02
