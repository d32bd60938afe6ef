"""outputGas's analytic line luminosities (forbVol and HbetaVol,
output_mod.f90:302-520) restated in the reference's FP32 (oracle,
mco_line_lum_f32) against the unmodified reference on the state after one
balance pass (tests/golden/opacity_*.npz: lines_L, lines_hb), and the FP64
restatement the GPU kernel follows.  No GPU."""
import numpy as np
import pytest

import mc_oracle as O
from conftest import balance_fixture, golden, grid_from_fixture


def _inputs(name):
    g = golden("opacity_" + name)
    b = balance_fixture(name)[0]
    axes, active, nc, nb, fl, _ = grid_from_fixture(g)
    dV, cells = O.cell_volumes(*axes, active, fl["symmetric"])
    ion = g["ionDen"].reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0)
    return g, b, dV, cells, ion


@pytest.mark.parametrize("name", ["c2mini", "dust_pe"])
def test_fp32_restatement_matches_reference(name):
    g, b, dV, cells, ion = _inputs(name)
    forb, hb = O.Balance(b).line_lum(g["Hden"], g["Te"], g["Ne"], ion, dV, cells, "f32")
    ref = g["lines_L"].reshape(-1, 17, 17)
    assert np.float32(hb) == g["lines_hb"][0]
    if b.ncel == 0:
        assert not np.any(ref)
        return
    # every line of every model ion, bit for bit
    assert np.array_equal(forb.view(np.uint32), ref.view(np.uint32))
    assert np.count_nonzero(ref) > 400
    # [OIII] 5007 (model ion O III, 1D2 -> 3P2) relative to Hbeta
    k = [tuple(m[:2]) for m in b.cel_meta].index((8, 3))
    assert 1.0 < forb[k, 3, 2] / hb < 10.0


def test_fp64_restatement_within_fp32_rounding():
    """The GPU's arithmetic (populations and sums in FP64) against the
    reference's REAL sums: lines above FP32's denormal range to a few ulp of
    FP32, Hbeta likewise."""
    g, b, dV, cells, ion = _inputs("c2mini")
    forb, hb = O.Balance(b).line_lum(g["Hden"], g["Te"], g["Ne"], ion, dV, cells, "f64")
    ref = g["lines_L"].reshape(-1, 17, 17).astype(np.float64)
    big = np.abs(ref) > 1e-30
    assert np.max(np.abs(forb[big] / ref[big] - 1)) < 1e-5
    assert abs(hb / float(g["lines_hb"][0]) - 1) < 1e-6
    # entries the reference holds as FP32 denormals or zero stay tiny
    assert np.all(np.abs(forb[~big]) < 1e-29)


def test_hden_enters_once_and_cells_add():
    """Hden multiplies each cell's emission once (populations depend on Ne,
    not Hden), and the sums over two disjoint cell sets add up."""
    g, b, dV, cells, ion = _inputs("c2mini")
    B = O.Balance(b)
    f1, h1 = B.line_lum(g["Hden"], g["Te"], g["Ne"], ion, dV, cells, "f64")
    f2, h2 = B.line_lum(g["Hden"] * 2, g["Te"], g["Ne"], ion, dV, cells, "f64")
    assert np.array_equal(f2, 2 * f1) and h2 == 2 * h1
    fa, ha = B.line_lum(g["Hden"], g["Te"], g["Ne"], ion, dV, cells[::2], "f64")
    fb, hb = B.line_lum(g["Hden"], g["Te"], g["Ne"], ion, dV, cells[1::2], "f64")
    assert np.allclose(fa + fb, f1, rtol=1e-12, atol=0) and abs((ha + hb) / h1 - 1) < 1e-12


@pytest.mark.parametrize("symmetric", [True, False])
def test_cell_volumes_tile_the_box(symmetric):
    """getVolume's widths on a uniform axis of spacing h: a symmetric grid
    covers (n - 1/2) h per axis (half a first cell), a full one n h."""
    n, h = 7, 2.e15
    ax = (np.arange(n) * h).astype(np.float32)
    active = np.arange(1, n ** 3 + 1, dtype=np.int32).reshape(n, n, n)
    dV, cells = O.cell_volumes(ax, ax, ax, active, symmetric)
    assert len(cells) == n ** 3 and cells[0] == 1 and cells[1] == 2
    side = (n - 0.5 if symmetric else n) * h / 1.e15
    assert abs(dV[1:].astype(np.float64).sum() / side ** 3 - 1) < 1e-6
    corner = (0.5 if symmetric else 1.0) * h / 1.e15
    assert dV[1] == np.float32(corner) ** 3
