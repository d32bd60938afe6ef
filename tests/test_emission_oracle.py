"""The re-emission tables (emissionDriver's gas part and setDiffusePDF's
dust term: recPDF rows and totalLines, emission_mod.f90:32-1311) restated
in the reference's mixed precision (oracle mco_emission_cell_f32) against
the unmodified reference on the state after one balance pass
(tests/golden/opacity_*.npz: em_recPDF, em_totalLines), and the variant
with FP64 level populations the GPU kernel follows.  No GPU."""
import numpy as np
import pytest

import mc_oracle as O
from conftest import balance_fixture, golden


def _inputs(name):
    g = golden("opacity_" + name)
    b = balance_fixture(name)[0]
    nc, nb = int(g["ints"][3]), int(g["ints"][4])
    ion = g["ionDen"].reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0)
    dust = None
    if "dust_abs" in g:
        nsz = int(g["dust_weight"].size)
        ns_ = int(g["dust_abun"].size)
        dust = dict(absx=g["dust_abs"].reshape(ns_, nsz, nb), weight=g["dust_weight"], gabun=g["dust_abun"],
                    tsub=g["dust_tsub"], Tdust=g["Tdust"].reshape(nc + 1, nsz + 1, -1), Ndust=g["dust_ndust"],
                    Hden=g["Hden"])
    return g, b, ion, dust


@pytest.mark.parametrize("name", ["c2mini", "dust_pe"])
def test_fp32_restatement_matches_reference(name):
    g, b, ion, dust = _inputs(name)
    E, keep = O.emission_tables(g)
    cells = g["cells"]
    rec, tl = O.Balance(b).emission(E, g["Te"], g["Ne"], ion, cells, "f32", dust=dust)
    # every recPDF entry and totalLines, bit for bit
    assert np.array_equal(rec.view(np.uint32), g["em_recPDF"].view(np.uint32))
    assert np.array_equal(tl.view(np.uint32), g["em_totalLines"][cells].view(np.uint32))
    # the rows are CDFs ending at 1, and some of the diffuse field ionises H
    assert np.all(np.diff(rec, axis=1) >= 0) and np.all(rec[:, -1] == 1.0)
    ip = int(g["em_ints"][4])
    assert np.all(rec[:, ip - 2] < 1.0)
    assert np.all((tl > 0) & (tl < 1))


def test_without_the_dust_term_the_rows_change():
    """The dust term is not a no-op on the dust deck: leaving it out moves
    the CDF and totalLines (a plausible omission the fixture catches)."""
    g, b, ion, dust = _inputs("dust_pe")
    E, keep = O.emission_tables(g)
    cells = g["cells"]
    rec, tl = O.Balance(b).emission(E, g["Te"], g["Ne"], ion, cells, "f32")
    assert np.max(np.abs(rec - g["em_recPDF"])) > 0.1
    assert not np.array_equal(tl, g["em_totalLines"][cells])


def test_fp64_populations_move_totallines_only():
    """With the level populations in FP64 (the GPU's), recPDF is unchanged
    (it holds no forbidden lines) and totalLines moves within 1e-6."""
    g, b, ion, dust = _inputs("c2mini")
    E, keep = O.emission_tables(g)
    cells = g["cells"]
    rec, tl = O.Balance(b).emission(E, g["Te"], g["Ne"], ion, cells, "f64")
    assert np.array_equal(rec, g["em_recPDF"])
    ref = g["em_totalLines"][cells]
    assert np.max(np.abs(tl / ref - 1)) < 1e-6
