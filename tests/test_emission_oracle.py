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


def test_cold_cells_restatement_matches_reference_nan_for_nan():
    """Cells at 1, 5, 19.554 and 60 K (temperatures iterateT's schedule
    leaves in barely lit cells; the balanced Ne and ionDen kept): the
    reference's own rows overflow into NaN from a bin on, and its totalLines
    is NaN.  The restatement forms the same rows: NaN where the reference has
    NaN, every other entry bit for bit."""
    g, b, ion, dust = _inputs("c2mini")
    E, keep = O.emission_tables(g)
    assert np.isnan(g["cold_recPDF"]).any()
    for k, c in enumerate(g["cold_cells"]):
        Te = g["Te"].copy()
        Te[c] = g["cold_Te"][k]
        rec, tl = O.Balance(b).emission(E, Te, g["Ne"], ion, [c], "f32")
        ref = g["cold_recPDF"][k]
        fin = np.isfinite(ref)
        assert np.array_equal(np.isnan(rec[0]), np.isnan(ref))
        assert np.array_equal(rec[0][fin].view(np.uint32), ref[fin].view(np.uint32))
        assert np.isnan(tl[0]) and np.isnan(g["cold_totalLines"][k])


def test_sampling_tables_draw_what_the_reference_scan_draws():
    """The rows the build writes for the reference's NaN rows (1 from the
    first NaN on; oracle.sampling_tables) give, through the transport's
    upper-bound search, the bin getNu2's literal scan gives on the raw row
    for every u; and a NaN totalLines never makes a line photon, as 0."""
    from test_oracle_golden import _getnu2_brute
    g, b, ion, dust = _inputs("c2mini")
    raw = np.concatenate([g["cold_recPDF"], g["em_recPDF"][:4]])
    rows, tl = O.sampling_tables(raw, np.concatenate([g["cold_totalLines"], [0.25] * 4]))
    assert np.all(tl[:4] == 0) and np.all(tl[4:] == np.float32(0.25))
    assert np.array_equal(rows[4:], g["em_recPDF"][:4])          # finite rows unchanged
    nb = raw.shape[1]
    us = np.random.default_rng(5).random(300).astype(np.float32)
    for r, m in zip(raw, rows):
        assert np.all(np.diff(m) >= 0) and m[-1] == 1.0
        for u in us:
            got, used = O.getnu2(m, [u])
            assert got == _getnu2_brute(r, u), (u, got)
            assert 2 <= got <= nb - 1
    # random <= totalLines with random in (0, 1]: never for NaN, never for 0
    assert not np.any(np.float32(1e-7) <= g["cold_totalLines"])


def test_barely_lit_chain_leaves_finite_states_and_sampleable_tables():
    """The FP64 balance (the GPU's model) from the reference's initial state
    on the estimator scaled by 1e-8, then from that state on it scaled by
    1e-10: every cell keeps a finite Te >= 1 K (the schedule's floor), some
    end at 1 K and some at a few K; the re-emission restatement of that
    state overflows into NaN rows there, and their sampling tables are CDFs
    ending at 1 with line fractions in [0, 1]."""
    from conftest import balance_fixture
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture("c2mini")
    B = O.Balance(b)
    Te1, Ne1, ion1, fl1, n1 = B.run(J * np.float32(1e-8), Hden, Te0, Ne0, ion0, "f64")
    Te2, Ne2, ion2, fl2, n2 = B.run(J * np.float32(1e-10), Hden, Te1, Ne1, ion1, "f64")
    t = Te2[1:]
    assert np.all(np.isfinite(t)) and np.all(t >= 1.0) and np.all(np.isfinite(Ne2[1:]))
    assert np.sum(t == 1.0) > 10 and np.sum((t > 1.0) & (t < 60.0)) > 5
    g, _, _, _ = _inputs("c2mini")
    E, keep = O.emission_tables(g)
    f32 = lambda a: np.asarray(a, np.float32)
    cells = np.arange(1, Te2.size)
    rec, tl = B.emission(E, f32(Te2), f32(Ne2), f32(ion2), cells, "f64")
    assert np.isnan(rec).any() and np.isnan(tl).any()
    rows, tl2 = O.sampling_tables(rec, tl)
    assert np.all(np.diff(rows, axis=1) >= 0) and np.all(rows[:, -1] == 1.0) and np.all(rows >= 0)
    assert np.all((tl2 >= 0) & (tl2 <= 1))
