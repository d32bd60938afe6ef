"""The per-cell balance restatement (oracle/mc_balance_impl.h) pinned
against the unmodified reference's updateCell, and the FP64 oracle against
it and against closed forms.  No GPU."""
import numpy as np
import pytest

import mc_oracle as O
from conftest import balance_fixture

NAMES = ["c1mini", "c2mini", "c2mini_thick", "dust_bal"]


@pytest.mark.parametrize("name", NAMES)
def test_fp32_restatement_is_bit_identical_to_reference(name):
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture(name)
    B = O.Balance(b)
    Te, Ne, ion, flags, nT = B.run(J, Hden, Te0, Ne0, ion0, "f32")
    hit = (flags & 16) == 0
    assert hit[1:].sum() > 20
    assert np.array_equal(Te[1:].view(np.uint32), ref["Te"][1:].view(np.uint32))
    assert np.array_equal(Ne[1:].view(np.uint32), ref["Ne"][1:].view(np.uint32))
    assert np.array_equal(ion[1:].view(np.uint32), ref["ion"][1:].view(np.uint32))
    conv = (flags & 1) != 0
    black = (flags & 2) != 0
    assert np.array_equal(conv[1:][hit[1:]], ref["conv"][1:][hit[1:]] == 1)
    assert np.array_equal(black[1:], ref["black"][1:] == 1)


@pytest.mark.parametrize("name", NAMES)
def test_fp64_oracle_vs_reference(name):
    """Stated error of the FP64 oracle against the FP32 reference."""
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture(name)
    B = O.Balance(b)
    Te, Ne, ion, flags, nT = B.run(J, Hden, Te0, Ne0, ion0, "f64")
    c = slice(1, None)
    assert np.max(np.abs(Te[c] - ref["Te"][c]) / ref["Te"][c]) < 1e-3
    assert np.max(np.abs(Ne[c] - ref["Ne"][c]) / ref["Ne"][c]) < 1e-3
    big = ref["ion"][c] > 1e-4
    assert np.max(np.abs(ion[c][big] - ref["ion"][c][big]) / ref["ion"][c][big]) < 2e-3


def test_ionisation_fractions_sum_to_one_and_ne_is_charge():
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture("c2mini")
    B = O.Balance(b)
    Te, Ne, ion, flags, nT = B.run(J, Hden, Te0, Ne0, ion0, "f64")
    on = [z for z in range(1, 31) if b.elem_on[z - 1]]
    for z in on:
        k = int(b.xref[z - 1]) - 1
        nst = min(z + 1, b.nstages)
        s = ion[1:, k, :nst].sum(axis=1)
        # clamping to [1e-20, 0.99999] aside, fractions sum to one
        assert np.all(np.abs(s - 1) < 2e-5)
    charge = np.zeros(len(Te))
    for z in on:
        k = int(b.xref[z - 1]) - 1
        for st in range(2, min(z + 1, b.nstages) + 1):
            x = ion[:, k, st - 1]
            charge += np.where(x >= 1e-10, (st - 1) * float(b.abun[z - 1]) * x, 0.0)
    assert np.allclose(Ne[1:], charge[1:] * Hden[1:], rtol=1e-12)


def test_hydrogen_photoionisation_equilibrium_closed_form():
    """Hydrogen's ion balance in every reached cell is the photoionisation
    equilibrium x(H+) Ne alpha_B(Te) = x(H0) Gamma: Gamma is the integral of
    J sigma / (h nu) over the H channel (update_mod.f90:184-269), written here
    from J and the cross-sections, and alpha is Verner & Ferland's (1996,
    ApJS 103, 467) fit for H I, whose published coefficients the atomic data
    must hold.  The fixed point stops when X(H0) moves by under 1 %, so the
    equation holds to about that; collisional ionisation is below 1e-6 of
    Gamma at these temperatures.  A dropped J term, a wrong h nu, a wrong
    recombination fit or Ne from the wrong species fails it."""
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture("c1mini")
    # Verner & Ferland (1996) Table 1, H I: a, b, T0, T1
    assert np.allclose(np.asarray(b.rnew)[0, 0], [7.982e-11, 0.7480, 3.148, 7.036e5], rtol=1e-6)
    B = O.Balance(b)
    Te, Ne, ion, flags, nT = B.run(J, Hden, Te0, Ne0, ion0, "f64")
    m = b.ch_meta[0]
    assert (m[0], m[1]) == (1, 1)
    ip, hi = int(m[4]), int(m[5])
    sig = b.ch_sigma[0].astype(np.float64)[ip - 1:hi]
    sig[sig < 1e-35] = 0.0
    hnu = 2.1799153e-11 * b.nuArray.astype(np.float64)[ip - 1:hi]      # erg, h c Ryd nu
    checked = 0
    for c in range(1, J.shape[0]):
        if flags[c] & (4 | 8 | 16):
            continue
        j = J[c].astype(np.float64)[ip - 1:hi]
        gam = 1e-20 + np.sum(np.where(j > 0, j * sig / hnu, 0.0))
        t = float(Te[c])
        sq0, sq1 = np.sqrt(t / 3.148), np.sqrt(t / 7.036e5)
        alpha = 7.982e-11 / (sq0 * (1 + sq0) ** (1 - 0.748) * (1 + sq1) ** (1 + 0.748))
        x0, xp = ion[c, 0, 0], ion[c, 0, 1]
        assert abs(xp * Ne[c] * alpha - x0 * gam) <= 0.02 * x0 * gam, (c, xp * Ne[c] * alpha / (x0 * gam))
        checked += 1
    assert checked > 100
    assert np.all(ion[1:, 0, 1] > 0.99)        # matter-bounded, fully ionised (R_s > R_out)
