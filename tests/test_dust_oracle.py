"""Grain temperatures (getDustT): the FP32 restatement (oracle/mc_dust.c)
against the unmodified reference's updateCell on a dust+gas deck, and the
FP64 oracle against it and against what the definition fixes.  No GPU."""
import numpy as np

import mc_oracle as O
from conftest import golden


def _fixture():
    g = golden("balance_dust_bal")
    nc, nb = int(g["ints"][3]), int(g["ints"][4])
    J = np.zeros((nc + 1) * nb, np.float32)
    J[g["J_idx"]] = g["J_val"]
    J = J.reshape(nb, nc + 1).T
    nS, nSz, nT, nsm = (int(v) for v in g["dust_ints"][:4])
    shp = lambda a: a.reshape(nc + 1, nSz + 1, nsm + 1)
    return g, J, shp(g["Tdust0"]), shp(g["Tdust"])


def test_fp32_restatement_is_bit_identical_to_reference():
    g, J, T0, Tref = _fixture()
    T = O.dust_temps(J, T0, g, "f32")
    hit = np.any(J > 0, axis=1)
    assert hit[1:].sum() > 100
    assert np.array_equal(T[1:].view(np.uint32), Tref[1:].view(np.uint32))
    # cells no packet reached keep their temperatures
    assert np.array_equal(T[1:][~hit[1:]], T0[1:][~hit[1:]])


def test_fp64_oracle_vs_reference():
    g, J, T0, Tref = _fixture()
    T = O.dust_temps(J, T0, g, "f64")
    assert np.max(np.abs(T[1:] - Tref[1:]) / np.maximum(Tref[1:], 1)) < 1e-4


def test_dust_temperature_inverts_the_emission_integral():
    """At the returned T the emission integral (linearly interpolated)
    equals the absorption integral; the size and species means are the
    weighted sums."""
    g, J, T0, Tref = _fixture()
    T = O.dust_temps(J, T0, g, "f64")
    nS, nSz, nT, nsm, nSp = (int(v) for v in g["dust_ints"][:5])
    em = g["dust_emint"].reshape(nT, nSz, nSp).astype(np.float64)   # (t, ai, s)
    absx = g["dust_abs"].reshape(nS, nSz, -1).astype(np.float64)
    w = g["dust_weight"].astype(np.float64)
    ab = g["dust_abun"].astype(np.float64)
    for c in range(1, J.shape[0], 7):
        if not np.any(J[c] > 0):
            continue
        rf = J[c].astype(np.float64) / np.float64(np.float32(3.141592654))
        for s in range(1, nS + 1):
            for ai in range(1, nSz + 1):
                x = float(np.sum(absx[s - 1, ai - 1] * rf))
                t = T[c, ai, s]
                if 1 < t < nT:
                    e = np.interp(t, np.arange(1, nT + 1), em[:, ai - 1, s - 1])
                    assert abs(e - x) <= 1e-9 * abs(x)
            assert abs(T[c, 0, s] - np.sum(T[c, 1:, s] * w)) <= 1e-12 * T[c, 0, s]
        assert abs(T[c, 0, 0] - np.sum(T[c, 0, 1:nS + 1] * ab)) <= 1e-12 * T[c, 0, 0]


def _pe_fixture():
    from conftest import balance_fixture
    g = golden("balance_dust_pe")
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture("dust_pe")
    nc = J.shape[0] - 1
    nS, nSz, nT, nsm = (int(v) for v in g["dust_ints"][:4])
    shp = lambda a: a.reshape(nc + 1, nSz + 1, nsm + 1)
    return g, b, J, Hden, Te0, Ne0, ion0, ref, shp(g["Tdust0"]), shp(g["Tdust"])


def test_photoelectric_fp32_restatement_is_bit_identical_to_reference():
    """Grain charge, photoelectric heating and gas-grain collisions
    (update_mod.f90:527-932) inside the temperature iteration, and their
    terms in getDustT: Te, Ne, every ionic fraction and every Tdust entry of
    the reference's updateCell, bit for bit."""
    g, b, J, Hden, Te0, Ne0, ion0, ref, T0, Tref = _pe_fixture()
    assert int(g["dust_pe"][0]) == 1
    B = O.Balance(b)
    Te, Ne, ion, flags, nIt, gdc, peh, pot = B.run_grains(J, Hden, Te0, Ne0, ion0, g, T0, "f32")
    hit = (flags & 16) == 0
    assert hit[1:].sum() > 200
    assert np.array_equal(Te[1:].view(np.uint32), ref["Te"][1:].view(np.uint32))
    assert np.array_equal(Ne[1:].view(np.uint32), ref["Ne"][1:].view(np.uint32))
    assert np.array_equal(ion[1:].view(np.uint32), ref["ion"][1:].view(np.uint32))
    assert np.array_equal(((flags & 1) != 0)[1:][hit[1:]], ref["conv"][1:][hit[1:]] == 1)
    T = O.dust_temps(J, T0, g, "f32", gdc=gdc, peh=peh)
    assert np.array_equal(T[1:].view(np.uint32), Tref[1:].view(np.uint32))
    # the terms matter: without them Te and Tdust move by percents
    T_off = O.dust_temps(J, T0, g, "f32")
    assert np.max(np.abs(T_off[1:] - Tref[1:]) / Tref[1:].clip(1)) > 1e-2
    # grain potentials: positive (photoelectric charging), below the photon energies
    assert np.all(pot[1:][hit[1:]] > 0) and np.all(pot[1:][hit[1:]] < 2.0)


def test_photoelectric_fp64_oracle_vs_reference():
    g, b, J, Hden, Te0, Ne0, ion0, ref, T0, Tref = _pe_fixture()
    B = O.Balance(b)
    Te, Ne, ion, flags, nIt, gdc, peh, pot = B.run_grains(J, Hden, Te0, Ne0, ion0, g, T0, "f64")
    c = slice(1, None)
    assert np.max(np.abs(Te[c] - ref["Te"][c]) / ref["Te"][c]) < 1e-5
    big = ref["ion"][c] > 1e-4
    assert np.max(np.abs(ion[c][big] - ref["ion"][c][big]) / ref["ion"][c][big]) < 1e-5
    T = O.dust_temps(J, T0, g, "f64", gdc=gdc, peh=peh)
    assert np.max(np.abs(T[c] - Tref[c]) / Tref[c].clip(1)) < 1e-5


def test_collision_potentials_are_the_channel_edges():
    """setDustGasCollHeatCool's ionisation potential of (elem, stage) is the
    edge of the photoionisation channel (elem, stage-1) the balance uses."""
    g = golden("balance_dust_pe")
    ns = int(g["at_ch_meta"].reshape(-1, 8).shape[0])
    ipc = g["dust_ipcoll"].reshape(-1, 30).T
    for m in g["at_ch_meta"].reshape(-1, 8):
        assert ipc[m[0] - 1, m[1]] == m[4]
    assert ns >= 3
