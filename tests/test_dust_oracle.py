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
