"""The CPU oracle pinned against the unmodified reference and against closed
forms.  No GPU.  Golden vectors: tests/golden/*.npz (tests/golden/make_fixtures.py)."""
import math

import numpy as np
import pytest

import mc_oracle as O
from conftest import (golden, grid_from_fixture, j_map_from_segments, oracle_problem,
                      reference_estimators, walk_packets)


# ------------------------------------------------------------------ RNG
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors published with Random123 (kat_vectors)."""
    assert O.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                           [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_is_24bit_in_unit_interval():
    u = np.array([O.uniform(7, 3, p, k) for p in range(50) for k in range(9)], np.float32)
    assert np.all(u >= 0) and np.all(u < 1)
    assert np.all((u * np.float32(2 ** 24)) == np.floor(u * np.float32(2 ** 24)))
    w = O.philox4x32_10([1, 5, 0, 3], [7, 0])
    assert O.uniform(7, 3, 5, 6) == np.float32(w[2] >> 8) * np.float32(2 ** -24)


# ----------------------------------------------------------- getNu2
def _getnu2_brute(cdf, u):
    # photon_mod.f90:745-757 read literally
    nuP = 1
    for i, c in enumerate(cdf, start=1):
        if u >= c:
            nuP = i
        else:
            break
    if nuP < len(cdf) - 1:
        nuP += 1
    return nuP


@pytest.mark.parametrize("seed", range(5))
def test_getnu2_matches_scan_and_closed_form(seed):
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(3, 40))
    w = rng.random(nb).astype(np.float32)
    w[rng.random(nb) < 0.3] = 0            # flat runs
    cdf = np.cumsum(w).astype(np.float32)
    cdf /= cdf[-1]
    cdf[cdf > np.float32(0.999998)] = 1.
    us = list(rng.random(200).astype(np.float32)) + list(cdf[:-1])   # exact CDF values too
    for u in us:
        if u in (0.0, 1.0, np.float32(0.9999999)):
            continue
        got, used = O.getnu2(cdf, [u])
        k = int(np.searchsorted(cdf, u, side="right"))     # closed form: #{cdf <= u}
        closed = max(k, 1)
        closed = closed + 1 if closed < nb - 1 else closed
        assert got == _getnu2_brute(cdf, u) == closed and used == 1


def test_getnu2_redraws_forbidden_values():
    cdf = np.linspace(0.1, 1.0, 10).astype(np.float32)
    got, used = O.getnu2(cdf, [0.0, np.float32(0.9999999), 0.55])
    assert used == 3 and got == _getnu2_brute(cdf, np.float32(0.55))


# ------------------------------------------------------ elementary functions
def test_own_math_is_correctly_rounded():
    rng = np.random.default_rng(3)
    for k in rng.integers(0, 1 << 24, 20000):
        u = np.float32(k) * np.float32(2 ** -24)
        x = np.float32(1) - u
        assert O.logf(float(x)) == np.float32(math.log(float(x)))
        a = np.float32(3.141592654) * (np.float32(2) * u - np.float32(1))
        s, c = O.sincosf(float(a))
        assert (s, c) == (np.float32(math.sin(float(a))), np.float32(math.cos(float(a))))
    assert O.logf(1.0) == 0.0 and O.logf(0.5) == np.float32(math.log(0.5))


def test_own_math_vs_libm_differ_by_at_most_one_ulp():
    import ctypes as C
    out = (C.c_int64 * 6)()
    O.lib().mco_math_mismatch(out)
    n = 1 << 24
    assert out[3] <= 1 and out[4] <= 1 and out[5] <= 1
    # glibc's float functions are faithfully, not correctly, rounded
    assert out[0] / n < 0.02 and out[1] / n < 0.05 and out[2] / n < 0.05


# ------------------------------------------------------------ walk geometry
@pytest.mark.parametrize("deck", ["c1mini", "full_mini"])
def test_oracle_walk_matches_reference_bit_exact(deck):
    g = golden("walk_" + deck)
    axes, active, nc, nb, fl, _ = grid_from_fixture(g)
    sc = g["scalars"]
    P = O.Problem(xAxis=axes[0], yAxis=axes[1], zAxis=axes[2], active=active,
                  opacity=np.zeros((nc + 1, nb), np.float32, order="F"),
                  nuArray=np.ones(nb, np.float32), specCDF=np.ones(nb, np.float32),
                  R_out=sc[0], geoCorr=sc[1:4], ionEdge1=sc[4], deltaE=1.0,
                  starPos=(0, 0, 0), starIdx=(1, 1, 1), **fl)
    pos, dirv, cell, maps, esc = walk_packets(g)
    n_ok = 0
    for i in range(len(pos)):
        cells, ds, st = P.walk(pos[i], dirv[i], cell[i])
        assert st in (1, 2), (i, st)
        got = j_map_from_segments(cells, ds, axes, active, fl["symmetric"])
        assert got == maps[i], "packet %d of %s" % (i, deck)
        assert (st == 1) == (esc[i] == 1.0), i          # escape recorded or silently dropped
        # invariants: one axis changes by one per segment, lengths positive
        steps = np.abs(np.diff(cells, axis=0)).sum(axis=1)
        assert np.all(steps <= 1) and np.all(ds > 0)
        n_ok += 1
    assert n_ok == len(pos) > 400


# ------------------------------------------------------------ full transport
@pytest.mark.parametrize("name", ["c1mini", "c1mini_thick", "full_mini_thick", "dust_mini",
                                  "dust_mini_thick", "c3mini"])
def test_oracle_transport_matches_reference_bit_exact(name):
    """libm mode + FP32 accumulation: identical bits to the reference's Jste,
    escapedPackets, Qphot and absInt under the same per-packet stream."""
    g = golden("transport_" + name)
    P = oracle_problem(g)
    J, E, cnt = P.transport(0, int(g["npackets"]), int(g["seed"]), 1, O.MATH_LIBM, O.J_FP32)
    Jr, Er = reference_estimators(g)
    assert np.array_equal(J.view(np.uint32), Jr.view(np.uint32))
    assert np.array_equal(E.view(np.uint32), Er.view(np.uint32))
    assert np.float32(cnt["qphot"]) == g["counters"][0]
    assert cnt["n_abs"] == g["counters"][1]
    assert cnt["n_sca"] == g["counters"][2]
    if name.startswith("dust") or name == "c3mini":
        assert cnt["n_sca"] > 50            # Henyey-Greenstein scatterings exercised


@pytest.mark.parametrize("name", ["full_mini_view", "c1mini_view"])
def test_oracle_viewing_angle_planes_match_reference_bit_exact(name):
    """escapedPackets(:,:,0:nAngleBins) of a run with two inclinations: the
    theta/phi bins and the plane rules of every escape site
    (photon_mod.f90:1582-1666), bit for bit in libm mode."""
    from conftest import reference_escapes_all, view_from_fixture
    g = golden("transport_" + name)
    P = oracle_problem(g, view=view_from_fixture(g))
    J, E, cnt = P.transport(0, int(g["npackets"]), int(g["seed"]), 1, O.MATH_LIBM, O.J_FP32)
    Er = reference_escapes_all(g)
    assert E.shape == Er.shape
    assert np.array_equal(E.view(np.uint32), Er.view(np.uint32))
    Jr, E0 = reference_estimators(g)
    assert np.array_equal(J.view(np.uint32), Jr.view(np.uint32))
    assert np.array_equal(E[:, :, 0].view(np.uint32), E0.view(np.uint32))
    # both viewpoints collect escapes
    assert (Er[:, :, 1].sum() > 0) and (Er[:, :, 2].sum() > 0)


def test_own_acos_atan_are_correctly_rounded():
    """The deterministic acos/atan (FP64 evaluation, one rounding) against
    numpy's FP64 functions rounded once, over a sweep of inputs; and within
    one ulp of the C library's float functions."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-1, 1, 4000), [-1.0, -0.5, 0.0, 0.5, 1.0, 1e-8, -1e-8]]).astype(np.float32)
    ts = np.concatenate([rng.standard_normal(2000) * 10.0, rng.uniform(-1e-3, 1e-3, 500),
                         [0.0, 1.0, -1.0, 1e30, -1e30]]).astype(np.float32)
    L = O.lib()
    for f, xsamp, ref in (("mco_acosf", xs, np.arccos), ("mco_atanf", ts, np.arctan)):
        own = np.array([getattr(L, f)(O.MATH_OWN, float(x)) for x in xsamp], np.float32)
        want = ref(xsamp.astype(np.float64)).astype(np.float32)
        assert np.array_equal(own, want), f
        libm = np.array([getattr(L, f)(O.MATH_LIBM, float(x)) for x in xsamp], np.float32)
        ulp = np.abs(own.view(np.int32).astype(np.int64) - libm.view(np.int32).astype(np.int64))
        assert ulp.max() <= 1, f


@pytest.mark.parametrize("name", ["c1mini_thick", "full_mini_thick"])
def test_energy_bookkeeping(name):
    """Every packet ends exactly once: escape, line packet, drop or trap."""
    g = golden("transport_" + name)
    P = oracle_problem(g)
    n = int(g["npackets"])
    J, E, cnt = P.transport(0, n, int(g["seed"]), 1, O.MATH_OWN, O.J_FP64)
    ends = cnt["n_esc"] + cnt["n_line"] + cnt["n_lost"] + cnt["n_trapped"]
    assert ends == n
    assert np.isclose(E.sum(), cnt["n_esc"] * P.deltaE, rtol=1e-12)
    assert np.all(J >= 0)


@pytest.mark.parametrize("name", ["c1mini_thick", "full_mini_thick"])
def test_own_math_statistically_equals_libm(name):
    """The deterministic math (used on the GPU) and glibc differ by <= 1 ulp
    on ~1% of arguments; the estimators agree within Monte Carlo noise."""
    g = golden("transport_" + name)
    P = oracle_problem(g)
    n = int(g["npackets"])
    Ja, Ea, ca = P.transport(0, n, int(g["seed"]), 1, O.MATH_LIBM, O.J_FP64)
    Jb, Eb, cb = P.transport(0, n, int(g["seed"]), 1, O.MATH_OWN, O.J_FP64)
    # most packets follow identical histories
    assert abs(ca["n_abs"] - cb["n_abs"]) <= 0.05 * ca["n_abs"]
    assert abs(Ja.sum() - Jb.sum()) <= 0.05 * Ja.sum()
    assert abs(Ea.sum() - Eb.sum()) <= 0.05 * Ea.sum()


def test_optically_thin_J_is_point_source_dilution():
    """kappa = 0 everywhere: the path-length estimator of each cell converges
    to L * <1/(4 pi r^2)> over the cell (Lucy 1999); checked on the summed J
    of all frequency bins against a Monte Carlo volume average of the exact
    point-source flux, in one symmetric octant (x8)."""
    g = golden("transport_c1mini")
    axes, active, nc, nb, fl, sidx = grid_from_fixture(g)
    P = oracle_problem(g, opacity=np.zeros((nc + 1, nb), np.float32, order="F"))
    n = 40000
    J, E, cnt = P.transport(0, n, 11, 1, O.MATH_OWN, O.J_FP64)
    Jc = J.sum(axis=1)                                  # sum of dS*dE/dV over bins
    # packets below the H edge escape at birth (photon_mod.f90:370); the
    # walking ones carry L = n_flights * deltaE  [1e36 erg/s]
    L = cnt["n_flights"] * P.deltaE
    rng = np.random.default_rng(0)
    checked = 0
    for (i, j, k) in [(3, 3, 3), (5, 2, 4), (6, 6, 6), (2, 7, 3)]:
        a = active[i - 1, j - 1, k - 1]
        if a <= 0:
            continue
        lo = [(ax[q - 2] + ax[q - 1]) / 2 for ax, q in zip(axes, (i, j, k))]
        hi = [(ax[q - 1] + ax[q]) / 2 for ax, q in zip(axes, (i, j, k))]
        pts = rng.uniform(lo, hi, size=(200000, 3))
        inv = np.mean(1.0 / (4 * np.pi * (pts ** 2).sum(axis=1)))
        # J_raw = sum dE[1e36 erg/s] dS[cm] / dV[1e45 cm^3]; all packets are
        # folded into the octant, so it collects 8x the flux of a full sphere
        expect = 8.0 * L * inv * 1e45
        assert abs(Jc[a] - expect) < 0.1 * expect, (a, Jc[a], expect)
        checked += 1
    assert checked >= 3
