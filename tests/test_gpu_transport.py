"""GPU parity of the transport path, through the C ABI (libmocassin_amd.so).

Bars:
  * geometry replay: the visited cells and FP32 segment lengths produce the
    reference's own per-packet Jste bits (golden vectors from the unmodified
    reference) and equal the oracle's walk bit for bit;
  * transport: same packets and streams as the oracle (deterministic math,
    FP64 accumulation): every event counter equal, J and escapes equal to
    1e-12 relative (only the FP64 summation order of the atomics differs),
    with one packet per lane and with two;
  * against the reference itself: the reference draws glibc's faithfully
    rounded log/sin/cos while the kernel draws correctly rounded ones, so
    ~1-3% of histories differ after a 1-ulp direction change; totals agree
    within Monte Carlo noise;
  * full size (45^3 octant, 800 bins): packet bookkeeping exact, results
    independent of launch partition, and a sampled packet range equal to
    the oracle.
"""
import numpy as np
import pytest

import mc_oracle as O
from conftest import (golden, grid_from_fixture, j_map_from_segments, oracle_problem,
                      reference_estimators, walk_packets)

pytestmark = pytest.mark.gpu

# The transport kernels: one packet per lane, two (the automatic choice at
# most 3 workgroups of 256 per CU), and one per lane with a pool of 32
# packets per wavefront in LDS (the library default, 4 per CU).
KERNELS = {"one": {"blocks_per_cu": 4, "threads": 256, "emit_min": 64, "walk_steps": 20, "kernel": "one"},
           "two": {"blocks_per_cu": 3, "threads": 256, "emit_min": 64, "walk_steps": 16, "kernel": "auto"},
           "pool": {"blocks_per_cu": 4, "threads": 256, "emit_min": 16, "walk_steps": 4, "kernel": "pool"}}


def transport_from_problem(P, seed, **kw):
    from mocassin_amd import Transport
    T = Transport(P.xAxis, P.yAxis, P.zAxis, P.active, P.nCells, P.nbins, P.nuArray,
                  P.R_out, P.geoCorr, P.ionEdge1, [P.starPos], [P.starIdx], [P.deltaE],
                  symmetric=P.symmetric, dust=P.dust, gas=P.gas, isotropic=P.isotropic,
                  multistars=P.multistars, seed=seed, **kw)
    T.upload_tables(P.opacity, P.specCDF.reshape(-1, 1), P.recPDF, P.totalLines,
                    P.scaOpac, P.gSca)
    return T


@pytest.mark.parametrize("deck", ["c1mini", "full_mini"])
def test_walk_replay_bit_exact_vs_reference(deck):
    g = golden("walk_" + deck)
    axes, active, nc, nb, fl, _ = grid_from_fixture(g)
    sc = g["scalars"]
    P = O.Problem(xAxis=axes[0], yAxis=axes[1], zAxis=axes[2], active=active,
                  opacity=np.zeros((nc + 1, nb), np.float32, order="F"),
                  nuArray=np.linspace(0.5, 2.0, nb).astype(np.float32),
                  specCDF=np.linspace(0.1, 1.0, nb).astype(np.float32),
                  R_out=sc[0], geoCorr=sc[1:4], ionEdge1=sc[4], deltaE=1.0,
                  starPos=(0, 0, 0), starIdx=(1, 1, 1),
                  recPDF=np.zeros((nc + 1, nb), np.float32, order="F") + 1,
                  totalLines=np.zeros(nc + 1, np.float32), **fl)
    pos, dirv, cell, maps, esc = walk_packets(g)
    print("walk replay", deck, len(pos), "packets", flush=True)
    T = transport_from_problem(P, 1)
    seq, ds, ln, st = T.walk_replay(pos, dirv, cell, max_len=1024)
    T.close()
    for i in range(len(pos)):
        n = int(ln[i])
        got = j_map_from_segments(seq[i, :n], ds[i, :n], axes, active, fl["symmetric"])
        assert got == maps[i], "packet %d" % i
        assert (st[i] == 1) == (esc[i] == 1.0)
        oc, ods, ost = P.walk(pos[i], dirv[i], cell[i])
        assert ost == st[i] and np.array_equal(oc, seq[i, :n])
        assert np.array_equal(ods.view(np.uint32), ds[i, :n].view(np.uint32))


def _cmp_estimators(J, E, Jo, Eo, rtol=1e-12):
    scale = max(float(np.abs(Jo).max()), 1e-300)
    assert np.allclose(J, Jo, rtol=rtol, atol=rtol * scale * 1e-6)
    assert np.allclose(E, Eo, rtol=rtol, atol=rtol * float(np.abs(Eo).max() + 1e-300) * 1e-6)
    assert np.array_equal(J > 0, Jo > 0)


def _check_pass_counters(c, npackets):
    """The scheduling counters bound the event counters: a walking lane
    ends its pass with a segment or as lost (ST_LOST; safeLimit losses
    follow a segment), an emitting lane starts a flight or ends its packet
    in a line or an escape, every packet is emitted at least once, and no
    pass has more than 64 lanes."""
    wl, wp, el, ep = c["n_walk_lanes"], c["n_walk_passes"], c["n_emit_lanes"], c["n_emit_passes"]
    assert c["n_seg"] <= wl <= c["n_seg"] + c["n_lost"]
    assert c["n_flights"] <= el <= c["n_flights"] + c["n_line"] + c["n_esc"]
    assert el >= npackets
    assert 0 < wl <= 64 * wp and 0 < el <= 64 * ep


@pytest.mark.parametrize("kernel", sorted(KERNELS))
@pytest.mark.parametrize("name", ["c1mini", "c1mini_thick", "full_mini_thick", "dust_mini",
                                  "dust_mini_thick", "c3mini"])
def test_transport_equals_oracle(name, kernel):
    g = golden("transport_" + name)
    P = oracle_problem(g)
    n, seed = int(g["npackets"]), int(g["seed"])
    print("transport vs oracle", name, kernel, n, flush=True)
    Jo, Eo, co = P.transport(0, n, seed, 1, O.MATH_OWN, O.J_FP64)
    T = transport_from_problem(P, seed)
    if KERNELS[kernel]:
        T.set_tuning(**KERNELS[kernel])
    c = T.energy_packet_driver(1, n, 0)
    J, E = T.fetch()
    T.close()
    for k in ("n_seg", "n_abs", "n_sca", "n_esc", "n_lost", "n_line", "n_trapped", "n_flights"):
        assert c[k] == co[k], k
    assert np.isclose(c["qphot"], co["qphot"], rtol=1e-12)
    _check_pass_counters(c, n)
    _cmp_estimators(J, E, Jo, Eo)


@pytest.mark.parametrize("kernel", sorted(KERNELS))
@pytest.mark.parametrize("name", ["full_mini_view", "c1mini_view"])
def test_viewing_angle_planes_equal_oracle(name, kernel):
    """escapedPackets(:,:,0:2) with two inclinations (phi on, on the full
    grid; phi off, on the symmetric octant): the GPU equals the oracle packet
    for packet; the oracle equals the reference bit for bit in libm mode
    (test_oracle_golden.py)."""
    from mocassin_amd import McError
    from conftest import reference_escapes_all, view_from_fixture
    g = golden("transport_" + name)
    view = view_from_fixture(g)
    P = oracle_problem(g, view=view)
    n, seed = int(g["npackets"]), int(g["seed"])
    print("viewing angles vs oracle", name, kernel, n, flush=True)
    Jo, Eo, co = P.transport(0, n, seed, 1, O.MATH_OWN, O.J_FP64)
    T = transport_from_problem(P, seed)
    if KERNELS[kernel]:
        T.set_tuning(**KERNELS[kernel])
    T.set_view_angles(view)
    assert T._lib.mc_esc_planes(T._h) == 3
    c = T.energy_packet_driver(1, n, 0)
    J, E = T.fetch()
    assert E.shape == Eo.shape == (P.nCells + 1, P.nbins + 1, 3)
    for k in ("n_seg", "n_abs", "n_sca", "n_esc", "n_lost", "n_line", "n_trapped", "n_flights"):
        assert c[k] == co[k], k
    _cmp_estimators(J, E, Jo, Eo)
    Er = reference_escapes_all(g)
    for a in range(3):
        assert (E[:, :, a].sum() > 0) == (Er[:, :, a].sum() > 0), a
    bad = dict(view, viewPointPtheta=np.where(np.asarray(view["viewPointPtheta"]) > 0, 3, 0))
    with pytest.raises(McError):
        T.set_view_angles(bad)
    T.set_view_angles(None)
    assert T._lib.mc_esc_planes(T._h) == 1
    J2, E2 = T.fetch()
    assert E2.shape == (P.nCells + 1, P.nbins + 1) and not E2.any()
    T.close()


@pytest.mark.parametrize("name", ["c1mini_thick", "full_mini_thick", "dust_mini"])
def test_transport_vs_reference_within_noise(name):
    g = golden("transport_" + name)
    P = oracle_problem(g)
    n, seed = int(g["npackets"]), int(g["seed"])
    Jr, Er = reference_estimators(g)
    T = transport_from_problem(P, seed)
    c = T.energy_packet_driver(1, n, 0)
    J, E = T.fetch()
    T.close()
    assert abs(c["n_abs"] - float(g["counters"][1])) <= 0.05 * float(g["counters"][1])
    assert abs(J.sum() - Jr.astype(np.float64).sum()) <= 0.05 * Jr.sum()
    assert abs(E.sum() - Er.astype(np.float64).sum()) <= 0.05 * Er.sum()
    # cells with many crossings agree cell by cell within 10%
    js, jrs = J.sum(axis=1), Jr.astype(np.float64).sum(axis=1)
    big = jrs > np.percentile(jrs[jrs > 0], 50)
    assert np.all(np.abs(js[big] - jrs[big]) <= 0.1 * jrs[big])


def test_launch_partition_independence():
    g = golden("transport_c1mini_thick")
    P = oracle_problem(g)
    n, seed = int(g["npackets"]), int(g["seed"])
    T = transport_from_problem(P, seed)
    T.energy_packet_driver(1, n, 0)
    J1, E1 = T.fetch()
    T.zero_estimators()
    T.energy_packet_driver(1, n // 3, 0)
    T.energy_packet_driver(1, n - n // 3, n // 3)
    J2, E2 = T.fetch()
    T.close()
    _cmp_estimators(J2, E2, J1, E1)


@pytest.mark.parametrize("kernel", sorted(KERNELS))
def test_ragged_counts_and_wide_packet_ids_equal_oracle(kernel):
    """Packet counts that do not fill a wavefront (1, 63, 65), none at all,
    and packet ranges across 2^32 and far above it: every packet id is a
    64-bit Philox counter, so each range equals the oracle's packet for
    packet, and a sweep of no packets leaves the estimators zero."""
    g = golden("transport_dust_mini_thick")
    P = oracle_problem(g)
    seed = int(g["seed"])
    T = transport_from_problem(P, seed)
    if KERNELS[kernel]:
        T.set_tuning(**KERNELS[kernel])
    try:
        c = T.energy_packet_driver(1, 0, 0)
        J, E = T.fetch()
        assert c["n_seg"] == 0 and c["n_flights"] == 0 and c["qphot"] == 0.0
        assert not J.any() and not E.any()
        for first, n in ((0, 1), (17, 63), (4096, 65), (2**32 - 40, 100), (2**40 + 3, 130),
                         (2**63 + 11, 70)):
            print("ragged/wide ids", first, n, flush=True)
            T.zero_estimators()
            c = T.energy_packet_driver(1, n, first)
            J, E = T.fetch()
            Jo, Eo, co = P.transport(first, n, seed, 1, O.MATH_OWN, O.J_FP64)
            for k in ("n_seg", "n_abs", "n_sca", "n_esc", "n_lost", "n_line", "n_trapped", "n_flights"):
                assert c[k] == co[k], (first, n, k)
            assert c["n_flights"] >= n
            _cmp_estimators(J, E, Jo, Eo)
    finally:
        T.close()


@pytest.fixture(scope="module")
def c2_problem():
    from mocassin_amd import synth
    p = synth.workload("c2", n_packets=1e8)
    return O.Problem(xAxis=p["xAxis"], yAxis=p["yAxis"], zAxis=p["zAxis"], active=p["active"],
                     opacity=p["opacity"], nuArray=p["nuArray"], specCDF=p["specCDF"][:, 0],
                     R_out=p["R_out"], geoCorr=p["geoCorr"], ionEdge1=p["ionEdge1"],
                     deltaE=p["deltaE"], starPos=p["starPos"], starIdx=p["starIdx"],
                     recPDF=p["recPDF"], totalLines=p["totalLines"])


@pytest.mark.parametrize("kernel", ["one", "two", "pool"])
def test_full_size_bookkeeping_and_sampled_parity(c2_problem, kernel):
    P = c2_problem
    print("c2 full size", kernel, P.nCells, P.nbins, flush=True)
    T = transport_from_problem(P, 2024)
    if KERNELS[kernel]:
        T.set_tuning(**KERNELS[kernel])
    n = 2_000_000
    c = T.energy_packet_driver(1, n, 0)
    J, E = T.fetch()
    assert c["n_esc"] + c["n_line"] + c["n_lost"] + c["n_trapped"] == n
    assert np.isclose(E.sum(), c["n_esc"] * float(P.deltaE), rtol=1e-9)
    assert np.all(J >= 0) and np.all(E >= 0)
    # a sampled packet range at full size equals the oracle
    T.zero_estimators()
    first, m = 1_234_567, 3000
    cs = T.energy_packet_driver(1, m, first)
    Js, Es = T.fetch()
    T.close()
    Jo, Eo, co = P.transport(first, m, 2024, 1, O.MATH_OWN, O.J_FP64)
    for k in ("n_seg", "n_abs", "n_esc", "n_line", "n_flights"):
        assert cs[k] == co[k], k
    _cmp_estimators(Js, Es, Jo, Eo)


def test_results_independent_of_launch_shape():
    g = golden("transport_full_mini_thick")
    P = oracle_problem(g)
    n, seed = int(g["npackets"]), int(g["seed"])
    T = transport_from_problem(P, seed)
    ref = None
    for tune in ({"blocks_per_cu": 1, "threads": 64, "emit_min": 1, "walk_steps": 1, "kernel": "auto"},
                 {"blocks_per_cu": 8, "threads": 256, "emit_min": 16, "walk_steps": 4, "kernel": "auto"},
                 {"blocks_per_cu": 16, "threads": 128, "emit_min": 64, "walk_steps": 64, "kernel": "auto"},
                 {"blocks_per_cu": 8, "threads": 256, "emit_min": 16, "walk_steps": 4, "kernel": "one"},
                 {"blocks_per_cu": 1, "threads": 64, "emit_min": 1, "walk_steps": 1, "kernel": "pool"},
                 {"blocks_per_cu": 8, "threads": 128, "emit_min": 64, "walk_steps": 64, "kernel": "pool"},
                 {"blocks_per_cu": 2, "threads": 192, "emit_min": 33, "walk_steps": 7, "kernel": "pool"}):
        T.set_tuning(**tune)
        T.zero_estimators()
        c = T.energy_packet_driver(1, n, 0)
        J, E = T.fetch()
        if ref is None:
            ref = (c, J, E)
            continue
        for k in ("n_seg", "n_abs", "n_esc", "n_line", "n_flights"):
            assert c[k] == ref[0][k], (tune, k)
        _cmp_estimators(J, E, ref[1], ref[2])
    T.close()


def test_tuning_kernel_field():
    """mc_tuning.kernel: out-of-range codes are refused and leave the shape
    unchanged; the names map to the header's codes; every kernel runs the
    same packets to the same counters."""
    from mocassin_amd import McError
    from mocassin_amd._capi import KERNELS as CODES
    g = golden("transport_c1mini_thick")
    P = oracle_problem(g)
    n, seed = 3000, int(g["seed"])
    T = transport_from_problem(P, seed)
    try:
        assert T.get_tuning() == {"blocks_per_cu": 4, "threads": 256, "emit_min": 16, "walk_steps": 16,
                                  "kernel": CODES["auto"]}
        for bad in (-1, 4, 99):
            with pytest.raises(McError):
                T.set_tuning(kernel=bad)
        assert T.get_tuning()["kernel"] == CODES["auto"]
        ref = None
        for name in ("auto", "one", "spare", "pool"):
            T.set_tuning(kernel=name)
            assert T.get_tuning()["kernel"] == CODES[name]
            T.zero_estimators()
            c = T.energy_packet_driver(1, n, 0)
            if ref is None:
                ref = c
            for k in ("n_seg", "n_abs", "n_esc", "n_line", "n_flights"):
                assert c[k] == ref[k], (name, k)
    finally:
        T.close()


def test_source_on_cell_walls_equals_oracle():
    """A source placed exactly on the corner where the walls of its cell meet
    those of the cells above it: every stellar packet moving up an axis
    starts with pathSegment's wall step (a packet within 1e-10 of a wall is
    moved into the next cell, photon_mod.f90:1263-1379) before its first
    segment, whose opacity is then the next cell's.  Clumpy density on a
    20^3 full grid, opacities scaled so that many first segments end in an
    absorption: the device equals the oracle packet for packet."""
    from mocassin_amd import synth
    n = 20
    dens = synth.clumpy_density(n, 5, corr=1.6, n_clumps=8)
    p = synth.build_problem(n, 3.e18, 3.e18, 300., 4.e48, 4.e4, 160, synth.H_HE, symmetric=False,
                            n_packets=40000, density=dens)
    x = np.asarray(p["xAxis"], np.float32)
    c = n // 2                                        # the source's cell (1-based) on each axis
    wall = (x[c - 1] + x[c]) / np.float32(2)          # the wall above it
    P = O.Problem(xAxis=p["xAxis"], yAxis=p["yAxis"], zAxis=p["zAxis"], active=p["active"],
                  opacity=p["opacity"] * np.float32(30), nuArray=p["nuArray"], specCDF=p["specCDF"][:, 0],
                  R_out=p["R_out"], geoCorr=p["geoCorr"], ionEdge1=p["ionEdge1"], deltaE=p["deltaE"],
                  starPos=(wall, wall, wall), starIdx=(c, c, c), recPDF=p["recPDF"],
                  totalLines=p["totalLines"], symmetric=False)
    print("source on walls", P.nCells, flush=True)
    T = transport_from_problem(P, 99)
    cg = T.energy_packet_driver(1, 40000, 0)
    J, E = T.fetch()
    T.close()
    Jo, Eo, co = P.transport(0, 40000, 99, 1, O.MATH_OWN, O.J_FP64)
    for k in ("n_seg", "n_abs", "n_esc", "n_lost", "n_line", "n_trapped", "n_flights"):
        assert cg[k] == co[k], k
    assert cg["n_abs"] > 1000
    _cmp_estimators(J, E, Jo, Eo)
