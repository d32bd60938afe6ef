"""The workloads of the clumpy learned warm start and the batched serving,
at reduced size, on the device:

  clumpy  the lognormal-plus-clumps H+He full grid (mocassin_amd.synth.
          clumpy_density) at 23^3, the opacity and re-emission tables built on
          the device from the reference's cold start, one sweep; the plain
          balance (the reference's schedule) equals the FP64 restatement cell
          by cell, and the warm-started balance (the network trained on a
          second seeded variant) reaches the restatement's temperatures within
          1 % in every cell both converge;
  batch   octant models with edges 20..80 cells (mocassin_amd.nebulae's
          specification, packets scaled down) in one batched launch, equal to
          their separate launches packet for packet;
  dust    the dust and gas workload (mocassin_amd.dustgas: lognormal nH,
          astronomical-silicate grains, Henyey-Greenstein scattering, 50,000 K
          source) at 21^3: the opacity, scattering and re-emission tables built
          on the device from the cold start feed a sweep that equals the sweep
          on the same tables through host memory, and the balance with grain
          charge, photoelectric heating and gas-grain collisions on that J
          equals the FP64 restatement cell by cell.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, golden, gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def _atomic_c1():
    from mocassin_amd.atomic import AtomicBundle
    atom = dict(np.load(os.path.join(ROOT, "tests", "golden", "atomic_c1.npz")))
    return atom, AtomicBundle.from_npz(atom)


def _clumpy(seed, atom, b, npk):
    from mocassin_amd import Transport, synth
    from mocassin_amd.atomic import emission_tables
    from mocassin_amd.converge import cold_state
    n = 23
    dens = synth.clumpy_density(n, seed, corr=8.0 * n / 100, n_clumps=12)
    p = synth.build_problem(n, 3.e18, 3.e18, 300., 4.e48, 4.e4, 400, synth.H_HE, symmetric=False,
                            n_packets=npk, density=dens, nu=b.nuArray)
    T = Transport(p["xAxis"], p["yAxis"], p["zAxis"], p["active"], p["nCells"], p["nbins"], p["nuArray"],
                  p["R_out"], p["geoCorr"], p["ionEdge1"], [p["starPos"]], [p["starIdx"]], [p["deltaE"]],
                  symmetric=False, seed=700 + seed, device=0)
    T.balance_setup(b)
    T.opacity_setup(atom["op_seg"], atom["op_sigma"], float(atom["op_brems1"][0]))
    T.emission_setup(**emission_tables(atom))
    Te0, Ne0, ion0, Hd = cold_state(b, atom["Te0"][0], atom["ion0"], p["Hden"])
    T.opacity_build(Hd, Te0, Ne0, ion0, host=False)
    T.emission_build(Hd, Te0, Ne0, ion0, host=False)
    T.upload_tables(None, p["specCDF"])
    T.zero_estimators()
    T.energy_packet_driver(1, npk, 0)
    return T, p, (Hd, Te0, Ne0, ion0)


def test_clumpy_balance_plain_and_warm_vs_fp64_restatement():
    import mc_oracle as O
    from mocassin_amd import mlp as MP
    atom, b = _atomic_c1()
    js = 1e-9
    print("clumpy: training variant", flush=True)
    Ttr, ptr, (Hd_t, Te_t, Ne_t, ion_t) = _clumpy(11, atom, b, 400000)
    try:
        r = Ttr.update_cells(Hd_t, Te_t, Ne_t, ion_t, jscale=js, solver="secant", tol=1e-5, max_iter=100,
                             ion_tol=1e-4)
        X, hit = Ttr.mlp_features(Hd_t, jscale=js)
        cells = np.nonzero(hit[1:] & ((r[3][1:] & (4 | 8)) == 0))[0] + 1
        Tg, M = MP.targets(b, r[0], r[2], cells)
        M[:, 0] *= 5.0
        MP.train(Ttr, X[cells - 1], Tg, M, steps=2000, batch=1024, lr=2e-3, seed=3)
        params, scale = Ttr.get_mlp()
    finally:
        Ttr.close()
    print("clumpy: judged variant", flush=True)
    T, p, (Hd, Te0, Ne0, ion0) = _clumpy(12, atom, b, 400000)
    try:
        Tp, Np, ionp, flp, nIp, _ = T.update_cells(Hd, Te0, Ne0, ion0, jscale=js)
        T.set_mlp(params, scale)
        Tw, Nw, ionw, flw, nIw, _ = T.update_cells(Hd, Te0, Ne0, ion0, jscale=js, solver="secant", tol=1e-3,
                                                   max_iter=60, ion_tol=1e-3, warm_start=True)
        J, _ = T.fetch()
    finally:
        T.close()
    B = O.Balance(b)
    To, No, iono, flo, nIo = B.run(J * js, Hd, Te0, Ne0, ion0, "f64")
    reached = (flo[1:] & 16) == 0
    assert reached.sum() > 2000
    same = (nIp[1:] == nIo[1:]) & (flp[1:] == flo[1:])
    rel = np.abs(Tp[1:] / To[1:] - 1)
    print("clumpy plain: %d cells, %d reached, identical schedule in %.4f of them, Te max rel diff %.2e "
          "(same schedule), %.2e (all)" % (Tp.size - 1, reached.sum(), same.mean(), rel[same].max(), rel.max()))
    assert same.mean() >= 0.999 and rel[same].max() < 1e-9
    assert np.all(rel[~same] < 0.01)
    conv = reached & ((flo[1:] & 4) == 0) & ((flw[1:] & 4) == 0)
    relw = np.abs(Tw[1:][conv] / To[1:][conv] - 1)
    restarted = np.count_nonzero(flw[1:][conv] & 32)
    print("clumpy warm: %d cells converged in both, Te max rel diff %.2e, mean T-iterations %.2f (plain %.2f), "
          "%d restarted" % (conv.sum(), relw.max(), nIw[1:][conv].mean(), nIp[1:][conv].mean(), restarted))
    assert conv.sum() > 0.9 * reached.sum() and relw.max() < 0.01


def test_batch_of_edges_20_to_80_equals_separate_runs():
    from mocassin_amd import transport_batch
    from mocassin_amd.nebulae import build, model_specs
    atom, b = _atomic_c1()
    specs = model_specs(64)
    # the smallest and the largest edge and four between, packets / 200
    order = np.argsort([s["edge"] for s in specs])
    pick = [specs[i] for i in order[np.linspace(0, 63, 6).astype(int)]]
    for s in pick:
        s["packets"] = max(1000, s["packets"] // 200)
    edges = [s["edge"] for s in pick]
    print("batch of edges %s" % edges, flush=True)
    assert min(edges) <= 21 and max(edges) >= 79
    models = build(pick, b.nuArray, log=lambda m: print(m, flush=True))
    try:
        npk = [s["packets"] for s in pick]
        for T in models:
            T.zero_estimators()
        cb, ms = transport_batch(models, [1] * len(models), npk, None,
                                 {"waves_per_cu": 16, "emit_min": 32, "walk_steps": 8, "chunk": 4096})
        got = [T.fetch() for T in models]
        for i, T in enumerate(models):
            T.zero_estimators()
            c = T.energy_packet_driver(1, npk[i], 0)
            J, E = T.fetch()
            for k in ("n_seg", "n_abs", "n_esc", "n_line", "n_flights", "n_lost", "n_trapped"):
                assert cb[i][k] == c[k], (edges[i], k)
            assert np.allclose(got[i][0], J, rtol=1e-12, atol=0)
            assert np.allclose(got[i][1], E, rtol=1e-12, atol=0)
    finally:
        for T in models:
            T.close()


def test_dust_gas_lognormal_device_tables_and_grain_balance():
    import mc_oracle as O
    from mocassin_amd import dustgas
    from mocassin_amd.iterate import jscale_of
    n, npk = 21, 400000
    print("dust and gas %d^3: device tables, sweep, grain balance" % n, flush=True)
    W = dustgas.open_context(n, corr=8.0 * n / 60, packets=npk)
    T, p, Td0 = W["T"], W["p"], W["Tdust0"]
    Hd, Te0, Ne0, ion0 = W["state"]
    js = jscale_of(False)
    try:
        op, sca, _ = T.opacity_build(Hd, Te0, Ne0, ion0, Tdust=Td0)
        rec, tl, _ = T.emission_build(Hd, Te0, Ne0, ion0, Tdust=Td0)
        runs = []
        for host in (False, True):
            if host:
                T.upload_tables(op, p["specCDF"], rec, tl, scaOpac=sca, gSca=W["gSca"])
            else:
                T.upload_tables(None, p["specCDF"], gSca=W["gSca"])
            T.zero_estimators()
            c = T.energy_packet_driver(1, npk, 0)
            runs.append((c, T.fetch()[0]))
        Tg, fl, nI, ms, Te, Ne, ion = T.update_grains(Td0, jscale=js, Hden=Hd, Te=Te0, Ne=Ne0, ionDen=ion0)
    finally:
        T.close()
    (c1, J1), (c2, J2) = runs
    for k in ("n_seg", "n_abs", "n_sca", "n_esc", "n_line", "n_flights", "n_lost", "n_trapped"):
        assert c1[k] == c2[k], k
    print("dust and gas: %d cells, %.1f segments, %.2f scatterings, %.2f absorptions per packet"
          % (p["nCells"], c1["n_seg"] / npk, c1["n_sca"] / npk, c1["n_abs"] / npk))
    assert c1["n_sca"] > 0.1 * npk and c1["n_abs"] > 0.1 * npk and c1["n_lost"] == 0
    assert np.allclose(J1, J2, rtol=1e-12, atol=0)
    # the FP64 restatement of the balance with grains on the same J
    g = dict(golden("balance_dust_pe"))
    g["dust_ndust"] = W["Ndust"]
    hit = np.nonzero((fl[1:] & 16) == 0)[0] + 1
    assert hit.size > 0.5 * p["nCells"]
    cells = hit[np.linspace(0, hit.size - 1, 400).astype(int)]
    J = J2 * js
    B = O.Balance(W["bundle"])
    Teo, Neo, iono, flo, nIo, gdc, peh, pot = B.run_grains(J, Hd, Te0, Ne0, ion0, g, Td0, "f64", cells=cells)
    To = O.dust_temps(J, Td0, g, "f64", gdc=gdc, peh=peh)
    assert np.array_equal(fl[cells], flo[cells]) and np.array_equal(nI[cells], nIo[cells])
    assert np.max(np.abs(Te[cells] / Teo[cells] - 1)) < 1e-9
    assert np.max(np.abs(Ne[cells] / Neo[cells] - 1)) < 1e-9
    on = iono[cells] > 0
    assert np.max(np.abs(ion[cells][on] / iono[cells][on] - 1)) < 1e-8
    assert np.max(np.abs(Tg[cells] - To[cells]) / To[cells].clip(1)) < 1e-9
    print("dust and gas: balance of %d sampled cells equals the FP64 restatement; Te %.0f-%.0f K, "
          "grains %.0f-%.0f K, %.1f ms" % (cells.size, Te[cells].min(), Te[cells].max(),
                                          Tg[cells][:, 1:, 1].min(), Tg[cells][:, 1:, 1].max(), ms))
