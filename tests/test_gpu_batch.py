"""Batched transport: many nebulae of uneven grids in one persistent launch
give each nebula exactly what its own launch gives (every packet keeps its
random stream)."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def _models(specs):
    from mocassin_amd import Transport, synth
    out = []
    for i, (n, npk, teff, nh) in enumerate(specs):
        p = synth.build_problem(n, 2.0e18, 2.0e18, nh, 1.e48, teff, 120, synth.H_HE, n_packets=npk)
        T = Transport(p["xAxis"], p["yAxis"], p["zAxis"], p["active"], p["nCells"], p["nbins"],
                      p["nuArray"], p["R_out"], p["geoCorr"], p["ionEdge1"], [p["starPos"]],
                      [p["starIdx"]], [p["deltaE"]], symmetric=True, seed=100 + i, device=0)
        T.upload_tables(p["opacity"], p["specCDF"], p["recPDF"], p["totalLines"])
        out.append(T)
    return out


SPECS = [(9, 30000, 3.e4, 100.), (15, 0, 4.e4, 1.e3), (21, 120000, 6.e4, 10.), (11, 777, 5.e4, 1.e4),
         (17, 65000, 3.5e4, 300.)]


@pytest.mark.parametrize("tuning", [None, {"waves_per_cu": 4, "emit_min": 16, "walk_steps": 4, "chunk": 64},
                                    {"waves_per_cu": 32, "emit_min": 48, "walk_steps": 16, "chunk": 4096}])
def test_batch_equals_separate_runs(tuning):
    from mocassin_amd import transport_batch
    print("batch vs separate runs", tuning, flush=True)
    models = _models(SPECS)
    try:
        npk = [s[1] for s in SPECS]
        offs = [0, 0, 12345, 2**33 + 7, 0]                  # one range far above 2^32
        for T in models:
            T.zero_estimators()
        cb, ms = transport_batch(models, [1] * len(models), npk, offs, tuning)
        got = [T.fetch() for T in models]
        for i, T in enumerate(models):
            T.zero_estimators()
            c = T.energy_packet_driver(1, npk[i], offs[i])
            J, E = T.fetch()
            for k in ("n_seg", "n_abs", "n_esc", "n_line", "n_flights", "n_lost", "n_trapped"):
                assert cb[i][k] == c[k], (i, k, cb[i][k], c[k])
            assert np.allclose(got[i][0], J, rtol=1e-12, atol=0)
            assert np.allclose(got[i][1], E, rtol=1e-12, atol=0)
            assert abs(cb[i]["qphot"] - c["qphot"]) <= 1e-12 * max(1.0, abs(c["qphot"]))
        assert cb[1]["n_flights"] == 0 and np.all(got[1][0] == 0)
    finally:
        for T in models:
            T.close()


def test_batch_argument_errors():
    from mocassin_amd import McError, transport_batch
    models = _models(SPECS[:2])
    try:
        with pytest.raises(McError, match="MC_EINVAL"):
            transport_batch(models, [1, 1], [10, -1])
        with pytest.raises(McError, match="MC_EINVAL"):
            transport_batch(models, [1, 2], [10, 10])
        with pytest.raises(McError, match="MC_EINVAL"):
            transport_batch(models, [1, 1], [10, 10], tuning={"waves_per_cu": 0, "emit_min": 8,
                                                             "walk_steps": 4, "chunk": 1024})
        with pytest.raises(McError, match="MC_EINVAL"):
            transport_batch([models[0], models[0]], [1, 1], [10, 10])
    finally:
        for T in models:
            T.close()


def test_batch_with_dust_and_gas_model_equals_separate_runs():
    """A dust and gas model (full grid, silicate grains, Henyey-Greenstein
    scattering, its tables built on the device) between two gas octants in
    one launch: every model equals its own launch."""
    from mocassin_amd import dustgas, transport_batch
    print("batch with a dust and gas model", flush=True)
    gas = _models([SPECS[0], SPECS[2]])
    W = dustgas.open_context(15, corr=2.0, packets=40000)
    models = [gas[0], W["T"], gas[1]]
    try:
        Hd, Te, Ne, ion = W["state"]
        W["T"].opacity_build(Hd, Te, Ne, ion, Tdust=W["Tdust0"], host=False)
        W["T"].emission_build(Hd, Te, Ne, ion, Tdust=W["Tdust0"], host=False)
        W["T"].upload_tables(None, W["p"]["specCDF"], gSca=W["gSca"])
        npk = [SPECS[0][1], 40000, SPECS[2][1]]
        for T in models:
            T.zero_estimators()
        cb, ms = transport_batch(models, [1] * 3, npk, None, None)
        got = [T.fetch() for T in models]
        for i, T in enumerate(models):
            T.zero_estimators()
            c = T.energy_packet_driver(1, npk[i], 0)
            J, E = T.fetch()
            for k in ("n_seg", "n_abs", "n_sca", "n_esc", "n_line", "n_flights", "n_lost", "n_trapped"):
                assert cb[i][k] == c[k], (i, k, cb[i][k], c[k])
            assert np.allclose(got[i][0], J, rtol=1e-12, atol=0)
            assert np.allclose(got[i][1], E, rtol=1e-12, atol=0)
        assert cb[1]["n_sca"] > 0.1 * npk[1]
    finally:
        for T in models:
            T.close()
