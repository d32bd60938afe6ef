"""The iterated state: the device's iteration (mocassin_amd/iterate.py:
tables of the state, sweep, balance from the previous state, as iterateMC,
iteration_mod.f90:81-1186) from the reference's cold start against the
unmodified reference's MCIterationDriver run to its own stopping point on
the same deck (tests/golden/converge_*.npz: four seeds of the reference,
eight MPI ranks each).

Both sides are Monte Carlo estimates; each side's state is the mean over its
four seeded runs.  In cells with at least 1e3 segments walked in a sweep:
converged Te within 1 %, ionic fractions within 2 % (every fraction of at
least 0.01; smaller fractions are reported), and the forbidden-line
luminosities of outputGas relative to Hbeta (lines above 1e-3 Hbeta) within
2 %."""
import numpy as np
import pytest

from conftest import golden, gpu_available, grid_from_fixture

pytestmark = pytest.mark.gpu

SEEDS = 4


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def _open(g):
    from mocassin_amd.atomic import AtomicBundle, emission_tables
    from mocassin_amd.transport import Transport
    axes, active, nc, nb, fl, star_idx = grid_from_fixture(g)
    sc = g["scalars"]
    b = AtomicBundle.from_npz(g)
    T = Transport(*axes, active, nc, nb, g["nuArray"], R_out=sc[0], geoCorr=sc[1:4], ionEdge1=sc[4],
                  starPos=sc[6:9], starIdx=star_idx, deltaE=sc[5], symmetric=fl["symmetric"], gas=True,
                  seed=4242, device=0)
    T.balance_setup(b)
    T.opacity_setup(g["op_seg"], g["op_sigma"], float(g["op_brems1"][0]))
    T.emission_setup(**emission_tables(g))
    shp = lambda a: a.reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0).astype(np.float64)
    state = (g["Hden"].astype(np.float64), g["Te0"].astype(np.float64), g["Ne0"].astype(np.float64),
             shp(g["ionDen0"]))
    return T, b, state, fl["symmetric"], shp


def _deck_packets(name):
    return {"c1": 10000000, "c2mini": 2000000}[name]


@pytest.mark.parametrize("name", ["c1", "c2mini"])
def test_iterated_state_matches_reference(name):
    from mocassin_amd.iterate import iterate
    print("converged state %s" % name, flush=True)
    g = golden("converge_" + name)
    T, b, state, sym, shp = _open(g)
    nc = T.nCells
    npk = _deck_packets(name)
    iters = int(g["iterations"])
    Te_d, ion_d, cross, hb_d, forb_d, n_done = [], [], [], [], [], []
    try:
        for s in range(SEEDS):
            st, flags, recs, cr = iterate(T, state, npk, g["specCDF"].reshape(-1, 1), sym, max_iter=iters,
                                          min_convergence=95.0, packet_offset=(s + 1) << 36, crossings=True)
            n_done.append(len(recs))
            Te_d.append(st[1])
            ion_d.append(st[3])
            cross.append(cr)
            if b.ncel > 0:
                forb, hb, _ = T.line_luminosities()          # the device state this run left
                forb_d.append(forb)
                hb_d.append(hb)
        # the reference's converged states, through the same line sums
        ref_forb, ref_hb = [], []
        if b.ncel > 0:
            for k in range(len(g["seeds"])):
                f, h, _ = T.line_luminosities(g["Hden"], g["ref_Te"][k], g["ref_Ne"][k], shp(g["ref_ionDen"][k]),
                                              lgBlack=g["ref_black"][k])
                ref_forb.append(f)
                ref_hb.append(h)
    finally:
        T.close()
    ref_Te = g["ref_Te"].astype(np.float64).mean(axis=0)
    ref_ion = np.mean([shp(a) for a in g["ref_ionDen"]], axis=0)
    dev_Te = np.mean(Te_d, axis=0)
    dev_ion = np.mean(ion_d, axis=0)
    well = np.zeros(nc + 1, bool)
    well[1:] = np.mean(cross, axis=0)[1:] >= 1000
    assert well.sum() > 0.5 * nc
    sd_ref = g["ref_Te"].astype(np.float64).std(axis=0)[well] / ref_Te[well]
    sd_dev = np.std(Te_d, axis=0)[well] / dev_Te[well]
    dT = np.abs(dev_Te[well] / ref_Te[well] - 1)
    print("%s: %d iterations (reference %d), %d well-sampled cells; Te: max rel diff %.4f, median %.5f; "
          "seed-to-seed sd reference %.4f, device %.4f (medians)"
          % (name, max(n_done), iters, well.sum(), dT.max(), np.median(dT), np.median(sd_ref), np.median(sd_dev)))
    assert dT.max() <= 0.01
    big = (ref_ion[well] >= 0.01)
    dI = np.abs(dev_ion[well][big] / ref_ion[well][big] - 1)
    small = (ref_ion[well] > 1e-6) & ~big
    dS = np.abs(dev_ion[well][small] / ref_ion[well][small] - 1)
    print("%s: ionic fractions >= 0.01: max rel diff %.4f (%d values); 1e-6..0.01: median %.4f, p95 %.4f"
          % (name, dI.max(), big.sum(), np.median(dS) if dS.size else 0, np.percentile(dS, 95) if dS.size else 0))
    assert dI.max() <= 0.02
    if b.ncel > 0:
        rf = np.mean([f / h for f, h in zip(ref_forb, ref_hb)], axis=0)
        df = np.mean([f / h for f, h in zip(forb_d, hb_d)], axis=0)
        lines = rf > 1e-3
        dL = np.abs(df[lines] / rf[lines] - 1)
        print("%s: %d forbidden lines above 1e-3 Hbeta: ratio to Hbeta max rel diff %.4f, median %.5f"
              % (name, lines.sum(), dL.max(), np.median(dL)))
        assert lines.sum() > 20 and dL.max() <= 0.02
