"""GPU per-cell balance (FP64) through the C ABI, against the FP64 oracle
and the unmodified reference's updateCell outputs (golden fixtures)."""
import numpy as np
import pytest

from conftest import balance_fixture, gpu_available

pytestmark = pytest.mark.gpu

NAMES = ["c1mini", "c2mini", "c2mini_thick", "dust_bal"]


def _context(nc, nb, nu, permute=False, seed=0):
    """A context whose active map holds cells 1..nc (the balance reads only
    the cell -> position map, not the geometry)."""
    from mocassin_amd.transport import Transport
    n = 2
    while n ** 3 < nc + 1:
        n += 1
    pos = np.arange(n ** 3)
    if permute:
        pos = np.random.default_rng(seed).permutation(n ** 3)
    act = np.zeros(n ** 3, np.int32)
    act[pos[:nc]] = np.arange(1, nc + 1, dtype=np.int32)
    ax = np.linspace(0.0, 1.e17, n).astype(np.float32)
    return Transport(ax, ax, ax, act.reshape(n, n, n, order="F"), nc, nb, nu, R_out=1.e17,
                     geoCorr=(0, 0, 0), ionEdge1=1.0, starPos=(0, 0, 0), starIdx=(1, 1, 1),
                     deltaE=1.0)


def _gpu_run(name, permute=False, J_mult=1.0, jscale=1.0, J_edit=None):
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture(name)
    T = _context(J.shape[0] - 1, b.nbins, b.nuArray, permute=permute)
    try:
        T.balance_setup(b)
        Jd = J.astype(np.float64) * J_mult
        if J_edit is not None:
            J_edit(Jd)
        T.set_estimators(Jd)
        out = T.update_cells(Hden, Te0, Ne0, ion0, jscale=jscale)
    finally:
        T.close()
    return out


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("name", NAMES)
def test_gpu_balance_matches_fp64_oracle(name):
    import mc_oracle as O
    print("balance vs FP64 oracle", name, flush=True)
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture(name)
    Te_o, Ne_o, ion_o, fl_o, nT_o = O.Balance(b).run(J, Hden, Te0, Ne0, ion0, "f64")
    Te, Ne, ion, fl, nT, ms = _gpu_run(name)
    c = slice(1, None)
    assert np.array_equal(fl[c], fl_o[c])
    assert np.array_equal(nT[c], nT_o[c])
    hit = (fl[c] & 16) == 0
    assert hit.sum() > 20
    assert np.max(np.abs(Te[c] - Te_o[c]) / Te_o[c]) < 1e-9
    assert np.max(np.abs(Ne[c] - Ne_o[c]) / Ne_o[c]) < 1e-9
    big = ion_o[c] > 1e-15
    assert np.max(np.abs(ion[c][big] - ion_o[c][big]) / ion_o[c][big]) < 1e-8
    # cells not reached keep their inputs exactly
    nh = ~hit
    assert np.array_equal(Te[c][nh], Te0[c][nh]) and np.array_equal(ion[c][nh], ion0[c][nh].astype(np.float64))


@pytest.mark.parametrize("name", NAMES)
def test_gpu_balance_vs_reference(name):
    """Parity bar against the reference's FP32 updateCell: Te within 1%,
    ionic fractions above 1e-4 within 2%, same convergence/black flags."""
    print("balance vs reference", name, flush=True)
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture(name)
    Te, Ne, ion, fl, nT, ms = _gpu_run(name)
    c = slice(1, None)
    hit = (fl[c] & 16) == 0
    assert np.max(np.abs(Te[c] - ref["Te"][c]) / ref["Te"][c]) < 0.01
    big = ref["ion"][c] > 1e-4
    assert np.max(np.abs(ion[c][big] - ref["ion"][c][big]) / ref["ion"][c][big]) < 0.02
    assert np.array_equal(((fl[c] & 1) != 0)[hit], ref["conv"][c][hit] == 1)
    assert np.array_equal((fl[c] & 2) != 0, ref["black"][c] == 1)


def test_gpu_balance_independent_of_cell_positions():
    print("balance permuted active map", flush=True)
    a = _gpu_run("c2mini")
    p = _gpu_run("c2mini", permute=True)
    for x, y in zip(a[:5], p[:5]):
        assert np.array_equal(x, y)


def test_gpu_balance_jscale():
    """J/2 with jscale 2 is the same J (exact in binary)."""
    print("balance jscale", flush=True)
    a = _gpu_run("c1mini")
    s = _gpu_run("c1mini", J_mult=0.5, jscale=2.0)
    for x, y in zip(a[:5], s[:5]):
        assert np.array_equal(x, y)


def test_gpu_balance_unreached_cells():
    print("balance unreached cells", flush=True)

    def cut(J):
        J[5:40, :] = 0.0
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture("c2mini")
    Te, Ne, ion, fl, nT, ms = _gpu_run("c2mini", J_edit=cut)
    assert np.all(fl[5:40] == 16) and np.all(nT[5:40] == 0)
    assert np.array_equal(Te[5:40], Te0[5:40]) and np.array_equal(Ne[5:40], Ne0[5:40])
    assert np.array_equal(ion[5:40], ion0[5:40].astype(np.float64))


def test_gpu_balance_errors():
    from mocassin_amd._capi import McError
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture("c1mini")
    T = _context(J.shape[0] - 1, b.nbins, b.nuArray)
    try:
        T._nE, T._ns = b.nElemUsed, b.nstages
        with pytest.raises(McError, match="MC_ESTATE"):
            T.update_cells(Hden, Te0, Ne0, ion0)
        T.balance_setup(b)
        T.set_estimators(J.astype(np.float64))
        bad = np.array(Te0, np.float64)
        bad[3] = -1.0
        with pytest.raises(McError, match="MC_EINVAL"):
            T.update_cells(Hden, bad, Ne0, ion0)
        with pytest.raises(McError, match="MC_EINVAL"):
            T.update_cells(Hden, Te0, Ne0, ion0, jscale=0.0)
    finally:
        T.close()
