"""GPU warm-start network: BF16-MFMA forward and backward kernels against
the CPU restatement with the same rounding points (oracle/mlp_oracle.py),
training on the device, and what a trained network gains the balance."""
import numpy as np
import pytest

import mlp_oracle as MO
from conftest import balance_fixture, gpu_available
from mocassin_amd import mlp as MP

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.skip("no GPU")


def _ctx(nc=8, nb=64, nu=None):
    from mocassin_amd.transport import Transport
    n = 2
    while n ** 3 < nc + 1:
        n += 1
    act = np.zeros(n ** 3, np.int32)
    act[:nc] = np.arange(1, nc + 1)
    ax = np.linspace(0.0, 1.e17, n).astype(np.float32)
    if nu is None:
        nu = np.geomspace(1e-5, 25, nb).astype(np.float32)
    return Transport(ax, ax, ax, act.reshape(n, n, n, order="F"), nc, len(nu), nu, R_out=1.e17,
                     geoCorr=(0, 0, 0), ionEdge1=1.0, starPos=(0, 0, 0), starIdx=(1, 1, 1), deltaE=1.0)


def _random_net(seed):
    rng = np.random.default_rng(seed)
    p = MP.init_params(seed)
    o = 0
    for k, n in MP.LAYERS:
        p[o + n * k:o + n * k + n] = rng.normal(0, 0.1, n)
        o += n * k + n
    p[MO.OFFS[3][1] + 31] = 0.0
    scale = np.concatenate([rng.normal(0, 1, 72), rng.uniform(0.5, 2, 72), rng.normal(0, 1, 31),
                            rng.uniform(0.5, 2, 31)]).astype(np.float32)
    return p, scale


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1.5, (n, 72)).astype(np.float32)
    T = rng.normal(0, 1, (n, 31)).astype(np.float32)
    M = (rng.random((n, 31)) > 0.15).astype(np.float32)
    return X, T, M


def test_forward_matches_oracle():
    print("mlp forward vs oracle", flush=True)
    p, scale = _random_net(1)
    X, _, _ = _data(2, 1000)
    T = _ctx()
    try:
        T.set_mlp(p, scale)
        Yg = T.mlp_predict(X)
    finally:
        T.close()
    Ys, _ = MO.forward(p, MO.standardise(X, scale), exact=False)
    Yo = Ys[:, :31] * scale[175:206] + scale[144:175]
    Ye_s, _ = MO.forward(p, MO.standardise(X, scale), exact=True)
    Ye = Ye_s[:, :31] * scale[175:206] + scale[144:175]
    err = np.abs(Yg - Yo) / (np.abs(Yo) + 1)
    assert np.max(err) < 2e-2 and np.mean(err) < 1e-3
    # rounding to BF16 costs about 1 % against the exact network
    assert np.mean(np.abs(Yg - Ye)) / np.mean(np.abs(Ye)) < 3e-2


def test_gradient_matches_oracle():
    print("mlp gradient vs oracle", flush=True)
    p, scale = _random_net(3)
    X, Tg, M = _data(4, 2048)
    T = _ctx()
    try:
        T.set_mlp(p, scale)
        Lg, gg = T.mlp_loss_grad(X, Tg, M)
    finally:
        T.close()
    X0 = MO.standardise(X, scale)
    Ts = MO.scale_targets(Tg, scale)
    Lo, go = MO.loss_grad(p, X0, Ts, M, exact=False)
    Le, ge = MO.loss_grad(p, X0, Ts, M, exact=True)
    assert abs(Lg - Lo) / Lo < 1e-3
    for (ow, ob), (k, n) in zip(MO.OFFS, MO.LAYERS):
        for a, b in ((ow, ow + n * k), (ob, ob + n)):
            rel = np.linalg.norm(gg[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
            assert rel < 1e-2, (a, rel)
    assert np.linalg.norm(gg - ge) / np.linalg.norm(ge) < 5e-2
    # padded parameters receive no gradient
    assert np.all(gg[MO.OFFS[0][0]:MO.OFFS[0][1]].reshape(128, 96)[:, 72:] == 0)
    assert np.all(gg[MO.OFFS[3][0]:MO.OFFS[3][1]].reshape(32, 128)[31] == 0)


def test_training_fits_a_smooth_map():
    print("mlp training", flush=True)
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (20000, 72)).astype(np.float32)
    A = rng.normal(0, 0.3, (72, 31))
    Tg = np.tanh(X @ A).astype(np.float32) + 0.2 * (X[:, :31] ** 2)
    M = np.ones_like(Tg)
    T = _ctx()
    try:
        loss, ms = MP.train(T, X, Tg, M, steps=1500, batch=512, lr=2e-3, seed=1)
        Yp = T.mlp_predict(X[:2000])
    finally:
        T.close()
    assert loss[:20].mean() > 5 * loss[-50:].mean()
    rel = np.mean((Yp - Tg[:2000]) ** 2) / np.mean((Tg[:2000] - Tg[:2000].mean(0)) ** 2)
    assert rel < 0.1
    print("training: loss %.3g -> %.3g, %.1f ms device" % (loss[:20].mean(), loss[-50:].mean(), ms))


def test_features_match_their_definition():
    """Network inputs of each cell from the device J: log10 of J summed over
    64 contiguous groups of bins, log10 nH, log10 abundances (-30 absent)."""
    print("mlp features", flush=True)
    T, b, J, Hden, Te0, Ne0, ion0 = _balance_ctx("c2mini")
    try:
        X, hit = T.mlp_features(Hden, jscale=2.0)
    finally:
        T.close()
    nb = b.nbins
    Jd = J.astype(np.float64) * 2.0
    for c in range(1, J.shape[0]):
        want = [np.log10(max(Jd[c, i * nb // 64:(i + 1) * nb // 64].sum(), 1e-30)) for i in range(64)]
        assert np.allclose(X[c - 1, :64], want, rtol=0, atol=1e-5)
        assert hit[c] == int(np.any(Jd[c] > 0))
        assert abs(X[c - 1, 64] - np.log10(Hden[c])) < 1e-5
    for k, z in enumerate(MP.SLOTS_Z):
        want = np.log10(b.abun[z - 1]) if b.elem_on[z - 1] else -30.0
        assert np.allclose(X[:, 65 + k], want, atol=1e-5)


def _balance_ctx(name):
    from test_gpu_balance import _context
    b, J, Hden, Te0, Ne0, ion0, ref = balance_fixture(name)
    T = _context(J.shape[0] - 1, b.nbins, b.nuArray)
    T.balance_setup(b)
    T.set_estimators(J.astype(np.float64))
    return T, b, J, Hden, Te0, Ne0, ion0


def test_warm_start_gains_over_cold_start():
    """The warm start's mechanism and its safety net.  A network trained on
    the reached cells of two decks warm-starts their secant balance: far
    fewer temperature iterations than the cold start, the same temperatures
    within the solver's tolerance, and almost no restarts.  Trained on half
    of the cells only (a few hundred: the network then predicts the other
    half to several per cent), the held-out cells still gain and reach the
    same temperatures."""
    print("warm start gain", flush=True)
    Xs, Ts, Ms, ctxs = [], [], [], []
    ref = {}
    for name in ("c2mini", "c2mini_thick"):
        T, b, J, Hden, Te0, Ne0, ion0 = _balance_ctx(name)
        ctxs.append((T, b, Hden, Te0, Ne0, ion0))
        Te_r, Ne_r, ion_r, fl_r, nI_r, _ = T.update_cells(Hden, Te0, Ne0, ion0, solver="secant", tol=1e-5,
                                                          max_iter=100, ion_tol=1e-4)
        X, hit = T.mlp_features(Hden)
        cells = np.nonzero(hit[1:] & ((fl_r[1:] & 4) == 0))[0] + 1
        Tt, Mt = MP.targets(b, Te_r, ion_r, cells)
        Xs.append(X[cells - 1])
        Ts.append(Tt)
        Ms.append(Mt)
        ref[name] = (Te_r, cells)
    X, Tg, M = np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms)
    M[:, 0] *= 5.0                      # log10 Te weighed above each fraction
    half = np.random.default_rng(0).random(len(X)) < 0.5
    kw = dict(solver="secant", tol=1e-3, max_iter=60, ion_tol=1e-3)

    def judge(params, scale, pick):
        off = 0
        cold_it, warm_it, worst, n_restart = [], [], 0.0, 0
        for (T, b, Hden, Te0, Ne0, ion0), name in zip(ctxs, ("c2mini", "c2mini_thick")):
            Te_r, cells = ref[name]
            sel = cells[pick[off:off + len(cells)]]
            off += len(cells)
            T.set_mlp(params, scale)
            Tc, _, _, flc, nc_, _ = T.update_cells(Hden, Te0, Ne0, ion0, **kw)
            Tw, _, _, flw, nw_, _ = T.update_cells(Hden, Te0, Ne0, ion0, warm_start=True, **kw)
            cold_it += list(nc_[sel])
            warm_it += list(nw_[sel])
            worst = max(worst, np.max(np.abs(Tw[sel] - Te_r[sel]) / Te_r[sel]))
            n_restart += int(np.count_nonzero(flw[sel] & 32))
            assert np.all((flw[sel] & 4) == 0)
        return np.mean(cold_it), np.mean(warm_it), worst, n_restart, len(cold_it)

    try:
        T0 = ctxs[0][0]
        MP.train(T0, X, Tg, M, steps=3000, batch=256, lr=2e-3, seed=2)
        c, w, worst, nr, n = judge(*T0.get_mlp(), np.ones(len(X), bool))
        print("all cells: cold %.2f, warm %.2f iterations per cell; worst Te error %.2e; %d of %d restarted"
              % (c, w, worst, nr, n))
        assert w < 0.5 * c and worst < 0.01 and nr <= max(1, 0.02 * n)
        MP.train(T0, X[half], Tg[half], M[half], steps=3000, batch=256, lr=2e-3, seed=2)
        c, w, worst, nr, n = judge(*T0.get_mlp(), ~half)
        print("held-out cells: cold %.2f, warm %.2f iterations per cell; worst Te error %.2e; %d of %d restarted"
              % (c, w, worst, nr, n))
        assert w < 0.7 * c and worst < 0.01 and nr <= max(1, 0.02 * n)
    finally:
        for c_ in ctxs:
            c_[0].close()
