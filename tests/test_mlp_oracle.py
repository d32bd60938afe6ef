"""The warm-start network's CPU restatement (oracle/mlp_oracle.py) pinned by
mathematics: finite differences of its FP64 loss, a linear closed form, and
BF16 rounding on known values.  No GPU."""
import numpy as np

import mlp_oracle as MO
from mocassin_amd import mlp as MP


def _problem(seed, n=6):
    rng = np.random.default_rng(seed)
    p = MP.init_params(seed)
    for (ow, ob), (k, nn) in zip(MO.OFFS, MO.LAYERS):
        p[ob:ob + nn] = rng.normal(0, 0.1, nn).astype(np.float32)
    p[MO.OFFS[3][1] + 31] = 0.0
    X0 = np.zeros((n, MO.N_IN_P))
    X0[:, :MO.N_IN] = rng.normal(0, 1, (n, MO.N_IN))
    T = rng.normal(0, 1, (n, MO.N_OUT))
    M = (rng.random((n, MO.N_OUT)) > 0.2).astype(np.float64)
    return p.astype(np.float64), X0, T, M


def test_layout_matches_the_product():
    assert MO.N_PARAMS == MP.n_params() == 49568
    assert [(k, n) for k, n in MO.LAYERS] == MP.LAYERS


def test_exact_gradient_matches_finite_differences():
    p, X0, T, M = _problem(3)
    L, g = MO.loss_grad(p, X0, T, M, exact=True)
    rng = np.random.default_rng(0)
    idx = []
    for (ow, ob), (k, n) in zip(MO.OFFS, MO.LAYERS):
        w = rng.integers(0, n * k, 12)
        idx += [ow + int(i) for i in w if p[ow + int(i)] != 0 or (int(i) % k) < MO.N_IN]
        idx += [ob + int(i) for i in rng.integers(0, min(n, 31), 4)]
    h = 1e-6
    worst = 0.0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        # unpack() rounds to float32: evaluate in float64 by bypassing it
        Lp = _loss64(pp, X0, T, M)
        Lm = _loss64(pm, X0, T, M)
        fd = (Lp - Lm) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(1e-8, abs(g[i]) + abs(fd)))
    assert worst < 1e-5


def _loss64(p, X0, T, M):
    """FP64 loss with FP64 parameters (no float32 cast)."""
    A = np.asarray(X0, np.float64)
    for l, ((k, n), (ow, ob)) in enumerate(zip(MO.LAYERS, MO.OFFS)):
        W = p[ow:ow + n * k].reshape(n, k)
        Z = A @ W.T + p[ob:ob + n]
        A = np.maximum(Z, 0) if l < 3 else Z
    d = A[:, :MO.N_OUT] - T
    return float((M * d * d).sum() / M.sum())


def test_linear_regime_closed_form():
    """Non-negative weights, inputs and biases keep every ReLU open: the
    network is the composed affine map W3 W2 W1 W0 x + (...)."""
    rng = np.random.default_rng(5)
    p = np.zeros(MO.N_PARAMS)
    Ws, bs = [], []
    for (ow, ob), (k, n) in zip(MO.OFFS, MO.LAYERS):
        W = rng.random((n, k)) * 0.05
        b = rng.random(n) * 0.1
        p[ow:ow + n * k] = W.ravel()
        p[ob:ob + n] = b
        Ws.append(W)
        bs.append(b)
    X0 = rng.random((4, MO.N_IN_P))
    Y, _ = MO.forward(p.astype(np.float32), X0.astype(np.float32), exact=True)
    W32 = [W.astype(np.float32).astype(np.float64) for W in Ws]
    b32 = [b.astype(np.float32).astype(np.float64) for b in bs]
    Wt = W32[3] @ W32[2] @ W32[1] @ W32[0]
    bt = W32[3] @ (W32[2] @ (W32[1] @ b32[0] + b32[1]) + b32[2]) + b32[3]
    ref = X0.astype(np.float32).astype(np.float64) @ Wt.T + bt
    assert np.allclose(Y, ref, rtol=1e-12, atol=1e-12)


def test_bf16_rounding_known_values():
    one = 1.0
    cases = {one: one, one + 2 ** -8: one, one + 3 * 2 ** -8: one + 2 ** -6,
             one + 2 ** -8 + 2 ** -20: one + 2 ** -7, -2.5: -2.5, 0.0: 0.0,
             3.0e38: float(np.float32(2.9938e38))}
    for x, want in cases.items():
        got = MO.bf16(np.array([x]))[0]
        assert got == np.float64(np.float32(want)) or abs(got - want) <= abs(want) * 2 ** -8, (x, got, want)
    assert MO.bf16(np.array([one + 2 ** -8]))[0] == one
    assert MO.bf16(np.array([one + 3 * 2 ** -8]))[0] == one + 2 ** -6


def test_rounded_gradient_close_to_exact():
    p, X0, T, M = _problem(11, n=64)
    Le, ge = MO.loss_grad(p, X0, T, M, exact=True)
    Lr, gr = MO.loss_grad(p, X0, T, M, exact=False)
    assert abs(Lr - Le) / Le < 2e-2
    assert np.linalg.norm(gr - ge) / np.linalg.norm(ge) < 5e-2
