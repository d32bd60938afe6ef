"""Host side of the balance's warm-start network: the data it is trained on,
its initialisation and scaling, and the training loop (all arithmetic runs
on the device through the C ABI: mc_mlp_features, mc_mlp_train,
mc_mlp_predict, and mc_balance_run with warm_start).

Training pairs come from the method's own solves: for every reached cell,
the features of its J (64 band sums, nH, abundances) and the balance's
converged log10 Te and log10 ionic fractions (floored at 1e-8) of H, He, C,
N, O, Ne, S (stages 1..min(Z+1, 5)).  Outputs of elements or stages a model
lacks carry zero weight in the loss.
"""
import numpy as np

N_IN, N_OUT, N_HID = 72, 31, 128
SLOTS_Z = (1, 2, 6, 7, 8, 10, 16)
LAYERS = [(96, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, 32)]
FRACTION_FLOOR = 1e-8


def n_params():
    return sum(n * k + n for k, n in LAYERS)


def slot_layout():
    """[(Z, stage, output column)] of the 30 ionic-fraction outputs."""
    out, col = [], 1
    for z in SLOTS_Z:
        for st in range(1, min(z + 1, 5) + 1):
            out.append((z, st, col))
            col += 1
    return out


def targets(bundle, Te, ion, cells):
    """Targets T [n][31] and mask M [n][31] for `cells` (1-based ids) from a
    balance state Te (nCells+1,), ion (nCells+1, nElemUsed, nstages).  The
    loss is sum M (y - t)^2 / sum M, so scaling a column of M weighs that
    output."""
    cells = np.asarray(cells)
    T = np.zeros((cells.size, N_OUT), np.float32)
    M = np.zeros((cells.size, N_OUT), np.float32)
    T[:, 0] = np.log10(Te[cells])
    M[:, 0] = 1.0
    for z, st, col in slot_layout():
        if not bundle.elem_on[z - 1] or st > min(z + 1, bundle.nstages):
            continue
        k = int(bundle.xref[z - 1]) - 1
        T[:, col] = np.log10(np.clip(ion[cells, k, st - 1], FRACTION_FLOOR, 1.0))
        M[:, col] = 1.0
    return T, M


def init_params(seed):
    """He-normal weights, zero biases; the padded input columns and output
    row stay zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(n_params(), np.float32)
    o = 0
    for l, (k, n) in enumerate(LAYERS):
        kin = N_IN if l == 0 else k
        W = rng.normal(0.0, np.sqrt(2.0 / kin), (n, k)).astype(np.float32)
        if l == 0:
            W[:, N_IN:] = 0
        if l == 3:
            W[N_OUT:, :] = 0
            W *= 0.5
        p[o:o + n * k] = W.ravel()
        o += n * k + n
    return p


def fit_scale(X, T, M):
    """Per-column mean and std of the features and of the weighted targets
    (std 1 where a column is constant)."""
    X = np.asarray(X, np.float64)
    xm, xs = X.mean(axis=0), X.std(axis=0)
    xs[xs < 1e-6] = 1.0
    T = np.asarray(T, np.float64)
    M = np.asarray(M, np.float64)
    w = M.sum(axis=0)
    ym = np.where(w > 0, (T * M).sum(axis=0) / np.maximum(w, 1), 0.0)
    yv = np.where(w > 0, (M * (T - ym) ** 2).sum(axis=0) / np.maximum(w, 1), 1.0)
    ys = np.sqrt(yv)
    ys[ys < 1e-6] = 1.0
    return np.concatenate([xm, xs, ym, ys]).astype(np.float32)


def train(ctx, X, T, M, steps=2000, batch=1024, lr=2e-3, seed=0, init_seed=1, params=None,
          decay=(1.0,)):
    """Fit the network on (X, T, M) with Adam on the device: `steps` steps
    in equal phases at lr times each factor of `decay`.  ctx: a
    mocassin_amd.Transport.  Returns (loss per step, device ms)."""
    X = np.asarray(X, np.float32)
    scale = fit_scale(X, T, M)
    ctx.set_mlp(init_params(init_seed) if params is None else params, scale)
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    order = np.concatenate([rng.permutation(n) for _ in range((steps * batch + n - 1) // n)])
    order = order[:steps * batch].astype(np.int32)
    losses, ms = [], 0.0
    per = steps // len(decay)
    for i, f in enumerate(decay):
        s0 = i * per
        s1 = steps if i == len(decay) - 1 else s0 + per
        l, m = ctx.mlp_train(X, T, M, order[s0 * batch:s1 * batch], batch, lr * f)
        losses.append(l)
        ms += m
    return np.concatenate(losses), ms
