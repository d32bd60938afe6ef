"""CPU restatement of the warm-start network's arithmetic (mocassin_amd/csrc/
mc_mlp.hip).  TEST INFRASTRUCTURE ONLY: the checker of the GPU forward and
backward kernels; nothing in the product imports it.

The network has no counterpart in the reference.  This file states it twice:
  exact=True   plain FP64 (the mathematical network: the gradient is checked
               against finite differences and a linear closed form);
  exact=False  FP64 arithmetic with the GPU's BF16 rounding points: the
               weights, the standardised inputs, every hidden activation and
               every back-propagated row are rounded to BF16 (nearest even),
               biases and accumulations stay wide.
Parameter layout: per layer W[out][in] then b[out]; layer sizes
(96->128), (128->128), (128->128), (128->32); 72 of the 96 inputs and 31 of
the 32 outputs are used.
"""
import numpy as np

N_IN, N_IN_P, N_HID, N_OUT, N_OUT_P = 72, 96, 128, 31, 32
LAYERS = [(N_IN_P, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT_P)]


def offsets():
    out, o = [], 0
    for k, n in LAYERS:
        out.append((o, o + n * k))
        o += n * k + n
    return out, o


OFFS, N_PARAMS = offsets()


def bf16(x):
    """Round to BF16, nearest even, returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p, exact):
    p = np.asarray(p, np.float32)
    Ws, bs = [], []
    for (k, n), (ow, ob) in zip(LAYERS, OFFS):
        W = p[ow:ow + n * k].reshape(n, k).astype(np.float64)
        Ws.append(W if exact else bf16(W))
        bs.append(p[ob:ob + n].astype(np.float64))
    return Ws, bs


def standardise(X, scale):
    """(X - mean) / std in FP32 as the device computes it, zero-padded to 96."""
    s = np.asarray(scale, np.float32)
    Xs = (np.asarray(X, np.float32) - s[:N_IN]) / s[N_IN:2 * N_IN]
    out = np.zeros((Xs.shape[0], N_IN_P), np.float32)
    out[:, :N_IN] = Xs
    return out


def scale_targets(T, scale):
    s = np.asarray(scale, np.float32)
    return (np.asarray(T, np.float32) - s[2 * N_IN:2 * N_IN + N_OUT]) / s[2 * N_IN + N_OUT:]


def forward(p, X0, exact=False):
    """X0 [n][96] standardised inputs -> (Y [n][32], activations list)."""
    Ws, bs = unpack(p, exact)
    A = np.asarray(X0, np.float64)
    if not exact:
        A = bf16(A)
    acts = [A]
    for l in range(4):
        Z = A @ Ws[l].T + bs[l]
        if l < 3:
            A = np.maximum(Z, 0.0)
            if not exact:
                A = bf16(A)
            acts.append(A)
        else:
            A = Z
    return A, acts


def loss_grad(p, X0, T, M, exact=False):
    """L = sum M (y - t)^2 / sum M over the 31 outputs, and dL/dp (FP64)."""
    Ws, bs = unpack(p, exact)
    Y, acts = forward(p, X0, exact)
    T = np.asarray(T, np.float64)
    M = np.asarray(M, np.float64)
    S = M.sum()
    d = Y[:, :N_OUT] - T
    L = float((M * d * d).sum() / S)
    g = np.zeros(N_PARAMS)
    dY = np.zeros((Y.shape[0], N_OUT_P))
    dY[:, :N_OUT] = 2.0 * M * d
    if not exact:
        dY = bf16(dY)
    for l in range(3, -1, -1):
        (k, n), (ow, ob) = LAYERS[l], OFFS[l]
        g[ow:ow + n * k] = (dY.T @ acts[l]).ravel()
        g[ob:ob + n] = dY.sum(axis=0)
        if l == 0:
            break
        dA = (dY @ Ws[l]) * (acts[l] > 0)
        dY = dA if exact else bf16(dA)
    return L, g / S
