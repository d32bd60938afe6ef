"""Batched serving of many independent nebulae in one persistent launch.

The workload: 64 octant models (seeded) with edge drawn uniformly from
20-80 cells, a blackbody source of Teff uniform in 30,000-60,000 K, uniform
nH log-uniform in 1e1-1e4 cm^-3, H+He gas, the 400-bin mesh of the H+He
atomic data, and log-uniform 1e6-1e8 packets per model.  `measure` times
one batched launch of all models (mc_transport_batch, launch shape chosen by
the autotuner on a tenth of the packets) against the same models' separate
launches back to back, and checks that both give the same counters.

Then the balance of every model on its sweep's J, one mc_balance_run per
model, from the reference's cold start: the reference's schedule for all
models, and the secant solve at the reference's tolerance plain and
warm-started on the 48 models the network was not trained on (it is trained
on the device on the balances of 16 models spread over the nH range).  The balance is not
batched across models: it is one wavefront per cell, so each model's launch
already holds thousands of independent wavefronts, and a batched launch
could save only the launch gaps (host wall minus device time, reported).
"""
import os
import time

import numpy as np

from . import autotune as AT
from . import synth
from .transport import Transport, transport_batch


def model_specs(n=64, seed=2024):
    rng = np.random.default_rng(seed)
    edge = rng.integers(20, 81, n)
    teff = rng.uniform(3.e4, 6.e4, n)
    nh = 10.0 ** rng.uniform(1.0, 4.0, n)
    npk = np.round(10.0 ** rng.uniform(6.0, 8.0, n)).astype(np.int64)
    return [dict(edge=int(e), teff=float(t), nh=float(h), packets=int(p)) for e, t, h, p in zip(edge, teff, nh, npk)]


def build(specs, nu, log=print):
    """One Transport context per model (grid 2.5e18 cm octant, Q(H) = 1e49 s^-1)."""
    models = []
    t = time.perf_counter()
    for i, s in enumerate(specs):
        p = synth.build_problem(s["edge"], 2.5e18, 2.5e18, s["nh"], 1.e49, s["teff"], len(nu), synth.H_HE,
                                n_packets=s["packets"], nu=nu)
        T = Transport(p["xAxis"], p["yAxis"], p["zAxis"], p["active"], p["nCells"], p["nbins"], p["nuArray"],
                      p["R_out"], p["geoCorr"], p["ionEdge1"], [p["starPos"]], [p["starIdx"]], [p["deltaE"]],
                      symmetric=True, seed=5000 + i, device=0)
        T.upload_tables(p["opacity"], p["specCDF"], p["recPDF"], p["totalLines"])
        models.append(T)
    log("nebulae: %d models built and uploaded in %.1f s" % (len(models), time.perf_counter() - t))
    return models


def _balances(models, specs, here, n_train, log, steps=6000, decay=(1.0,), te_weight=5.0):
    """Per-model balance of the J each model's last sweep left, from the
    reference's cold start: the reference's schedule for every model; then,
    on the models the network was not trained on, the secant solve at the
    reference's tolerance (1e-2) plain and warm-started.  The network is
    trained on the reference-schedule balances of n_train models spread
    over the nH range."""
    from . import mlp as MP
    from .atomic import AtomicBundle
    from .converge import cold_state
    from .iterate import jscale_of
    atom = dict(np.load(os.path.join(here, "..", "tests", "golden", "atomic_c1.npz")))
    bundle = AtomicBundle.from_npz(atom)
    js = jscale_of(True)
    colds, plain = [], []
    t = time.perf_counter()
    for m, s in zip(models, specs):
        m.balance_setup(bundle)
        hden = np.full(m.nCells + 1, s["nh"])
        hden[0] = 0.0
        Te0, Ne0, ion0, Hd = cold_state(bundle, atom["Te0"][0], atom["ion0"], hden)
        colds.append((Hd, Te0, Ne0, ion0))
        plain.append(m.update_cells(Hd, Te0, Ne0, ion0, jscale=js))
    wall_p = time.perf_counter() - t
    by_nh = np.argsort([s["nh"] for s in specs])
    train = sorted(by_nh[np.linspace(0, len(models) - 1, n_train).astype(int)].tolist())
    held = [i for i in range(len(models)) if i not in train]
    Xs, Ts, Ms = [], [], []
    for i in train:
        r = plain[i]
        X, hit = models[i].mlp_features(colds[i][0], jscale=js)
        cells = np.nonzero(hit[1:] & ((r[3][1:] & (4 | 8)) == 0))[0] + 1
        Tg, M = MP.targets(bundle, r[0], r[2], cells)
        M[:, 0] *= te_weight
        Xs.append(X[cells - 1])
        Ts.append(Tg)
        Ms.append(M)
    X, Tg, M = np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms)
    loss, train_ms = MP.train(models[train[0]], X, Tg, M, steps=steps, batch=4096, lr=2e-3, decay=decay)
    params, scale = models[train[0]].get_mlp()
    sec = dict(solver="secant", tol=1e-2, max_iter=60, ion_tol=0.01)
    runs = {"plain": {}, "warm": {}}
    acc = []
    for i in held:
        m, st = models[i], colds[i]
        m.set_mlp(params, scale)
        Xh, hit = m.mlp_features(st[0], jscale=js)
        c = np.nonzero(hit[1:])[0] + 1
        acc.append(np.abs(10.0 ** m.mlp_predict(Xh[c - 1])[:, 0].astype(np.float64) / plain[i][0][c] - 1))
        runs["plain"][i] = m.update_cells(*st, jscale=js, **sec)
        runs["warm"][i] = m.update_cells(*st, jscale=js, warm_start=True, **sec)
    acc = np.concatenate(acc)
    res = {}
    for k, rr in runs.items():
        its = np.concatenate([rr[i][4][1:][(rr[i][3][1:] & 16) == 0] for i in held])
        res[k] = {"device_ms": float(sum(rr[i][5] for i in held)), "mean_T_iterations": float(its.mean()),
                  "restarted": float(np.mean(np.concatenate([(rr[i][3][1:] & 32) != 0 for i in held])))}
    rel = []
    for i in held:
        a, w = runs["plain"][i], runs["warm"][i]
        ok = ((a[3][1:] & (4 | 16)) == 0) & ((w[3][1:] & (4 | 16)) == 0)
        rel.append(np.abs(w[0][1:][ok] / a[0][1:][ok] - 1))
    rel = np.concatenate(rel)
    out = {"reference_schedule": {"models": len(models), "device_ms": float(sum(r[5] for r in plain)),
                                  "host_wall_ms": wall_p * 1e3,
                                  "mean_T_iterations": float(np.concatenate(
                                      [r[4][1:][(r[3][1:] & 16) == 0] for r in plain]).mean())},
           "training": {"models": train, "pairs": int(X.shape[0]), "device_ms": train_ms, "steps": steps,
                        "lr_decay": list(decay), "te_weight": te_weight,
                        "loss_first": float(loss[:20].mean()), "loss_last": float(loss[-50:].mean())},
           "held_out_models": len(held),
           "prediction_Te_within_1pct": float(np.mean(acc <= 0.01)),
           "prediction_Te_within_5pct": float(np.mean(acc <= 0.05)),
           "secant_1e-2": res, "speedup_device": res["plain"]["device_ms"] / res["warm"]["device_ms"],
           "warm_vs_plain_Te_within_1pct": float(np.mean(rel <= 0.01))}
    log("nebulae balance: %d models, reference schedule %.1f ms device (%.1f ms host wall); held-out %d, secant "
        "1e-2: plain %.1f ms (%.2f T-iterations), warm %.1f ms (%.2f), %.2fx; prediction within 1%% / 5%%: "
        "%.2f / %.2f" % (len(models), out["reference_schedule"]["device_ms"], wall_p * 1e3, len(held),
                         res["plain"]["device_ms"], res["plain"]["mean_T_iterations"], res["warm"]["device_ms"],
                         res["warm"]["mean_T_iterations"], out["speedup_device"], out["prediction_Te_within_1pct"],
                         out["prediction_Te_within_5pct"]))
    return out


def measure(n=64, balance=True, train_models=16, train_steps=6000, decay=(1.0,), te_weight=5.0, log=print):
    here = os.path.dirname(os.path.abspath(__file__))
    nu = np.load(os.path.join(here, "..", "tests", "golden", "atomic_c1.npz"))["at_nuArray"]
    specs = model_specs(n)
    npk = [s["packets"] for s in specs]
    models = build(specs, nu, log)
    try:
        best, trials = AT.autotune_batch(models, [max(1, k // 10) for k in npk])
        log("nebulae: batch launch shape %s (%d candidates, %.1f..%.1f ms)"
            % (best, len(trials), min(t[1] for t in trials), max(t[1] for t in trials)))
        for m in models:
            m.zero_estimators()
        t = time.perf_counter()
        cb, ms_b = transport_batch(models, [1] * n, npk, None, best)
        wall_b = time.perf_counter() - t
        seq_ms, seq = 0.0, []
        t = time.perf_counter()
        for m, k in zip(models, npk):
            m.zero_estimators()
            c = m.energy_packet_driver(1, k, 0)
            seq.append(c)
            seq_ms += c["kernel_ms"]
        wall_s = time.perf_counter() - t
        same = all(a[k] == b[k] for a, b in zip(cb, seq) for k in ("n_seg", "n_abs", "n_esc", "n_flights"))
        tot = float(sum(npk))
        nseg = float(sum(c["n_seg"] for c in cb))
        out = {"models": n, "packets": tot, "cells": int(sum(m.nCells for m in models)),
               "segments": nseg, "segments_per_s_batched": nseg / (ms_b * 1e-3),
               "edges": [s["edge"] for s in specs], "tuning": best,
               "batched": {"device_ms": ms_b, "host_wall_ms": wall_b * 1e3, "packets_per_s": tot / (ms_b * 1e-3)},
               "separate": {"device_ms": seq_ms, "host_wall_ms": wall_s * 1e3,
                            "packets_per_s": tot / (seq_ms * 1e-3)},
               "speedup_device": seq_ms / ms_b, "speedup_wall": wall_s / wall_b,
               "counters_identical": bool(same),
               "autotune": [{"tuning": c, "ms": m_} for c, m_ in trials]}
        log("nebulae: %d models, %.3g packets: batched %.1f ms (%.3g packets/s), separate %.1f ms (%.3g), "
            "counters identical: %s" % (n, tot, ms_b, tot / (ms_b * 1e-3), seq_ms, tot / (seq_ms * 1e-3), same))
        if balance:
            out["balance"] = _balances(models, specs, here, train_models, log, train_steps, decay, te_weight)
        return out
    finally:
        for m in models:
            m.close()


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="batched transport and per-model balances of 64 nebulae")
    ap.add_argument("--train-models", type=int, default=16, help="models the warm-start network is trained on")
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--decay", default="1", help="learning-rate factors of equal phases, e.g. 1,0.3,0.1")
    ap.add_argument("--te-weight", type=float, default=5.0, help="weight of log10 Te in the loss")
    a = ap.parse_args()
    print(json.dumps(measure(train_models=a.train_models, train_steps=a.train_steps,
                             decay=tuple(float(v) for v in a.decay.split(",")), te_weight=a.te_weight,
                             log=lambda m: print(m, file=sys.stderr, flush=True))))
