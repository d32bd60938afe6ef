"""Batched serving of many independent nebulae in one persistent launch.

The workload: 64 octant models (seeded) with edge drawn uniformly from
20-80 cells, a blackbody source of Teff uniform in 30,000-60,000 K, uniform
nH log-uniform in 1e1-1e4 cm^-3, H+He gas, the 400-bin mesh of the H+He
atomic data, and log-uniform 1e6-1e8 packets per model.  `measure` times
one batched launch of all models (mc_transport_batch, launch shape chosen by
the autotuner on a tenth of the packets) against the same models' separate
launches back to back, and checks that both give the same counters.
"""
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


def measure(n=64, log=print):
    import os
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
        out = {"models": n, "packets": tot, "cells": int(sum(m.nCells for m in models)),
               "edges": [s["edge"] for s in specs], "tuning": best,
               "batched": {"device_ms": ms_b, "host_wall_ms": wall_b * 1e3, "packets_per_s": tot / (ms_b * 1e-3)},
               "separate": {"device_ms": seq_ms, "host_wall_ms": wall_s * 1e3,
                            "packets_per_s": tot / (seq_ms * 1e-3)},
               "speedup_device": seq_ms / ms_b, "speedup_wall": wall_s / wall_b,
               "counters_identical": bool(same),
               "autotune": [{"tuning": c, "ms": m_} for c, m_ in trials]}
        log("nebulae: %d models, %.3g packets: batched %.1f ms (%.3g packets/s), separate %.1f ms (%.3g), "
            "counters identical: %s" % (n, tot, ms_b, tot / (ms_b * 1e-3), seq_ms, tot / (seq_ms * 1e-3), same))
        return out
    finally:
        for m in models:
            m.close()


if __name__ == "__main__":
    import json
    import sys
    print(json.dumps(measure(log=lambda m: print(m, file=sys.stderr, flush=True))))
