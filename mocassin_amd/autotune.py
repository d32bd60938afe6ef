"""Autotuner for the transport kernels' launch shapes.

The transport kernel has five free parameters (mc_tuning in
include/mocassin_amd.h): resident workgroups per CU, workgroup size, the
number of waiting lanes that triggers a batched emission pass, the number
of walk passes between emission checks, and the kernel (auto: at most 3
workgroups of 256 per CU select the kernel with two packets per lane; pool:
one packet per lane and a pool of packets per wavefront in LDS).  The
default candidates are the auto grid and a pool grid.  The batched kernel
(mc_batch_tuning) has resident wavefronts per CU, the same two scheduling
parameters, and the packet-batch size handed to a wavefront at a time.
Results do not depend on any of them (every packet draws from its own
counter-based stream), so the tuner only times: it runs a short transport
of the real problem for each candidate and keeps the fastest by device time.
"""
import itertools
import json
import os

DB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_db.json")

DEFAULT_GRID = {
    "blocks_per_cu": (2, 3, 4, 8),
    "threads": (256,),
    "emit_min": (32, 48, 64),
    "walk_steps": (8, 12, 16, 20, 24, 32),
    "kernel": ("auto",),
}

POOL_GRID = {
    "blocks_per_cu": (4,),
    "threads": (256,),
    "emit_min": (16, 24, 32),
    "walk_steps": (8, 16),
    "kernel": ("pool",),
}


def candidates(grid=None):
    """The launch shapes to time: DEFAULT_GRID then POOL_GRID, or
    DEFAULT_GRID with the entries of `grid` replaced."""
    grids = [DEFAULT_GRID, POOL_GRID] if grid is None else [dict(DEFAULT_GRID, **grid)]
    for g in grids:
        keys = list(g)
        for vals in itertools.product(*(g[k] for k in keys)):
            yield dict(zip(keys, vals))


def autotune(transport, npackets, istar=1, grid=None, packet_offset=1 << 40, repeats=1):
    """Time every candidate on `npackets` packets of source `istar`; set and
    return (best tuning, list of (tuning, kernel_ms))."""
    results = []
    for i, cand in enumerate(candidates(grid)):
        transport.set_tuning(**cand)
        best_ms = None
        for r in range(repeats):
            transport.zero_estimators()
            c = transport.energy_packet_driver(istar, npackets,
                                               packet_offset + (i * repeats + r) * npackets)
            best_ms = c["kernel_ms"] if best_ms is None else min(best_ms, c["kernel_ms"])
        results.append((cand, best_ms))
    best = min(results, key=lambda t: t[1])[0]
    transport.set_tuning(**best)
    transport.zero_estimators()
    return best, results


BATCH_GRID = {
    "waves_per_cu": (8, 12, 16),
    "emit_min": (32, 64),
    "walk_steps": (8, 16, 32),
    "chunk": (1024, 4096),
}


def autotune_batch(models, npackets, istar=None, grid=None, packet_offset=1 << 40):
    """Time every batched launch shape (mc_batch_tuning) on npackets[i]
    packets of each model in one launch; return (best, [(tuning, ms)])."""
    from .transport import transport_batch
    g = dict(BATCH_GRID)
    g.update(grid or {})
    keys = list(g)
    ist = istar if istar is not None else [1] * len(models)
    results = []
    for i, vals in enumerate(itertools.product(*(g[k] for k in keys))):
        cand = dict(zip(keys, vals))
        for m in models:
            m.zero_estimators()
        _, ms = transport_batch(models, ist, npackets, [packet_offset + i * max(npackets)] * len(models), cand)
        results.append((cand, ms))
    for m in models:
        m.zero_estimators()
    return min(results, key=lambda t: t[1])[0], results


def load(workload):
    """Tuning stored for a workload name, or None."""
    try:
        return json.load(open(DB_PATH)).get(workload)
    except (OSError, ValueError):
        return None


def save(workload, tuning, note=""):
    try:
        db = json.load(open(DB_PATH))
    except (OSError, ValueError):
        db = {}
    db[workload] = dict(tuning)
    if note:
        db[workload]["note"] = note
    with open(DB_PATH, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
