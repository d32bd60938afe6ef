"""Transport throughput on one MI355X: packets/s of the full-abundance
45^3-octant workload (800 frequency bins, 1e8 packets per sweep), with the
wall time to converged Te of the plain and the warm-started iteration.

Setup (untimed): the workload's grid and source; from the reference's cold
start (TeStart = 1e4 K and its initial ionisation) the device iterates
tables -> sweep -> balance (mocassin_amd/iterate.py) a few times, then
builds the opacity and re-emission tables of the state it reached.  Those
tables stay resident in HBM.

A step is one transport sweep on them: zero the estimators, transport all
packets of the source; J_nu and the escape SED stay in HBM.  Prints one JSON
line with
  roofline      of the transport kernel (HIP-event timed on its own stream):
                algorithmic HBM bytes and, under "atomic", the J atomics per
                second against the probe of random-address FP64 atomics;
                "hbm_copy" is a 2 GiB device copy (the HBM rate this box
                reaches)
  cpu_baseline  the C restatement of the same method (oracle/) on a bounded
                sample of the same workload, on every host core the process
                may use, with the reference's measured rate relative to it
  stand_in      the same sweep on the synthetic stand-in tables
  dust_gas      the dust and gas workload at 60^3 (mocassin_amd/dustgas.py):
                sweeps on its device-built tables and its grain balance
  config1       the 19^3 H+He octant: one sweep of 1e6 packets and one
                balance
  balance, te_convergence, ...  secondary legs; a leg that fails records
                {"error": ...} under its key and the line is still printed

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--packets N]
"""
import argparse
import json
import math
import os
import sys
import time
import traceback

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6                 # MI355X dense FP64 vector
VALU_CLOCK_HZ = 2.4e9                   # MI355X peak engine clock
SEED = 20241202
# packets/s of the compiled reference (MOCASSIN 2.02, amdflang -O2 + MPICH,
# `mcref_harness sweep`, 8 ranks) over the C restatement's (oracle, FP64 J,
# own math, 8 threads) on the same deck's tables and the same 8 cores (Intel
# Xeon @ 2.70 GHz), transport phase only
REFERENCE_OVER_PORT = {"c1": 5.37e6 / 8.91e6, "c2": 4.39e5 / 3.13e5}


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def shape_of(t):
    """A launch shape as text: blocks_per_cu,threads,emit_min,walk_steps and
    the kernel's name unless the shape picks it (auto)."""
    from mocassin_amd._capi import KERNELS
    s = ",".join(str(t[k]) for k in ("blocks_per_cu", "threads", "emit_min", "walk_steps"))
    name = {v: k for k, v in KERNELS.items()}.get(t.get("kernel", 0), "auto")
    return s if name == "auto" else s + "," + name


def kernel_name(t):
    """The device kernel a launch shape runs on the 45^3 octant (a grid the
    pool kernel takes), as rocprofv3 names it."""
    from mocassin_amd._capi import KERNELS
    k = t.get("kernel", 0)
    if k == KERNELS["auto"]:
        k = KERNELS["spare"] if t["blocks_per_cu"] <= 3 else KERNELS["pool"]
    return {KERNELS["one"]: "mcd::transport_kernel<false>", KERNELS["spare"]: "mcd::transport_kernel<true>",
            KERNELS["pool"]: "mcd::pool_kernel<8, 32>"}[k]


def lanes_per_pass(c):
    """Busy lanes of a 64-lane wavefront per walk pass and per emission
    pass of one launch, from the kernel's own pass counters."""
    return {"walk": c["n_walk_lanes"] / max(c["n_walk_passes"], 1.0),
            "emit": c["n_emit_lanes"] / max(c["n_emit_passes"], 1.0),
            "walk_passes": c["n_walk_passes"], "emit_passes": c["n_emit_passes"]}


from mocassin_amd.transport import algorithmic_bytes  # noqa: E402


def host_cores():
    """(threads to use, description): the CPUs this process may run on,
    capped by a cgroup CPU quota where one is set."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    n = min(aff, quota) if quota else aff
    return max(1, n), {"affinity_cpus": aff, "cgroup_quota_cpus": quota, "os_cpu_count": os.cpu_count(),
                       "cpu_model": model}


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mc_oracle as O
    return O


def cpu_baseline(p, tables, budget_s=15.0):
    """The oracle (scalar C restatement, one problem copy per thread) timed
    on disjoint packet ranges of the same workload and tables."""
    O = oracle()
    from concurrent.futures import ThreadPoolExecutor
    opac, rec, tl = tables
    P = O.Problem(xAxis=p["xAxis"], yAxis=p["yAxis"], zAxis=p["zAxis"], active=p["active"],
                  opacity=opac, nuArray=p["nuArray"], specCDF=p["specCDF"][:, 0],
                  R_out=p["R_out"], geoCorr=p["geoCorr"], ionEdge1=p["ionEdge1"],
                  deltaE=p["deltaE"], starPos=p["starPos"], starIdx=p["starIdx"],
                  recPDF=rec, totalLines=tl)
    cores, desc = host_cores()
    t = time.perf_counter()
    P.transport(10 ** 9, 20000, SEED, 1, O.MATH_OWN, O.J_FP64)
    rate1 = 20000 / (time.perf_counter() - t)
    per_thread = max(20000, int(rate1 * budget_s / cores))

    def work(i):
        P.transport(2 * 10 ** 9 + i * per_thread, per_thread, SEED, 1, O.MATH_OWN, O.J_FP64)

    t = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(work, range(cores)))
    wall = time.perf_counter() - t
    value = cores * per_thread / wall
    return {"value": value, "unit": "packets/s", "cores": cores, "kind": "port",
            "sample": "%d packets of the same workload and tables (%d threads x %d), scalar C restatement, "
                      "FP64 J" % (cores * per_thread, cores, per_thread),
            "one_thread": rate1, "host": desc,
            "reference_over_port": REFERENCE_OVER_PORT["c2"],
            "reference_equivalent": value * REFERENCE_OVER_PORT["c2"],
            "note": "reference_equivalent scales the port's rate by the compiled reference's measured rate "
                    "relative to it on the same cores (BASELINE.md)"}


def balance_cpu_baseline(bundle, J, Hd, Te, Ne, ion, jscale, cells=400):
    """updateCell in FP64 (the C restatement the balance kernel follows) on
    `cells` evenly spaced reached cells of the same J and starting state,
    one thread: cells per second."""
    import numpy as np
    O = oracle()
    hit = np.nonzero(J[1:].sum(axis=1) > 0)[0] + 1
    pick = hit[np.linspace(0, hit.size - 1, min(cells, hit.size)).astype(np.int64)]
    rows = np.concatenate([[0], pick])
    B = O.Balance(bundle)
    Js = np.ascontiguousarray(J[rows] * jscale, np.float64)
    t = time.perf_counter()
    r = B.run(Js, Hd[rows], Te[rows], Ne[rows], ion[rows], "f64")
    dt = time.perf_counter() - t
    return {"value": pick.size / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "mean_iterations": float(r[4][1:].mean()),
            "sample": "%d reached cells of the same J and starting state, FP64 C restatement" % pick.size}


def opacity_cpu_baseline(atom, bundle, nb, nu, Hden, Te, Ne, ion, cells=3000):
    """ionizationDriver/addOpacity in the reference's FP32 (the C
    restatement, oracle/mc_opacity.c) on the first `cells` cells of the same
    state, one thread: cells per second."""
    import numpy as np
    O = oracle()
    n = min(cells, Te.size - 1)
    sl = slice(0, n + 1)
    m4 = np.asarray(atom["op_seg"], np.int32)
    m6 = np.zeros((m4.shape[0], 6), np.int32)
    m6[:, 0], m6[:, 1], m6[:, 3], m6[:, 4] = m4[:, 0], m4[:, 1], m4[:, 2], m4[:, 3]
    lens = m4[:, 3] - m4[:, 2] + 1
    g = dict(ints=np.array([0, 0, 0, n, nb, 0, 0, bundle.nElemUsed * 100 + bundle.nstages]), nuArray=nu,
             Hden=np.asarray(Hden[sl], np.float32), Te=np.asarray(Te[sl], np.float32),
             Te0=np.asarray(Te[sl], np.float32), Ne=np.asarray(Ne[sl], np.float32),
             ionDen=np.asfortranarray(np.asarray(ion[sl], np.float32)).ravel(order="F"),
             elem_on=np.asarray(bundle.elem_on, np.int32), xref=np.asarray(bundle.xref, np.int32),
             abun=np.asarray(bundle.abun, np.float32), seg_meta=m6, seg_sigma=atom["op_sigma"],
             seg_off=np.concatenate([[0], np.cumsum(lens)]).astype(np.int64), brems1=atom["op_brems1"])
    t = time.perf_counter()
    O.opacity_tables(g, gaunt_cache=True)
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": "%d cells of the same state, scalar C restatement" % n}


def lines_cpu_baseline(bundle, Hden, Te, Ne, ion, dV, cells=500):
    """outputGas's forbVol/HbetaVol sums (the FP64 C restatement the GPU
    kernel follows) over the first `cells` cells of the same state, one
    thread: cells per second."""
    import numpy as np
    O = oracle()
    n = min(cells, Te.size - 1)
    f32 = lambda v: np.asarray(v, np.float32).astype(np.float64)
    B = O.Balance(bundle)
    t = time.perf_counter()
    B.line_lum(f32(Hden), f32(Te), f32(Ne), f32(ion), f32(dV), np.arange(1, n + 1), "f64")
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": "%d cells of the same state, scalar C restatement" % n}


def emission_cpu_baseline(atom, bundle, nb, nu, Te, Ne, ion, cells=200):
    """emissionDriver's gas part in the reference's precision (the C
    restatement) on the first `cells` cells of the same state, one thread:
    cells per second."""
    import numpy as np
    O = oracle()
    n = min(cells, Te.size - 1)
    g = dict(atom)
    g["ints"] = np.array([0, 0, 0, n, nb], np.int32)
    g["nuArray"] = nu
    E, keep = O.emission_tables(g)
    B = O.Balance(bundle)
    t = time.perf_counter()
    B.emission(E, np.asarray(Te, np.float32), np.asarray(Ne, np.float32), np.asarray(ion, np.float32),
               np.arange(1, n + 1), "f32")
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": "%d cells of the same state, scalar C restatement" % n}


def fp64_rate(kind, t_iterations, device_ms):
    """FP64 rate of a balance stage: the FP64 operations per temperature
    iteration that rocprofv3 counted for this workload's balance
    (profiles/balance_c2_fp64.json) times this run's temperature iterations,
    plus the integrals, over the stage's device time."""
    prof = json.load(open(os.path.join(ROOT, "profiles", "balance_c2_fp64.json")))[kind]
    flops = prof["fp64_flops_per_temperature_iteration"] * t_iterations + prof["integrals_kernel_fp64_flops"]
    achieved = flops / (device_ms * 1e-3) / 1e12
    return {"achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
            "flops": flops, "source": "profiles/balance_c2_fp64.json (FP64 lane operations issued, inactive "
                                      "lanes included: an upper bound)"}


def leg(out, key, fn):
    """Run one secondary measurement; an exception is recorded, not raised."""
    t = time.perf_counter()
    try:
        out[key] = fn()
    except Exception as e:                        # noqa: BLE001 - recorded in the JSON line
        out[key] = {"error": "%s: %s" % (type(e).__name__, e)}
        log("%s failed: %s" % (key, e))
        traceback.print_exc(file=sys.stderr)
    log("%s leg %.1f s" % (key, time.perf_counter() - t))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=float, default=1e8)
    ap.add_argument("--setup-iterations", type=int, default=6,
                    help="device iterations from the cold start before the tables are built")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-te", action="store_true", help="skip the Te-convergence measurement")
    ap.add_argument("--no-balance", action="store_true", help="skip the balance and table legs")
    ap.add_argument("--no-grains", action="store_true", help="skip the balance with grains")
    ap.add_argument("--no-atomics", action="store_true", help="skip the atomic-throughput probe")
    ap.add_argument("--no-stand-in", action="store_true", help="skip the stand-in tables' sweep")
    ap.add_argument("--no-dust-gas", action="store_true", help="skip the dust and gas 60^3 workload")
    ap.add_argument("--autotune", action="store_true",
                    help="time the launch-shape candidates on this workload and store the best "
                         "in mocassin_amd/tuning_db.json (default: use the stored shape)")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shape blocks_per_cu,threads,emit_min,walk_steps[,kernel] "
                         "(kernel: auto, one, spare, pool)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (traffic calibration "
                         "for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU sharding is a design, see the documentation)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from mocassin_amd import Transport, synth
    from mocassin_amd import autotune as at
    from mocassin_amd.atomic import AtomicBundle, emission_tables
    from mocassin_amd.converge import cold_state
    from mocassin_amd.iterate import iterate, jscale_of

    npk = int(a.packets)
    t0 = time.perf_counter()
    atom = dict(np.load(os.path.join(ROOT, "tests", "golden", "atomic_c2.npz")))
    bundle = AtomicBundle.from_npz(atom)
    p = synth.workload("c2", n_packets=npk, nu=bundle.nuArray)
    T = Transport(p["xAxis"], p["yAxis"], p["zAxis"], p["active"], p["nCells"], p["nbins"],
                  p["nuArray"], p["R_out"], p["geoCorr"], p["ionEdge1"], [p["starPos"]],
                  [p["starIdx"]], [p["deltaE"]], symmetric=p["symmetric"], seed=SEED, device=0)
    T.balance_setup(bundle)
    T.opacity_setup(atom["op_seg"], atom["op_sigma"], float(atom["op_brems1"][0]))
    T.emission_setup(**emission_tables(atom))
    tuning_src = "default"
    if a.tuning:
        f = a.tuning.split(",")
        T.set_tuning(blocks_per_cu=int(f[0]), threads=int(f[1]), emit_min=int(f[2]), walk_steps=int(f[3]),
                     kernel=f[4] if len(f) > 4 else "auto")
        tuning_src = "command line"
    elif at.load("c2") and not a.autotune:
        db = at.load("c2")
        T.set_tuning(**{k: db[k] for k in ("blocks_per_cu", "threads", "emit_min", "walk_steps", "kernel")
                        if k in db})
        tuning_src = "mocassin_amd/tuning_db.json"
    js = jscale_of(p["symmetric"])
    spec = p["specCDF"]

    # ---- setup: iterate from the cold start, then the tables of that state
    Hd0 = cold_state(bundle, atom["Te0"][0], atom["ion0"], p["Hden"])
    cold = (Hd0[3], Hd0[0], Hd0[1], Hd0[2])
    state, flags, records, _ = iterate(T, cold, npk, spec, p["symmetric"], max_iter=a.setup_iterations,
                                       min_convergence=101.0, packet_offset=1 << 40, log=log)
    # the tables of the state reached, with host copies for the CPU baseline
    # (built from the state as the host holds it: the same values)
    host_tables = None
    if a.no_cpu:
        T.opacity_build(host=False)
        T.emission_build(host=False)
    else:
        opac_h = T.opacity_build(*_f32_state(state), host=True)[0]
        rec_h, tl_h, _ = T.emission_build(*_f32_state(state), host=True)
        host_tables = (opac_h, rec_h, tl_h)
    T.upload_tables(None, spec)
    if a.autotune:
        best, trials = at.autotune(T, max(1, min(npk, 10_000_000)), repeats=2)
        at.save("c2", best, "bench.py --autotune on the iterated tables, %d candidates" % len(trials))
        tuning_src = "autotuned now"
        log("autotuned %s (%d candidates, %.2f..%.2f ms)" % (best, len(trials), min(t[1] for t in trials),
                                                             max(t[1] for t in trials)))
    tuning = dict(T.get_tuning(), source=tuning_src)
    setup_s = time.perf_counter() - t0
    log("workload c2, %d cells, %d bins, %d packets/step, setup %.1f s (%d iterations, %.1f%% converged)"
        % (p["nCells"], p["nbins"], npk, setup_s, len(records), records[-1]["converged_pct"]))

    if a.copy_mb:
        n = a.copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- the timed steps
    step = 0
    for _ in range(a.warmup):
        T.zero_estimators()
        T.energy_packet_driver(1, npk, step * npk)
        step += 1
    torch.cuda.synchronize()
    counters = []
    t = time.perf_counter()
    for _ in range(a.steps):
        T.zero_estimators()
        counters.append(T.energy_packet_driver(1, npk, step * npk))
        step += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    value = a.steps * npk / elapsed
    kms = sum(c["kernel_ms"] for c in counters) / len(counters)
    byts = sum(algorithmic_bytes(c, p["nbins"]) for c in counters) / len(counters)
    achieved = byts / (kms * 1e-3) / 1e9
    c0 = counters[-1]
    log("%.3e packets/s, kernel %.1f ms/launch, %.3g seg/packet, %.3g abs/packet"
        % (value, kms, c0["n_seg"] / npk, c0["n_abs"] / npk))
    # HBM bytes per launch: calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per
    # packet of this workload times this launch, when the profile's launch
    # shape is the one measured now
    traffic, traffic_note = None, "no calibrated counter profile of this workload"
    valu = None
    tp = os.path.join(ROOT, "profiles", "transport_c2_traffic.json")
    if os.path.exists(tp):
        try:
            tj = json.load(open(tp))
            shape = shape_of(tuning)
            if tj.get("tuning") == shape and tj.get("tables") == "iterated":
                traffic = tj["hbm_bytes_per_packet"] * npk
                traffic_note = "profiles/transport_c2_traffic.json (%s)" % tj.get("tuning")
                # the walk's instruction issue: wave64 VALU instructions of
                # one launch (rocprofv3 SQ_INSTS_VALU) over this run's kernel
                # time, against 4 SIMDs per CU issuing one wave64 FP32
                # instruction every 4 cycles at the 2.4 GHz peak clock
                insts = tj["sq_insts_valu_per_packet"] * npk
                peak = 4 * torch.cuda.get_device_properties(0).multi_processor_count * VALU_CLOCK_HZ / 4
                valu = {"achieved": insts / (kms * 1e-3), "peak": peak, "unit": "wave64 VALU instructions/s",
                        "frac": insts / (kms * 1e-3) / peak, "instructions": insts,
                        "source": "profiles/transport_c2_traffic.json (SQ_INSTS_VALU, %s)" % tj.get("tuning")}
            else:
                traffic_note = "profile of another launch shape or tables (%s, %s): not used" % (
                    tj.get("tuning"), tj.get("tables"))
                log("traffic profile is of %s/%s, the run is %s/iterated: traffic not reported"
                    % (tj.get("tuning"), tj.get("tables"), shape))
        except (OSError, ValueError, KeyError):
            pass
    out = {
        "metric": "packets/s/GPU; wall-time to converged Te (±1%) plain vs MLP warm-start",
        "value": value, "unit": "packets/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic grid of the workload's shape; tables built on the device from the state %d "
                "iterations from the cold start reached" % len(records),
        "config": {"workload": "full-abundance 45^3 octant, 800 bins, %d packets/sweep, opacity and "
                               "re-emission tables of the iterated state" % npk,
                   "cells": p["nCells"], "nbins": p["nbins"], "packets_per_step": npk,
                   "accumulators": "fp64", "tuning": tuning,
                   "setup_iterations": records,
                   "segments_per_packet": c0["n_seg"] / npk, "absorptions_per_packet": c0["n_abs"] / npk},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "algorithmic_bytes": byts, "kernel": kernel_name(tuning), "kernel_ms": kms,
                     "segments": c0["n_seg"], "absorptions": c0["n_abs"], "escapes": c0["n_esc"],
                     "lanes_per_pass": lanes_per_pass(c0)},
    }
    if valu is not None:
        out["roofline"]["valu"] = valu

    def atomic_leg():
        from mocassin_amd.transport import probe_atomics
        words = int(T.shape[0] * T.shape[1] * T.shape[2] * p["nbins"])
        req = int(c0["n_seg"])
        ms64, rate64 = probe_atomics(words, req, lanes=64)
        ms16, rate16 = probe_atomics(words, req, lanes=16)
        ach = c0["n_seg"] / (kms * 1e-3)
        return {"achieved": ach, "peak": rate64, "unit": "atomics/s", "frac": ach / rate64,
                "probe": "%d FP64 atomic adds to random words of a %d-word array (the J3 footprint), "
                         "64 active lanes: %.2f ms; 16 active lanes: %.2f ms (%.3g/s)" % (req, words, ms64, ms16,
                                                                                      rate16),
                "probe_16_lanes": rate16,
                "note": "adds of one instruction that fall in one 64-byte line go to memory as one request "
                        "(mc_probe_atomic_lines), and the J rings keep a packet's updates in one instruction, "
                        "so merged adds run above the random-word rate"}
    if not a.no_atomics:
        leg(out["roofline"], "atomic", atomic_leg)

        def copy_leg():
            # a STREAM-like copy of 2 GiB on the device: the HBM rate this
            # box reaches, beside the 8 TB/s the peak is quoted at
            n = (1 << 31) // 4
            src = torch.ones(n, dtype=torch.float32, device="cuda")
            dst = torch.empty_like(src)
            dst.copy_(src)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            best = None
            for _ in range(3):
                e0.record()
                dst.copy_(src)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1)
                best = ms if best is None else min(best, ms)
            del src, dst
            gbs = 2.0 * (1 << 31) / (best * 1e-3) / 1e9
            return {"bytes_moved": 2 * (1 << 31), "ms": best, "GB_per_s": gbs, "of_peak": gbs / HBM_PEAK_GBS,
                    "note": "device-to-device copy of 2 GiB (read + write), best of 3"}
        leg(out["roofline"], "hbm_copy", copy_leg)

    # ---- secondary legs on the iterated state
    J_last = None
    if not a.no_balance:
        def balance_leg():
            nonlocal J_last
            Hd, Te, Ne, ion = state
            J_last, _ = T.fetch()
            res = {}
            runs = []
            for _ in range(2):
                tb = time.perf_counter()
                r = T.update_cells(Hd, Te, Ne, ion, jscale=js)
                runs.append((r[5], time.perf_counter() - tb, r))
            ms_b, wall_b, r = min(runs, key=lambda x: x[0])
            hit = (r[3][1:] & 16) == 0
            res = {"cells": int(p["nCells"]), "reached": int(hit.sum()), "device_ms": ms_b,
                   "host_wall_ms": wall_b * 1e3, "cells_per_s": float(hit.sum()) / (ms_b * 1e-3),
                   "mean_iterations": float(r[4][1:][hit].mean()),
                   "solver": "reference schedule (iterateT), FP64, from the iterated state on the last sweep's J"}
            res["fp64"] = fp64_rate("gas", float(r[4][1:][hit].sum()), ms_b)
            Te0, Ne0, ion0, Hdc = Hd0
            rc = T.update_cells(Hdc, Te0, Ne0, ion0, jscale=js)
            hc = (rc[3][1:] & 16) == 0
            res["cold_start"] = {"device_ms": rc[5], "cells_per_s": float(hc.sum()) / (rc[5] * 1e-3),
                                 "mean_iterations": float(rc[4][1:][hc].mean())}
            if not a.no_cpu:
                cb = balance_cpu_baseline(bundle, J_last, Hd, Te, Ne, ion, js)
                res["cpu_baseline"] = cb
                res["gpu_over_cpu"] = res["cells_per_s"] / cb["value"]
            # leave the device state of the iterated state's balance for the table legs
            T.update_cells(Hd, Te, Ne, ion, jscale=js)
            log("balance %d cells in %.1f ms (%.2f T-iterations per cell); cold start %.1f ms"
                % (hit.sum(), ms_b, res["mean_iterations"], rc[5]))
            return res
        leg(out, "balance", balance_leg)

        def opacity_leg():
            Hd, Te, Ne, ion = state
            ms_o = min(T.opacity_build(host=False)[2] for _ in range(3))
            n3 = int(np.prod(np.shape(p["active"])))
            nslot = sum(min(z, bundle.nstages) for z in range(1, 31) if bundle.elem_on[z - 1])
            byts_o = 4.0 * n3 * p["nbins"] + 4.0 * n3 + 4.0 * (p["nCells"] + 1) * (3 + nslot)
            res = {"device_ms": ms_o, "cells_per_s": p["nCells"] / (ms_o * 1e-3),
                   "roofline": {"bound": "hbm", "achieved": byts_o / (ms_o * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": byts_o / (ms_o * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "algorithmic_bytes": byts_o, "kernel": "opacity_kernel"}}
            if not a.no_cpu:
                res["cpu_baseline"] = opacity_cpu_baseline(atom, bundle, p["nbins"], p["nuArray"], Hd, Te, Ne, ion)
                res["gpu_over_cpu"] = res["cells_per_s"] / res["cpu_baseline"]["value"]
            return res
        leg(out["balance"] if isinstance(out.get("balance"), dict) else out, "opacity", opacity_leg)

        def lines_leg():
            Hd, Te, Ne, ion = state
            ms_l = min(T.line_luminosities()[2] for _ in range(3))
            forb, hb, _ = T.line_luminosities()
            res = {"device_ms": ms_l, "cells_per_s": p["nCells"] / (ms_l * 1e-3), "model_ions": int(bundle.ncel),
                   "lines_above_1e-3_hbeta": int(np.count_nonzero(forb > 1e-3 * hb)),
                   "kernel": "lines_kernel + lines_reduce_kernel"}
            if not a.no_cpu:
                O = oracle()
                dV, _ = O.cell_volumes(p["xAxis"], p["yAxis"], p["zAxis"], np.asarray(p["active"]), p["symmetric"])
                res["cpu_baseline"] = lines_cpu_baseline(bundle, Hd, Te, Ne, ion, dV)
                res["gpu_over_cpu"] = res["cells_per_s"] / res["cpu_baseline"]["value"]
            return res
        leg(out["balance"] if isinstance(out.get("balance"), dict) else out, "lines", lines_leg)

        def emission_leg():
            Hd, Te, Ne, ion = state
            ms_e = min(T.emission_build(host=False)[2] for _ in range(3))
            res = {"device_ms": ms_e, "cells_per_s": p["nCells"] / (ms_e * 1e-3), "kernel": "emission_kernel",
                   "free_bound_channels": int(atom["em_ints"][9])}
            if not a.no_cpu:
                res["cpu_baseline"] = emission_cpu_baseline(atom, bundle, p["nbins"], p["nuArray"], Te, Ne, ion)
                res["gpu_over_cpu"] = res["cells_per_s"] / res["cpu_baseline"]["value"]
            return res
        leg(out["balance"] if isinstance(out.get("balance"), dict) else out, "emission", emission_leg)

        def iteration_leg():
            # one whole iteration with every table resident (iterateMC's loop body)
            nonlocal step
            st, fl, recs, _ = iterate(T, state, npk, spec, p["symmetric"], max_iter=2, min_convergence=101.0,
                                      packet_offset=step * npk)
            step += 2
            r = recs[-1]
            return dict(r, packets=npk, note="opacity and re-emission builds of the state, upload, sweep, "
                                             "balance from the previous state (reference schedule); device "
                                             "times per stage, host wall time of the whole")
        leg(out["balance"] if isinstance(out.get("balance"), dict) else out, "iteration", iteration_leg)

        if not a.no_grains:
            def grains_leg():
                Hd, Te, Ne, ion = state
                T.dust_setup(**synth.grain_tables(p["nuArray"], Hd))
                Td0 = np.full((p["nCells"] + 1, 11, 2), 50.0)
                runs = []
                for _ in range(2):
                    tb = time.perf_counter()
                    r = T.update_grains(Td0, jscale=js, Hden=Hd, Te=Te, Ne=Ne, ionDen=ion)
                    runs.append((r[3], time.perf_counter() - tb, r))
                ms_g, wall_g, r = min(runs, key=lambda x: x[0])
                hit = (r[1][1:] & 16) == 0
                Tg = r[0][1:, 0, 0][hit]
                return {"device_ms": ms_g, "host_wall_ms": wall_g * 1e3,
                        "cells_per_s": float(hit.sum()) / (ms_g * 1e-3),
                        "mean_iterations": float(r[2][1:][hit].mean()),
                        "fp64": fp64_rate("grains", float(r[2][1:][hit].sum()), ms_g),
                        "grains": "astronomical silicate, 10 sizes, photoelectric charge and gas-grain collisions",
                        "Tdust_K": [float(Tg.min()), float(np.median(Tg)), float(Tg.max())]}
            leg(out["balance"] if isinstance(out.get("balance"), dict) else out, "with_grains", grains_leg)

    if not a.no_cpu:
        def cpu_leg():
            # the tables of the timed sweeps, as the host holds them
            return cpu_baseline(p, host_tables)
        leg(out, "cpu_baseline", cpu_leg)
        if "value" in out["cpu_baseline"]:
            out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
            out["gpu_over_reference_equivalent"] = value / out["cpu_baseline"]["reference_equivalent"]

    if not a.no_dust_gas:
        def dust_gas_atomic(T3, p3, c, km):
            if a.no_atomics:
                return None
            from mocassin_amd.transport import probe_atomics
            words = int(T3.shape[0] * T3.shape[1] * T3.shape[2] * p3["nbins"])
            ms64, rate64 = probe_atomics(words, int(c["n_seg"]), lanes=64)
            ach = c["n_seg"] / (km * 1e-3)
            return {"achieved": ach, "peak": rate64, "unit": "atomics/s", "frac": ach / rate64,
                    "probe": "%d random FP64 adds into %d words (this J3's footprint): %.2f ms"
                             % (int(c["n_seg"]), words, ms64)}

        def dust_gas_leg():
            # the dust and gas workload at its own shape: 60^3 full grid,
            # lognormal nH, silicate grains with HG scattering, 50,000 K
            from mocassin_amd import dustgas
            W = dustgas.open_context(60, packets=npk)
            T3, p3 = W["T"], W["p"]
            try:
                # four iterations from the cold start (tables with the grain
                # terms, sweep, balance with grains), then the sweeps are
                # timed on the tables of the state they reached
                st3, _, recs3, _ = iterate(T3, W["state"], npk, p3["specCDF"], False, max_iter=4,
                                           min_convergence=101.0, packet_offset=1 << 43, Tdust=W["Tdust0"],
                                           gSca=W["gSca"])
                Hd3, Te3, Ne3, ion3, Td3 = st3
                ms_o = T3.opacity_build(host=False)[2]
                ms_e = T3.emission_build(host=False)[2]
                T3.upload_tables(None, p3["specCDF"], gSca=W["gSca"])
                cs = []
                for k in range(3):
                    T3.zero_estimators()
                    cs.append(T3.energy_packet_driver(1, npk, k * npk))
                km = sum(c["kernel_ms"] for c in cs[1:]) / 2
                bb = sum(algorithmic_bytes(c, p3["nbins"]) for c in cs[1:]) / 2
                c = cs[-1]
                rg = T3.update_grains(Td3, jscale=jscale_of(False), Hden=Hd3, Te=Te3, Ne=Ne3, ionDen=ion3)
                hit = (rg[1][1:] & 16) == 0
                res = {"workload": "60^3 full grid, lognormal nH (mean 300, sigma_ln 0.5, 8-cell correlation), "
                                   "H+He with astronomical-silicate grains (10 sizes), Henyey-Greenstein "
                                   "scattering, 50,000 K source, 160 bins, %d packets per sweep, tables the device "
                                   "builds from the state four iterations from the cold start reached" % npk,
                       "setup_iterations": recs3,
                       "cells": int(p3["nCells"]), "kernel_ms": km, "packets_per_s": npk / (km * 1e-3),
                       "segments_per_packet": c["n_seg"] / npk, "scatterings_per_packet": c["n_sca"] / npk,
                       "absorptions_per_packet": c["n_abs"] / npk, "J_atomics_per_s": c["n_seg"] / (km * 1e-3),
                       "hbm_frac": bb / (km * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "atomic": dust_gas_atomic(T3, p3, c, km),
                       "opacity_ms": ms_o, "emission_ms": ms_e,
                       "grain_balance": {"device_ms": rg[3], "reached": int(hit.sum()),
                                         "cells_per_s": float(hit.sum()) / (rg[3] * 1e-3),
                                         "mean_iterations": float(rg[2][1:][hit].mean()),
                                         "solver": "reference schedule from the iterated state, grain "
                                                   "charge, photoelectric heating and gas-grain collisions"}}
                log("dust and gas 60^3: %.3e packets/s (%.1f ms, %.1f seg, %.2f scatterings per packet); "
                    "grain balance %.1f ms" % (res["packets_per_s"], km, res["segments_per_packet"],
                                               res["scatterings_per_packet"], rg[3]))
                return res
            finally:
                T3.close()
        leg(out, "dust_gas", dust_gas_leg)

    if not a.no_balance:
        def config1_leg():
            # the 19^3 H+He octant (400 bins, 1e6 packets): tables of the
            # reference's cold start built on the device, one sweep, one
            # balance.  Its FP64 J (22 MB) lies in the Infinity Cache.
            at1 = dict(np.load(os.path.join(ROOT, "tests", "golden", "atomic_c1.npz")))
            b1 = AtomicBundle.from_npz(at1)
            n1 = 1000000
            p1 = synth.workload("c1", n_packets=n1, nu=b1.nuArray)
            T1 = Transport(p1["xAxis"], p1["yAxis"], p1["zAxis"], p1["active"], p1["nCells"], p1["nbins"],
                           p1["nuArray"], p1["R_out"], p1["geoCorr"], p1["ionEdge1"], [p1["starPos"]],
                           [p1["starIdx"]], [p1["deltaE"]], symmetric=p1["symmetric"], seed=SEED + 1, device=0)
            try:
                T1.balance_setup(b1)
                T1.opacity_setup(at1["op_seg"], at1["op_sigma"], float(at1["op_brems1"][0]))
                T1.emission_setup(**emission_tables(at1))
                Te1, Ne1, ion1, Hd1 = cold_state(b1, at1["Te0"][0], at1["ion0"], p1["Hden"])
                T1.opacity_build(Hd1, Te1, Ne1, ion1, host=False)
                T1.emission_build(Hd1, Te1, Ne1, ion1, host=False)
                T1.upload_tables(None, p1["specCDF"])

                def timed3(off):
                    cs = []
                    for k in range(4):
                        T1.zero_estimators()
                        cs.append(T1.energy_packet_driver(1, n1, off + k * n1))
                    return sum(c["kernel_ms"] for c in cs[1:]) / 3, cs[-1]
                km_default, _ = timed3(0)
                # a launch this short is shaped by its ramp and tail: the
                # autotuner picks the shape for this workload
                best, trials = at.autotune(T1, n1, grid={"blocks_per_cu": (1, 2, 3, 4, 8), "threads": (64, 128, 256)},
                                           packet_offset=1 << 40)
                km, c = timed3(4 * n1)
                r = T1.update_cells(Hd1, Te1, Ne1, ion1, jscale=jscale_of(p1["symmetric"]))
                hit = (r[3][1:] & 16) == 0
                res = {"workload": "19^3 H+He octant, 400 bins, %d packets, tables of the cold start built on the "
                                   "device, one sweep and one balance" % n1,
                       "cells": int(p1["nCells"]), "kernel_ms": km, "packets_per_s": n1 / (km * 1e-3),
                       "tuning": best, "candidates": len(trials), "kernel_ms_default_shape": km_default,
                       "segments_per_packet": c["n_seg"] / n1, "J_atomics_per_s": c["n_seg"] / (km * 1e-3),
                       "balance": {"device_ms": r[5], "reached": int(hit.sum()),
                                   "mean_iterations": float(r[4][1:][hit].mean())}}
                if not a.no_atomics:
                    from mocassin_amd.transport import probe_atomics
                    words = int(T1.shape[0] * T1.shape[1] * T1.shape[2] * p1["nbins"])
                    _, rate = probe_atomics(words, int(c["n_seg"]), lanes=64)
                    res["atomic"] = {"achieved": res["J_atomics_per_s"], "peak": rate, "frac": res["J_atomics_per_s"] / rate,
                                     "probe": "random FP64 adds into this J3's %d words (Infinity Cache resident)" % words}
                log("19^3 octant: %.3e packets/s (%.2f ms per 1e6 packets, %.1f seg/packet); balance %.2f ms"
                    % (res["packets_per_s"], km, res["segments_per_packet"], r[5]))
                return res
            finally:
                T1.close()
        leg(out, "config1", config1_leg)

    if not a.no_stand_in:
        def stand_in_leg():
            T.upload_tables(p["opacity"], spec, p["recPDF"], p["totalLines"])
            cs = []
            for k in range(3):
                T.zero_estimators()
                cs.append(T.energy_packet_driver(1, npk, (step + k) * npk))
            km = sum(c["kernel_ms"] for c in cs[1:]) / 2
            bb = sum(algorithmic_bytes(c, p["nbins"]) for c in cs[1:]) / 2
            return {"kernel_ms": km, "packets_per_s": npk / (km * 1e-3), "segments_per_packet": cs[-1]["n_seg"] / npk,
                    "absorptions_per_packet": cs[-1]["n_abs"] / npk, "hbm_frac": bb / (km * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "tables": "synthetic: hydrogenic cross-sections, on-the-spot ionisation, totalLines 0.35"}
        leg(out, "stand_in", stand_in_leg)
    T.close()

    if not a.no_te:
        def te_leg():
            from mocassin_amd.converge import loop_convergence
            return loop_convergence(log=log)
        leg(out, "te_convergence", te_leg)
    print(json.dumps(out), flush=True)


def _f32_state(state):
    import numpy as np
    Hd, Te, Ne, ion = state
    return (np.asarray(Hd, np.float32), np.asarray(Te, np.float32), np.asarray(Ne, np.float32),
            np.asarray(ion, np.float32))


if __name__ == "__main__":
    main()
