"""Transport throughput on one MI355X: packets/s for the full-abundance
45^3-octant workload (800 frequency bins, 1e8 packets per sweep), and the
wall time to converged Te of the plain and the warm-started balance.

A step is one transport sweep: zero the estimators, transport all packets of
the source, keep J_nu and the escape SED resident in HBM (the host copy is
not part of the step).  Tables are synthetic of the workload's shape
(mocassin_amd/synth.py) on the frequency mesh of the workload's atomic data.
Prints one JSON line with
  roofline      of the transport kernel (HIP-event timed on its own stream)
  cpu_baseline  the C restatement of the same method (oracle/) on a bounded
                sample of the same workload, on the host's cores
  balance       the per-cell balance of the 45^3 workload on the last
                sweep's J (reference schedule, FP64): cells/s; with_grains:
                the same with silicate grains (grain charge, photoelectric
                heating, gas-grain collisions) and their temperatures
  te_convergence  plain vs warm-started balance on the clumpy 100^3 H+He
                workload (mocassin_amd/converge.py; --no-te skips it)

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--packets N]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(c, nbins):
    """24 B per segment (4 active + 4 opacity + 16 FP64 J read-modify-write),
    8 + 4*ceil(log2 nbins) B per absorption (line fraction, u-test operand,
    binary search of the re-emission CDF), 16 B per escape (FP64 RMW)."""
    return (24.0 * c["n_seg"] + (8.0 + 4.0 * math.ceil(math.log2(nbins))) * c["n_abs"]
            + 16.0 * c["n_esc"])


def cpu_baseline(p, seed, budget_s=15.0):
    """The oracle (scalar C restatement, one problem copy per thread) timed
    on disjoint packet ranges of the same workload, threads = host cores
    (at most 16)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mc_oracle as O
    from concurrent.futures import ThreadPoolExecutor
    P = O.Problem(xAxis=p["xAxis"], yAxis=p["yAxis"], zAxis=p["zAxis"], active=p["active"],
                  opacity=p["opacity"], nuArray=p["nuArray"], specCDF=p["specCDF"][:, 0],
                  R_out=p["R_out"], geoCorr=p["geoCorr"], ionEdge1=p["ionEdge1"],
                  deltaE=p["deltaE"], starPos=p["starPos"], starIdx=p["starIdx"],
                  recPDF=p["recPDF"], totalLines=p["totalLines"])
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(16, cores))
    t = time.perf_counter()
    P.transport(10 ** 9, 20000, seed, 1, O.MATH_OWN, O.J_FP64)
    rate1 = 20000 / (time.perf_counter() - t)
    per_thread = max(20000, int(rate1 * budget_s / cores))

    def work(i):
        t0 = time.perf_counter()
        P.transport(2 * 10 ** 9 + i * per_thread, per_thread, seed, 1, O.MATH_OWN, O.J_FP64)
        return time.perf_counter() - t0

    t = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(work, range(cores)))
    wall = time.perf_counter() - t
    return {"value": cores * per_thread / wall, "unit": "packets/s", "cores": cores,
            "kind": "port",
            "sample": "%d packets of the same workload (%d threads x %d), scalar C oracle"
                      % (cores * per_thread, cores, per_thread)}


def opacity_cpu_baseline(atom, bundle, nb, nu, Hden, Te, Ne, ion, cells=3000):
    """ionizationDriver/addOpacity in the reference's FP32 (the C
    restatement, oracle/mc_opacity.c) on the first `cells` cells of the same
    state, one thread: cells per second."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mc_oracle as O
    import numpy as np
    n = min(cells, Te.size - 1)
    sl = slice(0, n + 1)
    m4 = np.asarray(atom["op_seg"], np.int32)
    m6 = np.zeros((m4.shape[0], 6), np.int32)
    m6[:, 0], m6[:, 1], m6[:, 3], m6[:, 4] = m4[:, 0], m4[:, 1], m4[:, 2], m4[:, 3]
    lens = m4[:, 3] - m4[:, 2] + 1
    g = dict(ints=np.array([0, 0, 0, n, nb, 0, 0, bundle.nElemUsed * 100 + bundle.nstages]), nuArray=nu,
             Hden=np.asarray(Hden[sl], np.float32), Te=np.asarray(Te[sl], np.float32), Te0=np.asarray(Te[sl], np.float32),
             Ne=np.asarray(Ne[sl], np.float32),
             ionDen=np.asfortranarray(np.asarray(ion[sl], np.float32)).ravel(order="F"),
             elem_on=np.asarray(bundle.elem_on, np.int32), xref=np.asarray(bundle.xref, np.int32),
             abun=np.asarray(bundle.abun, np.float32), seg_meta=m6, seg_sigma=atom["op_sigma"],
             seg_off=np.concatenate([[0], np.cumsum(lens)]).astype(np.int64), brems1=atom["op_brems1"])
    t = time.perf_counter()
    O.opacity_tables(g, gaunt_cache=True)
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": "%d cells of the workload's state after the balance, scalar C restatement" % n}


def lines_cpu_baseline(bundle, Hden, Te, Ne, ion, dV, cells=500):
    """outputGas's forbVol/HbetaVol sums (the FP64 C restatement the GPU
    kernel follows, oracle mco_line_lum_f64) over the first `cells` cells of
    the same state, one thread: cells per second."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mc_oracle as O
    import numpy as np
    n = min(cells, Te.size - 1)
    f32 = lambda v: np.asarray(v, np.float32).astype(np.float64)
    B = O.Balance(bundle)
    t = time.perf_counter()
    B.line_lum(f32(Hden), f32(Te), f32(Ne), f32(ion), f32(dV), np.arange(1, n + 1), "f64")
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": "%d cells of the workload's state after the balance, scalar C restatement" % n}


def emission_cpu_baseline(atom, bundle, nb, nu, Te, Ne, ion, cells=200):
    """emissionDriver's gas part in the reference's precision (the C
    restatement, oracle mco_emission_cell_f32) on the first `cells` cells of
    the same state, one thread: cells per second."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mc_oracle as O
    import numpy as np
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
            "sample": "%d cells of the workload's state after the balance, scalar C restatement" % n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=float, default=1e8)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-te", action="store_true", help="skip the Te-convergence measurement")
    ap.add_argument("--no-balance", action="store_true", help="skip the balance measurement")
    ap.add_argument("--no-grains", action="store_true", help="skip the balance with grains")
    ap.add_argument("--no-lines", action="store_true", help="skip the line luminosities")
    ap.add_argument("--no-emission", action="store_true", help="skip the re-emission tables")
    ap.add_argument("--no-opacity", action="store_true", help="skip the opacity-table build")
    ap.add_argument("--autotune", action="store_true",
                    help="time the launch-shape candidates on this workload and store the best "
                         "in mocassin_amd/tuning_db.json (default: use the stored shape)")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shape blocks_per_cu,threads,emit_min,walk_steps")
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
    from mocassin_amd.atomic import AtomicBundle

    npk = int(a.packets)
    t0 = time.perf_counter()
    atom_path = os.path.join(ROOT, "tests", "golden", "atomic_%s.npz" % a.workload)
    atom = dict(np.load(atom_path)) if os.path.exists(atom_path) else None
    bundle = AtomicBundle.from_npz(atom) if atom else None
    p = synth.workload(a.workload, n_packets=npk, nu=None if bundle is None else bundle.nuArray)
    setup_s = time.perf_counter() - t0
    seed = 20241202
    T = Transport(p["xAxis"], p["yAxis"], p["zAxis"], p["active"], p["nCells"], p["nbins"],
                  p["nuArray"], p["R_out"], p["geoCorr"], p["ionEdge1"], [p["starPos"]],
                  [p["starIdx"]], [p["deltaE"]], symmetric=p["symmetric"], seed=seed, device=0)
    T.upload_tables(p["opacity"], p["specCDF"], p["recPDF"], p["totalLines"])
    from mocassin_amd import autotune as at
    tuning_src = "default"
    if a.tuning:
        b, th, e, w = (int(v) for v in a.tuning.split(","))
        T.set_tuning(blocks_per_cu=b, threads=th, emit_min=e, walk_steps=w)
        tuning_src = "command line"
    elif a.autotune:
        best, trials = at.autotune(T, max(1, min(npk, 10_000_000)), repeats=2)
        at.save(a.workload, best, "bench.py --autotune, %d candidates" % len(trials))
        tuning_src = "autotuned now"
        print("bench: autotuned %s (%d candidates, %.2f..%.2f ms)"
              % (best, len(trials), min(t[1] for t in trials), max(t[1] for t in trials)),
              file=sys.stderr, flush=True)
    elif at.load(a.workload):
        db = at.load(a.workload)
        T.set_tuning(**{k: db[k] for k in ("blocks_per_cu", "threads", "emit_min", "walk_steps")})
        tuning_src = "mocassin_amd/tuning_db.json"
    tuning = T.get_tuning()
    tuning["source"] = tuning_src
    print("bench: workload %s, %d cells, %d bins, %d packets/step, setup %.1f s"
          % (a.workload, p["nCells"], p["nbins"], npk, setup_s), file=sys.stderr, flush=True)

    if a.copy_mb:
        n = a.copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
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
    balance = None
    if bundle is not None and not a.no_balance:
        from mocassin_amd.converge import cold_state
        T.balance_setup(bundle)
        Te0, Ne0, ion0, Hd = cold_state(bundle, atom["Te0"][0], atom["ion0"], p["Hden"])
        js = 1e-9 / (8.0 if p["symmetric"] else 1.0)
        runs = []
        for _ in range(2):
            tb = time.perf_counter()
            r = T.update_cells(Hd, Te0, Ne0, ion0, jscale=js)
            runs.append((r[5], time.perf_counter() - tb, r))
        ms_b, wall_b, r = min(runs, key=lambda x: x[0])
        hit = (r[3][1:] & 16) == 0
        balance = {"cells": int(p["nCells"]), "reached": int(hit.sum()), "device_ms": ms_b,
                   "host_wall_ms": wall_b * 1e3, "cells_per_s": float(hit.sum()) / (ms_b * 1e-3),
                   "mean_iterations": float(r[4][1:][hit].mean()),
                   "solver": "reference schedule (iterateT), FP64, cold start TeStart=1e4 K"}
        print("bench: balance %d cells in %.1f ms (%.2f T-iterations per cell)"
              % (hit.sum(), ms_b, balance["mean_iterations"]), file=sys.stderr, flush=True)
        if not a.no_opacity and "op_seg" in atom:
            # the opacity tables of the balanced state, built where it lies
            T.opacity_setup(atom["op_seg"], atom["op_sigma"], float(atom["op_brems1"][0]))
            ms_o = min(T.opacity_build(host=False)[2] for _ in range(3))
            n3 = int(np.prod(np.shape(p["active"])))
            nslot = sum(min(z, bundle.nstages) for z in range(1, 31) if bundle.elem_on[z - 1])
            byts = 4.0 * n3 * p["nbins"] + 4.0 * n3 + 4.0 * (p["nCells"] + 1) * (3 + nslot)
            opacity = {"device_ms": ms_o, "cells_per_s": p["nCells"] / (ms_o * 1e-3),
                       "roofline": {"bound": "hbm", "achieved": byts / (ms_o * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": byts / (ms_o * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                    "algorithmic_bytes": byts, "kernel": "opacity_kernel"}}
            if not a.no_cpu:
                r2 = r
                opacity["cpu_baseline"] = opacity_cpu_baseline(atom, bundle, p["nbins"], p["nuArray"], Hd, r2[0],
                                                               r2[1], r2[2])
                opacity["gpu_over_cpu"] = opacity["cells_per_s"] / opacity["cpu_baseline"]["value"]
            balance["opacity"] = opacity
            print("bench: opacity tables %.3f ms (%.2f TB/s of table writes)"
                  % (ms_o, opacity["roofline"]["achieved"] / 1e3), file=sys.stderr, flush=True)
        if not a.no_lines and bundle.ncel > 0:
            # outputGas's analytic line luminosities of the balanced state
            ms_l = min(T.line_luminosities()[2] for _ in range(3))
            forb, hb, _ = T.line_luminosities()
            lines = {"device_ms": ms_l, "cells_per_s": p["nCells"] / (ms_l * 1e-3), "model_ions": int(bundle.ncel),
                     "lines_above_1e-3_hbeta": int(np.count_nonzero(forb > 1e-3 * hb)),
                     "kernel": "lines_kernel + lines_reduce_kernel"}
            if not a.no_cpu:
                sys.path.insert(0, os.path.join(ROOT, "oracle"))
                import mc_oracle as O
                dV, _ = O.cell_volumes(p["xAxis"], p["yAxis"], p["zAxis"], np.asarray(p["active"]),
                                       p["symmetric"])
                lines["cpu_baseline"] = lines_cpu_baseline(bundle, Hd, r[0], r[1], r[2], dV)
                lines["gpu_over_cpu"] = lines["cells_per_s"] / lines["cpu_baseline"]["value"]
            balance["lines"] = lines
            print("bench: line luminosities %.3f ms (%d model ions, %d lines above 1e-3 Hbeta)"
                  % (ms_l, bundle.ncel, lines["lines_above_1e-3_hbeta"]), file=sys.stderr, flush=True)
        if not a.no_emission and "em_ints" in atom:
            # the re-emission tables of the balanced state, built where it lies
            from mocassin_amd.atomic import emission_tables
            T.emission_setup(**emission_tables(atom))
            ms_e = min(T.emission_build(host=False)[2] for _ in range(3))
            em = {"device_ms": ms_e, "cells_per_s": p["nCells"] / (ms_e * 1e-3),
                  "kernel": "emission_kernel", "free_bound_channels": int(atom["em_ints"][9])}
            if not a.no_cpu:
                em["cpu_baseline"] = emission_cpu_baseline(atom, bundle, p["nbins"], p["nuArray"], r[0], r[1], r[2])
                em["gpu_over_cpu"] = em["cells_per_s"] / em["cpu_baseline"]["value"]
            balance["emission"] = em
            print("bench: re-emission tables %.3f ms (%d cells)" % (ms_e, p["nCells"]), file=sys.stderr, flush=True)
        if "emission" in balance and "opacity" in balance:
            # one whole iteration with every table resident: the sweep, the
            # balance of its J (from the state the last balance left), the
            # opacity and re-emission tables of the new state, and the upload
            # that keeps them on the device (iterateMC's loop body)
            stages = {}
            J_keep = T.fetch()                  # the legs below read the timed sweeps' J
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            T.zero_estimators()
            c = T.energy_packet_driver(1, npk, step * npk)
            step += 1
            stages["transport_ms"] = c["kernel_ms"]
            rb = T.update_cells(Hd, r[0], r[1], r[2], jscale=js)
            stages["balance_ms"] = rb[5]
            stages["opacity_ms"] = T.opacity_build(host=False)[2]
            stages["emission_ms"] = T.emission_build(host=False)[2]
            T.upload_tables(None, p["specCDF"])
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            T.set_estimators(*J_keep)
            del J_keep
            balance["iteration"] = dict(stages, host_wall_ms=wall * 1e3, packets=npk,
                                        note="transport, balance (reference schedule from the previous "
                                             "state), opacity and re-emission builds, table upload without "
                                             "host tables; device times per stage, wall time of the whole")
            print("bench: one iteration on the device: %.1f ms wall (transport %.1f, balance %.1f, opacity %.2f, "
                  "re-emission %.1f ms)" % (wall * 1e3, stages["transport_ms"], stages["balance_ms"],
                                           stages["opacity_ms"], stages["emission_ms"]), file=sys.stderr, flush=True)
        if not a.no_grains:
            # the same balance with silicate grains, their charge and the
            # photoelectric and collision terms, then the grain temperatures
            T.dust_setup(**synth.grain_tables(p["nuArray"], Hd))
            Td0 = np.full((p["nCells"] + 1, 11, 2), 50.0)
            runs = []
            for _ in range(2):
                tb = time.perf_counter()
                r = T.update_grains(Td0, jscale=js, Hden=Hd, Te=Te0, Ne=Ne0, ionDen=ion0)
                runs.append((r[3], time.perf_counter() - tb, r))
            ms_g, wall_g, r = min(runs, key=lambda x: x[0])
            hit = (r[1][1:] & 16) == 0
            Tg = r[0][1:, 0, 0][hit]
            balance["with_grains"] = {
                "device_ms": ms_g, "host_wall_ms": wall_g * 1e3, "cells_per_s": float(hit.sum()) / (ms_g * 1e-3),
                "mean_iterations": float(r[2][1:][hit].mean()),
                "grains": "astronomical silicate, 10 sizes, photoelectric charge and gas-grain collisions",
                "Tdust_K": [float(Tg.min()), float(np.median(Tg)), float(Tg.max())]}
            print("bench: balance with grains %.1f ms (%.2f T-iterations per cell), Tdust %.0f-%.0f K"
                  % (ms_g, balance["with_grains"]["mean_iterations"], Tg.min(), Tg.max()),
                  file=sys.stderr, flush=True)
    T.close()

    value = a.steps * npk / elapsed
    kms = sum(c["kernel_ms"] for c in counters) / len(counters)
    byts = sum(algorithmic_bytes(c, p["nbins"]) for c in counters) / len(counters)
    achieved = byts / (kms * 1e-3) / 1e9
    c0 = counters[-1]
    print("bench: %.3e packets/s, kernel %.1f ms/launch, %.3g seg/packet, %.3g abs/packet"
          % (value, kms, c0["n_seg"] / npk, c0["n_abs"] / npk), file=sys.stderr, flush=True)
    # HBM bytes per launch: calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per
    # packet of this workload (profiles/transport_c2_pmc.md) times this launch
    traffic = None
    tp = os.path.join(ROOT, "profiles", "transport_%s_traffic.json" % a.workload)
    if os.path.exists(tp):
        try:
            traffic = json.load(open(tp))["hbm_bytes_per_packet"] * npk
        except (OSError, ValueError, KeyError):
            traffic = None
    out = {
        "metric": "packets/s/GPU; wall-time to converged Te (\u00b11%) plain vs MLP warm-start",
        "value": value, "unit": "packets/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "full-abundance 45^3 octant, 800 bins, %d packets/sweep, "
                               "tables synthetic of that shape" % npk if a.workload == "c2"
                   else a.workload,
                   "cells": p["nCells"], "nbins": p["nbins"], "packets_per_step": npk,
                   "accumulators": "fp64", "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": byts,
                     "kernel": "transport_kernel", "kernel_ms": kms,
                     "segments": c0["n_seg"], "absorptions": c0["n_abs"], "escapes": c0["n_esc"]},
    }
    if balance is not None:
        out["balance"] = balance
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, seed)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    if not a.no_te:
        from mocassin_amd.converge import te_convergence
        tc = te_convergence(log=lambda m: print("bench: " + m, file=sys.stderr, flush=True))
        out["te_convergence"] = {
            "workload": tc["workload"], "well_sampled_cells": tc["well_sampled_cells"],
            "plain_ms": tc["runs"]["plain"]["device_ms"], "warm_ms": tc["runs"]["warm"]["device_ms"],
            "speedup": tc["speedup_device"], "speedup_at_tol_0.01": tc["speedup_device_tol_0.01"],
            "plain_iterations": tc["runs"]["plain"]["mean_iterations"],
            "warm_iterations": tc["runs"]["warm"]["mean_iterations"],
            "frac_within_1pct": {k: v["frac_within_1pct"] for k, v in tc["runs"].items()},
            "training_ms": tc["training"]["device_ms"],
            "note": "device time of the balance step (warm: + network inference); tolerance 1e-3 "
                    "on |heat-cool|/heat; Te_ref from a 1e-5 solve"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
