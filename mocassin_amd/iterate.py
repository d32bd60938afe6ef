"""The Monte Carlo iteration of iterateMC on one device.

Each iteration does what iterateMC does between two convergence checks
(iteration_mod.f90:81-939):
  1. the opacity tables of the current state (ionizationDriver per cell,
     117-228) and its re-emission tables (emissionDriver per cell, 229-400),
     built on the device where the state lies;
  2. one transport sweep of every source (energyPacketDriver, 458-724) into
     zeroed estimators;
  3. the per-cell balance of every cell from the state the previous
     iteration left (updateCell, 744-939), with J scaled by 1e-9 (and 1/8 on
     a symmetric octant);
  4. the share of converged cells: lgConverged counted over the active
     cells, unreached cells included as not converged (953-1001); the loop
     stops when it reaches minConvergence per cent, or after maxIterateMC
     iterations (1123-1186).
The state stays where the host keeps it (the reference's grid arrays): the
balance takes and returns it; the tables never leave the device.

iterate() runs the loop on an open Transport context that has its balance,
opacity and re-emission inputs set up (balance_setup, opacity_setup,
emission_setup).  With warm_start the balance of each iteration starts every
reached cell from the network's prediction (set_mlp first), falling back to
the previous state as the plain balance starts.
"""
import time

import numpy as np


def jscale_of(symmetric):
    """J units of the balance: 1e-9 (e36 erg/s over e45 cm^3 times cm), and
    1/8 on a symmetric octant (iteration_mod.f90:700-722)."""
    return 1e-9 / (8.0 if symmetric else 1.0)


def converged_percent(flags, active_cells):
    """convPercent of iterateMC: lgConverged (flags bit 0) over the active
    cells, unreached cells counted as not converged."""
    return 100.0 * float(np.count_nonzero(flags[1:] & 1)) / float(active_cells)


def iterate(T, state, packets, specCDF, symmetric, max_iter=30, min_convergence=95.0, packet_offset=0,
            balance=None, crossings=False, stop=None, log=None, Tdust=None, gSca=None):
    """Iterate transport and balance on the device from `state` = (Hden, Te,
    Ne, ionDen) (reference layouts, (nCells+1,) and (nCells+1, nElemUsed,
    nstages)).  packets: per sweep of source 1; packet_offset: the first
    packet id (iteration k uses [offset + k packets, offset + (k+1) packets)).
    balance: keyword options of Transport.update_cells (solver, tol,
    warm_start, ...).  crossings: count the segments walked in each cell in
    the last sweep.  stop(k, state, flags) -> bool ends the loop early (after
    its balance).  Tdust (nCells+1, nSizes+1, nSpeciesMax+1): grains (the
    context's dust_setup, with gSca for Henyey-Greenstein scattering); the
    tables then carry the dust terms, and each balance is updateCell with
    grains (update_grains), Tdust included in the state passed to stop and
    returned as state[4].  Returns (state, flags, records, crossings or
    None); one record per iteration with its device times, host wall time,
    counters and convergence."""
    Hden, Te, Ne, ion = state[:4]
    dust = Tdust is not None
    if dust:
        Tdust = np.array(Tdust, np.float64)
    Te = np.array(Te, np.float64)
    Ne = np.array(Ne, np.float64)
    ion = np.array(ion, np.float64)
    js = jscale_of(symmetric)
    opts = dict(balance or {})
    records = []
    flags = None
    cross = None
    spec = np.asarray(specCDF, np.float32)
    if crossings:
        T.count_crossings(True)
    for k in range(max_iter):
        t0 = time.perf_counter()
        host = k == 0
        st = (Hden, Te, Ne, ion) if host else (None, None, None, None)
        td = Tdust if (dust and host) else None
        ms_o = T.opacity_build(*st, Tdust=td, host=False)[2]
        ms_e = T.emission_build(*st, Tdust=td, host=False)[2]
        T.upload_tables(None, spec, gSca=gSca)
        T.zero_estimators()
        c = T.energy_packet_driver(1, packets, packet_offset + k * packets)
        if dust:
            Tdust, flags, nIt, ms_b, Te, Ne, ion = T.update_grains(Tdust, jscale=js, Hden=Hden, Te=Te, Ne=Ne,
                                                                   ionDen=ion, **opts)
        else:
            Te, Ne, ion, flags, nIt, ms_b = T.update_cells(Hden, Te, Ne, ion, jscale=js, **opts)
        wall = time.perf_counter() - t0
        hit = (flags[1:] & 16) == 0
        rec = {"iteration": k + 1, "opacity_ms": ms_o, "emission_ms": ms_e, "transport_ms": c["kernel_ms"],
               "balance_ms": ms_b, "wall_ms": wall * 1e3, "converged_pct": converged_percent(flags, T.nCells),
               "no_hit_pct": 100.0 * float(np.count_nonzero(~hit)) / T.nCells,
               "mean_T_iterations": float(nIt[1:][hit].mean()) if hit.any() else 0.0,
               "segments_per_packet": c["n_seg"] / packets, "absorptions_per_packet": c["n_abs"] / packets,
               "escapes_per_packet": c["n_esc"] / packets}
        records.append(rec)
        if log:
            log("iteration %d: %.1f%% converged, %.1f ms wall (transport %.1f, balance %.1f, opacity %.2f, "
                "re-emission %.2f ms)" % (k + 1, rec["converged_pct"], rec["wall_ms"], rec["transport_ms"],
                                         ms_b, ms_o, ms_e))
        cur = (Hden, Te, Ne, ion, Tdust) if dust else (Hden, Te, Ne, ion)
        if stop is not None and stop(k + 1, cur, flags):
            break
        if rec["converged_pct"] >= min_convergence:
            break
    if crossings:
        cross = T.crossings()
        T.count_crossings(False)
    return ((Hden, Te, Ne, ion, Tdust) if dust else (Hden, Te, Ne, ion)), flags, records, cross
