"""Writes the committed golden vectors of the transport path.

Runs ONLY in the build container, where /root/reference exists: it drives the
unmodified reference (MOCASSIN 2.02, compiled by `make -C oracle ref` into
oracle/_ref/mcref_harness) on the small decks of tests/golden/decks.py and
stores inputs and reference outputs as compressed .npz files:

  walk_<deck>.npz       per packet: origin, direction and start cell (as
                        float/int bits), and the reference's Jste entries of
                        that packet alone with deltaE = 1 and zero opacity,
                        i.e. the sum over its visits of dS/dV per cell
                        (photon_mod.f90:1822-1833), plus its escape flag.
  transport_<name>.npz  one problem's tables (opacity, recPDF, totalLines,
                        specCDF, nuArray, axes, active, scalars) and the
                        reference's raw FP32 Jste and escapedPackets after
                        N packets with the Philox stream pinned per packet.

usage: python tests/golden/make_fixtures.py [walk|transport|view|converge|all]
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, HERE)
import decks  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "mcref_harness")
WORK = os.path.join(ROOT, "oracle", "_ref", "work")


def run_harness(deck, mode, args, env_extra=None, tag=None, overrides=None):
    wd = os.path.join(WORK, tag or deck)
    os.makedirs(os.path.join(wd, "out"), exist_ok=True)
    os.makedirs(os.path.join(wd, "output"), exist_ok=True)
    decks.write_deck(wd, deck, overrides)
    env = dict(os.environ)
    env.update(env_extra or {})
    with open(os.path.join(wd, "log_%s.txt" % mode), "w") as log:
        subprocess.run([HARNESS, mode, "out"] + [str(a) for a in args], cwd=wd, env=env,
                       stdout=log, stderr=subprocess.STDOUT, check=True)
    return os.path.join(wd, "out")


def read_problem(d):
    rd = lambda n, dt=np.float32: np.fromfile(os.path.join(d, n + ".bin"), dt)
    ints = rd("ints", np.int32)
    nx, ny, nz, nc, nb = (int(v) for v in ints[:5])
    out = dict(ints=ints, scalars=rd("scalars"), xAxis=rd("xAxis"), yAxis=rd("yAxis"),
               zAxis=rd("zAxis"), active=rd("active", np.int32), nuArray=rd("nuArray"),
               specCDF=rd("specCDF"))
    for n in ("opacity", "recPDF", "totalLines", "scaOpac", "gSca", "widFlx"):
        p = os.path.join(d, n + ".bin")
        if os.path.exists(p):
            out[n] = rd(n)
    return out, (nx, ny, nz, nc, nb)


def walk_origins(deck, prob):
    """Origins exercising the walk's special cases (as float32 bits)."""
    x, y, z = prob["xAxis"], prob["yAxis"], prob["zAxis"]
    f = np.float32
    mid = lambda a, i: (a[i + 1] + a[i]) / f(2)        # (a(i+1)+a(i))/2 in FP32
    rng = np.random.default_rng(2024)
    o = []
    if deck.startswith("c1"):                      # symmetric octant, source at (0,0,0)
        o.append((x[0], y[0], z[0], 1))            # the central source, folded
        o.append((x[0], y[0], z[0], 0))            # unfolded: immediate reflections
        o.append((x[3], y[2], z[4], 0))            # on grid points
        o.append((mid(x, 3), y[5], mid(z, 1), 0))  # on cell walls
        o.append((x[-1] * f(0.999), y[1], z[6], 0))
        o.append((x[-1], y[-1], z[-1], 0))         # outer corner
        for _ in range(3):
            p = rng.uniform(0, 1, 3).astype(np.float32) * np.array([x[-1], y[-1], z[-1]], np.float32)
            o.append((p[0], p[1], p[2], 0))
    else:                                          # full grid, source at the centre
        c = len(x) // 2
        o.append((x[c], y[c], z[c], 1))
        o.append((x[2], y[c], z[7], 1))
        o.append((mid(x, 4), mid(y, 6), z[3], 1))
        o.append((x[0], y[0], z[0], 1))            # lower corner
        o.append((x[-1], y[-1], z[-1], 1))         # upper corner
        for _ in range(3):
            p = rng.uniform(-1, 1, 3).astype(np.float32) * np.array([x[-1], y[-1], z[-1]], np.float32)
            o.append((p[0], p[1], p[2], 1))
    return o


def make_walk(deck, npk=60, seed=777):
    d = run_harness(deck, "tables", [])
    prob, (nx, ny, nz, nc, nb) = read_problem(d)
    origins = walk_origins(deck, prob)
    wd = os.path.join(WORK, deck)
    with open(os.path.join(wd, "origins.txt"), "w") as f:
        for ox, oy, oz, fold in origins:
            f.write("%.9e %.9e %.9e %d\n" % (ox, oy, oz, fold))
    # a bin above the H edge so no packet escapes before walking
    nu_bin = int(np.nonzero(prob["nuArray"] > prob["scalars"][4] * 1.5)[0][0]) + 1
    d = run_harness(deck, "walk", [seed, npk, nu_bin, os.path.join(wd, "origins.txt")])
    rows = [l.split() for l in open(os.path.join(d, "walk_dirs.txt"))]
    jrows = [l.split() for l in open(os.path.join(d, "walk_J.txt"))]
    hx = lambda s: np.uint32(int(s, 16))
    dirs = np.array([[int(r[0]), int(r[1])] + [hx(v) for v in r[2:8]] + [int(v) for v in r[8:12]]
                     + [hx(r[12])] for r in rows], np.int64)
    jm = np.array([[int(r[0]), int(r[1]), int(r[2]), hx(r[3])] for r in jrows], np.int64)
    # origins written in decimal: the harness parsed them; keep the bits it used
    out = os.path.join(HERE, "walk_%s.npz" % deck)
    np.savez_compressed(
        out, xAxis=prob["xAxis"], yAxis=prob["yAxis"], zAxis=prob["zAxis"],
        active=prob["active"], ints=prob["ints"], scalars=prob["scalars"],
        nu_bin=np.int32(nu_bin),
        origin_id=dirs[:, 0].astype(np.int32), packet=dirs[:, 1].astype(np.int32),
        pos_bits=dirs[:, 2:5].astype(np.uint32), dir_bits=dirs[:, 5:8].astype(np.uint32),
        cell=dirs[:, 8:11].astype(np.int32), n_cells=dirs[:, 11].astype(np.int32),
        esc_bits=dirs[:, 12].astype(np.uint32),
        j_origin=jm[:, 0].astype(np.int32), j_packet=jm[:, 1].astype(np.int32),
        j_cell=jm[:, 2].astype(np.int32), j_bits=jm[:, 3].astype(np.uint32))
    print("wrote", out, len(rows), "packets", len(jrows), "J entries")


def make_transport(deck, npk, seed, name, opacity_scale=None, overrides=None):
    env = {"MCREF_OPACITY_SCALE": repr(float(opacity_scale))} if opacity_scale else None
    d = run_harness(deck, "transport", [seed, npk], env_extra=env, tag=name, overrides=overrides)
    prob, (nx, ny, nz, nc, nb) = read_problem(d)
    J = np.fromfile(os.path.join(d, "Jste_raw.bin"), np.float32)
    E = np.fromfile(os.path.join(d, "esc_raw.bin"), np.float32)
    q = np.fromfile(os.path.join(d, "transport_counters.bin"), np.float32)
    nzJ = np.nonzero(J)[0]
    nzE = np.nonzero(E)[0]
    out = os.path.join(HERE, "transport_%s.npz" % name)
    view = {}
    if os.path.exists(os.path.join(d, "esc_all.bin")):
        # viewing-angle planes escapedPackets(:,:,0:nAngleBins) and their tables
        rd = lambda n, dt=np.float32: np.fromfile(os.path.join(d, n + ".bin"), dt)
        Ea = rd("esc_all")
        nzA = np.nonzero(Ea)[0]
        view = dict(Eall_idx=nzA.astype(np.int32), Eall_val=Ea[nzA], view_ints=rd("view_ints", np.int32),
                    view_real=rd("view_real"), view_Ptheta=rd("view_Ptheta", np.int32),
                    view_Pphi=rd("view_Pphi", np.int32), view_theta=rd("view_theta"), view_phi=rd("view_phi"))
    np.savez_compressed(out, seed=np.uint64(seed), npackets=np.int64(npk), **prob,
                        J_idx=nzJ.astype(np.int32), J_val=J[nzJ],
                        E_idx=nzE.astype(np.int32), E_val=E[nzE], counters=q, **view)
    print("wrote", out, "J nonzero", len(nzJ), "E nonzero", len(nzE), "counters", q)


def make_balance(deck, npk, seed, name, opacity_scale=None):
    """Reference transport (J) + one updateCell pass over every cell."""
    sys.path.insert(0, ROOT)
    from mocassin_amd.atomic import AtomicBundle
    env = {"MCREF_OPACITY_SCALE": repr(float(opacity_scale))} if opacity_scale else None
    d = run_harness(deck, "balance", [seed, npk], env_extra=env, tag="bal_" + name)
    prob, (nx, ny, nz, nc, nb) = read_problem(d)
    b = AtomicBundle.from_harness_dump(d, "/root/reference/data/radrec.dat")
    rd = lambda n, dt=np.float32: np.fromfile(os.path.join(d, n + ".bin"), dt)
    J = rd("Jste_scaled")
    nzJ = np.nonzero(J)[0]
    out = os.path.join(HERE, "balance_%s.npz" % name)
    dust = {}
    if os.path.exists(os.path.join(d, "dust_ints.bin")):
        dust = dict(dust_ints=rd("dust_ints", np.int32), dust_abs=rd("dust_abs"),
                    dust_weight=rd("dust_weight"), dust_abun=rd("dust_abun"),
                    dust_emint=rd("dust_emint"), Tdust0=rd("Tdust0"), Tdust=rd("bal_Tdust"),
                    dust_pe=rd("dust_pe", np.int32), dust_vn=rd("dust_vn"), dust_msurf=rd("dust_msurf"),
                    dust_radius=rd("dust_radius"), dust_ndust=rd("dust_ndust"),
                    dust_aweight=rd("dust_aweight"), dust_ipcoll=rd("dust_ipcoll", np.int32))
    np.savez_compressed(
        out, ints=prob["ints"], nuArray=prob["nuArray"], **b.to_npz_dict(), **dust,
        Hden=rd("Hden"), Te0=rd("Te"), Ne0=rd("Ne"), ionDen0=rd("ionDen"),
        J_idx=nzJ.astype(np.int32), J_val=J[nzJ],
        Te=rd("bal_Te"), Ne=rd("bal_Ne"), ionDen=rd("bal_ionDen"),
        conv=rd("bal_conv", np.int32), black=rd("bal_black", np.int32))
    print("wrote", out, "cells", nc, "J nonzero", len(nzJ))


def make_opacity(deck, npk, seed, name, keep_cells=96):
    """Reference transport + one updateCell pass, then the opacity tables of
    the updated state (ionizationDriver in iterateMC's order, and the dust
    add): the state, the cross-section segments addOpacity adds, and the
    resulting opacities of `keep_cells` evenly spaced cells."""
    d = run_harness(deck, "opacity", [seed, npk], tag="op_" + name)
    prob, (nx, ny, nz, nc, nb) = read_problem(d)
    rd = lambda n, dt=np.float32: np.fromfile(os.path.join(d, n + ".bin"), dt)
    seg = rd("op_seg", np.int32).reshape(-1, 6)
    xs = rd("op_xsec")
    bp, nseg, _ = (int(v) for v in rd("op_ints", np.int32))
    # the entries inOpacity reads: bins lo..max(lo, min(hi, nbins))
    sig, off = [], [0]
    for e, n, sh, lo, hi, xp in seg:
        iup = max(lo, min(hi, nb))
        sig.append(xs[xp - 1:xp - 1 + iup - lo + 1])
        off.append(off[-1] + iup - lo + 1)
    cells = np.unique(np.linspace(1, nc, min(keep_cells, nc)).astype(np.int32))
    pick = lambda a: a.reshape(nb, nc + 1).T[cells].copy()
    out = dict(ints=prob["ints"], nuArray=prob["nuArray"], Hden=rd("Hden"), Te0=rd("Te"),
               Te=rd("op_Te"), Ne=rd("op_Ne"), ionDen=rd("op_ionDen"),
               elem_on=rd("at_elem_on", np.int32), xref=rd("at_xref", np.int32), abun=rd("at_abun"),
               at_ints=rd("at_ints", np.int32), seg_meta=seg, seg_sigma=np.concatenate(sig),
               seg_off=np.array(off, np.int64), brems1=xs[bp - 1:bp], cells=cells, op_gas=pick(rd("op_gas")),
               xAxis=prob["xAxis"], yAxis=prob["yAxis"], zAxis=prob["zAxis"], active=prob["active"],
               lines_L=rd("lines_L"), lines_hb=rd("lines_hb"),
               Ne0=rd("Ne"), ionDen0=rd("ionDen"))
    J = rd("Jste_scaled")
    nzJ = np.nonzero(J)[0]
    out.update(J_idx=nzJ.astype(np.int32), J_val=J[nzJ])
    # the state files writeGrid wrote, and the state resetGrid reads back
    # from them (a second run of the harness in the same directory)
    od = os.path.join(os.path.dirname(d), "output")
    for f in ("grid0", "grid1", "grid2", "dustGrid"):
        fp = os.path.join(od, f + ".out")
        if os.path.exists(fp):
            out["txt_" + f] = np.frombuffer(open(fp, "rb").read(), np.uint8)
    d2 = run_harness(deck, "resetgrid", [], tag="op_" + name)
    rg = lambda n, dt=np.float32: np.fromfile(os.path.join(d2, n + ".bin"), dt)
    out.update(rg_ints=rg("rg_ints", np.int32), rg_active=rg("rg_active", np.int32), rg_conv=rg("rg_conv", np.int32),
               rg_black=rg("rg_black", np.int32), rg_Te=rg("rg_Te"), rg_Ne=rg("rg_Ne"), rg_Hden=rg("rg_Hden"),
               rg_ionDen=rg("rg_ionDen"), rg_xAxis=rg("rg_xAxis"))
    if os.path.exists(os.path.join(d2, "rg_Tdust.bin")):
        out.update(rg_Ndust=rg("rg_Ndust"), rg_Tdust=rg("rg_Tdust"))
    # re-emission tables of the balanced state and the module data they use
    ei = rd("em_ints", np.int32)
    nfb = int(ei[9])
    out.update(em_ints=ei, em_lyP=rd("em_lyP", np.int32), em_tk=rd("em_tk"), em_gHI=rd("em_gHI"),
               em_gHeI=rd("em_gHeI"), em_gHeII=rd("em_gHeII"), em_eHI=rd("em_eHI", np.int32),
               em_eHeI=rd("em_eHeI", np.int32), em_eHeII=rd("em_eHeII", np.int32), em_gff=rd("em_gff"),
               em_gffHeII=rd("em_gffHeII"), em_sHI=rd("em_sHI"), em_sHeI=rd("em_sHeI"), em_sHeII=rd("em_sHeII"),
               em_fbm=rd("em_fbm", np.int32)[:4 * nfb], em_fbg=rd("em_fbg"), em_fbs=rd("em_fbs"),
               em_ydat=rd("em_ydat"), em_Aydat=rd("em_Aydat"), em_lyData=rd("em_lyData"), em_lyNu=rd("em_lyNu"),
               em_HeIc=rd("em_HeIc"), em_HIrec=rd("em_HIrec", np.float64), em_HeIIrec=rd("em_HeIIrec", np.float64),
               widFlx=rd("widFlx"), em_totalLines=rd("em_totalLines"),
               em_recPDF=pick(rd("em_recPDF")))
    if os.path.exists(os.path.join(d, "cold_cells.bin")):
        # emissionDriver of four cells at 1, 5, 19.554 and 60 K (raw rows,
        # NaN entries included)
        out.update(cold_cells=rd("cold_cells", np.int32), cold_Te=rd("cold_Te"),
                   cold_recPDF=rd("cold_recPDF").reshape(nb, -1).T.copy(), cold_totalLines=rd("cold_totalLines"))
    if os.path.exists(os.path.join(d, "op_total.bin")):
        out.update(Tdust=rd("op_Tdust"), op_total=pick(rd("op_total")), op_sca=pick(rd("op_sca")),
                   dust_ints=rd("dust_ints", np.int32), dust_abs=rd("dust_abs"), dust_sca=rd("dust_sca"),
                   dust_weight=rd("dust_weight"), dust_abun=rd("dust_abun"), dust_tsub=rd("dust_tsub"),
                   dust_ndust=rd("dust_ndust"))
    path = os.path.join(HERE, "opacity_%s.npz" % name)
    np.savez_compressed(path, **out)
    print("wrote", path, "cells", nc, "segments", nseg, "bins", nb)


def make_atomic(deck, name):
    """Atomic inputs of the balance for a workload's frequency mesh and
    composition, with the reference's initial cell state (TeStart and its
    initial ionisation fractions of one cell) as the cold start."""
    sys.path.insert(0, ROOT)
    from mocassin_amd.atomic import AtomicBundle
    d = run_harness(deck, "opacity", [5, 2000], tag="atom_" + name)
    prob, (nx, ny, nz, nc, nb) = read_problem(d)
    b = AtomicBundle.from_harness_dump(d, "/root/reference/data/radrec.dat")
    rd = lambda n, dt=np.float32: np.fromfile(os.path.join(d, n + ".bin"), dt)
    ion0 = rd("ionDen").reshape(b.nstages, b.nElemUsed, nc + 1)[:, :, 1].T.copy()
    # the opacity segments of this mesh and composition (addOpacity's order)
    seg = rd("op_seg", np.int32).reshape(-1, 6)
    xs = rd("op_xsec")
    bp = int(rd("op_ints", np.int32)[0])
    last = np.maximum(seg[:, 3], np.minimum(seg[:, 4], nb))
    sig = np.concatenate([xs[xp - 1:xp - 1 + l - lo + 1] for (_, _, _, lo, _, xp), l in zip(seg, last)])
    out = os.path.join(HERE, "atomic_%s.npz" % name)
    np.savez_compressed(out, **b.to_npz_dict(), Te0=rd("Te")[1:2], ion0=ion0,
                        op_seg=np.stack([seg[:, 0], seg[:, 1], seg[:, 3], last], axis=1).astype(np.int32),
                        op_sigma=sig, op_brems1=xs[bp - 1:bp], **_emission_data(rd))
    print("wrote", out, "nbins", nb, "channels", b.nch, "CEL ions", b.ncel)


def _emission_data(rd):
    """The module data of the re-emission tables on this mesh (harness dump)."""
    ei = rd("em_ints", np.int32)
    return dict(em_ints=ei, em_lyP=rd("em_lyP", np.int32), em_tk=rd("em_tk"), em_gHI=rd("em_gHI"),
                em_gHeI=rd("em_gHeI"), em_gHeII=rd("em_gHeII"), em_eHI=rd("em_eHI", np.int32),
                em_eHeI=rd("em_eHeI", np.int32), em_eHeII=rd("em_eHeII", np.int32), em_gff=rd("em_gff"),
                em_gffHeII=rd("em_gffHeII"), em_sHI=rd("em_sHI"), em_sHeI=rd("em_sHeI"), em_sHeII=rd("em_sHeII"),
                em_fbm=rd("em_fbm", np.int32)[:4 * int(ei[9])], em_fbg=rd("em_fbg"), em_fbs=rd("em_fbs"),
                em_ydat=rd("em_ydat"), em_Aydat=rd("em_Aydat"), em_lyData=rd("em_lyData"), em_lyNu=rd("em_lyNu"),
                em_HeIc=rd("em_HeIc"), em_HIrec=rd("em_HIrec", np.float64),
                em_HeIIrec=rd("em_HeIIrec", np.float64), widFlx=rd("widFlx"))


def make_converge(deck, name, seeds, overrides, ranks=8):
    """The reference's MCIterationDriver on `deck` (with `overrides`, e.g.
    maxIterateMC) to its own stopping point, once per seed under mpiexec with
    the random stream of rank r pinned to (seed, 0, r): the final Te, Ne,
    ionDen, lgConverged and lgBlack of every run, as the reference's
    resetGrid reads them back from the grid files writeGrid left.  With them
    the grid, source, initial state and atomic data the device loop starts
    from (one harness pass of the same deck)."""
    sys.path.insert(0, ROOT)
    from mocassin_amd.atomic import AtomicBundle
    d = run_harness(deck, "opacity", [5, 2000], tag="conv_atom_" + name, overrides=overrides)
    prob, (nx, ny, nz, nc, nb) = read_problem(d)
    b = AtomicBundle.from_harness_dump(d, "/root/reference/data/radrec.dat")
    rd = lambda n, dt=np.float32: np.fromfile(os.path.join(d, n + ".bin"), dt)
    seg = rd("op_seg", np.int32).reshape(-1, 6)
    xs = rd("op_xsec")
    bp = int(rd("op_ints", np.int32)[0])
    last = np.maximum(seg[:, 3], np.minimum(seg[:, 4], nb))
    sig = np.concatenate([xs[xp - 1:xp - 1 + l - lo + 1] for (_, _, _, lo, _, xp), l in zip(seg, last)])
    runs = {k: [] for k in ("Te", "Ne", "ionDen", "conv", "black")}
    dust = {}
    if os.path.exists(os.path.join(d, "dust_ints.bin")):
        # grains: the tables getDustT, the grain charge and the dust add
        # read, and the deck's starting grain temperatures
        dust = dict(dust_ints=rd("dust_ints", np.int32), dust_abs=rd("dust_abs"), dust_sca=rd("dust_sca"),
                    dust_weight=rd("dust_weight"), dust_abun=rd("dust_abun"), dust_tsub=rd("dust_tsub"),
                    dust_ndust=rd("dust_ndust"), dust_emint=rd("dust_emint"), dust_pe=rd("dust_pe", np.int32),
                    dust_vn=rd("dust_vn"), dust_msurf=rd("dust_msurf"), dust_radius=rd("dust_radius"),
                    dust_aweight=rd("dust_aweight"), dust_ipcoll=rd("dust_ipcoll", np.int32), Tdust0=rd("Tdust0"))
        runs["Tdust"] = []
    for seed in seeds:
        wd = os.path.join(WORK, "conv_%s_%d" % (name, seed))
        os.makedirs(os.path.join(wd, "out"), exist_ok=True)
        os.makedirs(os.path.join(wd, "output"), exist_ok=True)
        decks.write_deck(wd, deck, overrides)
        with open(os.path.join(wd, "log_iterate.txt"), "w") as log:
            # the reference stops in outputGas on its missing e1bx.d data
            # after writeGrid has written the converged state
            subprocess.run(["/opt/conda/bin/mpiexec", "-n", str(ranks), HARNESS, "iterate", "out", str(seed)],
                           cwd=wd, stdout=log, stderr=subprocess.STDOUT, timeout=3600)
        with open(os.path.join(wd, "log_resetgrid.txt"), "w") as log:
            subprocess.run([HARNESS, "resetgrid", "out"], cwd=wd, stdout=log, stderr=subprocess.STDOUT, check=True)
        rg = lambda n, dt=np.float32: np.fromfile(os.path.join(wd, "out", n + ".bin"), dt)
        runs["Te"].append(rg("rg_Te"))
        runs["Ne"].append(rg("rg_Ne"))
        runs["ionDen"].append(rg("rg_ionDen"))
        runs["conv"].append(rg("rg_conv", np.int32))
        runs["black"].append(rg("rg_black", np.int32))
        if dust:
            runs["Tdust"].append(rg("rg_Tdust"))
        summ = open(os.path.join(wd, "output", "summary.out")).read().count("energy packets used")
        print("converged run", name, "seed", seed, "iterations", summ, flush=True)
    out = os.path.join(HERE, "converge_%s.npz" % name)
    prob = {k: v for k, v in prob.items() if k != "widFlx"}       # _emission_data holds it
    np.savez_compressed(out, **prob, **b.to_npz_dict(), Hden=rd("Hden"), Te0=rd("Te"), Ne0=rd("Ne"),
                        ionDen0=rd("ionDen"), seeds=np.asarray(seeds, np.int64), ranks=np.int32(ranks),
                        iterations=np.int32(summ),
                        op_seg=np.stack([seg[:, 0], seg[:, 1], seg[:, 3], last], axis=1).astype(np.int32),
                        op_sigma=sig, op_brems1=xs[bp - 1:bp], **_emission_data(rd), **dust,
                        **{"ref_" + k: np.stack(v) for k, v in runs.items()})
    print("wrote", out, "cells", nc, "bins", nb, "seeds", len(seeds))


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("walk", "all"):
        make_walk("c1mini")
        make_walk("full_mini")
    if what in ("balance", "all"):
        make_balance("c1mini", 20000, 31, "c1mini")
        make_balance("c2mini", 20000, 4242, "c2mini")
        make_balance("c2mini", 20000, 77, "c2mini_thick", opacity_scale=300.0)
    if what in ("atomic", "all"):
        make_atomic("c1atom", "c1")
        make_atomic("c2atom", "c2")
    if what in ("dust", "all"):
        make_balance("dust_bal", 20000, 1618, "dust_bal")
        make_balance("dust_pe", 20000, 1618, "dust_pe")
        make_transport("dust_mini", 8000, 2718, "dust_mini")
        make_transport("dust_mini", 4000, 3141, "dust_mini_thick", opacity_scale=30.0)
        # config 3 at reduced size: lognormal density file, dust + gas, HG
        make_transport("c3mini", 8000, 6060, "c3mini")
    if what in ("opacity", "all"):
        make_opacity("c2mini", 20000, 4243, "c2mini")
        make_opacity("dust_pe", 20000, 1619, "dust_pe")
    if what in ("view", "all"):
        # two inclinations with phi (non-symmetric grid), and phi switched off
        # by a negative phi on the symmetric octant (grid_mod.f90:419-427)
        make_transport("full_mini", 8000, 4343, "full_mini_view", opacity_scale=100.0,
                       overrides={"inclination": "2 0.6 0.9 2.2 5.5"})
        make_transport("c1mini", 8000, 4444, "c1mini_view", opacity_scale=300.0,
                       overrides={"inclination": "2 0.4 -1. 1.3 -1."})
    if what in ("converge", "all"):
        # iterated to the reference's stopping point: 20 iterations of 1e7
        # packets on the 19^3 H+He octant, 12 of 2e6 on the 9^3 full-abundance
        # one (per-cell Monte Carlo noise of Te near 0.2 % per run)
        make_converge("c1", "c1", [101, 202, 303, 404], {"maxIterateMC": "20 95.", "nPhotons": "10000000"})
        make_converge("c2mini", "c2mini", [101, 202, 303, 404], {"maxIterateMC": "12 95.", "nPhotons": "2000000"})
    if what in ("converge", "converge_dust", "all"):
        # the dust and gas deck (15^3 lognormal densityFile, silicate grains
        # with the photoelectric effect, Henyey-Greenstein scattering)
        make_converge("c3mini", "c3mini", [101, 202, 303, 404, 505, 606, 707, 808],
                      {"maxIterateMC": "12 95.", "nPhotons": "4000000"})
    if what in ("transport", "all"):
        make_transport("c1mini", 20000, 12345, "c1mini")
        make_transport("c1mini", 8000, 999, "c1mini_thick", opacity_scale=300.0)
        make_transport("full_mini", 8000, 4242, "full_mini_thick", opacity_scale=100.0)
