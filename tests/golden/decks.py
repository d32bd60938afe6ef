"""Input decks (MOCASSIN `input/input.in` keyword files) for the parity and
benchmark workloads.

The keywords and their meaning are those parsed by the reference's readInput
(set_input_mod.f90:143-573).  `write_deck(dirpath, name)` writes
`input/input.in` and the abundance file under `dirpath`; the reference
harness and the build's host read the same files.
"""
import os

import numpy as np

# log10(X/H) by number for H, He, C, N, O, Ne, S (other elements absent)
FULL_GAS = {1: 0.0, 2: -1.0, 6: -3.6, 7: -4.0, 8: -3.3, 10: -4.0, 16: -5.0}
H_HE = {1: 0.0, 2: -1.0}


def abundance_lines(abun):
    out = []
    for z in range(1, 31):
        out.append("%.6e" % (10.0 ** abun[z]) if z in abun else "0.")
    return "\n".join(out) + "\n"


# name -> (keywords, abundances)
DECKS = {
    # H+He octant, one sweep: the CPU-runnable reference workload
    "c1": ([
        "symmetricXYZ", "nx 19", "ny 19", "nz 19",
        "edges 3.e18 3.e18 3.e18", "Rin 0.", "Rout 3.e18",
        "Hdensity 100.", "contShape blackbody", "TStellar 40000.",
        "LPhot 4.e12", "nbins 400", "nuMin 1.e-5", "nuMax 25.",
        "nPhotons 1000000", "maxIterateMC 1 100.", "TeStart 10000.",
    ], H_HE),
    # small H+He octant used for committed golden vectors
    "c1mini": ([
        "symmetricXYZ", "nx 9", "ny 9", "nz 9",
        "edges 3.e18 3.e18 3.e18", "Rin 0.", "Rout 3.e18",
        "Hdensity 100.", "contShape blackbody", "TStellar 40000.",
        "LPhot 4.e12", "nbins 160", "nuMin 1.e-5", "nuMax 25.",
        "nPhotons 20000", "maxIterateMC 1 100.", "TeStart 10000.",
    ], H_HE),
    # full-abundance ionisation-bounded octant (the packets/s workload)
    "c2": ([
        "symmetricXYZ", "nx 45", "ny 45", "nz 45",
        "edges 2.5e18 2.5e18 2.5e18", "Rin 0.", "Rout 2.5e18",
        "Hdensity 1000.", "contShape blackbody", "TStellar 40000.",
        "LPhot 1.e13", "nbins 800", "nuMin 1.e-5", "nuMax 25.", "nstages 5",
        "nPhotons 100000000", "maxIterateMC 30 95.", "TeStart 10000.",
    ], FULL_GAS),
    # small full-abundance octant for golden vectors
    "c2mini": ([
        "symmetricXYZ", "nx 9", "ny 9", "nz 9",
        "edges 2.5e18 2.5e18 2.5e18", "Rin 0.", "Rout 2.5e18",
        "Hdensity 1000.", "contShape blackbody", "TStellar 40000.",
        "LPhot 1.e13", "nbins 200", "nuMin 1.e-5", "nuMax 25.", "nstages 5",
        "nPhotons 20000", "maxIterateMC 1 95.", "TeStart 10000.",
    ], FULL_GAS),
    # the atomic data of the c1 and c2 workloads (frequency mesh, composition,
    # stages) taken from a 5^3 grid of the same keywords
    "c1atom": ([
        "symmetricXYZ", "nx 5", "ny 5", "nz 5",
        "edges 3.e18 3.e18 3.e18", "Rin 0.", "Rout 3.e18",
        "Hdensity 100.", "contShape blackbody", "TStellar 40000.",
        "LPhot 4.e12", "nbins 400", "nuMin 1.e-5", "nuMax 25.",
        "nPhotons 2000", "maxIterateMC 1 100.", "TeStart 10000.",
    ], H_HE),
    "c2atom": ([
        "symmetricXYZ", "nx 5", "ny 5", "nz 5",
        "edges 2.5e18 2.5e18 2.5e18", "Rin 0.", "Rout 2.5e18",
        "Hdensity 1000.", "contShape blackbody", "TStellar 40000.",
        "LPhot 1.e13", "nbins 800", "nuMin 1.e-5", "nuMax 25.", "nstages 5",
        "nPhotons 2000", "maxIterateMC 1 95.", "TeStart 10000.",
    ], FULL_GAS),
    # full (non-symmetric) grid, odd size: exercises the non-symmetric walk
    "full_mini": ([
        "nx 11", "ny 11", "nz 11",
        "edges 3.e18 3.e18 3.e18", "Rin 0.", "Rout 3.e18",
        "Hdensity 100.", "contShape blackbody", "TStellar 40000.",
        "LPhot 4.e12", "nbins 160", "nuMin 1.e-5", "nuMax 25.",
        "nPhotons 20000", "maxIterateMC 1 100.", "TeStart 10000.",
    ], H_HE),
}


def grain_sizes(n=10, amin=0.005, amax=0.25, p=3.5):
    """Grain-size file (makeDustXsec reads: n, then index, radius [um], weight):
    n log-spaced radii with weights a^-p."""
    step = (np.log10(amax) - np.log10(amin)) / (n - 1)
    rows = ["%d sizes" % n]
    for i in range(n):
        a = 10.0 ** (np.log10(amin) + i * step)
        rows.append("%d %.6e %.6e" % (i + 1, a, a ** (-p)))
    return "\n".join(rows) + "\n"


SILICATE = "1\n'dustData/sil-dlaor.nk' 1.0\n"

# dust + gas decks: keywords, abundances, extra input files
DUST_DECKS = {
    # small full grid, H+He gas with astronomical silicate (10 sizes, a^-3.5),
    # dust-to-gas mass 0.01, Henyey-Greenstein scattering
    "dust_mini": ([
        "nx 9", "ny 9", "nz 9",
        "edges 3.e18 3.e18 3.e18", "Rin 0.", "Rout 3.e18",
        "Hdensity 300.", "contShape blackbody", "TStellar 50000.",
        "LPhot 4.e12", "nbins 160", "nuMin 1.e-5", "nuMax 25.",
        "nPhotons 20000", "maxIterateMC 1 100.", "TeStart 10000.",
        "dustFile 'input/species.dat' 'input/sizes.dat'", "MdMg constant 0.01",
    ], H_HE, {"species.dat": SILICATE, "sizes.dat": None}),
    # the same without the grain photoelectric effect (grain charge and
    # photoelectric heating are outside the balance path)
    "dust_bal": ([
        "nx 9", "ny 9", "nz 9",
        "edges 3.e18 3.e18 3.e18", "Rin 0.", "Rout 3.e18",
        "Hdensity 300.", "contShape blackbody", "TStellar 50000.",
        "LPhot 4.e12", "nbins 160", "nuMin 1.e-5", "nuMax 25.",
        "nPhotons 20000", "maxIterateMC 1 100.", "TeStart 10000.",
        "dustFile 'input/species.dat' 'input/sizes.dat'", "MdMg constant 0.01", "noPhotoelectric",
    ], H_HE, {"species.dat": SILICATE, "sizes.dat": None}),
    # the same with the reference's default grain photoelectric effect: grain
    # charge, photoelectric heating and gas-grain collisions
    "dust_pe": ([
        "nx 9", "ny 9", "nz 9",
        "edges 3.e18 3.e18 3.e18", "Rin 0.", "Rout 3.e18",
        "Hdensity 300.", "contShape blackbody", "TStellar 50000.",
        "LPhot 4.e12", "nbins 160", "nuMin 1.e-5", "nuMax 25.",
        "nPhotons 20000", "maxIterateMC 1 100.", "TeStart 10000.",
        "dustFile 'input/species.dat' 'input/sizes.dat'", "MdMg constant 0.01",
    ], H_HE, {"species.dat": SILICATE, "sizes.dat": None}),
}


def density_file(n, edge, seed, mean=300.0, sigma_ln=0.5, corr=8.0):
    """densityFile contents (set_input_mod.f90:298-310, read by setMotherGrid,
    grid_mod.f90:956-1010): a '# nx ny nz' header, then 'x y z nH' for every
    lattice point, x outer and z inner; the axes are the automatic full-grid
    axes (grid_mod.f90:582-595) and nH the seeded lognormal field of
    mocassin_amd.synth.lognormal_density."""
    import sys
    sys.path.insert(0, os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "..")))
    from mocassin_amd import synth
    ax = synth.axes_full(n, edge)
    dens = synth.lognormal_density(n, seed, mean, sigma_ln, corr)
    rows = ["# %d %d %d" % (n, n, n)]
    for i in range(n):
        for j in range(n):
            for k in range(n):
                rows.append("%.9e %.9e %.9e %.9e" % (ax[i], ax[j], ax[k], dens[i, j, k]))
    return "\n".join(rows) + "\n"


# config 3 at reduced size: the 60^3 lognormal field's 8-cell correlation
# length scaled to the 15^3 grid
C3MINI_DENSITY = dict(n=15, edge=3.e18, seed=60, corr=2.0)
DUST_DECKS["c3mini"] = ([
    "nx 15", "ny 15", "nz 15",
    "densityFile 'input/density.dat'", "Rin 0.", "Rout 3.e18",
    "contShape blackbody", "TStellar 50000.",
    "LPhot 4.e12", "nbins 160", "nuMin 1.e-5", "nuMax 25.",
    "nPhotons 20000", "maxIterateMC 1 100.", "TeStart 10000.",
    "dustFile 'input/species.dat' 'input/sizes.dat'", "MdMg constant 0.01",
], H_HE, {"species.dat": SILICATE, "sizes.dat": None,
          "density.dat": lambda: density_file(**C3MINI_DENSITY)})


def write_deck(dirpath, name, overrides=None):
    files = {}
    if name in DUST_DECKS:
        keys, abun, files = DUST_DECKS[name]
    else:
        keys, abun = DECKS[name]
    keys = list(keys)
    if overrides:
        for k, v in overrides.items():
            keys = [x for x in keys if x.split()[0] != k]
            if v is not None:
                keys.append("%s %s" % (k, v))
    os.makedirs(os.path.join(dirpath, "input"), exist_ok=True)
    os.makedirs(os.path.join(dirpath, "output"), exist_ok=True)
    with open(os.path.join(dirpath, "input", "abun.in"), "w") as f:
        f.write(abundance_lines(abun))
    with open(os.path.join(dirpath, "input", "input.in"), "w") as f:
        f.write("\n".join(keys + ['nebComposition "input/abun.in"']) + "\n")
    for fn, text in files.items():
        with open(os.path.join(dirpath, "input", fn), "w") as f:
            f.write(grain_sizes() if text is None else (text() if callable(text) else text))
    return os.path.join(dirpath, "input", "input.in")
