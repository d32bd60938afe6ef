"""The dust and gas workload: an H+He nebula with astronomical-silicate grains
on a full Cartesian grid with a lognormal density field, lit by a 50,000 K
blackbody.

Grid: `n`^3 lattice points over [-3e18, 3e18] cm, active within 3e18 cm of
the centre, nH lognormal (mean 300 cm^-3, sigma_ln 0.5, correlation length
`corr` cells, seeded; mocassin_amd.synth.lognormal_density).  Source at the
centre with Q(H) = 4e48 s^-1.

The physics data are those the reference reads for its dust deck on its
160-bin mesh (tests/golden, written from the unmodified reference):
  balance_dust_pe.npz  atomic data of the balance; grain absorption
                       cross-sections, emission integrals, radii, work
                       functions and masses (10 sizes, a^-3.5, 0.005-0.25 um)
  opacity_dust_pe.npz  photoionisation cross-section segments, free-free
                       factor, re-emission module data, grain scattering
                       cross-sections and sublimation temperature
  transport_c3mini.npz the Henyey-Greenstein asymmetry factor gSca per bin
                       (ph_mod.f90:1514-1537 for this grain mix)
The grain number density follows nH in the deck's proportion (its
dust-to-gas mass 0.01).  The opacity and re-emission tables of a state are
built on the device (mc_opacity_build, mc_emission_build, both with the
grain terms); the balance runs with the grain charge, photoelectric heating
and gas-grain collisions (mc_update_cells).
"""
import os

import numpy as np

from . import synth
from .atomic import AtomicBundle, emission_tables
from .converge import cold_state
from .transport import Transport

_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "tests", "golden")


def _load(name):
    return dict(np.load(os.path.join(_GOLDEN, name + ".npz")))


def segments(g):
    """Opacity segments (elem, ion, first bin, last bin) of a fixture's
    addOpacity order, the last bin clamped to the mesh."""
    nb = int(g["ints"][4])
    m = g["seg_meta"]
    last = np.maximum(m[:, 3], np.minimum(m[:, 4], nb))
    return np.stack([m[:, 0], m[:, 1], m[:, 3], last], axis=1).astype(np.int32), g["seg_sigma"], float(g["brems1"][0])


def grid(n, seed=3, packets=1e8, corr=8.0):
    """The grid and source of the workload (synth.build_problem's dict) on
    the 160-bin mesh of the reference's dust deck."""
    bg = _load("balance_dust_pe")
    dens = synth.lognormal_density(n, seed, mean=300.0, sigma_ln=0.5, corr=corr)
    return synth.build_problem(n, 3.e18, 3.e18, 300., 4.e48, 5.e4, 160, synth.H_HE, symmetric=False,
                               n_packets=packets, density=dens, nu=bg["nuArray"])


def open_context(n, seed=3, packets=1e8, corr=8.0, device=0, rng_seed=3003):
    """An open Transport with the balance, grain, opacity and re-emission
    inputs set up.  Returns a dict: T (the context; the caller closes it),
    p (the grid), state = (Hden, Te, Ne, ionDen) of the reference's cold
    start (TeStart 10,000 K and its initial ionisation), Tdust0 (nCells+1,
    nSizes+1, nSpeciesMax+1) at the deck's starting grain temperature, gSca,
    bundle (the balance's atomic data) and Ndust (nCells+1)."""
    bg, og, tg = _load("balance_dust_pe"), _load("opacity_dust_pe"), _load("transport_c3mini")
    b = AtomicBundle.from_npz(bg)
    p = grid(n, seed, packets, corr)
    if not np.array_equal(np.asarray(p["nuArray"], np.float32), np.asarray(og["nuArray"], np.float32)):
        raise ValueError("the fixtures' frequency meshes differ")
    nc = p["nCells"]
    T = Transport(p["xAxis"], p["yAxis"], p["zAxis"], p["active"], nc, p["nbins"], p["nuArray"], p["R_out"],
                  p["geoCorr"], p["ionEdge1"], [p["starPos"]], [p["starIdx"]], [p["deltaE"]], symmetric=False,
                  dust=True, gas=True, seed=rng_seed, device=device)
    try:
        T.balance_setup(b)
        T.opacity_setup(*segments(og))
        nS, nSz, nT, nsm, nSp = (int(v) for v in bg["dust_ints"][:5])
        ratio = float(np.median(bg["dust_ndust"][1:] / bg["Hden"][1:]))
        Ndust = (np.asarray(p["Hden"], np.float64) * ratio).astype(np.float32)
        Ndust[0] = 0.0
        T.dust_setup(absXsec=bg["dust_abs"].reshape(nS, nSz, -1), grainWeight=bg["dust_weight"],
                     grainAbun=bg["dust_abun"], emIntegral=bg["dust_emint"].reshape(nT, nSz, nSp).transpose(2, 1, 0),
                     nTemps=nT, nSpeciesMax=nsm, photoelectric=True, grainRadius=bg["dust_radius"],
                     grainVn=bg["dust_vn"], MsurfAtom=bg["dust_msurf"], aWeight=bg["dust_aweight"], Ndust=Ndust,
                     scaXsec=og["dust_sca"].reshape(nS, nSz, -1), TdustSublime=og["dust_tsub"])
        T.emission_setup(**emission_tables(og))
        ion0 = bg["ionDen0"].reshape(b.nstages, b.nElemUsed, -1)[:, :, 1].T
        Te0, Ne0, ion, Hd = cold_state(b, float(bg["Te0"][1]), ion0, p["Hden"])
        Td0 = np.full((nc + 1, nSz + 1, nsm + 1), float(bg["Tdust0"][(nSz + 1) * (nsm + 1)]))
    except Exception:
        T.close()
        raise
    return dict(T=T, p=p, state=(Hd, Te0, Ne0, ion), Tdust0=Td0, gSca=np.asarray(tg["gSca"], np.float32),
                bundle=b, Ndust=Ndust)
