"""Atomic inputs of the per-cell balance, in the layout the balance kernel
(and the CPU oracle) consume.

The Fortran host already holds all of these after its setup (readData,
makeElements, initXSecArray): photoionisation cross sections on the
frequency mesh and their index ranges, outer-shell statistical weights,
radiative and dielectronic recombination fits, charge-exchange fits and the
CEL model ions (levels, A values, collision strengths).  This module packs
them as numpy arrays:

    ch_meta  int32 [nch, 8]     elem, ion, g0, g1, IPnuP, hiRate, hiHeat, outShell
    ch_sigma f32   [nch, nbins] cross sections on nuArray (0 outside)
    rrec     f32   [30, 30, 2]  radiative recombination power laws rrec(1:2, z, n)
    rnew     f32   [30, 30, 4]  Verner & Ferland fits rnew(1:4, z, n)
    fe       f32   [10, 3]      Fe (n = 4..13) fits of radRecFit
    dr_int/dr_real, ap_int/ap_real   dielectronic fit tables (25 and 167 rows)
    cel_meta int32 [ncel, 6]    elem, ion, ii, jj, nlev, ntemp  (forLines order)
    cel_dbl  f64  concatenated e(nl), A(nl,nl), qom(nt,nl,nl) per model ion
    cel_flt  f32  concatenated logT(nt), g(nl), br, z, a_r(4), a_d(5)

`read_radrec` parses the host's radrec.dat in the record order radRecFit
reads it (update_mod.f90:1704-1752).
"""
import os

import numpy as np

# Fe I-X radiative recombination fits of radRecFit (update_mod.f90:1737-1743),
# rows n = 4..13 (number of electrons), columns: a, b, c of a/t^(b + c log10 t)
FE_FITS = np.array([
    [4.33e-10, 3.91e-10, 3.49e-10, 3.16e-10, 2.96e-10, 2.59e-10, 2.24e-10, 1.91e-10, 1.68e-10, 1.46e-10],
    [0.531, 0.523, 0.521, 0.534, 0.557, 0.567, 0.579, 0.601, 0.602, 0.597],
    [5.77e-02, 6.15e-02, 6.22e-02, 6.02e-02, 5.79e-02, 5.65e-02, 5.49e-02, 5.10e-02, 5.07e-02, 5.22e-02],
], np.float32).T.copy()


def read_radrec(path):
    """rrec [30,30,2] and rnew [30,30,4] (z-1, n-1, coefficient) from radrec.dat,
    read in radRecFit's record order."""
    rows = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line:
                rows.append([float(v) for v in line.replace(",", " ").split()])
    it = iter(rows)
    rrec = np.zeros((30, 30, 2), np.float32)
    rnew = np.zeros((30, 30, 4), np.float32)
    for n in range(4, 31):
        if n != 11:
            for z in range(n, 31):
                if z != 26 or n >= 14:
                    rrec[z - 1, n - 1, :] = next(it)[:2]
    for n in range(1, 4):
        for z in range(n, 31):
            rnew[z - 1, n - 1, :] = next(it)[:4]
    for z in range(11, 31):
        rnew[z - 1, 10, :] = next(it)[:4]
    for n in range(4, 11):
        for z in range(max(6, n), 9):
            rnew[z - 1, n - 1, :] = next(it)[:4]
        rnew[9, n - 1, :] = next(it)[:4]
    for n in range(12, 27):
        rnew[25, n - 1, :] = next(it)[:4]
    return rrec, rnew


class AtomicBundle:
    """Atomic inputs of the balance for one frequency mesh and composition."""

    FIELDS = ("nuArray", "elem_on", "xref", "abun", "ch_meta", "ch_sigma", "rrec", "rnew", "fe",
              "dr_int", "dr_real", "ap_int", "ap_real", "cel_meta", "cel_dbl", "cel_flt")

    def __init__(self, nstages, nElemUsed, XHILimit=0.05, **arrays):
        self.nstages, self.nElemUsed = int(nstages), int(nElemUsed)
        self.XHILimit = float(XHILimit)
        for k in self.FIELDS:
            setattr(self, k, arrays[k])
        self.nbins = int(self.nuArray.size)
        self.nch = int(self.ch_meta.shape[0])
        self.ncel = int(self.cel_meta.shape[0])

    @classmethod
    def from_npz(cls, d, prefix="at_"):
        kw = {k: d[prefix + k] for k in cls.FIELDS}
        sc = d[prefix + "scalars"]
        return cls(int(sc[0]), int(sc[1]), float(sc[2]), **kw)

    def to_npz_dict(self, prefix="at_"):
        out = {prefix + k: getattr(self, k) for k in self.FIELDS}
        out[prefix + "scalars"] = np.array([self.nstages, self.nElemUsed, self.XHILimit])
        return out

    @classmethod
    def from_harness_dump(cls, d, radrec_path, XHILimit=0.05):
        rd = lambda n, dt=np.float32: np.fromfile(os.path.join(d, n + ".bin"), dt)
        hdr = rd("at_ints", np.int32)
        nstages, nElemUsed, nch, ncel = int(hdr[0]), int(hdr[1]), int(hdr[3]), int(hdr[4])
        nu = rd("nuArray")
        nb = nu.size
        rrec, rnew = read_radrec(radrec_path)
        return cls(nstages, nElemUsed, XHILimit,
                   nuArray=nu, elem_on=rd("at_elem_on", np.int32), xref=rd("at_xref", np.int32),
                   abun=rd("at_abun"), ch_meta=rd("ch_meta", np.int32).reshape(nch, 8),
                   ch_sigma=rd("ch_sigma").reshape(nch, nb), rrec=rrec, rnew=rnew, fe=FE_FITS,
                   dr_int=rd("dr_int", np.int32), dr_real=rd("dr_real"),
                   ap_int=rd("ap_int", np.int32), ap_real=rd("ap_real"),
                   cel_meta=rd("cel_meta", np.int32).reshape(ncel, 6),
                   cel_dbl=rd("cel_dbl", np.float64), cel_flt=rd("cel_flt"))


def emission_tables(d):
    """Keyword arguments of Transport.emission_setup (mc_emission_tables)
    from arrays saved with an em_ prefix: the module data emissionDriver
    reads, as the reference's setup leaves them (tkGamma, the log gamma
    tables and their edge bins, Gaunt factors, the Milne cross-sections from
    their first bins, the heavy-ion free-bound channels, the two-photon and
    recombination-line data) and widFlx."""
    ei = d["em_ints"]
    nfb = int(ei[9])
    return dict(tkGamma=d["em_tk"], logGammaHI=d["em_gHI"], logGammaHeI=d["em_gHeI"], logGammaHeII=d["em_gHeII"],
                HINuEdgeP=d["em_eHI"], HeINuEdgeP=d["em_eHeI"], HeIINuEdgeP=d["em_eHeII"], gauntFF=d["em_gff"],
                gauntFFHeII=d["em_gffHeII"], HIPnuP=int(ei[4]), HeIPnuP=int(ei[5]), HeIIPnuP=int(ei[6]),
                xSecHI=d["em_sHI"], xSecHeI=d["em_sHeI"], xSecHeII=d["em_sHeII"], fb_meta=d["em_fbm"][:4 * nfb],
                fb_gratio=d["em_fbg"][:max(nfb, 1)], fb_sigma=d["em_fbs"], y_dat=d["em_ydat"], Ay_dat=d["em_Aydat"],
                HIRecLineData=d["em_HIrec"], HeIIRecLineData=d["em_HeIIrec"], HeIrecLineCoeff=d["em_HeIc"],
                HeIILymanData=d["em_lyData"], HeIILymanNu=d["em_lyNu"], widFlx=d["widFlx"])
