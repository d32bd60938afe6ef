"""Host-side mirror of MOCASSIN's transport step over the C ABI.

`Transport` plays the role of energyPacketDriver (photon_mod.f90:26) for one
grid: the host builds it from the same arrays the reference keeps in
`grid_type` (common_mod.f90:241-302) and module globals, uploads the
per-sweep tables, and calls `energy_packet_driver(istar, n)` once per source
per sweep, exactly where iterateMC calls energyPacketDriver
(iteration_mod.f90:474-496).  The estimator post-processing of iterateMC
(x1e-9, /8 in symmetric mode; iteration_mod.f90:700-722) is `scale_J`.

All arrays use the reference's Fortran layouts:
    active(nx,ny,nz) int32 order='F'
    opacity/recPDF/scaOpac (nCells+1, nbins) float32 order='F'
    specCDF (nbins, nStars) float32 order='F'
"""
import ctypes as C
import math

import numpy as np

from . import _capi


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F"))


class Transport:
    """One device context holding a grid, its tables and FP64 estimators."""

    def __init__(self, xAxis, yAxis, zAxis, active, nCells, nbins, nuArray,
                 R_out, geoCorr, ionEdge1, starPos, starIdx, deltaE,
                 symmetric=True, dust=False, gas=True, isotropic=False, multistars=False,
                 seed=12345, device=0):
        self._lib = _capi.load()
        self.nCells, self.nbins = int(nCells), int(nbins)
        active = np.asfortranarray(active, np.int32)
        self.shape = active.shape
        starPos = np.asarray(starPos, np.float32).reshape(-1, 3)
        starIdx = np.asarray(starIdx, np.int32).reshape(-1, 3)
        deltaE = np.asarray(deltaE, np.float32).reshape(-1)
        self.nStars = starPos.shape[0]
        # keep the host copies alive for the duration of the create call
        self._keep = [_f32(xAxis), _f32(yAxis), _f32(zAxis), active.ravel(order="F").copy(),
                      _f32(nuArray), _f32(starPos), np.ascontiguousarray(starIdx.ravel()),
                      _f32(deltaE)]
        k = self._keep
        d = _capi.McDesc()
        d.abi_version = _capi.MC_ABI_VERSION
        d.nx, d.ny, d.nz = self.shape
        d.xAxis, d.yAxis, d.zAxis = k[0].ctypes.data, k[1].ctypes.data, k[2].ctypes.data
        d.active = k[3].ctypes.data
        d.nCells, d.nbins = self.nCells, self.nbins
        d.flags = ((_capi.MC_SYMMETRIC if symmetric else 0) | (_capi.MC_DUST if dust else 0) |
                   (_capi.MC_GAS if gas else 0) | (_capi.MC_ISOTROPIC if isotropic else 0) |
                   (_capi.MC_MULTISTARS if multistars else 0))
        d.R_out = float(R_out)
        d.geoCorr[:] = [float(g) for g in geoCorr]
        d.ionEdge1 = float(ionEdge1)
        d.nuArray = k[4].ctypes.data
        d.nStars = self.nStars
        d.starPos, d.starIdx, d.deltaE = k[5].ctypes.data, k[6].ctypes.data, k[7].ctypes.data
        d.seed = int(seed)
        d.device = int(device)
        h = C.c_void_p()
        rc = self._lib.mc_ctx_create(C.byref(d), C.byref(h))
        if rc:
            raise _capi.McError(rc, self._lib.mc_ctx_last_error(None).decode())
        self._h = h
        self._keep = None

    def _check(self, rc):
        if rc:
            raise _capi.McError(rc, self._lib.mc_ctx_last_error(self._h).decode())

    def upload_tables(self, opacity, specCDF, recPDF=None, totalLines=None,
                      scaOpac=None, gSca=None):
        arrs = [None if opacity is None else _f32(opacity), None if scaOpac is None else _f32(scaOpac),
                None if recPDF is None else _f32(recPDF),
                None if totalLines is None else _f32(totalLines), _f32(specCDF),
                None if gSca is None else _f32(gSca)]
        n = (self.nCells + 1) * self.nbins
        for a, want in zip(arrs, (n, n, n, self.nCells + 1, self.nbins * self.nStars, self.nbins)):
            if a is not None and a.size != want:
                raise ValueError("table of size %d, expected %d" % (a.size, want))
        ptr = [None if a is None else a.ctypes.data for a in arrs]
        self._check(self._lib.mc_upload_tables(self._h, *ptr))

    def zero_estimators(self):
        self._check(self._lib.mc_zero_estimators(self._h))

    def energy_packet_driver(self, istar, npackets, packet_offset=0):
        """Transport npackets packets of source istar; returns the counters."""
        c = _capi.McCounters()
        self._check(self._lib.mc_transport_run(self._h, int(istar), int(npackets),
                                               int(packet_offset), None, None, C.byref(c)))
        return c.as_dict()

    def set_view_angles(self, view=None):
        """Viewing-angle planes of the escapes (mc_set_view_angles): view is
        None (plane 0 only) or dict(nAngleBins, totAngleBinsTheta,
        totAngleBinsPhi, dTheta, dPhi, viewPointPtheta (0:totT), viewPointPphi
        (0:totP), viewPointTheta, viewPointPhi (0:nAngleBins)).  Zeroes the
        escapes."""
        n = 0 if view is None else int(view["nAngleBins"])
        if n == 0:
            self._check(self._lib.mc_set_view_angles(self._h, None))
            self._planes = 1
            return
        keep = [np.ascontiguousarray(view[k], np.int32) for k in ("viewPointPtheta", "viewPointPphi")] + \
               [np.ascontiguousarray(view[k], np.float32) for k in ("viewPointTheta", "viewPointPhi")]
        v = _capi.McViewAngles()
        v.nAngleBins = n
        v.totAngleBinsTheta, v.totAngleBinsPhi = int(view["totAngleBinsTheta"]), int(view["totAngleBinsPhi"])
        v.dTheta, v.dPhi = float(np.float32(view["dTheta"])), float(np.float32(view["dPhi"]))
        v.viewPointPtheta, v.viewPointPphi, v.viewPointTheta, v.viewPointPhi = (a.ctypes.data for a in keep)
        self._check(self._lib.mc_set_view_angles(self._h, C.byref(v)))
        self._planes = n + 1

    def _esc_shape(self):
        p = getattr(self, "_planes", 1)
        return (self.nCells + 1, self.nbins + 1) + ((p,) if p > 1 else ())

    def fetch(self):
        """(J (nCells+1,nbins), esc (nCells+1,nbins+1[,nAngleBins+1])) FP64,
        unscaled, order='F'."""
        J = np.zeros((self.nCells + 1, self.nbins), np.float64, order="F")
        E = np.zeros(self._esc_shape(), np.float64, order="F")
        self._check(self._lib.mc_fetch_estimators(self._h, J.ctypes.data, E.ctypes.data))
        return J, E

    def count_crossings(self, enable=True):
        self._check(self._lib.mc_count_crossings(self._h, 1 if enable else 0))

    def crossings(self):
        """Segments walked in each cell (nCells+1,) since the last zeroing."""
        out = np.zeros(self.nCells + 1, np.uint32)
        self._check(self._lib.mc_fetch_crossings(self._h, out.ctypes.data))
        return out

    def set_estimators(self, J, esc=None):
        """Overwrite the device J (nCells+1, nbins) [and esc (nCells+1, nbins+1)]."""
        J = np.asarray(J, np.float64)
        if J.shape != (self.nCells + 1, self.nbins):
            raise ValueError("J of shape %s" % (J.shape,))
        Jf = np.asfortranarray(J)
        Ef = None if esc is None else np.asfortranarray(np.asarray(esc, np.float64))
        if Ef is not None and Ef.shape != self._esc_shape():
            raise ValueError("esc of shape %s" % (Ef.shape,))
        self._check(self._lib.mc_set_estimators(self._h, Jf.ctypes.data,
                                                None if Ef is None else Ef.ctypes.data))

    def balance_setup(self, bundle):
        """Upload the atomic inputs of the balance (mocassin_amd.atomic.AtomicBundle)."""
        b = bundle
        keep = {}

        def arr(name, dt):
            a = np.ascontiguousarray(getattr(b, name), dt)
            keep[name] = a
            return a.ctypes.data
        A = _capi.McAtomic()
        A.nstages, A.nElemUsed = b.nstages, b.nElemUsed
        A.elemOn, A.xref, A.abun = arr("elem_on", np.int32), arr("xref", np.int32), arr("abun", np.float32)
        A.nch, A.ch_meta, A.ch_sigma = b.nch, arr("ch_meta", np.int32), arr("ch_sigma", np.float32)
        A.rrec, A.rnew, A.fe = arr("rrec", np.float32), arr("rnew", np.float32), arr("fe", np.float32)
        A.dr_int, A.dr_real = arr("dr_int", np.int32), arr("dr_real", np.float32)
        A.ap_int, A.ap_real = arr("ap_int", np.int32), arr("ap_real", np.float32)
        A.ncel, A.cel_meta = b.ncel, arr("cel_meta", np.int32)
        A.cel_dbl, A.cel_flt = arr("cel_dbl", np.float64), arr("cel_flt", np.float32)
        A.XHILimit = b.XHILimit
        if keep["ch_sigma"].size != b.nch * self.nbins:
            raise ValueError("cross sections not on this frequency mesh")
        self._nE, self._ns, self._ncel = b.nElemUsed, b.nstages, b.ncel
        self._check(self._lib.mc_balance_setup(self._h, C.byref(A)))

    @staticmethod
    def _opts(solver=None, tol=1e-3, max_iter=60, ion_tol=0.01, warm_start=False, first_step=0.0,
              first_slope=0.0, warm_iter=0):
        if solver is None and not warm_start and ion_tol == 0.01:
            return None
        opts = _capi.McBalanceOpts()
        opts.solver = _capi.MC_SOLVER_SECANT if solver == "secant" else _capi.MC_SOLVER_REFERENCE
        opts.max_iter, opts.tol, opts.ion_tol = int(max_iter), float(tol), float(ion_tol)
        opts.warm_start = 1 if warm_start else 0
        opts.first_step = float(first_step)
        opts.first_slope = float(first_slope)
        opts.warm_iter = int(warm_iter)
        return opts

    def update_cells(self, Hden, Te, Ne, ionDen, jscale=1.0, solver=None, tol=1e-3, max_iter=60,
                     ion_tol=0.01, warm_start=False, first_step=0.0, first_slope=0.0, warm_iter=0):
        """updateCell for every active cell from the device J times jscale.
        Hden/Te/Ne (nCells+1,), ionDen (nCells+1, nElemUsed, nstages).
        solver None: the reference's schedule; "secant": secant iteration to
        |heat-cool| <= tol heat.  warm_start: start reached cells from the
        network's prediction (set_mlp first), falling back to the given state.
        Returns (Te, Ne, ionDen, flags, nIter, kernel_ms), ionDen in
        column-major order (the layout the host and the library keep); flags bit0
        converged, bit1 black, bit2 no Te balance, bit3 no ion balance, bit4
        not hit, bit5 restarted from the given state."""
        nc1 = self.nCells + 1
        Te = np.array(Te, np.float64).reshape(nc1)
        Ne = np.array(Ne, np.float64).reshape(nc1)
        Hd = np.ascontiguousarray(Hden, np.float64).reshape(nc1)
        # the library's layout is column-major: one copy, kept in that order
        ion = np.array(np.reshape(ionDen, (nc1, self._nE, self._ns)), np.float64, order="F")
        flags = np.zeros(nc1, np.int32)
        nIt = np.zeros(nc1, np.int32)
        ms = C.c_double()
        opts = self._opts(solver, tol, max_iter, ion_tol, warm_start, first_step, first_slope, warm_iter)
        self._check(self._lib.mc_balance_run(self._h, float(jscale), None if opts is None else C.byref(opts),
                                             Hd.ctypes.data, Te.ctypes.data,
                                             Ne.ctypes.data, ion.ctypes.data, flags.ctypes.data,
                                             nIt.ctypes.data, C.byref(ms)))
        return Te, Ne, ion, flags, nIt, ms.value

    # ---- grains
    def dust_setup(self, absXsec, grainWeight, grainAbun, emIntegral, nTemps, nSpeciesMax=None,
                   XHILimit=0.01, photoelectric=False, grainRadius=None, grainVn=None, MsurfAtom=None,
                   aWeight=None, Ndust=None, scaXsec=None, TdustSublime=None):
        """Upload the tables of getDustT for one dust chemistry.  absXsec
        (nSpecies, nSizes, nbins); emIntegral (nSpecies, nSizes, nTemps).
        photoelectric: the grain charge and photoelectric terms, which need
        grainRadius (nSizes), grainVn, MsurfAtom (nSpecies), aWeight (30)
        and Ndust (nCells+1).  scaXsec (nSpecies, nSizes, nbins) with
        TdustSublime (nSpecies) and Ndust: the dust add of opacity_build."""
        keep = [np.ascontiguousarray(absXsec, np.float32), np.ascontiguousarray(grainWeight, np.float32),
                np.ascontiguousarray(grainAbun, np.float32),
                np.asfortranarray(np.asarray(emIntegral, np.float32))]
        pe = [None] * 5
        if photoelectric:
            pe = [np.ascontiguousarray(v, np.float32) for v in (grainRadius, grainVn, MsurfAtom, aWeight, Ndust)]
            if pe[3].size != 30 or pe[4].size != self.nCells + 1:
                raise ValueError("aWeight needs 30 entries and Ndust nCells+1")
        op = [None, None]
        if scaXsec is not None:
            op = [np.ascontiguousarray(scaXsec, np.float32), np.ascontiguousarray(TdustSublime, np.float32)]
            if pe[4] is None:
                pe[4] = np.ascontiguousarray(Ndust, np.float32)
        self._dust_keep = keep + pe + op
        nS, nSz, nb = keep[0].shape
        if nb != self.nbins or keep[3].shape != (nS, nSz, int(nTemps)):
            raise ValueError("dust tables do not match nbins / (nSpecies, nSizes, nTemps)")
        D = _capi.McDustTables()
        D.nSpecies, D.nSizes, D.nTemps = nS, nSz, int(nTemps)
        D.nSpeciesMax = nS if nSpeciesMax is None else int(nSpeciesMax)
        D.photoelectric = 1 if photoelectric else 0
        D.absXsec, D.grainWeight, D.grainAbun, D.emIntegral = (k.ctypes.data for k in keep)
        D.grainRadius, D.grainVn, D.MsurfAtom, D.aWeight, D.Ndust = (None if v is None else v.ctypes.data
                                                                     for v in pe)
        D.scaXsec, D.TdustSublime = (None if v is None else v.ctypes.data for v in op)
        D.XHILimit = float(XHILimit)
        self._check(self._lib.mc_dust_setup(self._h, C.byref(D)))
        self._dust_shape = (D.nSpeciesMax + 1, nSz + 1)

    def update_grains(self, Tdust, jscale=1.0, Hden=None, Te=None, Ne=None, ionDen=None, **balance):
        """updateCell with grains (mc_update_cells): Tdust (nCells+1, nSizes+1,
        nSpeciesMax+1) in, the new one out; with gas, also the balance of
        update_cells (its keyword options apply).  Returns (Tdust, flags,
        nIter, kernel_ms) and, with gas, (Te, Ne, ionDen) after them."""
        nc1 = self.nCells + 1
        ld, nsz = self._dust_shape
        T = np.array(Tdust, np.float64).reshape(nc1, nsz, ld)
        flags = np.zeros(nc1, np.int32)
        nIt = np.zeros(nc1, np.int32)
        ms = C.c_double()
        gas = [None] * 4
        opts = None
        if Hden is not None:
            gas = [np.ascontiguousarray(Hden, np.float64).reshape(nc1), np.array(Te, np.float64).reshape(nc1),
                   np.array(Ne, np.float64).reshape(nc1),
                   np.array(np.reshape(ionDen, (nc1, self._nE, self._ns)), np.float64, order="F")]
            opts = self._opts(**balance)
        ptr = [None if g is None else g.ctypes.data for g in gas]
        self._check(self._lib.mc_update_cells(self._h, float(jscale), None if opts is None else C.byref(opts),
                                              *ptr, T.ctypes.data, flags.ctypes.data, nIt.ctypes.data,
                                              C.byref(ms)))
        out = (T, flags, nIt, ms.value)
        if Hden is not None:
            out += (gas[1], gas[2], gas[3])
        return out

    # ---- opacity tables
    def opacity_setup(self, seg_meta, seg_sigma, brems1):
        """Cross-section segments in addOpacity's order: seg_meta [nseg][4]
        (elem, ion, first bin, last bin), seg_sigma concatenated."""
        m = np.ascontiguousarray(seg_meta, np.int32).reshape(-1, 4)
        sg = np.ascontiguousarray(seg_sigma, np.float32)
        T = _capi.McOpacityTables(m.shape[0], m.ctypes.data, sg.ctypes.data, float(brems1))
        self._op_keep = (m, sg)
        self._check(self._lib.mc_opacity_setup(self._h, C.byref(T)))

    def opacity_build(self, Hden=None, Te=None, Ne=None, ionDen=None, Tdust=None, host=True):
        """Opacity tables on the device (they become the transport's).  The
        state as the host holds it (FP32), or None for the device state of
        the last update.  Returns (opacity, scaOpac, kernel_ms); the tables
        (nCells+1, nbins) are None when host is False."""
        nc1 = self.nCells + 1
        st = [None] * 4
        if Te is not None:
            st = [np.ascontiguousarray(Hden, np.float32).reshape(nc1), np.ascontiguousarray(Te, np.float32).reshape(nc1),
                  np.ascontiguousarray(Ne, np.float32).reshape(nc1),
                  np.asfortranarray(np.asarray(ionDen, np.float32).reshape(nc1, self._nE, self._ns))]
        td = None if Tdust is None else np.ascontiguousarray(Tdust, np.float32).ravel()
        if td is not None and td.size != nc1 * int(np.prod(getattr(self, "_dust_shape", (0,)))):
            raise ValueError("Tdust needs (nCells+1) x (nSizes+1) x (nSpeciesMax+1) entries (dust_setup first)")
        op = sca = None
        if host:
            op = np.zeros((self.nbins, nc1), np.float32)
            sca = np.zeros((self.nbins, nc1), np.float32)
        ms = C.c_double()
        ptr = lambda a: None if a is None else a.ctypes.data
        self._check(self._lib.mc_opacity_build(self._h, *(ptr(a) for a in st), ptr(td), ptr(op), ptr(sca),
                                               C.byref(ms)))
        if host:
            op, sca = op.T.copy(), sca.T.copy()
        return op, sca, ms.value

    def emission_setup(self, tkGamma, logGammaHI, logGammaHeI, logGammaHeII, HINuEdgeP, HeINuEdgeP, HeIINuEdgeP,
                       gauntFF, gauntFFHeII, HIPnuP, HeIPnuP, HeIIPnuP, xSecHI, xSecHeI, xSecHeII, fb_meta,
                       fb_gratio, fb_sigma, y_dat, Ay_dat, HIRecLineData, HeIIRecLineData, HeIrecLineCoeff,
                       HeIILymanData, HeIILymanNu, widFlx):
        """Upload the module data emissionDriver reads (mc_emission_tables);
        logGamma* (ntk, nlim), HIRecLineData (28, 7), HeIIRecLineData
        (28, 15), HeIrecLineCoeff (34, 3, 4) as the host holds them."""
        keep = {}

        def arr(name, a, dt):
            keep[name] = np.asfortranarray(a, dt).ravel(order="F")
            return keep[name].ctypes.data
        T = _capi.McEmissionTables()
        lg = [np.asarray(v, np.float32) for v in (logGammaHI, logGammaHeI, logGammaHeII)]
        T.ntk = int(np.size(tkGamma))
        T.nHI, T.nHeI, T.nHeII = (int(np.size(v) // T.ntk) for v in lg)
        T.tkGamma = arr("tk", tkGamma, np.float32)
        T.logGammaHI, T.logGammaHeI, T.logGammaHeII = (arr("lg%d" % k, v.reshape(T.ntk, -1, order="F"), np.float32)
                                                       for k, v in enumerate(lg))
        T.HINuEdgeP, T.HeINuEdgeP, T.HeIINuEdgeP = (arr("e%d" % k, v, np.int32)
                                                    for k, v in enumerate((HINuEdgeP, HeINuEdgeP, HeIINuEdgeP)))
        T.gauntFF, T.gauntFFHeII = arr("gff", gauntFF, np.float32), arr("gff2", gauntFFHeII, np.float32)
        T.HIPnuP, T.HeIPnuP, T.HeIIPnuP = int(HIPnuP), int(HeIPnuP), int(HeIIPnuP)
        T.xSecHI, T.xSecHeI, T.xSecHeII = (arr("x%d" % k, v, np.float32)
                                           for k, v in enumerate((xSecHI, xSecHeI, xSecHeII)))
        fbm = np.ascontiguousarray(fb_meta, np.int32).reshape(-1, 4)
        T.nfb = fbm.shape[0]
        keep["fbm"] = fbm
        T.fb_meta = fbm.ctypes.data if T.nfb else None
        T.fb_gratio = arr("fbg", np.asarray(fb_gratio, np.float32).reshape(-1)[:max(T.nfb, 1)], np.float32)
        T.fb_sigma = arr("fbs", np.asarray(fb_sigma, np.float32).reshape(-1), np.float32)
        T.y_dat, T.Ay_dat = arr("y", y_dat, np.float32), arr("Ay", Ay_dat, np.float32)
        T.HIRecLineData = arr("hi", np.asarray(HIRecLineData, np.float64).reshape(28, 7, order="F"), np.float64)
        T.HeIIRecLineData = arr("heii", np.asarray(HeIIRecLineData, np.float64).reshape(28, 15, order="F"),
                                np.float64)
        T.HeIrecLineCoeff = arr("hei", np.asarray(HeIrecLineCoeff, np.float32).reshape(34, 3, 4, order="F"),
                                np.float32)
        T.nHeIILyman = int(np.size(HeIILymanData))
        T.HeIILymanData, T.HeIILymanNu = arr("lyd", HeIILymanData, np.float32), arr("lyn", HeIILymanNu, np.float32)
        T.widFlx = arr("wid", widFlx, np.float32)
        self._check(self._lib.mc_emission_setup(self._h, C.byref(T)))

    def emission_build(self, Hden=None, Te=None, Ne=None, ionDen=None, Tdust=None, host=True):
        """The re-emission tables on the device (they become the
        transport's).  The REAL state, or None for the device state of the
        last update.  Returns (recPDF (nCells+1, nbins), totalLines
        (nCells+1,), kernel_ms); the tables are None when host is False."""
        nc1 = self.nCells + 1
        st = [None] * 4
        if Te is not None:
            st = [np.ascontiguousarray(Hden, np.float32).reshape(nc1), np.ascontiguousarray(Te, np.float32).reshape(nc1),
                  np.ascontiguousarray(Ne, np.float32).reshape(nc1),
                  np.asfortranarray(np.asarray(ionDen, np.float32).reshape(nc1, self._nE, self._ns))]
        td = None if Tdust is None else np.ascontiguousarray(Tdust, np.float32).ravel()
        rec = tl = None
        if host:
            rec = np.zeros((self.nbins, nc1), np.float32)
            tl = np.zeros(nc1, np.float32)
        ms = C.c_double()
        ptr = lambda a: None if a is None else a.ctypes.data
        self._check(self._lib.mc_emission_build(self._h, *(ptr(a) for a in st), ptr(td), ptr(rec), ptr(tl),
                                                C.byref(ms)))
        if host:
            rec = rec.T.copy()
        return rec, tl, ms.value

    def line_luminosities(self, Hden=None, Te=None, Ne=None, ionDen=None, lgBlack=None):
        """outputGas's analytic line luminosities: (forbVol [nCel, 17, 17]
        (model ion in forLines order, upper level, lower level), HbetaVol,
        kernel_ms), before the division by HbetaVol.  The state as the host
        holds it (FP32, with lgBlack or None), or None for the device state
        of the last update."""
        nc1 = self.nCells + 1
        st = [None] * 4
        bl = None
        if Te is not None:
            st = [np.ascontiguousarray(Hden, np.float32).reshape(nc1), np.ascontiguousarray(Te, np.float32).reshape(nc1),
                  np.ascontiguousarray(Ne, np.float32).reshape(nc1),
                  np.asfortranarray(np.asarray(ionDen, np.float32).reshape(nc1, self._nE, self._ns))]
            if lgBlack is not None:
                bl = np.ascontiguousarray(lgBlack, np.int32).reshape(nc1)
        forb = np.zeros((self._ncel, 17, 17), np.float64)
        hb = C.c_double()
        ms = C.c_double()
        ptr = lambda a: None if a is None else a.ctypes.data
        self._check(self._lib.mc_line_luminosities(self._h, *(ptr(a) for a in st), ptr(bl), forb.ctypes.data,
                                                   C.byref(hb), C.byref(ms)))
        return forb, hb.value, ms.value

    # ---- warm-start network
    def set_mlp(self, params, scale):
        p = np.ascontiguousarray(params, np.float32).ravel()
        sc = np.ascontiguousarray(scale, np.float32).ravel()
        if p.size != _capi.MC_MLP_NPARAMS or sc.size != _capi.MC_MLP_NSCALE:
            raise ValueError("params %d (want %d), scale %d (want %d)"
                             % (p.size, _capi.MC_MLP_NPARAMS, sc.size, _capi.MC_MLP_NSCALE))
        self._check(self._lib.mc_mlp_set(self._h, p.ctypes.data, sc.ctypes.data))

    def get_mlp(self):
        p = np.zeros(_capi.MC_MLP_NPARAMS, np.float32)
        sc = np.zeros(_capi.MC_MLP_NSCALE, np.float32)
        self._check(self._lib.mc_mlp_get(self._h, p.ctypes.data, sc.ctypes.data))
        return p, sc

    def mlp_features(self, Hden, jscale=1.0):
        """(X [nCells, 72], hit [nCells+1]) from the device J."""
        Hd = np.ascontiguousarray(Hden, np.float64).reshape(self.nCells + 1)
        X = np.zeros((self.nCells, _capi.MC_MLP_NIN), np.float32)
        hit = np.zeros(self.nCells + 1, np.int32)
        self._check(self._lib.mc_mlp_features(self._h, float(jscale), Hd.ctypes.data, X.ctypes.data,
                                              hit.ctypes.data))
        return X, hit

    def mlp_predict(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, _capi.MC_MLP_NIN)
        Y = np.zeros((X.shape[0], _capi.MC_MLP_NOUT), np.float32)
        self._check(self._lib.mc_mlp_predict(self._h, X.ctypes.data, X.shape[0], Y.ctypes.data))
        return Y

    def mlp_loss_grad(self, X, T, M):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, _capi.MC_MLP_NIN)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, _capi.MC_MLP_NOUT)
        M = np.ascontiguousarray(M, np.float32).reshape(-1, _capi.MC_MLP_NOUT)
        g = np.zeros(_capi.MC_MLP_NPARAMS, np.float32)
        loss = C.c_double()
        self._check(self._lib.mc_mlp_loss_grad(self._h, X.ctypes.data, T.ctypes.data, M.ctypes.data,
                                               X.shape[0], g.ctypes.data, C.byref(loss)))
        return loss.value, g

    def mlp_train(self, X, T, M, order, batch, lr):
        """Adam steps over batches order.reshape(-1, batch); returns (loss per step, device ms)."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, _capi.MC_MLP_NIN)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, _capi.MC_MLP_NOUT)
        M = np.ascontiguousarray(M, np.float32).reshape(-1, _capi.MC_MLP_NOUT)
        order = np.ascontiguousarray(order, np.int32).ravel()
        steps = order.size // int(batch)
        loss = np.zeros(steps, np.float32)
        ms = C.c_double()
        self._check(self._lib.mc_mlp_train(self._h, X.ctypes.data, T.ctypes.data, M.ctypes.data,
                                           X.shape[0], order.ctypes.data, steps, int(batch), float(lr),
                                           loss.ctypes.data, C.byref(ms)))
        return loss, ms.value

    def walk_replay(self, pos, dirv, cell, max_len=2048):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        dirv = np.ascontiguousarray(dirv, np.float32).reshape(-1, 3)
        cell = np.ascontiguousarray(cell, np.int32).reshape(-1, 3)
        n = pos.shape[0]
        seq = np.zeros((n, max_len, 3), np.int32)
        ds = np.zeros((n, max_len), np.float32)
        ln = np.zeros(n, np.int32)
        st = np.zeros(n, np.int32)
        self._check(self._lib.mc_walk_replay(self._h, n, pos.ctypes.data, dirv.ctypes.data,
                                             cell.ctypes.data, seq.ctypes.data, ds.ctypes.data,
                                             ln.ctypes.data, st.ctypes.data, int(max_len)))
        return seq, ds, ln, st

    def set_tuning(self, blocks_per_cu=None, threads=None, emit_min=None, walk_steps=None, kernel=None):
        """The transport launch shape (mc_tuning); kernel is a code or one of
        "auto", "one", "spare", "pool".  Fields left None keep
        their current values."""
        t = _capi.McTuning()
        self._check(self._lib.mc_get_tuning(self._h, C.byref(t)))
        if isinstance(kernel, str):
            kernel = _capi.KERNELS[kernel]
        for k, v in (("blocks_per_cu", blocks_per_cu), ("threads", threads),
                     ("emit_min", emit_min), ("walk_steps", walk_steps), ("kernel", kernel)):
            if v is not None:
                setattr(t, k, int(v))
        self._check(self._lib.mc_set_tuning(self._h, C.byref(t)))

    def get_tuning(self):
        t = _capi.McTuning()
        self._check(self._lib.mc_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def transport_batch(models, istar, npackets, offsets=None, tuning=None):
    """One persistent launch over several Transport contexts (one nebula
    each): npackets[i] packets of source istar[i] of models[i].  tuning: dict
    with waves_per_cu, emit_min, walk_steps, chunk (None: defaults).  Returns
    (counters per model, device ms of the launch)."""
    n = len(models)
    lib = _capi.load()
    hs = (C.c_void_p * n)(*[m._h for m in models])
    ist = np.ascontiguousarray(istar, np.int32).reshape(n)
    npk = np.ascontiguousarray(npackets, np.int64).reshape(n)
    off = None if offsets is None else np.ascontiguousarray(offsets, np.uint64).reshape(n)
    tu = None
    if tuning is not None:
        tu = _capi.McBatchTuning()
        for k, v in tuning.items():
            setattr(tu, k, int(v))
    cnt = (_capi.McCounters * n)()
    ms = C.c_double()
    rc = lib.mc_transport_batch(hs, n, ist.ctypes.data, npk.ctypes.data,
                                None if off is None else off.ctypes.data,
                                None if tu is None else C.byref(tu), cnt, C.byref(ms))
    if rc:
        raise _capi.McError(rc, lib.mc_ctx_last_error(models[0]._h).decode())
    return [c.as_dict() for c in cnt], ms.value


def scale_J(J, symmetric):
    """iterateMC's estimator post-processing: J *= 1e-9, then /8 in symmetric
    mode (iteration_mod.f90:706, 717-722); J in units of erg s^-1 cm^-2."""
    out = J * 1.e-9
    if symmetric:
        out = out / 8.
    return out



def algorithmic_bytes(c, nbins):
    """Algorithmic HBM bytes of one sweep from its counters: 24 B per
    segment (4 active + 4 opacity + 16 FP64 J read-modify-write), 8 +
    4*ceil(log2 nbins) B per absorption (line fraction, u-test operand,
    binary search of the re-emission CDF), 16 B per escape (FP64 RMW)."""
    return (24.0 * c["n_seg"] + (8.0 + 4.0 * math.ceil(math.log2(nbins))) * c["n_abs"]
            + 16.0 * c["n_esc"])

def probe_atomics(words, requests, lanes=64, fp64=True, device=0):
    """Atomic adds per second of the device (mc_probe_atomics): `requests`
    adds to random words of a `words`-element array, `lanes` active lanes
    per wavefront.  Returns (kernel_ms, adds per second)."""
    L = _capi.load()
    ms = C.c_double()
    rc = L.mc_probe_atomics(int(device), int(words), int(requests), int(lanes), 1 if fp64 else 0, C.byref(ms))
    if rc:
        raise _capi.McError(rc, "mc_probe_atomics")
    return ms.value, requests / (ms.value * 1e-3)


def probe_atomic_lines(words, requests, group, spread=False, device=0):
    """FP64 atomic adds per second when each instruction's lanes form sets
    of `group` adding to one 64-byte line (mc_probe_atomic_lines), the sets'
    lanes adjacent or spread across the wavefront.  Returns (kernel_ms, adds
    per second, line requests per second)."""
    L = _capi.load()
    ms = C.c_double()
    rc = L.mc_probe_atomic_lines(int(device), int(words), int(requests), int(group), 1 if spread else 0,
                                 C.byref(ms))
    if rc:
        raise _capi.McError(rc, "mc_probe_atomic_lines")
    rate = requests / (ms.value * 1e-3)
    return ms.value, rate, rate / group
