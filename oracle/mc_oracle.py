"""ctypes front end of the C restatement in oracle/mc_oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product never imports this module.

`Problem` holds one transport problem in the reference's Fortran layouts:
    active(nx,ny,nz)            int32, order='F'
    opacity/recPDF/scaOpac      float32 (nCells+1, nbins), order='F'
so arrays dumped by the reference harness load without reshuffling.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MATH_LIBM, MATH_OWN = 0, 1
J_FP32, J_FP64 = 0, 1


class _Problem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xAxis", C.c_void_p), ("yAxis", C.c_void_p), ("zAxis", C.c_void_p),
        ("active", C.c_void_p),
        ("nCells", C.c_int), ("nbins", C.c_int),
        ("opacity", C.c_void_p), ("scaOpac", C.c_void_p), ("recPDF", C.c_void_p),
        ("totalLines", C.c_void_p), ("specCDF", C.c_void_p), ("nuArray", C.c_void_p),
        ("gSca", C.c_void_p),
        ("R_out", C.c_float), ("geoCorrX", C.c_float), ("geoCorrY", C.c_float),
        ("geoCorrZ", C.c_float), ("ionEdge1", C.c_float), ("deltaE", C.c_float),
        ("starPos", C.c_float * 3), ("starIdx", C.c_int * 3),
        ("lgSymmetricXYZ", C.c_int), ("lgDust", C.c_int), ("lgGas", C.c_int),
        ("lgIsotropic", C.c_int), ("lgMultiStars", C.c_int),
        ("math_mode", C.c_int), ("j_mode", C.c_int),
        ("seed", C.c_uint64), ("stream", C.c_uint32),
        ("nAngleBins", C.c_int), ("totAngleBinsTheta", C.c_int), ("totAngleBinsPhi", C.c_int),
        ("dTheta", C.c_float), ("dPhi", C.c_float),
        ("viewPointPtheta", C.c_void_p), ("viewPointPphi", C.c_void_p),
        ("viewPointTheta", C.c_void_p), ("viewPointPhi", C.c_void_p),
    ]


class _Counters(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "n_seg", "n_abs", "n_sca", "n_esc", "n_lost", "n_line", "n_trapped",
        "n_flights", "qphot")] + [("error", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_mc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        L.mco_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 2 + [C.POINTER(C.c_uint32)]
        L.mco_uniform.restype = C.c_float
        L.mco_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32]
        L.mco_getnu2.restype = C.c_int
        L.mco_getnu2.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.mco_logf.restype = C.c_float
        L.mco_logf.argtypes = [C.c_int, C.c_float]
        for f in ("mco_acosf", "mco_atanf"):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [C.c_int, C.c_float]
        L.mco_sincosf.argtypes = [C.c_int, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.mco_walk.restype = C.c_int
        L.mco_walk.argtypes = [C.POINTER(_Problem), C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.mco_transport.restype = C.c_int
        L.mco_transport.argtypes = [C.POINTER(_Problem), C.c_uint64, C.c_uint64,
                                    C.c_void_p, C.c_void_p, C.POINTER(_Counters)]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().mco_philox4x32_10(c, k, o)
    return list(o)


def uniform(seed, stream, packet, k):
    return lib().mco_uniform(seed, stream, packet, k)


def logf(x, mode=MATH_OWN):
    return lib().mco_logf(mode, x)


def sincosf(a, mode=MATH_OWN):
    s, c = C.c_float(), C.c_float()
    lib().mco_sincosf(mode, a, C.byref(s), C.byref(c))
    return s.value, c.value


def sampling_tables(rec, totalLines):
    """The re-emission tables as the reference samples them when they hold
    NaN (rows of cells at a few tens of K, where emissionDriver's REAL sums
    overflow): getNu2's scan (photon_mod.f90:745-752) stops at the first
    entry u >= probDen fails, which a NaN entry fails as an entry of 1 does,
    so each row becomes 1 from its first NaN on; newPhotonPacket's test
    random <= totalLines (925-929) never passes a NaN, as with 0.  rec:
    rows [..., nbins]; returns copies (rec, totalLines)."""
    rec = np.array(rec, np.float32, copy=True)
    tl = np.array(totalLines, np.float32, copy=True)
    r2 = rec.reshape(-1, rec.shape[-1])
    for row in r2:
        bad = np.isnan(row)
        if bad.any():
            row[int(np.argmax(bad)):] = 1.0
    tl[np.isnan(tl)] = 0.0
    return r2.reshape(rec.shape), tl


def getnu2(cdf, u):
    cdf = np.ascontiguousarray(cdf, np.float32)
    u = np.ascontiguousarray(u, np.float32)
    used = C.c_int()
    r = lib().mco_getnu2(cdf.ctypes.data, int(cdf.size), u.ctypes.data, int(u.size), C.byref(used))
    return r, used.value


class Problem:
    """One transport problem (single grid, one source) in Fortran layouts."""

    def __init__(self, xAxis, yAxis, zAxis, active, opacity, nuArray, specCDF,
                 R_out, geoCorr, ionEdge1, deltaE, starPos, starIdx,
                 recPDF=None, totalLines=None, scaOpac=None, gSca=None,
                 symmetric=True, dust=False, gas=True, isotropic=False, multistars=False, view=None):
        """view: None, or dict(nAngleBins, totAngleBinsTheta, totAngleBinsPhi,
        dTheta, dPhi, viewPointPtheta (0:totT), viewPointPphi (0:totP),
        viewPointTheta, viewPointPhi (0:nAngleBins)) as initCartesianGrid
        sets them (grid_mod.f90:416-467)."""
        f32 = lambda a: None if a is None else np.asarray(a, np.float32)
        self.view = None
        if view is not None and int(view["nAngleBins"]) > 0:
            self.view = dict(view)
            for k in ("viewPointPtheta", "viewPointPphi"):
                self.view[k] = np.ascontiguousarray(view[k], np.int32)
            for k in ("viewPointTheta", "viewPointPhi"):
                self.view[k] = np.ascontiguousarray(view[k], np.float32)
        self.xAxis, self.yAxis, self.zAxis = f32(xAxis), f32(yAxis), f32(zAxis)
        self.active = np.asfortranarray(active, np.int32)
        self.nx, self.ny, self.nz = self.active.shape
        self.opacity = np.asfortranarray(opacity, np.float32)
        self.nCells = self.opacity.shape[0] - 1
        self.nbins = self.opacity.shape[1]
        self.nuArray, self.specCDF = f32(nuArray), f32(specCDF)
        self.recPDF = None if recPDF is None else np.asfortranarray(recPDF, np.float32)
        self.totalLines = f32(totalLines)
        self.scaOpac = None if scaOpac is None else np.asfortranarray(scaOpac, np.float32)
        self.gSca = f32(gSca)
        self.R_out = float(R_out)
        self.geoCorr = tuple(float(g) for g in geoCorr)
        self.ionEdge1 = float(ionEdge1)
        self.deltaE = float(deltaE)
        self.starPos = tuple(float(v) for v in starPos)
        self.starIdx = tuple(int(v) for v in starIdx)
        self.symmetric, self.dust, self.gas = bool(symmetric), bool(dust), bool(gas)
        self.isotropic, self.multistars = bool(isotropic), bool(multistars)

    def _struct(self, math_mode, j_mode, seed, stream):
        s = _Problem()
        s.nx, s.ny, s.nz = self.nx, self.ny, self.nz
        s.xAxis, s.yAxis, s.zAxis = _ptr(self.xAxis), _ptr(self.yAxis), _ptr(self.zAxis)
        s.active = _ptr(self.active)
        s.nCells, s.nbins = self.nCells, self.nbins
        s.opacity, s.scaOpac, s.recPDF = _ptr(self.opacity), _ptr(self.scaOpac), _ptr(self.recPDF)
        s.totalLines, s.specCDF, s.nuArray = _ptr(self.totalLines), _ptr(self.specCDF), _ptr(self.nuArray)
        s.gSca = _ptr(self.gSca)
        s.R_out = self.R_out
        s.geoCorrX, s.geoCorrY, s.geoCorrZ = self.geoCorr
        s.ionEdge1, s.deltaE = self.ionEdge1, self.deltaE
        s.starPos[:] = self.starPos
        s.starIdx[:] = self.starIdx
        s.lgSymmetricXYZ, s.lgDust, s.lgGas = int(self.symmetric), int(self.dust), int(self.gas)
        s.lgIsotropic, s.lgMultiStars = int(self.isotropic), int(self.multistars)
        s.math_mode, s.j_mode = math_mode, j_mode
        s.seed, s.stream = seed, stream
        if self.view is not None:
            v = self.view
            s.nAngleBins = int(v["nAngleBins"])
            s.totAngleBinsTheta, s.totAngleBinsPhi = int(v["totAngleBinsTheta"]), int(v["totAngleBinsPhi"])
            s.dTheta, s.dPhi = float(np.float32(v["dTheta"])), float(np.float32(v["dPhi"]))
            s.viewPointPtheta, s.viewPointPphi = _ptr(v["viewPointPtheta"]), _ptr(v["viewPointPphi"])
            s.viewPointTheta, s.viewPointPhi = _ptr(v["viewPointTheta"]), _ptr(v["viewPointPhi"])
        return s

    @property
    def nplanes(self):
        return 1 if self.view is None else int(self.view["nAngleBins"]) + 1

    def transport(self, first, n, seed, stream=1, math_mode=MATH_OWN, j_mode=J_FP64):
        """Returns (J, esc, counters dict); J (nCells+1, nbins), esc (nCells+1,
        nbins+1), or (nCells+1, nbins+1, nAngleBins+1) with viewing angles."""
        dt = np.float32 if j_mode == J_FP32 else np.float64
        J = np.zeros((self.nCells + 1, self.nbins), dt, order="F")
        E = np.zeros((self.nCells + 1, self.nbins + 1, self.nplanes), dt, order="F")
        cnt = _Counters()
        s = self._struct(math_mode, j_mode, seed, stream)
        rc = lib().mco_transport(C.byref(s), first, n, J.ctypes.data, E.ctypes.data, C.byref(cnt))
        if rc:
            raise RuntimeError("oracle transport stopped with code %d" % rc)
        if self.view is None:
            E = E[:, :, 0]
        return J, E, {f: getattr(cnt, f) for f, _ in _Counters._fields_}

    def walk(self, pos, dirv, cell, max_len=4096):
        """Zero-opacity geometric walk; returns (cells (n,3) 1-based, ds (n,), status)."""
        pos = np.ascontiguousarray(pos, np.float32)
        dirv = np.ascontiguousarray(dirv, np.float32)
        cell = np.ascontiguousarray(cell, np.int32)
        cells = np.zeros((max_len, 3), np.int32)
        ds = np.zeros(max_len, np.float32)
        st = C.c_int()
        s = self._struct(MATH_OWN, J_FP64, 0, 0)
        n = lib().mco_walk(C.byref(s), pos.ctypes.data, dirv.ctypes.data, cell.ctypes.data,
                           cells.ctypes.data, ds.ctypes.data, max_len, C.byref(st))
        return cells[:min(n, max_len)].copy(), ds[:min(n, max_len)].copy(), st.value


def load_harness_dump(d, **flag_overrides):
    """Problem from a reference-harness output directory (oracle/ref)."""
    rd = lambda n, dt=np.float32: np.fromfile(os.path.join(d, n + ".bin"), dt)
    ints = rd("ints", np.int32)
    nx, ny, nz, nc, nb = (int(v) for v in ints[:5])
    si = int(ints[5])
    star_idx = (si // 100000000, (si // 10000) % 10000, si % 10000)
    flags = int(ints[6])
    sc = rd("scalars")
    has = lambda n: os.path.exists(os.path.join(d, n + ".bin"))
    tab = lambda n: rd(n).reshape(nb, nc + 1).T if has(n) else None
    kw = dict(
        xAxis=rd("xAxis"), yAxis=rd("yAxis"), zAxis=rd("zAxis"),
        active=rd("active", np.int32).reshape(nz, ny, nx).transpose(2, 1, 0),
        opacity=tab("opacity"), nuArray=rd("nuArray"), specCDF=rd("specCDF"),
        R_out=sc[0], geoCorr=sc[1:4], ionEdge1=sc[4], deltaE=sc[5], starPos=sc[6:9],
        starIdx=star_idx, recPDF=tab("recPDF"),
        totalLines=rd("totalLines") if has("totalLines") else None,
        scaOpac=tab("scaOpac"), gSca=rd("gSca") if has("gSca") else None,
        symmetric=bool(flags & 1), dust=bool(flags & 2), gas=bool(flags & 4),
        isotropic=bool(flags & 8), multistars=bool(flags & 16))
    kw.update(flag_overrides)
    return Problem(**kw)


# ------------------------------------------------------------------ balance
class _Atomic(C.Structure):
    _fields_ = [
        ("nbins", C.c_int), ("nstages", C.c_int), ("nElemUsed", C.c_int),
        ("nuArray", C.c_void_p),
        ("elem_on_in", C.c_int32 * 30), ("xref_in", C.c_int32 * 30), ("abun_in", C.c_float * 30),
        ("nch", C.c_int), ("ch_meta", C.c_void_p), ("ch_sigma", C.c_void_p),
        ("rrec", C.c_void_p), ("rnew", C.c_void_p), ("fe", C.c_void_p),
        ("dr_int", C.c_void_p), ("dr_real", C.c_void_p), ("ap_int", C.c_void_p), ("ap_real", C.c_void_p),
        ("ncel", C.c_int), ("cel_meta", C.c_void_p), ("cel_dbl", C.c_void_p), ("cel_flt", C.c_void_p),
        ("XHILimit", C.c_float),
        ("elemOn", C.c_int * 31), ("xref", C.c_int * 31), ("abun", C.c_float * 31),
        ("chan_g", C.c_int32 * 620), ("chan_ip", C.c_int32 * 310),
        ("cel_dbl_off", C.c_size_t * 64), ("cel_flt_off", C.c_size_t * 64),
    ]


class _Grains(C.Structure):
    _fields_ = [("nS", C.c_int), ("nSizes", C.c_int), ("absx", C.c_void_p), ("weight", C.c_void_p),
                ("gabun", C.c_void_p), ("vn", C.c_void_p), ("msurf", C.c_void_p), ("radius", C.c_void_p),
                ("aweight", C.c_void_p)]


class _Emission(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nb", "ntk", "nHI", "nHeI", "nHeII", "ipHI", "ipHeI", "ipHeII", "i0HeI",
                                       "i0HeII", "nfb", "nly", "j2TsP")] + \
               [(n, C.c_void_p) for n in ("nu", "widFlx", "tk", "gHI", "gHeI", "gHeII", "eHI", "eHeI", "eHeII",
                                          "gff", "gffHeII", "sHI", "sHeI", "sHeII", "fbm", "fbg", "fbs", "ydat",
                                          "Aydat", "lyData", "lyP", "HeIc", "HIrec", "HeIIrec")]


class _EmissionDust(C.Structure):
    _fields_ = [("nS", C.c_int), ("nSizes", C.c_int), ("nsMax", C.c_int)] + \
               [(n, C.c_void_p) for n in ("absx", "weight", "gabun", "tsub", "Tdust")] + \
               [("Ndust", C.c_float), ("Hden", C.c_float)]


def emission_tables(g):
    """(mco_emission, arrays kept alive) from a fixture's em_* arrays."""
    ei = [int(v) for v in g["em_ints"]]
    keep = {}

    def a(name, dt):
        keep[name] = np.ascontiguousarray(g[name], dt)
        return keep[name].ctypes.data
    E = _Emission()
    E.nb = int(g["ints"][4])
    E.ntk, E.nHI, E.nHeI, E.nHeII, E.ipHI, E.ipHeI, E.ipHeII, E.i0HeI, E.i0HeII, E.nfb, E.nly, E.j2TsP = ei[:12]
    E.nu, E.widFlx, E.tk = a("nuArray", np.float32), a("widFlx", np.float32), a("em_tk", np.float32)
    E.gHI, E.gHeI, E.gHeII = a("em_gHI", np.float32), a("em_gHeI", np.float32), a("em_gHeII", np.float32)
    E.eHI, E.eHeI, E.eHeII = a("em_eHI", np.int32), a("em_eHeI", np.int32), a("em_eHeII", np.int32)
    E.gff, E.gffHeII = a("em_gff", np.float32), a("em_gffHeII", np.float32)
    E.sHI, E.sHeI, E.sHeII = a("em_sHI", np.float32), a("em_sHeI", np.float32), a("em_sHeII", np.float32)
    E.fbm, E.fbg, E.fbs = a("em_fbm", np.int32), a("em_fbg", np.float32), a("em_fbs", np.float32)
    E.ydat, E.Aydat, E.lyData = a("em_ydat", np.float32), a("em_Aydat", np.float32), a("em_lyData", np.float32)
    E.lyP, E.HeIc = a("em_lyP", np.int32), a("em_HeIc", np.float32)
    E.HIrec, E.HeIIrec = a("em_HIrec", np.float64), a("em_HeIIrec", np.float64)
    return E, keep


class Balance:
    """FP32 (the reference's precision) or FP64 restatement of updateCell
    for every cell of a grid, given an atomic bundle (mocassin_amd.atomic)."""

    def __init__(self, bundle):
        L = lib()
        if not hasattr(L, "_bal_declared"):
            for name, t in (("mco_balance_cell_f32", C.c_float), ("mco_balance_cell_f64", C.c_double)):
                f = getattr(L, name)
                f.restype = C.c_int
                f.argtypes = [C.POINTER(_Atomic), C.c_void_p, t, C.POINTER(t), C.POINTER(t),
                              C.c_void_p, C.POINTER(C.c_int)]
            L.mco_atomic_init.argtypes = [C.POINTER(_Atomic)]
            L.mco_atomic_init.restype = C.c_int
            L._bal_declared = True
        b = bundle
        keep = {}

        def arr(name, dt):
            a = np.ascontiguousarray(getattr(b, name), dt)
            keep[name] = a
            return a.ctypes.data
        A = _Atomic()
        A.nbins, A.nstages, A.nElemUsed = b.nbins, b.nstages, b.nElemUsed
        A.nuArray = arr("nuArray", np.float32)
        A.elem_on_in[:] = [int(v) for v in b.elem_on]
        A.xref_in[:] = [int(v) for v in b.xref]
        A.abun_in[:] = [float(v) for v in b.abun]
        A.nch, A.ch_meta, A.ch_sigma = b.nch, arr("ch_meta", np.int32), arr("ch_sigma", np.float32)
        A.rrec, A.rnew, A.fe = arr("rrec", np.float32), arr("rnew", np.float32), arr("fe", np.float32)
        A.dr_int, A.dr_real = arr("dr_int", np.int32), arr("dr_real", np.float32)
        A.ap_int, A.ap_real = arr("ap_int", np.int32), arr("ap_real", np.float32)
        A.ncel, A.cel_meta = b.ncel, arr("cel_meta", np.int32)
        A.cel_dbl, A.cel_flt = arr("cel_dbl", np.float64), arr("cel_flt", np.float32)
        A.XHILimit = b.XHILimit
        rc = L.mco_atomic_init(C.byref(A))
        if rc:
            raise RuntimeError("mco_atomic_init failed (%d)" % rc)
        self.A, self._keep, self.b = A, keep, b

    def run(self, J, Hden, Te, Ne, ionDen, precision="f64", cells=None):
        """J (nCells+1, nbins) scaled Jste; Hden/Te/Ne (nCells+1,);
        ionDen (nCells+1, nElemUsed, nstages).  Returns new Te, Ne, ionDen,
        flags (bit0 converged, bit1 black, bit2 no Te balance, bit3 no ion
        balance, bit4 not hit) and T-iteration counts."""
        L = lib()
        dt = np.float32 if precision == "f32" else np.float64
        ct = C.c_float if precision == "f32" else C.c_double
        f = L.mco_balance_cell_f32 if precision == "f32" else L.mco_balance_cell_f64
        nc1 = J.shape[0]
        Te = np.array(Te, dt)
        Ne = np.array(Ne, dt)
        ion = np.array(ionDen, dt)
        flags = np.zeros(nc1, np.int32)
        nT = np.zeros(nc1, np.int32)
        Jd = np.ascontiguousarray(J, dt)
        for c in (range(1, nc1) if cells is None else cells):
            te, ne = ct(Te[c]), ct(Ne[c])
            sl = np.ascontiguousarray(ion[c].T)          # [stage][elem]
            n = C.c_int()
            flags[c] = f(C.byref(self.A), Jd[c].ctypes.data, dt(Hden[c]), C.byref(te), C.byref(ne),
                         sl.ctypes.data, C.byref(n))
            Te[c], Ne[c] = te.value, ne.value
            ion[c] = sl.T
            nT[c] = n.value
        return Te, Ne, ion, flags, nT

    def line_lum(self, Hden, Te, Ne, ionDen, dV, cells, precision="f32"):
        """outputGas's forbVol [ncel, 17, 17] (model ion, upper, lower) and
        HbetaVol over `cells` in that order; Hden/Te/Ne/dV (nCells+1,),
        ionDen (nCells+1, nElemUsed, nstages).  f32: the reference's REAL
        sums; f64: all in double."""
        L = lib()
        if not hasattr(L, "_lines_declared"):
            for name in ("mco_line_lum_f32", "mco_line_lum_f64"):
                f = getattr(L, name)
                f.restype = None
                f.argtypes = [C.POINTER(_Atomic), C.c_int] + [C.c_void_p] * 7
            L._lines_declared = True
        dt = np.float32 if precision == "f32" else np.float64
        f = L.mco_line_lum_f32 if precision == "f32" else L.mco_line_lum_f64
        cells = np.asarray(cells, np.int64)
        col = lambda a: np.ascontiguousarray(np.asarray(a)[cells], dt)
        ion = np.ascontiguousarray(np.asarray(ionDen)[cells].transpose(0, 2, 1), dt)   # [cell][stage][elem]
        args = [col(Hden), col(Te), col(Ne), ion, col(dV)]
        forb = np.zeros((self.b.ncel, 17, 17), dt)
        hb = np.zeros(1, dt)
        f(C.byref(self.A), len(cells), *(a.ctypes.data for a in args), forb.ctypes.data, hb.ctypes.data)
        return forb, float(hb[0])

    def emission(self, E, Te, Ne, ionDen, cells, precision="f32", dust=None):
        """emissionDriver's gas part: (recPDF rows [len(cells), nbins],
        totalLines [len(cells)]) of `cells` from the REAL state Te/Ne
        (nCells+1,), ionDen (nCells+1, nElemUsed, nstages).  dust: dict
        absx [nS][nSizes][nbins], weight, gabun, tsub, Tdust (nCells+1,
        nSizes+1, nsMax+1), Ndust, Hden (nCells+1,), or None."""
        L = lib()
        if not hasattr(L, "_em_declared"):
            for name in ("mco_emission_cell_f32", "mco_emission_cell_f64"):
                f = getattr(L, name)
                f.restype = C.c_int
                f.argtypes = [C.POINTER(_Atomic), C.POINTER(_Emission), C.c_void_p, C.c_float, C.c_float,
                              C.c_void_p, C.c_void_p, C.c_void_p]
            L._em_declared = True
        f = L.mco_emission_cell_f32 if precision == "f32" else L.mco_emission_cell_f64
        cells = np.asarray(cells, np.int64)
        rec = np.zeros((len(cells), E.nb), np.float32)
        tl = np.zeros(len(cells), np.float32)
        ion = np.asarray(ionDen, np.float32)
        D, dk = None, {}
        if dust is not None:
            for k in ("absx", "weight", "gabun", "tsub"):
                dk[k] = np.ascontiguousarray(dust[k], np.float32)
            Td = np.asarray(dust["Tdust"], np.float32)
            D = _EmissionDust()
            D.nS, D.nSizes = dk["absx"].shape[0], dk["absx"].shape[1]
            D.nsMax = Td.shape[2] - 1
            D.absx, D.weight = dk["absx"].ctypes.data, dk["weight"].ctypes.data
            D.gabun, D.tsub = dk["gabun"].ctypes.data, dk["tsub"].ctypes.data
        for r, c in enumerate(cells):
            sl = np.ascontiguousarray(ion[c].T)          # [stage][elem]
            t = np.zeros(1, np.float32)
            row = np.zeros(E.nb, np.float32)
            if D is not None:
                blk = np.ascontiguousarray(Td[c])          # [size][species]
                D.Tdust = blk.ctypes.data
                D.Ndust, D.Hden = float(np.float32(dust["Ndust"][c])), float(np.float32(dust["Hden"][c]))
            if f(C.byref(self.A), C.byref(E), None if D is None else C.byref(D), float(np.float32(Te[c])),
                 float(np.float32(Ne[c])), sl.ctypes.data, row.ctypes.data, t.ctypes.data):
                raise RuntimeError("coincident gamma edges")
            rec[r], tl[r] = row, t[0]
        return rec, tl

    def run_grains(self, J, Hden, Te, Ne, ionDen, g, Tdust, precision="f64", cells=None):
        """run() with the grain charge, photoelectric and gas-grain collision
        terms of a dust chemistry (g: a balance fixture's dust_* arrays;
        Tdust (nCells+1, nSizes+1, nSpeciesMax+1) before the update).
        Returns run()'s outputs and gdc, peh, pot [nCells+1][nS*nSizes]."""
        L = lib()
        dt = np.float32 if precision == "f32" else np.float64
        ct = C.c_float if precision == "f32" else C.c_double
        f = L.mco_balance_cell_pe_f32 if precision == "f32" else L.mco_balance_cell_pe_f64
        f.restype = C.c_int
        f.argtypes = [C.POINTER(_Atomic), C.POINTER(_Grains), C.c_int, C.c_void_p, ct, C.POINTER(ct),
                      C.POINTER(ct), C.c_void_p, C.POINTER(C.c_int), ct, C.c_void_p, C.c_void_p, C.c_void_p,
                      C.c_void_p]
        nS, nSz, nT_, nsm = (int(v) for v in g["dust_ints"][:4])
        keep = [np.ascontiguousarray(g[k], np.float32) for k in
                ("dust_abs", "dust_weight", "dust_abun", "dust_vn", "dust_msurf", "dust_radius", "dust_aweight")]
        G = _Grains(nS, nSz, *(k.ctypes.data for k in keep))
        nc1 = J.shape[0]
        Te = np.array(Te, dt)
        Ne = np.array(Ne, dt)
        ion = np.array(ionDen, dt)
        nd = np.asarray(g["dust_ndust"], dt)
        Td = np.ascontiguousarray(np.asarray(Tdust, dt).reshape(nc1, nSz + 1, nsm + 1))
        flags = np.zeros(nc1, np.int32)
        nT = np.zeros(nc1, np.int32)
        K = nS * nSz
        gdc, peh, pot = (np.zeros((nc1, K), dt) for _ in range(3))
        Jd = np.ascontiguousarray(J, dt)
        for c in (range(1, nc1) if cells is None else cells):
            te, ne = ct(Te[c]), ct(Ne[c])
            sl = np.ascontiguousarray(ion[c].T)
            n = C.c_int()
            flags[c] = f(C.byref(self.A), C.byref(G), nsm, Jd[c].ctypes.data, dt(Hden[c]), C.byref(te),
                         C.byref(ne), sl.ctypes.data, C.byref(n), dt(nd[c]), Td[c].ctypes.data,
                         gdc[c].ctypes.data, peh[c].ctypes.data, pot[c].ctypes.data)
            Te[c], Ne[c] = te.value, ne.value
            ion[c] = sl.T
            nT[c] = n.value
        return Te, Ne, ion, flags, nT, gdc, peh, pot


def dust_temps(J, Tdust, g, precision="f64", gdc=None, peh=None):
    """getDustT for every cell (oracle/mc_dust.c).  J: scaled Jste
    (nCells+1, nbins); Tdust: (nCells+1, nSizes+1, nSpeciesMax+1) input
    temperatures; g: dict with dust_ints (nS, nSizes, nTemps, nsMax,
    nSpecies), dust_abs, dust_weight, dust_abun, dust_emint.  Returns the new
    Tdust in the same shape."""
    L = lib()
    dt = np.float32 if precision == "f32" else np.float64
    f = L.mco_dust_temps_f32 if precision == "f32" else L.mco_dust_temps_f64
    nS, nSz, nT, nsm, nSp = (int(v) for v in g["dust_ints"][:5])
    nc1, nb = J.shape
    Jf = np.asfortranarray(np.asarray(J, dt))
    T = np.ascontiguousarray(np.asarray(Tdust, dt).reshape(nc1, nSz + 1, nsm + 1))
    keep = [np.ascontiguousarray(g[k], np.float32) for k in ("dust_abs", "dust_weight", "dust_abun", "dust_emint")]
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    add = [None, None]
    if gdc is not None:
        add = [np.ascontiguousarray(gdc, dt), np.ascontiguousarray(peh, dt)]
    f(Jf.ctypes.data, nc1 - 1, nb, nS, nsm, nSz, nT, keep[0].ctypes.data, keep[1].ctypes.data,
      keep[2].ctypes.data, keep[3].ctypes.data, nSp, None if add[0] is None else add[0].ctypes.data,
      None if add[1] is None else add[1].ctypes.data, T.ctypes.data)
    return T


def cell_volumes(xAxis, yAxis, zAxis, active, symmetric):
    """getVolume (grid_mod.f90:2876-2964) of every cell of a 3-D grid in
    FP32, and the cells in outputGas's loop order (x outer, z inner).
    active (nx, ny, nz).  Returns (dV (nCells+1,), cells)."""
    def widths(v):
        v = np.asarray(v, np.float32)
        n = v.size
        w = np.zeros(n, np.float32)
        for i in range(1, n + 1):
            if 1 < i < n:
                w[i - 1] = np.float32(abs(v[i] - v[i - 2])) / np.float32(2.)
            elif i == 1:
                d = np.float32(abs(v[1] - v[0]))
                w[0] = d / np.float32(2.) if symmetric else d
            else:
                w[n - 1] = np.float32(abs(v[n - 1] - v[n - 2]))
        return w / np.float32(1.e15)
    wx, wy, wz = widths(xAxis), widths(yAxis), widths(zAxis)
    active = np.asarray(active)
    nc = int(active.max())
    dV = np.zeros(nc + 1, np.float32)
    cells = []
    nx, ny, nz = active.shape
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                c = int(active[i, j, k])
                if c > 0:
                    dV[c] = (wx[i] * wy[j]) * wz[k]
                    cells.append(c)
    return dV, np.array(cells, np.int64)


def opacity_tables(g, gaunt_cache=True, Te_cache=None):
    """Gas opacity (0:nCells, 1:nbins) of an opacity fixture's state in the
    reference's FP32 (oracle/mc_opacity.c), cells 1..nCells in order.
    gaunt_cache: the reference's Gaunt-factor reuse (starting from the
    uniform initial Te0, as after the initial table step), else exact per
    cell.  Returns (nCells+1, nbins) float32."""
    L = lib()
    f = L.mco_opacity_f32
    f.restype = None
    f.argtypes = [C.c_int] * 4 + [C.c_void_p] * 8 + [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                                     C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    ints = g["ints"]
    nc, nb = int(ints[3]), int(ints[4])
    nE, ns = int(ints[7]) // 100, int(ints[7]) % 100
    a = lambda k, dt=np.float32: np.ascontiguousarray(g[k], dt)
    keep = [a("nuArray"), a("Hden"), a("Te"), a("Ne"), a("ionDen"), a("elem_on", np.int32), a("xref", np.int32),
            a("abun"), a("seg_meta", np.int32), a("seg_sigma"), a("seg_off", np.int64)]
    out = np.zeros(((nc + 1) * nb), np.float32)
    tc = C.c_float(float(g["Te0"][1]) if Te_cache is None else Te_cache)
    gv = C.c_float(0.0)
    if gaunt_cache:
        # the initial table step: every cell at Te0, so one recomputation
        gv = C.c_float(L_gaunt(float(g["Te0"][1]), float(g["nuArray"][0])))
    f(nc, nb, ns, nE, *(k.ctypes.data for k in keep[:8]), int(g["seg_meta"].shape[0]), keep[8].ctypes.data,
      keep[9].ctypes.data, keep[10].ctypes.data, float(g["brems1"][0]), 1 if gaunt_cache else 0, C.byref(tc),
      C.byref(gv), out.ctypes.data)
    return out.reshape(nb, nc + 1).T.copy()


def L_gaunt(Te, nu1, z=0.0, g_prev=0.0):
    """getGauntFF at one frequency (oracle/mc_opacity.c), FP32."""
    L = lib()
    f = L.mco_gaunt_ff
    f.restype = C.c_float
    f.argtypes = [C.c_float] * 4
    return f(z, float(np.log10(np.float32(Te))), float(np.log10(np.float32(nu1))), g_prev)


def dust_opacity(g, opacity):
    """The dust add on top of a gas opacity (nCells+1, nbins); returns
    (opacity, scaOpac) in FP32."""
    L = lib()
    f = L.mco_dust_opacity_f32
    f.restype = None
    f.argtypes = [C.c_int] * 5 + [C.c_void_p] * 9
    ints = g["ints"]
    nc, nb = int(ints[3]), int(ints[4])
    nS, nSz, nT, nsm = (int(v) for v in g["dust_ints"][:4])
    a = lambda k: np.ascontiguousarray(g[k], np.float32)
    keep = [a("dust_abs"), a("dust_sca"), a("dust_weight"), a("dust_abun"), a("dust_tsub"), a("dust_ndust"),
            a("Tdust")]
    op = np.asfortranarray(np.asarray(opacity, np.float32)).copy(order="F")
    sca = np.zeros_like(op, order="F")
    f(nc, nb, nS, nsm, nSz, *(k.ctypes.data for k in keep), sca.ctypes.data, op.ctypes.data)
    return np.ascontiguousarray(op), np.ascontiguousarray(sca)
