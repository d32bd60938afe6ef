import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle"), GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def grid_from_fixture(g):
    """(axes, active(nx,ny,nz) order F, nCells, nbins, flags dict) of a fixture."""
    ints = g["ints"]
    nx, ny, nz, nc, nb = (int(v) for v in ints[:5])
    active = g["active"].reshape(nz, ny, nx).transpose(2, 1, 0)
    flags = int(ints[6])
    si = int(ints[5])
    star_idx = (si // 100000000, (si // 10000) % 10000, si % 10000)
    fl = dict(symmetric=bool(flags & 1), dust=bool(flags & 2), gas=bool(flags & 4),
              isotropic=bool(flags & 8), multistars=bool(flags & 16))
    return (g["xAxis"], g["yAxis"], g["zAxis"]), np.asfortranarray(active), nc, nb, fl, star_idx


def oracle_problem(g, **over):
    """oracle Problem from a transport fixture (Fortran layouts)."""
    import mc_oracle as O
    axes, active, nc, nb, fl, star_idx = grid_from_fixture(g)
    sc = g["scalars"]
    tab = lambda n: g[n].reshape(nb, nc + 1).T if n in g else None
    kw = dict(xAxis=axes[0], yAxis=axes[1], zAxis=axes[2], active=active,
              opacity=tab("opacity"), nuArray=g["nuArray"], specCDF=g["specCDF"],
              R_out=sc[0], geoCorr=sc[1:4], ionEdge1=sc[4], deltaE=sc[5], starPos=sc[6:9],
              starIdx=star_idx, recPDF=tab("recPDF"),
              totalLines=g.get("totalLines"), scaOpac=tab("scaOpac"), gSca=g.get("gSca"), **fl)
    kw.update(over)
    return O.Problem(**kw)


def view_from_fixture(g):
    """The viewing-angle tables of a transport fixture (dict for the oracle's
    Problem(view=...) and Transport.set_view_angles), or None."""
    if "view_ints" not in g:
        return None
    n, tt, tp = (int(v) for v in g["view_ints"])
    return dict(nAngleBins=n, totAngleBinsTheta=tt, totAngleBinsPhi=tp, dTheta=g["view_real"][0],
                dPhi=g["view_real"][1], viewPointPtheta=g["view_Ptheta"], viewPointPphi=g["view_Pphi"],
                viewPointTheta=g["view_theta"], viewPointPhi=g["view_phi"])


def reference_escapes_all(g):
    """Reference raw FP32 escapedPackets (nCells+1, nbins+1, nAngleBins+1)."""
    nc, nb, na = int(g["ints"][3]), int(g["ints"][4]), int(g["view_ints"][0])
    E = np.zeros((nc + 1) * (nb + 1) * (na + 1), np.float32)
    E[g["Eall_idx"]] = g["Eall_val"]
    return E.reshape(na + 1, nb + 1, nc + 1).transpose(2, 1, 0)


def reference_estimators(g):
    """Reference raw FP32 Jste (nCells+1,nbins) and escapedPackets (nCells+1,nbins+1)."""
    nc, nb = int(g["ints"][3]), int(g["ints"][4])
    J = np.zeros((nc + 1) * nb, np.float32)
    J[g["J_idx"]] = g["J_val"]
    E = np.zeros((nc + 1) * (nb + 1), np.float32)
    E[g["E_idx"]] = g["E_val"]
    return J.reshape(nb, nc + 1).T, E.reshape(nb + 1, nc + 1).T


def walk_packets(g):
    """Per packet: pos, dir (float32 (n,3)), start cell (n,3), reference J map dict."""
    pos = g["pos_bits"].view(np.float32)
    dirv = g["dir_bits"].view(np.float32)
    maps = {}
    for o, p, c, b in zip(g["j_origin"], g["j_packet"], g["j_cell"], g["j_bits"]):
        maps.setdefault((int(o), int(p)), {})[int(c)] = int(b)
    keys = list(zip(g["origin_id"].tolist(), g["packet"].tolist()))
    return pos, dirv, g["cell"], [maps.get(k, {}) for k in keys], g["esc_bits"].view(np.float32)


def j_map_from_segments(cells, ds, axes, active, symmetric, dE=np.float32(1.0)):
    """Sum (dS*dE)/dV per active cell in FP32, in segment order (what the
    reference's Jste holds for one packet, photon_mod.f90:1822-1833)."""
    f = np.float32
    dfac = []
    for a in axes:
        n = len(a)
        w = np.empty(n, np.float32)
        for i in range(1, n + 1):
            if 1 < i < n:
                w[i - 1] = np.abs(a[i] - a[i - 2]) / f(2)
            elif i == 1:
                w[i - 1] = np.abs(a[1] - a[0]) / f(2) if symmetric else np.abs(a[1] - a[0])
            else:
                w[i - 1] = np.abs(a[n - 1] - a[n - 2])
        dfac.append(w / f(1.e15))
    out = {}
    for (x, y, z), s in zip(cells, ds):
        dV = dfac[0][x - 1] * dfac[1][y - 1] * dfac[2][z - 1]
        a = int(active[x - 1, y - 1, z - 1])
        out[a] = f(out.get(a, f(0)) + (f(s) * dE) / dV)
    return {k: int(np.float32(v).view(np.uint32)) for k, v in out.items() if v != 0}


def gpu_available():
    """A HIP device is visible to the runtime libmocassin_amd.so links
    (asked directly: torch need not be loaded, and once the library has
    loaded the system HIP runtime torch's own reports no device)."""
    import ctypes
    try:
        from mocassin_amd import _capi
        _capi.load()
        hip = ctypes.CDLL("libamdhip64.so.7")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def balance_fixture(name):
    """(bundle, J, Hden, Te0, Ne0, ion0, ref) of a balance fixture; arrays
    indexed by cell 0..nCells, ion arrays (nCells+1, nElemUsed, nstages)."""
    from mocassin_amd.atomic import AtomicBundle
    g = golden("balance_" + name)
    b = AtomicBundle.from_npz(g)
    nc, nb = int(g["ints"][3]), int(g["ints"][4])
    J = np.zeros((nc + 1) * nb, np.float32)
    J[g["J_idx"]] = g["J_val"]
    J = J.reshape(nb, nc + 1).T
    shp = lambda a: a.reshape(b.nstages, b.nElemUsed, nc + 1).transpose(2, 1, 0)
    ref = dict(Te=g["Te"], Ne=g["Ne"], ion=shp(g["ionDen"]), conv=g["conv"], black=g["black"])
    return b, J, g["Hden"], g["Te0"], g["Ne0"], shp(g["ionDen0"]), ref
