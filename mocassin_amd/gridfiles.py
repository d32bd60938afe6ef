"""The grid state files of a MOCASSIN run (grid0.out, grid1.out, grid2.out,
dustGrid.out) through the library's writer and reader (mc_write_grid,
mc_read_grid): what writeGrid leaves in output/ and resetGrid reads back
for a warm start (grid_mod.f90:2646-2760, 2967-3568)."""
import ctypes as C

import numpy as np

from . import _capi


def _desc(xAxis, yAxis, zAxis, active, nCells, gas=True, dust=False, R_out=0.0, keep=None):
    active = np.asfortranarray(active, np.int32)
    k = [np.ascontiguousarray(xAxis, np.float32), np.ascontiguousarray(yAxis, np.float32),
         np.ascontiguousarray(zAxis, np.float32), active.ravel(order="F").copy()]
    keep.extend(k)
    d = _capi.McDesc()
    d.abi_version = _capi.MC_ABI_VERSION
    d.nx, d.ny, d.nz = active.shape
    d.xAxis, d.yAxis, d.zAxis, d.active = (a.ctypes.data for a in k)
    d.nCells = int(nCells)
    d.flags = (_capi.MC_GAS if gas else 0) | (_capi.MC_DUST if dust else 0)
    d.R_out = float(R_out)
    return d


def _atomic(bundle, keep):
    if bundle is None:
        return None
    on = np.ascontiguousarray(bundle.elem_on, np.int32)
    xr = np.ascontiguousarray(bundle.xref, np.int32)
    keep.extend([on, xr])
    A = _capi.McAtomic()
    A.nstages, A.nElemUsed = bundle.nstages, bundle.nElemUsed
    A.elemOn, A.xref = on.ctypes.data, xr.ctypes.data
    return A


def _call(fn, grid, bundle, path, st):
    L = _capi.load()
    keep = []
    d = _desc(keep=keep, **grid)
    A = _atomic(bundle, keep)
    rc = getattr(L, fn)(C.byref(d), None if A is None else C.byref(A), path.encode(), C.byref(st))
    if rc:
        raise _capi.McError(rc, L.mc_ctx_last_error(None).decode())


def write_grid(path, grid, state, bundle=None):
    """Write the state files into directory `path`.  grid: dict(xAxis, yAxis,
    zAxis, active (nx, ny, nz), nCells, gas, dust, R_out); state: dict with
    Te, Ne, Hden (nCells+1,), ionDen (nCells+1, nElemUsed, nstages),
    lgConverged, lgBlack (nCells+1,) and with grains Ndust, Tdust (nCells+1,
    nSizes+1, nSpeciesMax+1) and totalDustMass; bundle: the elements
    switched on (mocassin_amd.atomic.AtomicBundle), for gas."""
    nc1 = int(grid["nCells"]) + 1
    keep = {}

    def f32(name, shape=None, order="C"):
        if state.get(name) is None:
            return None
        a = np.asarray(state[name], np.float32)
        keep[name] = np.asfortranarray(a) if order == "F" else np.ascontiguousarray(a)
        return keep[name].ctypes.data
    st = _capi.McGridState()
    st.Te, st.Ne, st.Hden = f32("Te"), f32("Ne"), f32("Hden")
    st.ionDen = f32("ionDen", order="F")
    for k in ("lgConverged", "lgBlack"):
        keep[k] = np.ascontiguousarray(state[k], np.int32).reshape(nc1)
    st.lgConverged, st.lgBlack = keep["lgConverged"].ctypes.data, keep["lgBlack"].ctypes.data
    st.Ndust, st.Tdust = f32("Ndust"), f32("Tdust")
    if state.get("Tdust") is not None:
        st.nSizes, st.nSpeciesMax = (int(v) - 1 for v in np.shape(state["Tdust"])[1:])
        st.totalDustMass = float(state.get("totalDustMass", 0.0))
    _call("mc_write_grid", grid, bundle, str(path), st)


def read_grid(path, grid, bundle=None, nSizes=0, nSpeciesMax=0):
    """Read the state files of directory `path` as resetGrid does; returns
    the state dict of write_grid (ionDen (nCells+1, nElemUsed, nstages),
    Tdust (nCells+1, nSizes+1, nSpeciesMax+1) with grains)."""
    nc1 = int(grid["nCells"]) + 1
    out = {"lgConverged": np.zeros(nc1, np.int32), "lgBlack": np.zeros(nc1, np.int32)}
    st = _capi.McGridState()
    st.lgConverged, st.lgBlack = out["lgConverged"].ctypes.data, out["lgBlack"].ctypes.data
    if grid.get("gas", True):
        for k in ("Te", "Ne", "Hden"):
            out[k] = np.zeros(nc1, np.float32)
        out["ionDen"] = np.zeros((nc1, bundle.nElemUsed, bundle.nstages), np.float32, order="F")
        st.Te, st.Ne, st.Hden = (out[k].ctypes.data for k in ("Te", "Ne", "Hden"))
        st.ionDen = out["ionDen"].ctypes.data
    if grid.get("dust", False):
        out["Ndust"] = np.zeros(nc1, np.float32)
        out["Tdust"] = np.zeros((nc1, nSizes + 1, nSpeciesMax + 1), np.float32)
        st.Ndust, st.Tdust = out["Ndust"].ctypes.data, out["Tdust"].ctypes.data
        st.nSizes, st.nSpeciesMax = nSizes, nSpeciesMax
    _call("mc_read_grid", grid, bundle, str(path), st)
    if "ionDen" in out:
        out["ionDen"] = np.ascontiguousarray(out["ionDen"])
    return out
