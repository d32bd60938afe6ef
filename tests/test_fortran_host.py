"""The Fortran host interface (mocassin_amd/fortran/mocassin_amd_mod.f90):
its check program links against libmocassin_amd.so, reports C-side errors
without a device, and on the GPU produces exactly what the Python binding
of the same C ABI produces (transport J and escapes to the FP64 atomics'
summation-order rounding, balance outputs on the same J bit for bit after
narrowing to REAL), and stops a sweep into a J of the wrong extents; the
updateCell loop with grains likewise."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, balance_fixture, golden, gpu_available, oracle_problem

EXE = os.path.join(ROOT, "mocassin_amd", "fortran", "mc_fortran_check")


def _exe():
    if not os.path.exists(EXE):
        raise RuntimeError("%s is not built; run make -C mocassin_amd/csrc" % EXE)
    return EXE


def test_fortran_module_declares_every_entry_point():
    import re
    from test_capi_cpu import header_functions
    src = open(os.path.join(ROOT, "mocassin_amd", "fortran", "mocassin_amd_mod.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(mc_\w+)"\)', src))
    assert set(header_functions()) == bound


def test_fortran_selftest_reports_c_errors():
    r = subprocess.run([_exe(), "selftest"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "selftest ok" in r.stdout and "ABI" in r.stdout


def _write_problem(path, P, b, Hden, Te0, Ne0, ion0, npackets, seed):
    nc, nb = P.nCells, P.nbins
    nx, ny, nz = P.active.shape
    flags = 1 * P.symmetric | 4
    hdr = np.zeros(24, np.int64)
    hdr[:17] = [nx, ny, nz, nc, nb, flags, 1, npackets, seed, 1, b.nstages, b.nElemUsed, b.nch,
                b.ncel, b.cel_dbl.size, b.cel_flt.size, 0]
    f32 = lambda a: np.asarray(a, np.float32).ravel(order="F")
    i32 = lambda a: np.asarray(a, np.int32).ravel(order="F")
    parts = [hdr, f32(P.xAxis), f32(P.yAxis), f32(P.zAxis), i32(P.active),
             f32([P.R_out, *P.geoCorr, P.ionEdge1]), f32(P.nuArray), f32(P.starPos),
             i32(P.starIdx), f32([P.deltaE]), f32(P.opacity), f32(P.recPDF), f32(P.totalLines),
             f32(P.specCDF),
             i32(b.elem_on), i32(b.xref), f32(b.abun), np.ascontiguousarray(b.ch_meta, np.int32).ravel(),
             np.ascontiguousarray(b.ch_sigma, np.float32).ravel(),
             np.ascontiguousarray(b.rrec, np.float32).ravel(),
             np.ascontiguousarray(b.rnew, np.float32).ravel(), np.ascontiguousarray(b.fe, np.float32).ravel(),
             i32(b.dr_int), f32(b.dr_real), i32(b.ap_int), f32(b.ap_real),
             np.ascontiguousarray(b.cel_meta, np.int32).ravel(), np.asarray(b.cel_dbl, np.float64),
             f32(b.cel_flt), f32([b.XHILimit]),
             f32(Hden), f32(Te0), f32(Ne0), f32(ion0)]
    with open(path, "wb") as f:
        for a in parts:
            f.write(np.ascontiguousarray(a).tobytes())


@pytest.mark.gpu
def test_fortran_host_matches_python_binding(tmp_path):
    if not gpu_available():
        pytest.skip("no GPU")
    from mocassin_amd import Transport
    print("fortran host: c1mini transport + balance", flush=True)
    g = golden("transport_c1mini")
    P = oracle_problem(g)
    b, J0, Hden, Te0, Ne0, ion0, ref = balance_fixture("c1mini")
    assert np.array_equal(np.asarray(b.nuArray, np.float32), np.asarray(P.nuArray, np.float32))
    npk, seed = 200000, 4242
    _write_problem(str(tmp_path / "problem.bin"), P, b, Hden, Te0, Ne0, ion0, npk, seed)
    # "short": after the run, a sweep into a J one row short, which the
    # module's wrapper must stop before the library copies into it
    r = subprocess.run([_exe(), "run", str(tmp_path), "short"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "balance ok" in r.stdout
    assert "energyPacketDriver: J is" in r.stdout and "FAILED" not in r.stdout

    nc, nb = P.nCells, P.nbins
    Jf = np.fromfile(str(tmp_path / "J.bin"), np.float64).reshape(nb, nc + 1).T
    Ef = np.fromfile(str(tmp_path / "esc.bin"), np.float64).reshape(nb + 1, nc + 1).T
    raw = open(str(tmp_path / "balance.bin"), "rb").read()
    nE, ns = b.nElemUsed, b.nstages
    o = 0
    out = {}
    for k, n, dt in (("Te", nc + 1, np.float32), ("Ne", nc + 1, np.float32),
                     ("ion", (nc + 1) * nE * ns, np.float32), ("conv", nc + 1, np.int32),
                     ("black", nc + 1, np.int32), ("counts", 3, np.int32)):
        out[k] = np.frombuffer(raw, dt, n, o)
        o += n * np.dtype(dt).itemsize
    assert o == len(raw)

    T = Transport(P.xAxis, P.yAxis, P.zAxis, P.active, nc, nb, P.nuArray, P.R_out, P.geoCorr,
                  P.ionEdge1, [P.starPos], [P.starIdx], [P.deltaE], symmetric=P.symmetric,
                  seed=seed, device=0)
    try:
        T.upload_tables(P.opacity, P.specCDF.reshape(-1, 1), P.recPDF, P.totalLines)
        T.energy_packet_driver(1, npk, 0)
        J, E = T.fetch()
        # the balance is deterministic per cell: run it on the Fortran run's J
        T.set_estimators(Jf)
        T.balance_setup(b)
        Te, Ne, ion, fl, nIt, ms = T.update_cells(Hden, Te0.astype(np.float32), Ne0.astype(np.float32),
                                                  np.asarray(ion0, np.float32), jscale=1e-9 / 8)
    finally:
        T.close()
    # FP64 atomics sum in a run-dependent order: equal to rounding
    assert np.allclose(Jf, J, rtol=1e-12, atol=0) and np.allclose(Ef, E, rtol=1e-12, atol=0)
    assert np.count_nonzero(Jf) == np.count_nonzero(J) > 1000
    assert np.array_equal(out["Te"], Te.astype(np.float32))
    assert np.array_equal(out["Ne"], Ne.astype(np.float32))
    ion_f = out["ion"].reshape(ns, nE, nc + 1).transpose(2, 1, 0)
    assert np.array_equal(ion_f, ion.astype(np.float32))
    hit = (fl & 16) == 0
    assert np.array_equal(out["conv"][1:][hit[1:]], (fl[1:][hit[1:]] & 1))
    assert np.array_equal(out["black"][1:], (fl[1:] & 2) >> 1)
    assert out["counts"][0] == int(np.sum(~hit[1:]))


@pytest.mark.gpu
def test_fortran_grains_matches_python_binding(tmp_path):
    """gpu_update_cells_dust (updateCell with getDustT, the grain charge and
    the photoelectric terms) through the Fortran module on the reference's
    dust deck estimator: Tdust, Te, Ne and the ionic fractions equal the
    Python binding's after narrowing to REAL, and the flags follow the
    reference's bookkeeping."""
    import types
    if not gpu_available():
        pytest.skip("no GPU")
    from test_gpu_balance import _context
    from test_gpu_dust import _fixture, _pe_tables
    print("fortran host: grains of the dust deck", flush=True)
    g, b, J, Hden, Te0, Ne0, ion0, ref, T0, Tref = _fixture("dust_pe")
    nc, nb = J.shape[0] - 1, J.shape[1]
    n = 2
    while n ** 3 < nc + 1:
        n += 1
    act = np.zeros(n ** 3, np.int32)
    act[:nc] = np.arange(1, nc + 1, dtype=np.int32)
    ax = np.linspace(0.0, 1.e17, n).astype(np.float32)
    zt = np.zeros((nc + 1, nb), np.float32)
    P = types.SimpleNamespace(nCells=nc, nbins=nb, active=act.reshape(n, n, n, order="F"), symmetric=False,
                              xAxis=ax, yAxis=ax, zAxis=ax, R_out=1.e17, geoCorr=(0., 0., 0.), ionEdge1=1.0,
                              nuArray=b.nuArray, starPos=(0., 0., 0.), starIdx=(1, 1, 1), deltaE=1.0,
                              opacity=zt, recPDF=zt, totalLines=np.zeros(nc + 1, np.float32),
                              specCDF=np.linspace(0.0, 1.0, nb, dtype=np.float32))
    _write_problem(str(tmp_path / "problem.bin"), P, b, Hden, Te0, Ne0, ion0, 0, 1)
    nS, nSz, nT, nsm, nSp = (int(v) for v in g["dust_ints"][:5])
    hdr = np.zeros(8, np.int64)
    hdr[:5] = [nS, nSz, nT, nsm, 1]
    f32 = lambda a: np.ascontiguousarray(a, np.float32).ravel()
    with open(str(tmp_path / "grains.bin"), "wb") as f:
        for a in (hdr, f32(g["dust_abs"]), f32(g["dust_weight"]), f32(g["dust_abun"]), f32(g["dust_emint"]),
                  f32(g["dust_radius"]), f32(g["dust_vn"]), f32(g["dust_msurf"]), f32(g["dust_aweight"]),
                  f32(g["dust_ndust"]), f32([b.XHILimit]), f32(g["Tdust0"]),
                  np.asarray(J, np.float64).ravel(order="F")):
            f.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([_exe(), "grains", str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "grains ok" in r.stdout
    raw = open(str(tmp_path / "grains_out.bin"), "rb").read()
    nE, ns = b.nElemUsed, b.nstages
    o, out = 0, {}
    for k, m, dt in (("Td", (nc + 1) * (nSz + 1) * (nsm + 1), np.float32), ("Te", nc + 1, np.float32),
                     ("Ne", nc + 1, np.float32), ("ion", (nc + 1) * nE * ns, np.float32),
                     ("conv", nc + 1, np.int32), ("black", nc + 1, np.int32), ("counts", 3, np.int32)):
        out[k] = np.frombuffer(raw, dt, m, o)
        o += m * np.dtype(dt).itemsize
    assert o == len(raw)
    T = _context(nc, nb, b.nuArray)
    try:
        T.set_estimators(J.astype(np.float64))
        T.balance_setup(b)
        T.dust_setup(**_pe_tables(g))
        Tg, fl, nI, ms, Te, Ne, ion = T.update_grains(T0, Hden=Hden, Te=Te0, Ne=Ne0, ionDen=ion0)
    finally:
        T.close()
    assert np.array_equal(out["Td"], Tg.astype(np.float32).ravel())
    assert np.array_equal(out["Te"], Te.astype(np.float32)) and np.array_equal(out["Ne"], Ne.astype(np.float32))
    assert np.array_equal(out["ion"].reshape(ns, nE, nc + 1).transpose(2, 1, 0), ion.astype(np.float32))
    hit = (fl & 16) == 0
    assert np.array_equal(out["conv"][1:][hit[1:]], fl[1:][hit[1:]] & 1)
    assert np.array_equal(out["black"][1:], (fl[1:] & 2) >> 1)
    assert out["counts"][0] == int(np.sum(~hit[1:]))
    print("fortran grains: %d cells, Tdust %.1f-%.1f K" % (nc, Tg[1:, 1:, 1].min(), Tg[1:, 1:, 1].max()))
