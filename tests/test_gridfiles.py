"""Grid state files (mc_write_grid / mc_read_grid) against the unmodified
reference: the files its writeGrid wrote for the state after one balance
pass, and the state its resetGrid read back from them
(tests/golden/opacity_*.npz: txt_grid*, rg_*).  Host code; no GPU."""
import os

import numpy as np
import pytest

from conftest import balance_fixture, golden, grid_from_fixture
from mocassin_amd import gridfiles


def _grid(g):
    axes, active, nc, nb, fl, _ = grid_from_fixture(g)
    return dict(xAxis=axes[0], yAxis=axes[1], zAxis=axes[2], active=active, nCells=nc, gas=fl["gas"],
                dust=fl["dust"], R_out=float(np.max(np.abs(axes[0]))))


def _reference_state(g, b):
    """resetGrid's state, in write_grid's layout."""
    nx, ny, nz, nc = (int(v) for v in g["rg_ints"][:4])
    nE, ns, nsm, nsz = (int(v) for v in g["rg_ints"][4:8])
    st = dict(Te=g["rg_Te"], Ne=g["rg_Ne"], Hden=g["rg_Hden"], lgConverged=g["rg_conv"], lgBlack=g["rg_black"],
              ionDen=g["rg_ionDen"].reshape(ns, nE, nc + 1).transpose(2, 1, 0))
    if "rg_Tdust" in g:
        st.update(Ndust=g["rg_Ndust"], Tdust=g["rg_Tdust"].reshape(nc + 1, nsz + 1, nsm + 1))
        tail = bytes(g["txt_dustGrid"]).decode().split("[1.e45 g]:")[1].split()[0]
        st["totalDustMass"] = np.float32(tail)
    return st


def _put_reference_files(g, d):
    for k in g:
        if k.startswith("txt_"):
            with open(os.path.join(d, k[4:] + ".out"), "wb") as f:
                f.write(bytes(g[k]))


@pytest.mark.parametrize("name", ["c2mini", "dust_pe"])
def test_reader_matches_reference_resetgrid(name, tmp_path):
    g = golden("opacity_" + name)
    b = balance_fixture(name)[0]
    _put_reference_files(g, str(tmp_path))
    grid = _grid(g)
    ref = _reference_state(g, b)
    kw = {}
    if grid["dust"]:
        kw = dict(nSizes=ref["Tdust"].shape[1] - 1, nSpeciesMax=ref["Tdust"].shape[2] - 1)
    mine = gridfiles.read_grid(tmp_path, grid, b, **kw)
    for k in ("Te", "Ne", "Hden", "ionDen", "Ndust", "Tdust"):
        if k in ref:
            # bit for bit, including row 0 (the last inactive point's record)
            # and the stages above min(Z+1, nstages) resetGrid fills from
            # the previous element's record
            assert np.array_equal(np.asarray(mine[k]).view(np.uint32), np.asarray(ref[k], np.float32).view(np.uint32)), k
    assert np.array_equal(mine["lgConverged"], ref["lgConverged"])
    assert np.array_equal(mine["lgBlack"], ref["lgBlack"])
    if name == "c2mini":
        # the stale stages are there: H has two stages, its third is He's
        assert np.any(mine["ionDen"][1:, 0, 2] != 0)


@pytest.mark.parametrize("name", ["c2mini", "dust_pe"])
def test_writer_matches_reference_writegrid_byte_for_byte(name, tmp_path):
    g = golden("opacity_" + name)
    b = balance_fixture(name)[0]
    gridfiles.write_grid(tmp_path, _grid(g), _reference_state(g, b), b)
    for k in g:
        if k.startswith("txt_"):
            mine = open(os.path.join(str(tmp_path), k[4:] + ".out"), "rb").read()
            assert mine == bytes(g[k]), k


def test_round_trip_and_number_forms(tmp_path):
    """Values across the F/E boundaries survive write + read exactly.  The
    forms follow the rule that reproduces every number of the reference's
    files (F when the value rounded to one digit lies in [0.1, 1e7), 1P E
    else; digits as its runtime picks them); the boundary values beyond the
    files' range (near 1e7, FLT_MAX, denormals) follow the same rule but no
    reference file holds them."""
    g = golden("opacity_c2mini")
    b = balance_fixture("c2mini")[0]
    grid = _grid(g)
    st = _reference_state(g, b)
    vals = np.array([0.1, 0.099999994, 9999999.0, 1.0e7, -0.5, 1.0e-45, 3.4028235e38, 123.456, 1.0, 2.5e18, 0.095,
                     9499999.0], np.float32)
    st["Te"] = st["Te"].copy()
    st["Te"][1:1 + vals.size] = vals
    gridfiles.write_grid(tmp_path, grid, st, b)
    back = gridfiles.read_grid(tmp_path, grid, b)
    assert np.array_equal(back["Te"].view(np.uint32), st["Te"].view(np.uint32))
    toks = open(os.path.join(str(tmp_path), "grid1.out")).read().split()
    have = set(toks)
    for t in (".1", ".099999994", "9.999999E+06", "1.E+07", "-.5", "1.E-45", "123.456", "1.", "2.5E+18", "9.5E-02",
              "9499999."):
        assert t in have, t


def test_mismatched_grid_is_refused(tmp_path):
    from mocassin_amd._capi import McError
    g = golden("opacity_c2mini")
    b = balance_fixture("c2mini")[0]
    _put_reference_files(g, str(tmp_path))
    grid = _grid(g)
    grid["xAxis"] = np.asarray(grid["xAxis"], np.float32) * np.float32(2)
    with pytest.raises(McError):
        gridfiles.read_grid(tmp_path, grid, b)
