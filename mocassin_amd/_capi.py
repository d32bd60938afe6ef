"""ctypes declarations of include/mocassin_amd.h (libmocassin_amd.so).

The shared library is built in-tree by `make -C mocassin_amd/csrc` (or
__graft_entry__.build()).  Loading fails loudly when it is missing: there is
no fallback path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmocassin_amd.so")

MC_ABI_VERSION = 3
MC_SYMMETRIC, MC_DUST, MC_GAS, MC_ISOTROPIC, MC_MULTISTARS = 0x01, 0x02, 0x04, 0x08, 0x10
ERRORS = {0: "MC_OK", 1: "MC_EINVAL", 2: "MC_ENODEV", 3: "MC_ENOMEM", 4: "MC_ESTATE",
          5: "MC_ETABLE", 6: "MC_EPACKET", 7: "MC_EUNSUP"}

# every extern "C" symbol the header declares
EXPORTS = ("mc_ctx_create", "mc_upload_tables", "mc_zero_estimators", "mc_transport_run",
           "mc_fetch_estimators", "mc_walk_replay", "mc_set_tuning", "mc_get_tuning",
           "mc_set_estimators", "mc_balance_setup", "mc_balance_run",
           "mc_mlp_set", "mc_mlp_get", "mc_mlp_features", "mc_mlp_predict", "mc_mlp_loss_grad",
           "mc_mlp_train", "mc_count_crossings", "mc_fetch_crossings", "mc_transport_batch",
           "mc_transport_batch_noalloc",
           "mc_dust_setup", "mc_update_cells", "mc_opacity_setup", "mc_opacity_build",
           "mc_line_luminosities", "mc_emission_setup", "mc_emission_build", "mc_write_grid", "mc_read_grid",
           "mc_set_view_angles", "mc_esc_planes", "mc_ctx_shape", "mc_probe_atomics", "mc_probe_atomic_lines",
           "mc_ctx_last_error", "mc_ctx_destroy")

MC_SOLVER_REFERENCE, MC_SOLVER_SECANT = 0, 1
# mc_tuning.kernel
KERNELS = {"auto": 0, "one": 1, "spare": 2, "pool": 3}
MC_MLP_NIN, MC_MLP_NOUT, MC_MLP_NPARAMS, MC_MLP_NSCALE = 72, 31, 49568, 206


class McDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("xAxis", C.c_void_p), ("yAxis", C.c_void_p), ("zAxis", C.c_void_p),
        ("active", C.c_void_p), ("nCells", C.c_int32), ("nbins", C.c_int32),
        ("flags", C.c_int32), ("R_out", C.c_float), ("geoCorr", C.c_float * 3),
        ("ionEdge1", C.c_float), ("nuArray", C.c_void_p), ("nStars", C.c_int32),
        ("starPos", C.c_void_p), ("starIdx", C.c_void_p), ("deltaE", C.c_void_p),
        ("seed", C.c_uint64), ("device", C.c_int32),
    ]


class McCounters(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "n_seg", "n_abs", "n_sca", "n_esc", "n_lost", "n_line", "n_trapped", "n_flights",
        "qphot", "kernel_ms", "n_walk_passes", "n_walk_lanes", "n_emit_passes", "n_emit_lanes")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class McTuning(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int32), ("threads", C.c_int32),
                ("emit_min", C.c_int32), ("walk_steps", C.c_int32), ("kernel", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class McAtomic(C.Structure):
    _fields_ = [
        ("nstages", C.c_int32), ("nElemUsed", C.c_int32), ("elemOn", C.c_void_p),
        ("xref", C.c_void_p), ("abun", C.c_void_p), ("nch", C.c_int32), ("ch_meta", C.c_void_p),
        ("ch_sigma", C.c_void_p), ("rrec", C.c_void_p), ("rnew", C.c_void_p), ("fe", C.c_void_p),
        ("dr_int", C.c_void_p), ("dr_real", C.c_void_p), ("ap_int", C.c_void_p),
        ("ap_real", C.c_void_p), ("ncel", C.c_int32), ("cel_meta", C.c_void_p),
        ("cel_dbl", C.c_void_p), ("cel_flt", C.c_void_p), ("XHILimit", C.c_float),
    ]


class McBatchTuning(C.Structure):
    _fields_ = [("waves_per_cu", C.c_int32), ("emit_min", C.c_int32), ("walk_steps", C.c_int32),
                ("chunk", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class McBalanceOpts(C.Structure):
    _fields_ = [("solver", C.c_int32), ("max_iter", C.c_int32), ("tol", C.c_double),
                ("ion_tol", C.c_double), ("warm_start", C.c_int32), ("first_step", C.c_double),
                ("first_slope", C.c_double), ("warm_iter", C.c_int32)]


class McEmissionTables(C.Structure):
    _fields_ = [("ntk", C.c_int32), ("nHI", C.c_int32), ("nHeI", C.c_int32), ("nHeII", C.c_int32)] + \
               [(n, C.c_void_p) for n in ("tkGamma", "logGammaHI", "logGammaHeI", "logGammaHeII", "HINuEdgeP",
                                          "HeINuEdgeP", "HeIINuEdgeP", "gauntFF", "gauntFFHeII")] + \
               [("HIPnuP", C.c_int32), ("HeIPnuP", C.c_int32), ("HeIIPnuP", C.c_int32)] + \
               [(n, C.c_void_p) for n in ("xSecHI", "xSecHeI", "xSecHeII")] + [("nfb", C.c_int32)] + \
               [(n, C.c_void_p) for n in ("fb_meta", "fb_gratio", "fb_sigma", "y_dat", "Ay_dat", "HIRecLineData",
                                          "HeIIRecLineData", "HeIrecLineCoeff")] + \
               [("nHeIILyman", C.c_int32)] + [(n, C.c_void_p) for n in ("HeIILymanData", "HeIILymanNu", "widFlx")]


class McDustTables(C.Structure):
    _fields_ = [("nSpecies", C.c_int32), ("nSpeciesMax", C.c_int32), ("nSizes", C.c_int32),
                ("nTemps", C.c_int32), ("photoelectric", C.c_int32), ("absXsec", C.c_void_p),
                ("grainWeight", C.c_void_p), ("grainAbun", C.c_void_p), ("emIntegral", C.c_void_p),
                ("grainRadius", C.c_void_p), ("grainVn", C.c_void_p), ("MsurfAtom", C.c_void_p),
                ("aWeight", C.c_void_p), ("Ndust", C.c_void_p), ("scaXsec", C.c_void_p),
                ("TdustSublime", C.c_void_p), ("XHILimit", C.c_float)]


class McOpacityTables(C.Structure):
    _fields_ = [("nseg", C.c_int32), ("seg_meta", C.c_void_p), ("seg_sigma", C.c_void_p), ("brems1", C.c_float)]


class McViewAngles(C.Structure):
    _fields_ = [("nAngleBins", C.c_int32), ("totAngleBinsTheta", C.c_int32), ("totAngleBinsPhi", C.c_int32),
                ("dTheta", C.c_float), ("dPhi", C.c_float),
                ("viewPointPtheta", C.c_void_p), ("viewPointPphi", C.c_void_p),
                ("viewPointTheta", C.c_void_p), ("viewPointPhi", C.c_void_p)]


class McGridState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("Te", "Ne", "Hden", "ionDen", "lgConverged", "lgBlack", "Ndust",
                                          "Tdust")] + \
               [("nSpeciesMax", C.c_int32), ("nSizes", C.c_int32), ("totalDustMass", C.c_float)]


class McError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, str(code)), msg))
        self.code = code


_LIB = None


def load():
    """Load libmocassin_amd.so, declaring the C ABI; raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmocassin_amd.so is not built (%s); run "
                           "`make -C mocassin_amd/csrc` or __graft_entry__.build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
    L.mc_ctx_create.argtypes = [C.POINTER(McDesc), C.POINTER(vp)]
    L.mc_upload_tables.argtypes = [vp] * 7
    L.mc_zero_estimators.argtypes = [vp]
    L.mc_transport_run.argtypes = [vp, i32, i64, u64, vp, vp, C.POINTER(McCounters)]
    L.mc_fetch_estimators.argtypes = [vp, vp, vp]
    L.mc_walk_replay.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, i32]
    L.mc_set_tuning.argtypes = [vp, C.POINTER(McTuning)]
    L.mc_get_tuning.argtypes = [vp, C.POINTER(McTuning)]
    L.mc_set_estimators.argtypes = [vp, vp, vp]
    L.mc_balance_setup.argtypes = [vp, C.POINTER(McAtomic)]
    L.mc_balance_run.argtypes = [vp, C.c_double, C.POINTER(McBalanceOpts), vp, vp, vp, vp, vp, vp, vp]
    L.mc_mlp_set.argtypes = [vp, vp, vp]
    L.mc_mlp_get.argtypes = [vp, vp, vp]
    L.mc_mlp_features.argtypes = [vp, C.c_double, vp, vp, vp]
    L.mc_mlp_predict.argtypes = [vp, vp, i64, vp]
    L.mc_mlp_loss_grad.argtypes = [vp, vp, vp, vp, i64, vp, vp]
    L.mc_mlp_train.argtypes = [vp, vp, vp, vp, i64, vp, i32, i32, C.c_float, vp, vp]
    L.mc_transport_batch.argtypes = [C.POINTER(vp), i32, vp, vp, vp, C.POINTER(McBatchTuning),
                                     C.POINTER(McCounters), vp]
    L.mc_transport_batch_noalloc.argtypes = L.mc_transport_batch.argtypes
    L.mc_dust_setup.argtypes = [vp, C.POINTER(McDustTables)]
    L.mc_update_cells.argtypes = [vp, C.c_double, C.POINTER(McBalanceOpts)] + [vp] * 8
    L.mc_opacity_setup.argtypes = [vp, C.POINTER(McOpacityTables)]
    L.mc_opacity_build.argtypes = [vp] * 9
    L.mc_line_luminosities.argtypes = [vp] * 9
    L.mc_emission_setup.argtypes = [vp, C.POINTER(McEmissionTables)]
    L.mc_write_grid.argtypes = [C.POINTER(McDesc), vp, C.c_char_p, C.POINTER(McGridState)]
    L.mc_read_grid.argtypes = [C.POINTER(McDesc), vp, C.c_char_p, C.POINTER(McGridState)]
    L.mc_emission_build.argtypes = [vp] * 9
    L.mc_count_crossings.argtypes = [vp, i32]
    L.mc_set_view_angles.argtypes = [vp, C.POINTER(McViewAngles)]
    L.mc_esc_planes.argtypes = [vp]
    L.mc_ctx_shape.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.mc_fetch_crossings.argtypes = [vp, vp]
    L.mc_probe_atomics.argtypes = [i32, C.c_int64, C.c_int64, i32, i32, C.POINTER(C.c_double)]
    L.mc_probe_atomic_lines.argtypes = [i32, C.c_int64, C.c_int64, i32, i32, C.POINTER(C.c_double)]
    L.mc_ctx_last_error.argtypes = [vp]
    L.mc_ctx_last_error.restype = C.c_char_p
    L.mc_ctx_destroy.argtypes = [vp]
    L.mc_ctx_destroy.restype = None
    for name in EXPORTS[:-2]:
        getattr(L, name).restype = C.c_int
    _LIB = L
    return L
