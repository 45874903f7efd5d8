"""ctypes bindings of libpihm_amd.so (the HIP RHS, device N_Vector, integrator, driver).

The shared library is built in-tree (``make -C pihm-cuda_amd``); importing this module
fails loudly when it is missing: there is no CPU fallback for the product path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libpihm_amd.so")

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)

PIHM_OK = 0
PIHM_ERR_ARG = -1
PIHM_ERR_HIP = -2
PIHM_ERR_NAN = 1
PIHM_ERR_TOO_LARGE = -3


class MeshDesc(C.Structure):
    """pihm_mesh_desc (include/pihm_rhs.h)."""

    _fields_ = [
        ("nelem", C.c_int32), ("nriver", C.c_int32), ("surf_mode", C.c_int32),
        ("riv_mode", C.c_int32), ("stepsize", C.c_double),
        ("nabr", c_int32_p), ("bc_type", c_int32_p),
        ("x", c_double_p), ("y", c_double_p),
        ("area", c_double_p), ("zmin", c_double_p), ("zmax", c_double_p),
        ("edge", c_double_p), ("nabrdist", c_double_p), ("nabr_x", c_double_p), ("nabr_y", c_double_p),
        ("depth", c_double_p), ("ksath", c_double_p), ("ksatv", c_double_p), ("kinfv", c_double_p),
        ("dinf", c_double_p), ("alpha", c_double_p), ("beta", c_double_p), ("porosity", c_double_p),
        ("dmac", c_double_p), ("kmach", c_double_p), ("kmacv", c_double_p), ("areafv", c_double_p),
        ("areafh", c_double_p), ("rough", c_double_p), ("rzd", c_double_p),
        ("down", c_int32_p), ("leftele", c_int32_p), ("rightele", c_int32_p),
        ("riverbc_type", c_int32_p), ("intrpl_ord", c_int32_p),
        ("rarea", c_double_p), ("rzmin", c_double_p), ("rzmax", c_double_p), ("zbed", c_double_p),
        ("node_zmax", c_double_p), ("dist_left", c_double_p), ("dist_right", c_double_p),
        ("rdepth", c_double_p), ("coeff", c_double_p), ("length", c_double_p), ("width", c_double_p),
        ("rrough", c_double_p), ("cwr", c_double_p), ("rksath", c_double_p), ("rksatv", c_double_p),
        ("bedthick", c_double_p), ("rporosity", c_double_p),
    ]


class StepInputs(C.Structure):
    """pihm_step_inputs (include/pihm_rhs.h)."""

    _fields_ = [("pcpdrp", c_double_p), ("edir", c_double_p), ("ett", c_double_p),
                ("ws0_surf", c_double_p), ("elem_bc", c_double_p), ("river_bc", c_double_p)]


class RunConfig(C.Structure):
    """pihm_run_config (include/pihm_driver.h)."""

    _fields_ = [("t_start", C.c_double), ("t_end", C.c_double), ("stepsize", C.c_double),
                ("etstep", C.c_double), ("reltol", C.c_double), ("abstol", C.c_double),
                ("initstep", C.c_double), ("nncfn", C.c_double), ("nnimax", C.c_double),
                ("nnimin", C.c_double), ("decr", C.c_double), ("incr", C.c_double),
                ("stmin", C.c_double), ("maxl", C.c_int), ("max_steps_per_call", C.c_long),
                ("graphs", C.c_int), ("precond", C.c_int),
                ("rec_h", c_double_p), ("rec_q", c_int32_p), ("rec_iters", c_int32_p), ("rec_cap", C.c_long),
                ("rec_n", C.POINTER(C.c_long)),
                ("replay_h", c_double_p), ("replay_q", c_int32_p), ("replay_iters", c_int32_p), ("replay_n", C.c_long)]


class Forcing(C.Structure):
    """pihm_forcing (include/pihm_driver.h)."""

    _fields_ = [("nstation", C.c_int32), ("nrec", C.c_int32), ("rec_dt", C.c_double),
                ("prcp", c_double_p), ("edir", c_double_p), ("ett", c_double_p),
                ("station", c_int32_p)]


class RunStats(C.Structure):
    """pihm_run_stats (include/pihm_driver.h)."""

    _fields_ = [("nst", C.c_long), ("nfe", C.c_long), ("nfes", C.c_long), ("nni", C.c_long),
                ("nli", C.c_long), ("ncfn", C.c_long), ("netf", C.c_long),
                ("model_steps", C.c_long), ("wall_s", C.c_double), ("flag", C.c_int),
                ("nsetups", C.c_long), ("npe", C.c_long), ("nps", C.c_long)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class NoahSoilDesc(C.Structure):
    _fields_ = [("nsoil", c_int32_p), ("sldpth", c_double_p), ("smcmax", c_double_p), ("frzx", c_double_p)]


class LsmDesc(C.Structure):
    """pihm_lsm_desc (include/pihm_lsm.h)."""

    _fields_ = [(n, c_double_p) for n in (
        "albedomin", "albedomax", "shdfac", "cmcfactr", "cfactr", "zlvl_wind", "rzd",
        "rgl", "rsmin", "rsmax", "topt", "depth", "porosity", "smcmin", "smcwlt", "smcref", "zmax")] + [
        ("station", c_int32_p), ("cal_edir", C.c_double), ("cal_ec", C.c_double), ("cal_ett", C.c_double)]


PIHM_MET_NVAR = 5


class Ts(C.Structure):
    """pihm_ts (include/pihm_bc.h)."""

    _fields_ = [("nts", C.c_int32), ("nrec", C.c_int32), ("ftime", c_double_p), ("value", c_double_p)]


class Meteo(C.Structure):
    """pihm_meteo (include/pihm_lsm.h)."""

    _fields_ = [("nstation", C.c_int32), ("nrec", C.c_int32), ("ftime", c_double_p), ("data", c_double_p)]


NOAH_LSM_ARRAYS = ["shdfac", "cmcfactr", "laimin", "laimax", "emissmin", "emissmax", "albedomin", "albedomax",
                   "z0min", "z0max", "snup", "cfactr", "rgl", "rsmin", "rsmax", "topt", "hs", "smcref", "smcwlt",
                   "smcdry", "quartz", "zlvl_wind"]
NOAH_LSM_SCALARS = ["csoil", "sbeta", "salp", "fxexp", "czil", "lvcoef", "zbot", "tbot", "snoalb", "zlvl"]
PIHM_NOAH_NSCALAR = 10
PIHM_NOAH_NFLUX = 16
NOAH_FLUX_NAMES = ["edir", "ett", "ec", "pcpdrp", "drip", "esnow", "etp", "snomlt", "ssoil", "sheat", "eta",
                   "fdown", "sncovr", "beta", "rc", "pc"]
NOAH_STATE_NAMES = ["t1", "sneqv", "snowh", "sndens", "cmc", "ch", "cm", "snotime1", "albedo", "emissi"]


class NoahLsmDesc(C.Structure):
    """pihm_noah_lsm_desc (include/pihm_noahlsm.h)."""

    _fields_ = [("nroot", c_int32_p), ("station", c_int32_p)] + [(n, c_double_p) for n in NOAH_LSM_ARRAYS] + \
        [(n, C.c_double) for n in NOAH_LSM_SCALARS]


RhsFn = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p)
PrecSetupFn = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_double,
                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)
PrecSolveFn = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                          C.c_double, C.c_int, C.c_void_p, C.c_void_p)
PIHM_PREC_NONE = 0
PIHM_PREC_RIGHT = 2
PIHM_REC_ITERS = 8
PIHM_NOAH_MAXLYR = 11
PIHM_RHS_ONE_LAUNCH, PIHM_RHS_TWO_KERNELS = 0, 1     # pihm_ctx_set_rhs_path
PIHM_TUNE_SPANS = (0, 1, 2, 4, 8, 16, 32, 64)
PIHM_ENS_TUNE_G = (1, 2, 4, 8)


class EnsTuning(C.Structure):
    """pihm_ens_tuning (include/pihm_ensemble.h)."""

    _fields_ = [("tile_span", C.c_int), ("ops_chunks_per_wg", C.c_int),
                ("rhs_ms", C.c_double * 8), ("ops_ms", C.c_double * 4)]

    def as_dict(self):
        return {"tile_span": self.tile_span, "ops_chunks_per_wg": self.ops_chunks_per_wg,
                "rhs_ms_by_span": dict(zip(PIHM_TUNE_SPANS, list(self.rhs_ms))),
                "ops_ms_by_chunks_per_wg": dict(zip(PIHM_ENS_TUNE_G, list(self.ops_ms)))}
ForcingFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double)
SummaryFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)
StepEndFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_void_p)


class Backend(C.Structure):
    """pihm_backend (include/pihm_driver.h)."""

    _fields_ = [("rhs", C.c_void_p), ("rhs_data", C.c_void_p), ("apply_forcing", C.c_void_p),
                ("summary", C.c_void_p), ("model_step_end", C.c_void_p), ("user", C.c_void_p),
                ("apply_bc", C.c_void_p), ("prec_setup", C.c_void_p), ("prec_solve", C.c_void_p)]


_lib = None

# (name, restype, argtypes)
_PROTOS = [
    ("pihm_ctx_create", C.c_int, [C.POINTER(MeshDesc), C.POINTER(C.c_void_p)]),
    ("pihm_ctx_destroy", None, [C.c_void_p]),
    ("pihm_ctx_create_batch", C.c_int, [C.c_int, C.POINTER(C.POINTER(MeshDesc)), C.POINTER(C.c_void_p)]),
    ("pihm_ctx_create_slots", C.c_int, [C.c_int, c_int32_p, c_int32_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("pihm_ctx_load_member", C.c_int, [C.c_void_p, C.c_int, C.POINTER(MeshDesc)]),
    ("pihm_ctx_nmembers", C.c_int, [C.c_void_p]),
    ("pihm_ctx_member_info", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("pihm_ctx_set_active", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("pihm_ctx_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("pihm_stream_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("pihm_stream_destroy", C.c_int, [C.c_void_p]),
    ("pihm_ctx_set_step_inputs", C.c_int, [C.c_void_p, C.POINTER(StepInputs)]),
    ("pihm_ctx_set_noah", C.c_int, [C.c_void_p, c_double_p]),
    ("pihm_noah_lsm_create", C.c_int, [C.c_void_p, C.POINTER(NoahLsmDesc), C.POINTER(Meteo), C.POINTER(C.c_void_p)]),
    ("pihm_noah_lsm_destroy", None, [C.c_void_p]),
    ("pihm_noah_lsm_set_lai", C.c_int, [C.c_void_p, c_double_p]),
    ("pihm_noah_lsm_set_state", C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p]),
    ("pihm_noah_lsm_get_state", C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    ("pihm_noah_lsm_step", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    ("pihm_noah_lsm_get_fluxes", C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    ("pihm_noah_lsm_time", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double)]),
    ("pihm_run_gpu_flux", C.c_int, [C.c_void_p, C.POINTER(RunConfig), c_double_p, C.POINTER(RunStats)]),
    ("pihm_noah_soil_create", C.c_int, [C.c_void_p, C.POINTER(NoahSoilDesc), C.POINTER(C.c_void_p)]),
    ("pihm_noah_soil_destroy", None, [C.c_void_p]),
    ("pihm_noah_soil_set_state", C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p]),
    ("pihm_noah_soil_get_state", C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_int32_p, c_double_p,
                                           c_double_p, c_double_p, c_double_p]),
    ("pihm_noah_soil_step", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, c_double_p, c_double_p]),
    ("pihm_noah_soil_time", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    ("pihm_noah_soil_start", C.c_int, [C.c_void_p, C.c_void_p]),
    ("pihm_noah_soil_summary", C.c_int, [C.c_void_p, C.c_double, C.c_void_p]),
    ("pihm_gpu_backend_set_noah_soil", C.c_int, [C.c_void_p, C.c_void_p]),
    ("pihm_run_gpu_noah", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, c_double_p, C.c_void_p]),
    ("pihm_ctx_set_noah_step", C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, C.c_int, c_int32_p, c_int32_p]),
    ("pihm_ctx_neq", C.c_int, [C.c_void_p]),
    ("pihm_ctx_get_perm", C.c_int, [C.c_void_p, c_int32_p]),
    ("pihm_ctx_state_to_device", C.c_int, [C.c_void_p, c_double_p, C.c_void_p]),
    ("pihm_ctx_state_from_device", C.c_int, [C.c_void_p, C.c_void_p, c_double_p]),
    ("PIHM_ODE_Hip", C.c_int, [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pihm_ctx_set_nan_check", C.c_int, [C.c_void_p, C.c_int]),
    ("pihm_rhs_launch", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pihm_ctx_poll_nan", C.c_int, [C.c_void_p]),
    ("pihm_ctx_set_record", C.c_int, [C.c_void_p, C.c_int]),
    ("pihm_ctx_set_closure", C.c_int, [C.c_void_p, C.c_int]),
    ("pihm_ctx_sync", C.c_int, [C.c_void_p]),
    ("pihm_ctx_time_kernels", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_double_p,
                                        c_double_p]),
    ("pihm_ctx_time_rhs", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_double_p]),
    ("pihm_ctx_get_river_edge_ovlflow", C.c_int, [C.c_void_p, c_double_p]),
    ("pihm_ctx_set_river_edge_ovlflow", C.c_int, [C.c_void_p, c_double_p]),
    ("pihm_ctx_get_fluxes", C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                      c_double_p]),
    ("pihm_math_pow", C.c_int, [C.c_int, c_double_p, c_double_p, c_double_p, C.c_int]),
    ("N_VNew_Hip", C.c_void_p, [C.c_long]),
    ("N_VNewEmpty_Hip", C.c_void_p, [C.c_long]),
    ("N_VMake_Hip", C.c_void_p, [C.c_long, C.c_void_p]),
    ("N_VDestroy_Hip", None, [C.c_void_p]),
    ("N_VSetStream_Hip", None, [C.c_void_p, C.c_void_p]),
    ("N_VCopyToDevice_Hip", C.c_int, [C.c_void_p, c_double_p]),
    ("N_VCopyFromDevice_Hip", C.c_int, [C.c_void_p, c_double_p]),
    ("N_VGetDeviceArrayPointer_Hip", C.c_void_p, [C.c_void_p]),
    ("N_VGetLength_Hip", C.c_long, [C.c_void_p]),
    ("N_VGetVectorID_Hip", C.c_int, [C.c_void_p]),
    ("N_VClone_Hip", C.c_void_p, [C.c_void_p]),
    ("N_VCloneEmpty_Hip", C.c_void_p, [C.c_void_p]),
    ("N_VSpace_Hip", None, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    ("N_VGetArrayPointer_Hip", C.c_void_p, [C.c_void_p]),
    ("N_VSetArrayPointer_Hip", None, [C.c_void_p, C.c_void_p]),
    ("N_VLinearSum_Hip", None, [C.c_double, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]),
    ("N_VConst_Hip", None, [C.c_double, C.c_void_p]),
    ("N_VProd_Hip", None, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("N_VDiv_Hip", None, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("N_VScale_Hip", None, [C.c_double, C.c_void_p, C.c_void_p]),
    ("N_VAbs_Hip", None, [C.c_void_p, C.c_void_p]),
    ("N_VInv_Hip", None, [C.c_void_p, C.c_void_p]),
    ("N_VAddConst_Hip", None, [C.c_void_p, C.c_double, C.c_void_p]),
    ("N_VDotProd_Hip", C.c_double, [C.c_void_p, C.c_void_p]),
    ("pihm_nv_time_ops", C.c_int, [C.c_long, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("pihm_nv_dotprod_noalloc", C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("N_VMaxNorm_Hip", C.c_double, [C.c_void_p]),
    ("N_VWrmsNorm_Hip", C.c_double, [C.c_void_p, C.c_void_p]),
    ("N_VWrmsNormMask_Hip", C.c_double, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("N_VMin_Hip", C.c_double, [C.c_void_p]),
    ("N_VWL2Norm_Hip", C.c_double, [C.c_void_p, C.c_void_p]),
    ("N_VL1Norm_Hip", C.c_double, [C.c_void_p]),
    ("N_VCompare_Hip", None, [C.c_double, C.c_void_p, C.c_void_p]),
    ("N_VInvTest_Hip", C.c_int, [C.c_void_p, C.c_void_p]),
    ("N_VConstrMask_Hip", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("N_VMinQuotient_Hip", C.c_double, [C.c_void_p, C.c_void_p]),
    ("N_VLinearCombination_Hip", C.c_int, [C.c_int, c_double_p, C.POINTER(C.c_void_p), C.c_void_p]),
    ("N_VDotProdMulti_Hip", C.c_int, [C.c_int, C.c_void_p, C.POINTER(C.c_void_p), c_double_p]),
    ("N_VLinearSumWrms_Hip", C.c_double, [C.c_double, C.c_void_p, C.c_double, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
    ("PihmCVodeCreate", C.c_void_p, [C.c_int, C.c_int]),
    ("PihmCVodeInit", C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    ("PihmCVodeReInit", C.c_int, [C.c_void_p, C.c_double, C.c_void_p]),
    ("PihmCVodeSStolerances", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    ("PihmCVodeSetUserData", C.c_int, [C.c_void_p, C.c_void_p]),
    ("PihmCVodeSetInitStep", C.c_int, [C.c_void_p, C.c_double]),
    ("PihmCVodeSetMaxStep", C.c_int, [C.c_void_p, C.c_double]),
    ("PihmCVodeSetMinStep", C.c_int, [C.c_void_p, C.c_double]),
    ("PihmCVodeSetMaxNumSteps", C.c_int, [C.c_void_p, C.c_long]),
    ("PihmCVodeSetMaxOrd", C.c_int, [C.c_void_p, C.c_int]),
    ("PihmCVodeSetStopTime", C.c_int, [C.c_void_p, C.c_double]),
    ("PihmCVodeSetStabLimDet", C.c_int, [C.c_void_p, C.c_int]),
    ("PihmCVodeSetGraphRhs", C.c_int, [C.c_void_p, C.c_int]),
    ("PihmCVSpgmr", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("PihmCVSpilsSetPreconditioner", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("PihmCVodeSetStepRecord", C.c_int, [C.c_void_p, c_double_p, c_int32_p, c_int32_p, C.c_long, C.POINTER(C.c_long)]),
    ("PihmCVodeSetStepReplay", C.c_int, [C.c_void_p, c_double_p, c_int32_p, c_int32_p, C.c_long]),
    ("PihmCVodeGetNumLinSolvSetups", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVSpilsGetNumPrecEvals", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVSpilsGetNumPrecSolves", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVode", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.POINTER(C.c_double), C.c_int]),
    ("PihmCVodeGetDky", C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_void_p]),
    ("PihmCVodeFree", None, [C.POINTER(C.c_void_p)]),
    ("PihmCVodeGetNumSteps", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVodeGetNumRhsEvals", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVodeGetNumLinIters", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVodeGetNumJtimesRhsEvals", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVodeGetNumNonlinSolvIters", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVodeGetNumNonlinSolvConvFails", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVodeGetNumErrTestFails", C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    ("PihmCVodeGetLastStep", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("PihmCVodeGetLastOrder", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("PihmCVodeGetCurrentTime", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("PihmCVodeGetCurrentStep", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("pihm_drive", C.c_int, [C.POINTER(Backend), C.c_void_p, C.POINTER(RunConfig), C.POINTER(RunStats)]),
    ("pihm_gpu_backend_create", C.c_int, [C.c_void_p, C.POINTER(Forcing), C.POINTER(C.c_void_p)]),
    ("pihm_gpu_backend_destroy", None, [C.c_void_p]),
    ("pihm_gpu_backend_get", C.c_int, [C.c_void_p, C.POINTER(Backend)]),
    ("pihm_run_gpu", C.c_int, [C.c_void_p, C.POINTER(Forcing), C.POINTER(RunConfig), c_double_p,
                               C.POINTER(RunStats)]),
    ("pihm_ens_create", C.c_int, [C.c_void_p, C.POINTER(Forcing), C.POINTER(RunConfig), C.POINTER(C.c_void_p)]),
    ("pihm_ens_destroy", None, [C.c_void_p]),
    ("pihm_ens_set_state", C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    ("pihm_ens_get_state", C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    ("pihm_ens_set_member_enabled", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("pihm_ens_run", C.c_int, [C.c_void_p, C.c_double]),
    ("pihm_ens_member_stats", C.c_int, [C.c_void_p, C.c_int, C.POINTER(RunStats)]),
    ("pihm_ens_counters", C.c_int, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    ("pihm_ens_op_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("pihm_ens_create_slots", C.c_int, [C.c_void_p, c_int32_p, c_int32_p, C.POINTER(RunConfig), C.POINTER(C.c_void_p)]),
    ("pihm_ens_load_member", C.c_int, [C.c_void_p, C.c_int, C.POINTER(MeshDesc), C.POINTER(Forcing), c_double_p,
                                       C.c_double]),
    ("pihm_ens_advance", C.c_int, [C.c_void_p, C.c_long, C.POINTER(C.c_int)]),
    ("pihm_ens_member_status", C.c_int, [C.c_void_p, C.c_int]),
    ("pihm_ens_autotune", C.c_int, [C.c_void_p, C.c_int, C.POINTER(EnsTuning)]),
    ("pihm_output_create", C.c_int, [C.c_void_p, C.c_void_p, C.c_char_p, C.c_void_p, C.c_int, C.c_int,
                                     C.POINTER(C.c_void_p)]),
    ("pihm_output_destroy", None, [C.c_void_p]),
    ("pihm_output_update", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("pihm_output_start", C.c_int, [C.c_void_p, C.c_void_p]),
    ("pihm_output_print", C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    ("pihm_output_write_ic", C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
    ("pihm_run_gpu_producers", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(RunConfig), c_double_p,
                                         C.POINTER(RunStats)]),
    ("pihm_ens_set_tuning", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("pihm_ctx_set_rhs_path", C.c_int, [C.c_void_p, C.c_int]),
    ("pihm_ctx_get_rhs_path", C.c_int, [C.c_void_p]),
    ("pihm_ctx_set_tile_span", C.c_int, [C.c_void_p, C.c_int]),
    ("pihm_ctx_get_tile_span", C.c_int, [C.c_void_p]),
    ("pihm_ctx_autotune", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int), c_double_p]),
    ("pihm_lsm_create", C.c_int, [C.c_void_p, C.POINTER(LsmDesc), C.POINTER(Meteo), C.POINTER(C.c_void_p)]),
    ("pihm_lsm_destroy", None, [C.c_void_p]),
    ("pihm_lsm_set_veg", C.c_int, [C.c_void_p, c_double_p, c_double_p, C.c_double]),
    ("pihm_lsm_set_state", C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    ("pihm_lsm_get_state", C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    ("pihm_lsm_step", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_void_p]),
    ("pihm_lsm_get_fluxes", C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p]),
    ("pihm_run_gpu_lsm", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(RunConfig), c_double_p,
                                   C.POINTER(RunStats)]),
    ("pihm_bc_create", C.c_int, [C.c_void_p, c_int32_p, C.POINTER(Ts), c_int32_p, C.POINTER(Ts),
                                 C.POINTER(C.c_void_p)]),
    ("pihm_bc_destroy", None, [C.c_void_p]),
    ("pihm_bc_apply", C.c_int, [C.c_void_p, C.c_double]),
    ("pihm_bc_time", C.c_int, [C.c_void_p, C.c_double, C.c_int, c_double_p]),
    ("pihm_lsm_time", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_int, c_double_p]),
    ("pihm_ctx_get_step_inputs", C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                           c_double_p]),
]


def exported_symbols():
    """Names the C ABI exports (used by the symbol-table test)."""
    return [p[0] for p in _PROTOS]


def lib():
    """Load libpihm_amd.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libpihm_amd.so not found at {LIB_PATH}; build it with "
            "`make -C pihm-cuda_amd` (or __graft_entry__.build()). There is no CPU fallback.")
    L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    for name, res, args in _PROTOS:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
