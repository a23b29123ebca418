"""ctypes binding of libmsmpb_hip.so (the forcefield.h C ABI plus msmpb_hip.h).

torch is imported first on purpose: its bundled HIP runtime and the one the
library was linked against share the soname libamdhip64.so.7, so loading
torch first makes both use a single runtime in the process.
"""
import ctypes
import os

import torch  # noqa: F401  (see module docstring)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsmpb_hip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

MSM_PREC_FP64, MSM_PREC_MIXED = 0, 1
MSM_LAYOUT_FIXED, MSM_LAYOUT_ASIS = 0, 1
MSM_METHOD_MSM, MSM_METHOD_PME = 0, 1
# return codes of include/msmpb_hip.h
MSM_OK, MSM_ERR_ARG, MSM_ERR_STATE, MSM_ERR_HIP, MSM_ERR_ALLOC = 0, -1, -2, -3, -4


class FFStruct(ctypes.Structure):
    """Public part of FF (include/forcefield.h; code/forcefield.h:22-69)."""
    _fields_ = [
        ("N", ctypes.c_int),
        ("q", _dp),
        ("A", (ctypes.c_double * 3) * 3),
        ("Ai", (ctypes.c_double * 3) * 3),
        ("detA", ctypes.c_double),
        ("relCutoff", ctypes.c_double),
        ("errEst", ctypes.c_double),
        ("errEstMoore", ctypes.c_double),
        ("ntc", ctypes.c_int),
        ("tmax", ctypes.c_int),
        ("estErr", ctypes.c_double),
        ("orderAcc", ctypes.c_int),
        ("maxLevel", ctypes.c_int),
        ("topGridDim", ctypes.c_int * 3),
        ("nlist", _ip),
        ("nlist_len", ctypes.c_int),
        ("nlist_interaction_count", ctypes.c_int),
        ("FFT", ctypes.c_bool),
        ("fftw_in", ctypes.c_void_p),
        ("forward", ctypes.c_void_p),
        ("backward", ctypes.c_void_p),
        ("tau", _dp),
        ("dsigma", _dp),
        ("sigmad", ctypes.POINTER(_dp)),
        ("Q", _dp),
        ("J", _dp),
        ("omegap", ctypes.c_void_p),
        ("nLim", ctypes.c_int),
        ("aCut", _dp),
        ("cutoff", ctypes.c_double),
        ("cL", _dp * 3),
        ("khat", ctypes.POINTER(_dp)),
        ("coeff1", ctypes.c_double),
        ("coeff2", ctypes.c_double),
        ("time_partcl2partcl", ctypes.c_double),
        ("time_nlist", ctypes.c_double),
        ("time_grid2grid", ctypes.c_double * 10),
        ("time_stencil", ctypes.c_double * 10),
        ("time_anterpolation", ctypes.c_double),
        ("time_interpolation", ctypes.c_double),
        ("time_restriction", ctypes.c_double * 10),
        ("time_prolongation", ctypes.c_double * 10),
        ("time_padding", ctypes.c_double * 10),
    ]


FFP = ctypes.POINTER(FFStruct)

# every symbol include/forcefield.h and include/msmpb_hip.h declare
EXPORTS = [
    "FF_new", "FF_get_cutoff", "FF_set_cutoff", "FF_set_orderAcc", "FF_set_maxLevel",
    "FF_set_topGridDim", "FF_set_FFT", "FF_build", "FF_get_orderAcc", "FF_get_maxLevel",
    "FF_get_topGridDim", "FF_get_FFT", "FF_get_errEst", "FF_get_estErr", "FF_set_estErr",
    "FF_rebuild", "FF_energy", "FF_delete", "FF_get_deltaF", "msm4g_tictocmanager", "msm4g_tic",
    "msm4g_toc", "invert", "omegap", "prod", "prodT", "_FF_get_errEst", "grid2grid",
    "FF_last_error", "FF_set_precision", "FF_get_precision", "FF_set_layout_compat",
    "FF_set_stage_timing", "FF_get_energy_terms", "FF_energy_potential", "msm_eval_device",
    "msm_forward_device", "msm_adjoint_device", "msm_grid_dims", "msm_stencil_dims",
    "msm_copy_grid", "msm_copy_stencil", "msm_count_pairs", "msm_stage_times", "msm_stage_name",
    "FF_set_method", "FF_get_beta", "msm_grid2grid_device_noalloc",
    # include/msm_bridge.h
    "msmb_create", "msmb_build", "msmb_rebuild", "msmb_energy", "msmb_configure", "msmb_info",
    "msmb_tables", "msmb_times", "msmb_delete",
]

# every symbol include/pme.h and include/pme_hip.h declare (libpme_hip.so)
PME_EXPORTS = [
    "FF_new", "FF_set_cutoff", "FF_get_cutoff", "FF_set_orderAcc", "FF_set_topGridDim",
    "FF_set_tolDir", "FF_build", "FF_get_orderAcc", "FF_get_topGridDim", "FF_get_tolDir",
    "FF_get_errEst", "FF_rebuild", "FF_energy", "FF_delete", "msm4g_tictocmanager", "msm4g_tic",
    "msm4g_toc", "prod", "prodT", "pme_set_precision",
]

class PMEStruct(ctypes.Structure):
    """FF of include/pme.h (the reference's code/pme.h:15-44)."""
    _fields_ = [
        ("N", ctypes.c_int),
        ("q", _dp),
        ("A", (ctypes.c_double * 3) * 3),
        ("Ai", (ctypes.c_double * 3) * 3),
        ("detA", ctypes.c_double),
        ("cutoff", ctypes.c_double),
        ("orderAcc", ctypes.c_int),
        ("topGridDim", ctypes.c_int * 3),
        ("nlist", _ip),
        ("nlist_len", ctypes.c_int),
        ("nlist_interaction_count", ctypes.c_int),
        ("tolDir", ctypes.c_double),
        ("fftw_in", ctypes.c_void_p),
        ("forward", ctypes.c_void_p),
        ("backward", ctypes.c_void_p),
        ("beta", ctypes.c_double),
        ("Q", _dp),
        ("cL", _dp * 3),
        ("khat", _dp),
        ("coeff1", ctypes.c_double),
        ("coeff2", ctypes.c_double),
        ("time_partcl2partcl", ctypes.c_double),
        ("time_nlist", ctypes.c_double),
        ("time_grid2grid", ctypes.c_double),
        ("time_stencil", ctypes.c_double),
        ("time_anterpolation", ctypes.c_double),
        ("time_interpolation", ctypes.c_double),
    ]


PMEP = ctypes.POINTER(PMEStruct)
PME_LIB_PATH = os.path.join(os.path.dirname(LIB_PATH), "libpme_hip.so")

_lib = None
_pme = None


def pme_lib():
    """libpme_hip.so (the reference's pme.h API); fails loudly when absent."""
    global _pme
    if _pme is not None:
        return _pme
    lib()  # torch's HIP runtime first, then the MSM library it depends on
    if not os.path.exists(PME_LIB_PATH):
        raise RuntimeError(f"{PME_LIB_PATH} is missing: build it with `python -m msmpb_amd.build`")
    P = ctypes.CDLL(PME_LIB_PATH)
    M3 = (ctypes.c_double * 3) * 3
    P.FF_new.restype = PMEP
    P.FF_new.argtypes = [ctypes.c_int, _dp, ctypes.POINTER(M3)]
    P.FF_set_cutoff.argtypes = [PMEP, ctypes.c_double]
    P.FF_set_orderAcc.argtypes = [PMEP, ctypes.c_int]
    P.FF_set_topGridDim.argtypes = [PMEP, _ip]
    P.FF_set_tolDir.argtypes = [PMEP, ctypes.c_double]
    P.FF_build.argtypes = [PMEP, _dp]
    P.FF_rebuild.argtypes = [PMEP, ctypes.POINTER(M3), _dp]
    P.FF_energy.restype = ctypes.c_double
    P.FF_energy.argtypes = [PMEP, _dp, _dp, _dp]
    P.FF_get_errEst.restype = ctypes.c_double
    P.FF_get_errEst.argtypes = [PMEP, ctypes.c_int, _dp]
    P.FF_get_cutoff.restype = ctypes.c_double
    P.FF_get_cutoff.argtypes = [PMEP]
    P.FF_get_tolDir.restype = ctypes.c_double
    P.FF_get_tolDir.argtypes = [PMEP]
    P.FF_get_orderAcc.argtypes = [PMEP]
    P.FF_get_topGridDim.argtypes = [PMEP, _ip]
    P.FF_delete.argtypes = [PMEP]
    P.pme_set_precision.argtypes = [PMEP, ctypes.c_int]
    _pme = P
    return P


def lib():
    """Load the library; fails loudly when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m msmpb_amd.build` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    L.FF_new.restype = FFP
    L.FF_new.argtypes = [ctypes.c_int, _dp]
    for name in ("FF_set_orderAcc", "FF_set_maxLevel"):
        getattr(L, name).argtypes = [FFP, ctypes.c_int]
    L.FF_set_cutoff.argtypes = [FFP, ctypes.c_double]
    L.FF_set_estErr.argtypes = [FFP, ctypes.c_double]
    L.FF_set_topGridDim.argtypes = [FFP, _ip]
    L.FF_set_FFT.argtypes = [FFP, ctypes.c_bool]
    L.FF_get_FFT.argtypes = [FFP]
    L.FF_get_FFT.restype = ctypes.c_bool
    L.FF_build.argtypes = [FFP, ctypes.c_int, _dp, _dp, ctypes.c_double]
    L.FF_rebuild.argtypes = [FFP, _dp, _dp, ctypes.c_double]
    L.FF_energy.restype = ctypes.c_double
    L.FF_energy.argtypes = [FFP, _dp, _dp, _dp]
    L.FF_energy_potential.restype = ctypes.c_double
    L.FF_energy_potential.argtypes = [FFP, _dp, _dp, _dp]
    L.FF_delete.argtypes = [FFP]
    for name in ("FF_get_cutoff", "FF_get_errEst", "FF_get_estErr", "FF_get_deltaF"):
        getattr(L, name).restype = ctypes.c_double
        getattr(L, name).argtypes = [FFP]
    L._FF_get_errEst.restype = ctypes.c_double
    L._FF_get_errEst.argtypes = [FFP, ctypes.c_int]
    for name in ("FF_get_orderAcc", "FF_get_maxLevel", "FF_get_precision"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [FFP]
    L.FF_get_topGridDim.argtypes = [FFP, _ip]
    L.FF_last_error.restype = ctypes.c_char_p
    L.FF_last_error.argtypes = []
    for name in ("FF_set_precision", "FF_set_layout_compat", "FF_set_stage_timing"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [FFP, ctypes.c_int]
    L.FF_get_energy_terms.restype = ctypes.c_int
    L.FF_get_energy_terms.argtypes = [FFP, _dp]
    L.msm_eval_device.restype = ctypes.c_int
    L.msm_eval_device.argtypes = [FFP, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.msm_forward_device.restype = ctypes.c_int
    L.msm_forward_device.argtypes = [FFP, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int, ctypes.c_void_p]
    L.msm_adjoint_device.restype = ctypes.c_int
    L.msm_adjoint_device.argtypes = [FFP, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    for name in ("msm_grid_dims", "msm_stencil_dims"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [FFP, ctypes.c_int, _ip]
    L.msm_copy_grid.restype = ctypes.c_int
    L.msm_copy_grid.argtypes = [FFP, ctypes.c_int, ctypes.c_int, _dp]
    L.msm_copy_stencil.restype = ctypes.c_int
    L.msm_copy_stencil.argtypes = [FFP, ctypes.c_int, _dp]
    L.msm_count_pairs.restype = ctypes.c_int
    L.msm_count_pairs.argtypes = [FFP, _dp, ctypes.POINTER(ctypes.c_longlong)]
    L.msm_stage_times.restype = ctypes.c_int
    L.msm_stage_times.argtypes = [FFP, _dp, ctypes.c_int]
    L.msm_stage_name.restype = ctypes.c_char_p
    L.msm_stage_name.argtypes = [ctypes.c_int]
    L.msm4g_tictocmanager.restype = ctypes.c_double
    L.msm4g_tictocmanager.argtypes = [ctypes.c_int]
    L.msm4g_toc.restype = ctypes.c_double
    L.invert.restype = ctypes.c_double
    L.invert.argtypes = [_dp]
    L.FF_set_method.restype = ctypes.c_int
    L.FF_set_method.argtypes = [FFP, ctypes.c_int, ctypes.c_double]
    L.FF_get_beta.restype = ctypes.c_double
    L.FF_get_beta.argtypes = [FFP]
    _lib = L
    return L


def last_error():
    return lib().FF_last_error().decode()
