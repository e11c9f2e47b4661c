"""ctypes binding of the C ABI declared in include/pmnb.h.

The library is the HIP build in this package directory.  There is no CPU
fallback: if the library is missing the import of the force module fails
with a message naming the build step.
"""
import ctypes
import os

from . import _build

PMNB_OK = 0
PMNB_ERR_ARG = 1
PMNB_ERR_HIP = 2
PMNB_ERR_UNSUPPORTED = 3
PMNB_ERR_STATE = 4
PMNB_ERR_CAPACITY = 5

PMNB_ELEC_NONE = 0
PMNB_ELEC_COULOMB = 1
PMNB_ELEC_EWALD_REAL = 2
PMNB_SWITCH_C2 = 0
PMNB_SWITCH_C1 = 1
PMNB_PREC_FP64 = 0
PMNB_PREC_MIXED = 1
PMNB_FLAG_VIRIAL = 1
PMNB_FLAG_MOLVIRIAL = 2

# Every function declared in include/pmnb.h.
EXPORTED = (
    "pmnb_create", "pmnb_set_topology", "pmnb_set_box", "pmnb_evaluate",
    "pmnb_evaluate_device", "pmnb_pairset", "pmnb_force_pairs",
    "pmnb_kernel_timing", "pmnb_pair_path",
    "pmnb_last_pair_count",
    "pmnb_last_error", "pmnb_destroy", "pmnb_version",
    "pmnb_batch_create", "pmnb_batch_add", "pmnb_batch_size",
    "pmnb_batch_evaluate_device", "pmnb_batch_run_graph",
    "pmnb_batch_autotune", "pmnb_batch_set_shape", "pmnb_batch_force_pairs",
    "pmnb_batch_pair_count",
    "pmnb_batch_last_error", "pmnb_batch_destroy",
    "pmnb_set_parameters_device", "pmnb_adjoint_device",
    "pmnb_lf_create", "pmnb_lf_run_device", "pmnb_lf_set_bonded",
    "pmnb_lf_forces_device", "pmnb_lf_iterations", "pmnb_lf_destroy",
    "pmnb_lf_run_graph", "pmnb_batch_lf_create",
    "pmnb_evaluate_device_noalloc",
    "pmnb_ewald_create", "pmnb_ewald_info", "pmnb_ewald_evaluate_device",
    "pmnb_ewald_destroy", "pmnb_pme_create", "pmnb_pme_info",
    "pmnb_pme_evaluate_device", "pmnb_pme_destroy", "pmnb_lf_set_kspace",
)


class PmnbParams(ctypes.Structure):
    _fields_ = [
        ("elec", ctypes.c_int32),
        ("elec_switch", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("lj_switchon", ctypes.c_double),
        ("lj_cutoff", ctypes.c_double),
        ("elec_switchon", ctypes.c_double),
        ("elec_cutoff", ctypes.c_double),
        ("algorithm_cutoff", ctypes.c_double),
        ("cell_size", ctypes.c_double),
        ("ewald_alpha", ctypes.c_double),
    ]


class PmnbLfParams(ctypes.Structure):
    _fields_ = [
        ("timestep", ctypes.c_double),
        ("shake", ctypes.c_int32),
        ("rattle", ctypes.c_int32),
        ("shake_epsilon", ctypes.c_double),
        ("shake_max_iter", ctypes.c_int32),
        ("rattle_max_iter", ctypes.c_int32),
        ("rattle_epsilon", ctypes.c_double),
    ]


class PmnbError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("pmnb error %d: %s" % (code, message))
        self.code = code


_LIB = None

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_u32 = ctypes.c_uint32


def load():
    """Returns the loaded library; raises ImportError if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(_build.LIB_PATH):
        raise ImportError(
            "libpmnb.so is not built (%s); run "
            "`python -m protomoladdon_amd._build` or __graft_entry__.build()"
            % _build.LIB_PATH)
    lib = ctypes.CDLL(_build.LIB_PATH)
    lib.pmnb_create.argtypes = [ctypes.POINTER(PmnbParams), ctypes.POINTER(_p)]
    lib.pmnb_create.restype = ctypes.c_int
    lib.pmnb_set_topology.argtypes = [_p, _i32, _p, _p, _p, _i32, _p, _dbl,
                                      _i32, _p]
    lib.pmnb_set_topology.restype = ctypes.c_int
    lib.pmnb_set_box.argtypes = [_p, _p, _p, _p, _p]
    lib.pmnb_set_box.restype = ctypes.c_int
    lib.pmnb_evaluate.argtypes = [_p, _p, _p, _p, _p, _p, _p, _u32]
    lib.pmnb_evaluate.restype = ctypes.c_int
    lib.pmnb_evaluate_device.argtypes = [_p, _p, _p, _p, _p, _u32, _p]
    lib.pmnb_evaluate_device_noalloc.argtypes = [_p, _p, _p, _p, _p, _u32, _p]
    lib.pmnb_evaluate_device_noalloc.restype = ctypes.c_int
    lib.pmnb_evaluate_device.restype = ctypes.c_int
    lib.pmnb_pairset.argtypes = [_p, _p, _p, _i64, ctypes.POINTER(_i64)]
    lib.pmnb_pairset.restype = ctypes.c_int
    lib.pmnb_force_pairs.argtypes = [_p, _p, _p, _i64, ctypes.POINTER(_i64)]
    lib.pmnb_force_pairs.restype = ctypes.c_int
    lib.pmnb_batch_force_pairs.argtypes = [_p, _p, _p, _i64,
                                           ctypes.POINTER(_i64)]
    lib.pmnb_batch_force_pairs.restype = ctypes.c_int
    lib.pmnb_pair_path.argtypes = [_p, ctypes.POINTER(_i32)]
    lib.pmnb_pair_path.restype = ctypes.c_int
    lib.pmnb_kernel_timing.argtypes = [_p, _i32, _p, ctypes.POINTER(_i64)]
    lib.pmnb_kernel_timing.restype = ctypes.c_int
    lib.pmnb_last_pair_count.argtypes = [_p, ctypes.POINTER(_i64)]
    lib.pmnb_last_pair_count.restype = ctypes.c_int
    lib.pmnb_last_error.argtypes = [_p]
    lib.pmnb_last_error.restype = ctypes.c_char_p
    lib.pmnb_destroy.argtypes = [_p]
    lib.pmnb_destroy.restype = None
    lib.pmnb_version.argtypes = []
    lib.pmnb_version.restype = ctypes.c_int
    lib.pmnb_batch_create.argtypes = [ctypes.POINTER(PmnbParams), _i32, _p,
                                      _dbl, ctypes.POINTER(_p)]
    lib.pmnb_batch_create.restype = ctypes.c_int
    lib.pmnb_batch_add.argtypes = [_p, _i32, _p, _p, _p, _i32, _p, _p, _p, _p,
                                   _p, ctypes.POINTER(_i32)]
    lib.pmnb_batch_add.restype = ctypes.c_int
    lib.pmnb_batch_size.argtypes = [_p, ctypes.POINTER(_i32),
                                    ctypes.POINTER(_i64)]
    lib.pmnb_batch_size.restype = ctypes.c_int
    lib.pmnb_batch_evaluate_device.argtypes = [_p, _p, _p, _p, _u32, _p]
    lib.pmnb_batch_evaluate_device.restype = ctypes.c_int
    lib.pmnb_batch_run_graph.argtypes = [_p, _p, _p, _p, _u32, _i32, _p]
    lib.pmnb_batch_run_graph.restype = ctypes.c_int
    lib.pmnb_batch_autotune.argtypes = [_p, _p, _p, _p, _u32,
                                        ctypes.POINTER(_i32),
                                        ctypes.POINTER(_i32),
                                        ctypes.POINTER(_dbl)]
    lib.pmnb_batch_autotune.restype = ctypes.c_int
    lib.pmnb_batch_set_shape.argtypes = [_p, _i32, _i32]
    lib.pmnb_batch_set_shape.restype = ctypes.c_int
    lib.pmnb_batch_pair_count.argtypes = [_p, ctypes.POINTER(_i64)]
    lib.pmnb_batch_pair_count.restype = ctypes.c_int
    lib.pmnb_batch_last_error.argtypes = [_p]
    lib.pmnb_batch_last_error.restype = ctypes.c_char_p
    lib.pmnb_batch_destroy.argtypes = [_p]
    lib.pmnb_batch_destroy.restype = None
    lib.pmnb_set_parameters_device.argtypes = [_p, _p, _p, _p]
    lib.pmnb_set_parameters_device.restype = ctypes.c_int
    lib.pmnb_adjoint_device.argtypes = [_p, _p, _p, _p, _p, _p, _p]
    lib.pmnb_adjoint_device.restype = ctypes.c_int
    lib.pmnb_lf_create.argtypes = [_p, ctypes.POINTER(PmnbLfParams), _i32, _p,
                                   _i32, _p, _p, ctypes.POINTER(_p)]
    lib.pmnb_lf_create.restype = ctypes.c_int
    lib.pmnb_lf_run_device.argtypes = [_p, _p, _p, _p, _p, _i32, _u32, _p]
    lib.pmnb_lf_run_device.restype = ctypes.c_int
    lib.pmnb_lf_run_graph.argtypes = [_p, _p, _p, _p, _p, _i32, _i32, _u32, _p]
    lib.pmnb_lf_run_graph.restype = ctypes.c_int
    lib.pmnb_batch_lf_create.argtypes = [_p, ctypes.POINTER(PmnbLfParams), _i32,
                                         _p, _i32, _p, _p, ctypes.POINTER(_p)]
    lib.pmnb_batch_lf_create.restype = ctypes.c_int
    lib.pmnb_lf_iterations.argtypes = [_p, _i64, _p, _i64,
                                       ctypes.POINTER(_i64),
                                       ctypes.POINTER(_i64)]
    lib.pmnb_lf_iterations.restype = ctypes.c_int
    lib.pmnb_lf_set_bonded.argtypes = [_p, _i32, _p, _p, _p, _i32, _p, _p, _p,
                                       _p, _p]
    lib.pmnb_lf_set_bonded.restype = ctypes.c_int
    lib.pmnb_lf_forces_device.argtypes = [_p, _p, _p, _p, _u32, _p]
    lib.pmnb_lf_forces_device.restype = ctypes.c_int
    lib.pmnb_ewald_create.argtypes = [_p, _dbl, _dbl, _i32, _p, _dbl, _i32, _p,
                                      ctypes.POINTER(_p)]
    lib.pmnb_ewald_create.restype = ctypes.c_int
    lib.pmnb_ewald_info.argtypes = [_p, _p, _p, _p, _p]
    lib.pmnb_ewald_info.restype = ctypes.c_int
    lib.pmnb_ewald_evaluate_device.argtypes = [_p, _p, _p, _p, _u32, _p]
    lib.pmnb_ewald_evaluate_device.restype = ctypes.c_int
    lib.pmnb_ewald_destroy.argtypes = [_p]
    lib.pmnb_ewald_destroy.restype = None
    lib.pmnb_pme_create.argtypes = [_p, _i32, _i32, _i32, _i32, _dbl, _dbl,
                                    _dbl, _i32, _p, _dbl, _i32, _p,
                                    ctypes.POINTER(_p)]
    lib.pmnb_pme_create.restype = ctypes.c_int
    lib.pmnb_pme_info.argtypes = [_p, _p, _p]
    lib.pmnb_pme_info.restype = ctypes.c_int
    lib.pmnb_pme_evaluate_device.argtypes = [_p, _p, _p, _p, _u32, _p]
    lib.pmnb_pme_evaluate_device.restype = ctypes.c_int
    lib.pmnb_pme_destroy.argtypes = [_p]
    lib.pmnb_pme_destroy.restype = None
    lib.pmnb_lf_set_kspace.argtypes = [_p, _p, _p]
    lib.pmnb_lf_set_kspace.restype = ctypes.c_int
    lib.pmnb_lf_destroy.argtypes = [_p]
    lib.pmnb_lf_destroy.restype = None
    _LIB = lib
    return lib


def check(lib, ctx, code):
    if code != PMNB_OK:
        msg = lib.pmnb_last_error(ctx) if ctx else b""
        raise PmnbError(code, (msg or b"").decode(errors="replace"))


def check_batch(lib, batch, code):
    if code != PMNB_OK:
        msg = lib.pmnb_batch_last_error(batch) if batch else b""
        raise PmnbError(code, (msg or b"").decode(errors="replace"))
