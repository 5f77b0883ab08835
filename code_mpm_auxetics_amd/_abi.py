"""ctypes binding of the C ABI declared in ``include/mpm_auxetics.h``.

The library is loaded from this package directory (built in place by
``build.py``).  A missing library raises: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_auxetics.so")

# status codes
MPM_OK = 0
MPM_E_ARG = -1
MPM_E_HIP = -2
MPM_E_NAN = -3
MPM_E_STATE = -4
MPM_E_OUT_OF_GRID = -5
MPM_E_ARENA = -6
MPM_E_NOMEM = -7
MPM_E_NO_DEVICE = -8

MPM_FP64 = 0
MPM_FP32 = 1
MPM_PATH_PLAIN = 0
MPM_PATH_CONTACT = 1
MPM_LAUNCH_GRAPH = 0
MPM_LAUNCH_DIRECT = 1

MPM_ELASTIC, MPM_NEOHOOKEAN, MPM_VISCOELASTIC, MPM_PLASTIC, MPM_VONMISESHARDENING, MPM_GASS = range(6)
MPM_FIX_X, MPM_FIX_Y, MPM_FIX_Z = 1, 2, 4

MPM_F_X, MPM_F_V, MPM_F_F, MPM_F_STRESS, MPM_F_STRAIN, MPM_F_EPSP, MPM_F_ALPHA, MPM_F_VOL, MPM_F_KERNEL = (
    1 << k for k in range(9))
MPM_F_ALL = 0x1FF
MPM_F_J2 = 1 << 9
MPM_F_CELL = 1 << 10

# every symbol the header declares (checked by tests/test_abi.py)
EXPORTED = [
    "mpm_abi_version", "mpm_create", "mpm_upload_points", "mpm_step", "mpm_readback",
    "mpm_checks", "mpm_reaction_history", "mpm_get_time", "mpm_last_step_ms",
    "mpm_last_error", "mpm_destroy", "mpm_sort_keys", "mpm_kernel_times", "mpm_group_reaction_history",
    "mpm_diff_create", "mpm_diff_forward", "mpm_diff_backward", "mpm_diff_last_error", "mpm_diff_destroy",
    "mpm_report", "mpm_set_log", "mpm_constitutive", "mpm_step_async", "mpm_step_poll", "mpm_step_wait",
    "mpm_grid_nodes", "mpm_constitutive_noalloc",
]

# constitutive laws of mpm_constitutive (MaterialPoint.h:8-16 type codes; 16: calculateState_J2)
MPM_LAW_ELASTIC, MPM_LAW_NEOHOOKEAN, MPM_LAW_VISCOELASTIC, MPM_LAW_PLASTIC, MPM_LAW_VONMISES, MPM_LAW_IDEALGAS = range(6)
MPM_LAW_J2_STATE = 16

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)


class MpmConfig(C.Structure):
    _fields_ = [
        ("length", C.c_double * 3),
        ("cells", C.c_int32 * 3),
        ("dt_max", C.c_double),
        ("time_start", C.c_double),
        ("time_end", C.c_double),
        ("damping", C.c_double),
        ("mass_min", C.c_double),
        ("console_interval", C.c_double),
        ("precision", C.c_int32),
        ("strain_form", C.c_int32),
        ("neohookean", C.c_int32),
        ("track_strain", C.c_int32),
        ("n_timeline", C.c_int32),
        ("timeline_t", _dp),
        ("timeline_v", _dp),
        ("node_fixed", _u8p),
        ("brick", C.c_int32 * 3),
        ("drift_check", C.c_int32),
        ("friction", C.c_double),
        ("device", C.c_int32),
        ("verbose", C.c_int32),
        ("profile", C.c_int32),
        ("launch", C.c_int32),
        ("node_group", C.POINTER(C.c_uint16)),
        ("n_groups", C.c_int32),
        ("persistent", C.c_int32),
    ]


class MpmPoints(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("x", _dp), ("v", _dp), ("fext", _dp), ("F", _dp), ("strain", _dp), ("stress", _dp),
        ("eps_p", _dp), ("alpha", _dp), ("mass", _dp), ("vol", _dp), ("vol0", _dp),
        ("E", _dp), ("nu", _dp), ("sy", _dp), ("C0", _dp), ("C1", _dp),
        ("type", _i32p), ("dc", _u8p), ("surface", _u8p),
        ("kernel", _dp), ("kernel_grad", _dp),
        ("j2", _dp), ("cell", _i32p), ("order", _u32p),
    ]


class MpmNodes(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("active", _u8p), ("fixed", _u8p),
        ("x", _dp), ("mass", _dp), ("v", _dp), ("f", _dp), ("f_temp", _dp),
        ("kernel", _dp), ("kernel_grad", _dp),
    ]


class MpmCrBatch(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("E", _dp), ("nu", _dp), ("yield_", _dp), ("alpha", _dp), ("C0", _dp), ("C1", _dp),
        ("stress", _dp), ("strain_inc", _dp), ("F", _dp),
        ("stress_inc", _dp), ("plastic_inc", _dp), ("alpha_inc", _dp), ("j2", _dp),
    ]


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libmpm_auxetics.so is not built; run `python -m code_mpm_auxetics_amd.build` "
            "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.mpm_abi_version.restype = C.c_int
    L.mpm_create.argtypes = [C.POINTER(MpmConfig), C.POINTER(C.c_void_p)]
    L.mpm_upload_points.argtypes = [C.c_void_p, C.POINTER(MpmPoints)]
    L.mpm_step.argtypes = [C.c_void_p, C.c_double, C.c_int32, _i32p]
    L.mpm_step_async.argtypes = [C.c_void_p, C.c_double, C.c_int32]
    L.mpm_step_poll.argtypes = [C.c_void_p, _i32p]
    L.mpm_step_wait.argtypes = [C.c_void_p, _i32p]
    L.mpm_readback.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(MpmPoints)]
    L.mpm_grid_nodes.argtypes = [C.c_void_p, C.POINTER(MpmNodes)]
    L.mpm_checks.argtypes = [C.c_void_p, _dp]
    L.mpm_reaction_history.argtypes = [C.c_void_p, _dp, C.c_int32, _i32p]
    L.mpm_group_reaction_history.argtypes = [C.c_void_p, _dp, C.c_int32, _i32p]
    L.mpm_get_time.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.mpm_last_step_ms.argtypes = [C.c_void_p, _dp]
    L.mpm_last_error.argtypes = [C.c_void_p]
    L.mpm_last_error.restype = C.c_char_p
    L.mpm_destroy.argtypes = [C.c_void_p]
    L.mpm_destroy.restype = None
    L.mpm_sort_keys.argtypes = [_u32p, C.c_int32, _u32p, _u32p, C.c_int32]
    L.mpm_kernel_times.argtypes = [C.c_void_p, _dp]
    L.mpm_report.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int32, _i32p]
    L.mpm_set_log.argtypes = [C.c_void_p, C.c_char_p]
    L.mpm_constitutive.argtypes = [C.c_int32, C.POINTER(MpmCrBatch), C.c_int32]
    L.mpm_constitutive_noalloc.argtypes = [C.c_int32, C.POINTER(MpmCrBatch), C.c_int32]
    L.mpm_diff_create.argtypes = [C.POINTER(MpmConfig), C.POINTER(MpmPoints), C.c_int32, C.c_int32,
                                  C.POINTER(C.c_void_p)]
    L.mpm_diff_forward.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.mpm_diff_backward.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.mpm_diff_last_error.argtypes = [C.c_void_p]
    L.mpm_diff_last_error.restype = C.c_char_p
    L.mpm_diff_destroy.argtypes = [C.c_void_p]
    L.mpm_diff_destroy.restype = None
    for name in EXPORTED:
        if name not in ("mpm_last_error", "mpm_destroy", "mpm_abi_version", "mpm_diff_last_error",
                        "mpm_diff_destroy"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


class MpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"MPM error {code}: {msg}")
        self.code = code


def check(code, handle=None):
    if code != MPM_OK:
        msg = lib().mpm_last_error(handle)
        raise MpmError(code, msg.decode() if msg else "")
    return code
