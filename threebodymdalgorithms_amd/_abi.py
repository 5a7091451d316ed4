"""ctypes binding to libtb3.so, the C ABI declared in include/tb3.h.

The library is built in-tree (see threebodymdalgorithms_amd.build) and loaded
from this package directory.  There is no fallback: if the library is missing
or fails to load, every entry point raises TB3Error.
"""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# TB3_LIBRARY names another build of the same library (for example one made
# with different compile options); the default is the in-tree build.
LIB_PATH = os.environ.get("TB3_LIBRARY") or os.path.join(PKG_DIR, "libtb3.so")
CLI_PATH = os.path.join(PKG_DIR, "bin", "tb3")

OK, EINVAL, ENOMEM, EHIP, EUNSUPPORTED, ENODEV = range(6)
STATUS_NAMES = {OK: "TB3_OK", EINVAL: "TB3_EINVAL", ENOMEM: "TB3_ENOMEM", EHIP: "TB3_EHIP",
                EUNSUPPORTED: "TB3_EUNSUPPORTED", ENODEV: "TB3_ENODEV"}

# Same numbering as the reference's AlgorithmType (utility/enums.hpp:1).
NATA, P3BCA, AUTA = 0, 1, 2
ALGORITHMS = {"nata": NATA, "p3bca": P3BCA, "auta": AUTA}

FP64_EXACT, FP64, FP32_MIXED = 0, 1, 2
PRECISIONS = {"exact": FP64_EXACT, "fp64": FP64, "fp32": FP32_MIXED}


class TB3Error(RuntimeError):
    """A tb3_* call failed; `status` is the tb3_status code."""

    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_ctxp = ctypes.c_void_p

# name -> (restype, argtypes); kept in step with include/tb3.h
SIGNATURES = {
    "tb3_version": (ctypes.c_char_p, []),
    "tb3_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "tb3_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_ctxp)]),
    "tb3_destroy": (ctypes.c_int, [_ctxp]),
    "tb3_last_error": (ctypes.c_char_p, [_ctxp]),
    "tb3_configure": (ctypes.c_int, [_ctxp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int]),
    "tb3_set_particles": (ctypes.c_int, [_ctxp, ctypes.c_int64, _i32p, _dp, _dp, _dp, _dp]),
    "tb3_get_particles": (ctypes.c_int, [_ctxp, ctypes.c_int64, _i32p, _dp, _dp, _dp, _dp, _dp]),
    "tb3_step": (ctypes.c_int, [_ctxp, ctypes.c_double, _dp, _u64p, _u64p]),
    "tb3_run": (ctypes.c_int, [_ctxp, ctypes.c_int, ctypes.c_double, _dp, _u64p]),
    "tb3_batch_set": (ctypes.c_int, [_ctxp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "tb3_batch_run": (ctypes.c_int, [_ctxp, ctypes.c_int, ctypes.c_double, _dp, _u64p]),
    "tb3_batch_counts": (ctypes.c_int, [_ctxp, ctypes.c_int32, _u64p]),
    "tb3_batch_energy": (ctypes.c_int, [_ctxp, ctypes.c_int32, _dp, _u64p]),
    "tb3_compute_forces": (ctypes.c_int, [_ctxp, _u64p]),
    "tb3_energy": (ctypes.c_int, [_ctxp, _dp, _u64p]),
    "tb3_energy_grad": (ctypes.c_int, [_ctxp, _dp, _dp, _dp, _u64p]),
    "tb3_energy_dev": (ctypes.c_int, [_ctxp, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                      ctypes.c_void_p, _u64p]),
    "tb3_energy_grad_dev": (ctypes.c_int, [_ctxp, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _u64p]),
    "tb3_energy_grad_dev_noalloc": (ctypes.c_int, [_ctxp, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double,
                                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _u64p]),
    "tb3_autotune": (ctypes.c_int, [_ctxp, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int)]),
    "tb3_force_scale": (ctypes.c_int, [_ctxp, _dp]),
    "tb3_get_tuning": (ctypes.c_int, [_ctxp, _dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "tb3_set_tuning": (ctypes.c_int, [_ctxp, ctypes.c_double, ctypes.c_int]),
    "tb3_set_cell_slots": (ctypes.c_int, [_ctxp, ctypes.c_int]),
    "tb3_get_cell_slots": (ctypes.c_int, [_ctxp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "tb3_set_profiling": (ctypes.c_int, [_ctxp, ctypes.c_int]),
    "tb3_get_profile": (ctypes.c_int, [_ctxp, _dp, ctypes.c_int64]),
    "tb3_selftest_fp64": (ctypes.c_int, [_ctxp, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp]),
    "tb3_selftest_fp64_inrange": (ctypes.c_int, [_ctxp, ctypes.c_int64, _dp, _dp, _dp, _dp]),
}

_lib = None


def load():
    """Load libtb3.so (once).  Raises TB3Error if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TB3Error(ENODEV, f"{LIB_PATH} is not built; run threebodymdalgorithms_amd.build.build()")
    # torch carries its own HIP runtime.  Loaded first, it provides the
    # libamdhip64.so.7 the library needs, and the process has one runtime;
    # loaded after the library, it would bring a second HSA runtime into the
    # process, which then finds no GPU for torch.
    import torch  # noqa: F401
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise TB3Error(ENODEV, f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(ctx, status):
    if status != OK:
        msg = load().tb3_last_error(ctx)
        raise TB3Error(status, msg.decode() if msg else "")
