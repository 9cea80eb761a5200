"""ctypes binding of libddla_mi355x.so (the C ABI declared in include/ddla_mi355x.h).

Mirrors the reference's python/ddla/_c_api.py: status constants, exception
classes and check_status.  The shared library must have been built
(``python -c "import __graft_entry__ as g; g.build()"``); importing this module
without it raises ImportError -- there is no fallback implementation.
"""
import ctypes
import os

STATUS_SUCCESS = 0
STATUS_INVALID_HANDLE = 1
STATUS_INVALID_VALUE = 2
STATUS_ALLOC_FAILED = 3
STATUS_BACKEND_UNAVAILABLE = 4
STATUS_INTERNAL_ERROR = 5
STATUS_EXCEPTION = 6

BACKEND_CPU = 0
BACKEND_GPU = 1

DTYPE_, CTXT_, M_, N_, MB_, NB_, RSRC_, CSRC_, LLD_, DLEN_ = range(10)

LIBRARY_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libddla_mi355x.so")


class DdlaError(RuntimeError):
    """Base class of LibDDLA C API failures."""


class DdlaInvalidHandleError(DdlaError):
    pass


class DdlaInvalidValueError(DdlaError):
    pass


class DdlaAllocFailedError(DdlaError):
    pass


class DdlaBackendUnavailableError(DdlaError):
    pass


class DdlaInternalError(DdlaError):
    pass


class DdlaExceptionError(DdlaError):
    """A C++ exception escaped the library (status 6); the message is its what()."""


def _load():
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(
            f"{LIBRARY_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    # One HIP runtime per process.  torch ships its own libamdhip64 and
    # libhsa-runtime64 under the same sonames as the system ROCm ones this
    # library links; loaded first, they satisfy this library's dependencies,
    # whereas loading this library first would leave torch to load a second
    # runtime that then finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return ctypes.CDLL(LIBRARY_PATH)


lib = _load()

_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)
_i = ctypes.c_int
_c = ctypes.c_char
_ll = ctypes.c_longlong

_SIGNATURES = {
    "ddla_c_error_string": (ctypes.c_char_p, [_i]),
    "ddla_c_init_backend": (_i, [ctypes.POINTER(_vp), _i]),
    "ddla_c_set": (_i, [_vp, _i, _i, _i, _c]),
    "ddla_c_destroy": (_i, [ctypes.POINTER(_vp)]),
    "ddla_c_backend_available": (_i, [_i]),
    "ddla_c_get_backend": (_i, [_vp]),
    "ddla_c_grid_dims": (_i, [_vp, _ip, _ip]),
    "ddla_c_grid_coords": (_i, [_vp, _ip, _ip]),
    "ddla_c_rank": (_i, [_vp]),
    "ddla_c_size": (_i, [_vp]),
    "ddla_c_descinit": (_i, [_ip, _vp, _i, _i, _i, _i, _i, _i, _i]),
    "ddla_c_numloc": (_i, [_i, _i, _i, _i, _i]),
    "ddla_c_malloc": (_i, [_vp, ctypes.POINTER(_vp), _ll]),
    "ddla_c_free": (_i, [_vp, _vp]),
    "ddla_c_memcpy": (_i, [_vp, _vp, _vp, _ll]),
    "ddla_c_memset": (_i, [_vp, _vp, _i, _ll]),
    "ddla_c_memcpy2d": (_i, [_vp, _vp, _ll, _vp, _ll, _ll, _ll]),
    "ddla_c_set_wait_limit": (_i, [_i]),
    "ddla_c_set_panel_width": (_i, [_i]),
    "ddla_c_set_communication": (_i, [_vp, _i]),
    "ddla_c_get_communication": (_i, [_vp, ctypes.POINTER(_i)]),
    "ddla_c_communication_available": (_i, [_vp, _i]),
    "ddla_c_default_comm": (_i, []),
    "ddla_c_comm_rank": (_i, [_i]),
    "ddla_c_comm_size": (_i, [_i]),
    "ddla_c_bcast_dims": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i), _i]),
    "ddla_c_synchronize": (_i, [_vp]),
    "ddla_c_set_stream": (_i, [_vp, _vp]),
    "ddla_c_get_stream": (_vp, [_vp]),
    "ddla_c_fill_normal": (_i, [_vp, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_ulonglong, _ll, _vp]),
    "ddla_d_spd_generate": (_i, [_vp, _i, ctypes.c_ulonglong, ctypes.c_double, _vp, _i, _vp]),
    "ddla_d_pposv_mixed": (_i, [_vp, _c, _i, _i, _vp, _i, _i, _ip, _vp, _i, _i, _ip, _i, ctypes.c_double, _i, _ip,
                                _ip]),
    "ddla_d_pposv_noalloc": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _i, _i, _ip, _vp, _i, _i, _ip, _ip]),
    "ddla_d_pposv_mixed_workspace_size": (_i, [_vp, _c, _i, _i, _ip, _ip, _llp, _llp]),
    "ddla_d_pposv_mixed_run": (_i, [_vp, _c, _i, _i, _vp, _ip, _vp, _ip, _i, ctypes.c_double, _i, _ip, _ip, _vp,
                                    _ll]),
    "ddla_c_spd_generate_batch": (_i, [_vp, _i, _ip, ctypes.c_ulonglong, ctypes.c_double, _i, _vp, _ip, _i, _vp,
                                       _ip, _vp]),
    "ddla_c_graph_launch": (_i, [_vp, _vp]),
    "ddla_c_graph_destroy": (_i, [_vp]),
    "ddla_c_vbatched_schedule": (_i, [_vp, _i, _ip, _ip, _ip, _ip]),
    "ddla_c_vbatched_set_schedule": (_i, [_vp, _i, _i, _i, _i]),
    "ddla_c_vbatched_kernel": (_i, [_vp, _i, _ip]),
    "ddla_c_potrf_schedule": (_i, [_i, _i, _ip, _i, _ip]),
    "ddla_d_gp_batched_workspace_size": (_i, [_i, _llp]),
    "ddla_d_gp_batched_forward": (_i, [_vp, _i, _i, _vp, _ll, _vp, _ll, _vp, _vp, _vp]),  # info: device array
    "ddla_d_gp_batched_backward": (_i, [_vp, _i, _i, _vp, _vp, _vp, _ll, _vp, _ll]),
    "ddla_c_vbatched_set_kernel": (_i, [_vp, _i, _i]),
    "ddla_c_profile": (_i, [_vp, _i]),
    "ddla_c_profile_read": (_i, [_vp, _i, _ip, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double)]),
}
for _p in ("s", "d"):
    _SIGNATURES.update({
        f"ddla_{_p}_ppotrf": (_i, [_vp, _c, _i, _vp, _i, _i, _ip, _ip]),
        f"ddla_{_p}_ppotrs": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _ip, _vp, _ip]),
        f"ddla_{_p}_pposv": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _i, _i, _ip, _vp, _i, _i, _ip, _ip]),
        f"ddla_{_p}_ppotrf_workspace_size": (_i, [_vp, _c, _i, _ip, _llp, _llp]),
        f"ddla_{_p}_ppotrf_run": (_i, [_vp, _c, _i, _vp, _i, _i, _ip, _ip, _vp, _ll, _vp, _ll]),
        f"ddla_{_p}_ppotrs_workspace_size": (_i, [_vp, _c, _i, _i, _ip, _ip, _llp, _llp]),
        f"ddla_{_p}_ppotrs_run": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _ip, _vp, _ip, _vp, _ll, _vp, _ll]),
        f"ddla_{_p}_pposv_workspace_size": (_i, [_vp, _c, _c, _i, _i, _ip, _ip, _llp, _llp]),
        f"ddla_{_p}_pposv_run": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _i, _i, _ip, _vp, _i, _i, _ip, _ip, _vp, _ll,
                                      _vp, _ll]),
        f"ddla_{_p}_potrf": (_i, [_vp, _c, _i, _vp, _i, _vp]),
        f"ddla_{_p}_potrf_bottom_right": (_i, [_vp, _c, _i, _vp, _i, _ip]),
        f"ddla_{_p}_ppotrf_bottom_right": (_i, [_vp, _c, _i, _vp, _ip, _ip]),
        f"ddla_{_p}_ptrtrs": (_i, [_vp, _c, _c, _c, _c, _i, _i, _vp, _ip, _vp, _ip]),
        f"ddla_{_p}_trsm": (_i, [_vp, _c, _c, _c, _c, _i, _i, _vp, _vp, _i, _vp, _i]),
        f"ddla_{_p}_potrf_backward": (_i, [_vp, _c, _i, _vp, _i, _vp, _i, _vp, _i]),
        f"ddla_{_p}_potrs_backward": (_i, [_vp, _c, _i, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _vp, _i]),
        f"ddla_{_p}_logdet_backward": (_i, [_vp, _i, _vp, _i, ctypes.c_double, _vp, _i]),
        f"ddla_{_p}_potri": (_i, [_vp, _c, _i, _vp, _i, _vp, _i]),
        f"ddla_{_p}_pposv_vbatched": (_i, [_vp, _c, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _i, _i]),
        f"ddla_{_p}_pposv_vbatched_workspace_size": (_i, [_vp, _i, _i, _llp]),
        f"ddla_{_p}_pposv_vbatched_run": (_i, [_vp, _c, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _ll]),
        f"ddla_{_p}_pposv_vbatched_tune": (_i, [_vp, _c, _vp, _vp, _vp, _i, _vp, _vp, _i, _i, _ip, _ip]),
        f"ddla_{_p}_pposv_vbatched_graph": (_i, [_vp, _c, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _i, _i,
                                                 ctypes.POINTER(_vp)]),
        f"ddla_{_p}_diag_logdet": (_i, [_vp, _i, _vp, _i, _vp]),
        f"ddla_{_p}_pdiag_logdet": (_i, [_vp, _i, _vp, _ip, _vp]),
        f"ddla_{_p}_gemm_vbatched": (_i, [_vp, _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
        f"ddla_{_p}_gemm_vbatched2s": (_i, [_vp, _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                            _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _ip,
                                            _i]),
    })

for _p in ("c", "z"):
    _SIGNATURES.update({
        f"ddla_{_p}_ppotrf": (_i, [_vp, _c, _i, _vp, _i, _i, _ip, _ip]),
        f"ddla_{_p}_ppotrs": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _ip, _vp, _ip]),
        f"ddla_{_p}_pposv": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _i, _i, _ip, _vp, _i, _i, _ip, _ip]),
        f"ddla_{_p}_ppotrf_workspace_size": (_i, [_vp, _c, _i, _ip, _llp, _llp]),
        f"ddla_{_p}_ppotrf_run": (_i, [_vp, _c, _i, _vp, _i, _i, _ip, _ip, _vp, _ll, _vp, _ll]),
        f"ddla_{_p}_ppotrs_workspace_size": (_i, [_vp, _c, _i, _i, _ip, _ip, _llp, _llp]),
        f"ddla_{_p}_ppotrs_run": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _ip, _vp, _ip, _vp, _ll, _vp, _ll]),
        f"ddla_{_p}_pposv_workspace_size": (_i, [_vp, _c, _c, _i, _i, _ip, _ip, _llp, _llp]),
        f"ddla_{_p}_pposv_run": (_i, [_vp, _c, _c, _c, _i, _i, _vp, _i, _i, _ip, _vp, _i, _i, _ip, _ip, _vp, _ll,
                                      _vp, _ll]),
        f"ddla_{_p}_potrf": (_i, [_vp, _c, _i, _vp, _i, _vp]),
        f"ddla_{_p}_diag_logdet": (_i, [_vp, _i, _vp, _i, _vp]),
        f"ddla_{_p}_pdiag_logdet": (_i, [_vp, _i, _vp, _ip, _vp]),
        f"ddla_{_p}_gemm_vbatched": (_i, [_vp, _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
        f"ddla_{_p}_gemm_vbatched2s": (_i, [_vp, _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                            _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _ip,
                                            _i]),
        f"ddla_{_p}_gemm_vbatched_mixed": (_i, [_vp, _i, _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                _vp, _vp, _i]),
        f"ddla_{_p}_gemm_vbatched2s_mixed": (_i, [_vp, _i, _i, _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                  _vp, _vp, _vp, _c, _c, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                  _vp, _vp, _i, _i, _ip, _i]),
    })

for _name, (_res, _args) in _SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def check_status(status: int, where: str = ""):
    """Raise the DdlaError subclass matching a non-zero C status."""
    if status == STATUS_SUCCESS:
        return
    msg = lib.ddla_c_error_string(status)
    text = msg.decode("utf-8", errors="replace") if msg else ""
    ctx = f" after {where}" if where else ""
    cls = {
        STATUS_INVALID_HANDLE: DdlaInvalidHandleError,
        STATUS_INVALID_VALUE: DdlaInvalidValueError,
        STATUS_ALLOC_FAILED: DdlaAllocFailedError,
        STATUS_BACKEND_UNAVAILABLE: DdlaBackendUnavailableError,
        STATUS_INTERNAL_ERROR: DdlaInternalError,
        STATUS_EXCEPTION: DdlaExceptionError,
    }.get(status, DdlaError)
    raise cls(f"LibDDLA status {status}{ctx}: {text}")


def char(c) -> bytes:
    """A LAPACK flag as the single byte the C ABI takes (upper-cased)."""
    if isinstance(c, bytes):
        return c.upper()
    return str(c).upper().encode("ascii")
