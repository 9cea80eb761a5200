"""Handle lifecycle (mirrors the reference's python/ddla/backend.py).

One process drives one GPU through a 1 x 1 process grid.
"""
import ctypes
import enum

from ._capi import BACKEND_CPU, BACKEND_GPU, check_status, lib


class Backend(enum.IntEnum):
    CPU = BACKEND_CPU
    GPU = BACKEND_GPU


class Handle:
    """Owns a DdlaHandle_t: created with ddla_c_init_backend, gridded with ddla_c_set."""

    def __init__(self, backend: Backend = Backend.GPU, nprows: int = 1, npcols: int = 1, major: str = "R"):
        self._handle = ctypes.c_void_p()
        if backend is None:
            raise TypeError("backend must explicitly select CPU or GPU")
        check_status(lib.ddla_c_init_backend(ctypes.byref(self._handle), int(backend)), "ddla_c_init_backend")
        st = lib.ddla_c_set(self._handle, 0, nprows, npcols, major.encode("ascii"))
        if st != 0:
            lib.ddla_c_destroy(ctypes.byref(self._handle))
            check_status(st, "ddla_c_set")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def destroy(self):
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value is not None:
            lib.ddla_c_destroy(ctypes.byref(handle))
            self._handle = ctypes.c_void_p()

    def __del__(self):
        self.destroy()

    @property
    def ptr(self) -> ctypes.c_void_p:
        return self._handle

    @property
    def backend(self) -> Backend:
        return Backend(lib.ddla_c_get_backend(self._handle))

    @staticmethod
    def is_backend_available(backend: Backend) -> bool:
        return bool(lib.ddla_c_backend_available(int(backend)))

    @property
    def rank(self) -> int:
        return lib.ddla_c_rank(self._handle)

    @property
    def size(self) -> int:
        return lib.ddla_c_size(self._handle)

    @property
    def grid(self):
        r, c = ctypes.c_int(), ctypes.c_int()
        check_status(lib.ddla_c_grid_dims(self._handle, ctypes.byref(r), ctypes.byref(c)), "ddla_c_grid_dims")
        return r.value, c.value

    def synchronize(self):
        check_status(lib.ddla_c_synchronize(self._handle), "ddla_c_synchronize")

    @property
    def stream(self) -> int:
        return lib.ddla_c_get_stream(self._handle) or 0

    def set_stream(self, stream: int):
        """Replace the compute stream (0 = the default stream, shared with torch's default stream)."""
        check_status(lib.ddla_c_set_stream(self._handle, ctypes.c_void_p(stream)), "ddla_c_set_stream")

    # -- per-launch event profiler -------------------------------------------
    KERNEL_CLASSES = ("trailing_update", "panel", "solve", "trtri", "residual", "other", "lookahead_update")

    def profile(self, enable: bool):
        check_status(lib.ddla_c_profile(self._handle, 1 if enable else 0), "ddla_c_profile")

    def profile_read(self):
        """{class: (launches, ms, flops, bytes)} of the launches recorded since profile(True)."""
        out = {}
        for cls, name in enumerate(self.KERNEL_CLASSES):
            cnt = ctypes.c_int()
            ms, fl, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            check_status(lib.ddla_c_profile_read(self._handle, cls, ctypes.byref(cnt), ctypes.byref(ms),
                                                 ctypes.byref(fl), ctypes.byref(by)), "ddla_c_profile_read")
            out[name] = (cnt.value, ms.value, fl.value, by.value)
        return out
