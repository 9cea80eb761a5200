"""CPU checks of the C ABI: the library loads, exports what include/ddla_mi355x.h
declares, and its host-side argument handling behaves like the reference's
generated shim (fortran/ddla_c_api.cpp:43-240).  No kernel is launched."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import libddla_amd as dd
from libddla_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_names():
    text = open(os.path.join(ROOT, "include", "ddla_mi355x.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ddla_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_names()
    assert len(names) > 40
    out = subprocess.run(["nm", "-D", "--defined-only", dd.LIBRARY_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert [n for n in names if n not in exported] == []


def test_reference_entry_points_present():
    # The names a binding written for the reference's C ABI looks up.
    for name in ("ddla_c_init_backend", "ddla_c_set", "ddla_c_destroy", "ddla_c_descinit", "ddla_c_malloc",
                 "ddla_c_memcpy", "ddla_c_synchronize", "ddla_c_error_string", "ddla_s_ppotrf", "ddla_d_ppotrf",
                 "ddla_s_ppotrs", "ddla_d_ppotrs", "ddla_s_pposv", "ddla_d_pposv"):
        assert hasattr(lib, name)


def test_status_strings():
    assert lib.ddla_c_error_string(0) == b"DDLA_STATUS_SUCCESS"
    assert lib.ddla_c_error_string(4) == b"DDLA_STATUS_BACKEND_UNAVAILABLE"
    assert lib.ddla_c_error_string(99) == b"unknown LibDDLA C API status"


def test_backends():
    assert lib.ddla_c_backend_available(0) == 0
    assert lib.ddla_c_backend_available(1) == 1
    h = ctypes.c_void_p()
    assert lib.ddla_c_init_backend(ctypes.byref(h), 0) == 4  # CPU backend is not built
    assert lib.ddla_c_init_backend(ctypes.byref(h), 7) == 2
    assert lib.ddla_c_init_backend(ctypes.byref(h), 1) == 0  # host object only, no device call yet
    assert lib.ddla_c_get_backend(h) == 1
    assert lib.ddla_c_destroy(ctypes.byref(h)) == 0
    assert h.value is None
    assert lib.ddla_c_destroy(ctypes.byref(h)) == 0  # idempotent


def test_null_handle_is_status_1():
    desc = np.zeros(9, dtype=np.int32)
    info = ctypes.c_int(-1)
    dp = desc.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    assert lib.ddla_d_ppotrf(None, b"L", 4, None, 1, 1, dp, ctypes.byref(info)) == 1
    assert lib.ddla_s_pposv(None, b"L", b"L", b"N", 4, 1, None, 1, 1, dp, None, 1, 1, dp, ctypes.byref(info)) == 1
    assert lib.ddla_c_synchronize(None) == 1
    assert lib.ddla_c_descinit(dp, None, 4, 4, 2, 2, 0, 0, -1) == 1


def test_ungridded_handle_raises_through_status_6():
    h = ctypes.c_void_p()
    assert lib.ddla_c_init_backend(ctypes.byref(h), 1) == 0
    desc = np.array([1, 0, 4, 4, 2, 2, 0, 0, 4], dtype=np.int32)
    info = ctypes.c_int(-1)
    st = lib.ddla_d_ppotrf(h, b"L", 4, ctypes.c_void_p(16), 1, 1, desc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           ctypes.byref(info))
    assert st == 6
    assert b"process grid" in lib.ddla_c_error_string(6)
    with pytest.raises(dd.DdlaExceptionError):
        dd.check_status(st, "ppotrf")
    lib.ddla_c_destroy(ctypes.byref(h))


def numroc_bruteforce(n, nb, iproc, src, nprocs):
    return sum(1 for g in range(n) if (src + g // nb) % nprocs == iproc)


def test_numloc_matches_block_cyclic_count():
    for n in (0, 1, 7, 64, 100, 257):
        for nb in (1, 3, 16, 64):
            for nprocs in (1, 2, 3, 4):
                for src in range(nprocs):
                    for iproc in range(nprocs):
                        assert lib.ddla_c_numloc(n, nb, iproc, src, nprocs) == numroc_bruteforce(
                            n, nb, iproc, src, nprocs)


def test_wait_limit_setter_returns_previous_and_restores_default():
    """Process-wide spin limit of the persistent kernels' waits: no GPU call."""
    default = 1 << 23
    assert lib.ddla_c_set_wait_limit(5) == default
    assert lib.ddla_c_set_wait_limit(7) == 5
    assert lib.ddla_c_set_wait_limit(0) == 7  # <= 0 restores the default
    assert lib.ddla_c_set_wait_limit(-3) == default
    assert lib.ddla_c_set_wait_limit(0) == default


def test_panel_width_setter_rounds_clamps_and_restores():
    """Process-wide panel width override of the blocked factorization: no GPU
    call.  Returns the previous override (0 = the built-in rule)."""
    assert lib.ddla_c_set_panel_width(768) == 0
    assert lib.ddla_c_set_panel_width(700) == 768  # rounded down to 640
    assert lib.ddla_c_set_panel_width(5000) == 640  # clamped to 768
    assert lib.ddla_c_set_panel_width(10) == 768  # at least 64
    assert lib.ddla_c_set_panel_width(0) == 64
    assert lib.ddla_c_set_panel_width(-1) == 0


def test_single_process_communicator_entries():
    """The communicator helpers of a one-process build need no GPU."""
    comm = lib.ddla_c_default_comm()
    assert lib.ddla_c_comm_rank(comm) == 0 and lib.ddla_c_comm_size(comm) == 1
    assert lib.ddla_c_communication_available(None, 2) == 0
    assert lib.ddla_c_set_communication(None, 2) == 1
