"""GPU tests of the runtime entries beside the drivers
(fortran/ddla_c_api.cpp:134-163, :249-258, :304-349 of the reference):

  * ddla_c_memcpy2d: rectangles between pitched buffers in every direction
    (host <-> device, device -> device, host -> host) against numpy slicing,
    with the bytes outside the rectangle left as they were; the empty
    rectangle succeeds without touching null pointers; a pitch below the width,
    a null pointer with a nonempty rectangle and a null handle are rejected
    (include/ddla/ddla_handle_t.h:180-197);
  * communication selection: every valid mode is accepted and read back, an
    invalid one is status 2 and keeps the previous mode, the default is CCL,
    the query is false for invalid modes and null handles;
  * the single-process communicator helpers: rank 0 of a world of 1, and a
    dimension broadcast that leaves its arguments as they are;
  * the wait-limit error path of the persistent kernels, forced from the host
    with a spin limit of one poll (no kernel faults or hangs: every wait is
    bounded and every workgroup exits);
  * the three-phase ppotrs (workspace query, run on caller scratch) against
    the driver, bit for bit; the stream hand-over of ddla_c_set_stream; the
    device normal generator against the oracle's convention.
"""
import ctypes

import numpy as np
import pytest

import libddla_amd as dd

pytestmark = pytest.mark.gpu


def device_buffer(h, nbytes):
    p = ctypes.c_void_p()
    dd.check_status(dd.lib.ddla_c_malloc(h.ptr, ctypes.byref(p), nbytes), "ddla_c_malloc")
    return p


def test_memcpy2d_every_direction(handle):
    h = handle
    rng = np.random.RandomState(0)
    rows, width, spitch, dpitch = 37, 24, 40, 56  # bytes per row copied, source and destination pitches
    src = rng.randint(0, 256, size=(rows, spitch), dtype=np.uint8)
    # host -> device: a device buffer prefilled with a pattern
    dev = device_buffer(h, rows * dpitch)
    fill = np.full((rows, dpitch), 0xA5, dtype=np.uint8)
    dd.check_status(dd.lib.ddla_c_memcpy(h.ptr, dev, fill.ctypes.data, fill.nbytes))
    assert dd.lib.ddla_c_memcpy2d(h.ptr, dev, dpitch, src.ctypes.data, spitch, width, rows) == 0
    back = np.zeros((rows, dpitch), dtype=np.uint8)
    dd.check_status(dd.lib.ddla_c_memcpy(h.ptr, back.ctypes.data, dev, back.nbytes))
    h.synchronize()
    want = fill.copy()
    want[:, :width] = src[:, :width]
    assert np.array_equal(back, want)
    # device -> device, into a second device buffer with another pitch
    dev2 = device_buffer(h, rows * 32)
    zero = np.zeros((rows, 32), dtype=np.uint8)
    dd.check_status(dd.lib.ddla_c_memcpy(h.ptr, dev2, zero.ctypes.data, zero.nbytes))
    assert dd.lib.ddla_c_memcpy2d(h.ptr, dev2, 32, dev, dpitch, width, rows) == 0
    # device -> host
    out = np.full((rows, 30), 7, dtype=np.uint8)
    assert dd.lib.ddla_c_memcpy2d(h.ptr, out.ctypes.data, 30, dev2, 32, width, rows) == 0
    h.synchronize()
    want2 = np.full((rows, 30), 7, dtype=np.uint8)
    want2[:, :width] = src[:, :width]
    assert np.array_equal(out, want2)
    # host -> host finishes before returning
    hh = np.zeros((rows, width + 8), dtype=np.uint8)
    assert dd.lib.ddla_c_memcpy2d(h.ptr, hh.ctypes.data, width + 8, src.ctypes.data, spitch, width, rows) == 0
    assert np.array_equal(hh[:, :width], src[:, :width]) and not hh[:, width:].any()
    # empty rectangles, bad arguments
    assert dd.lib.ddla_c_memcpy2d(h.ptr, None, 0, None, 0, 0, 5) == 0
    assert dd.lib.ddla_c_memcpy2d(h.ptr, None, 0, None, 0, 5, 0) == 0
    assert dd.lib.ddla_c_memcpy2d(h.ptr, dev, 8, src.ctypes.data, spitch, 16, 2) == 2
    assert dd.lib.ddla_c_memcpy2d(h.ptr, None, dpitch, src.ctypes.data, spitch, width, 2) == 2
    assert dd.lib.ddla_c_memcpy2d(None, dev, dpitch, src.ctypes.data, spitch, width, 2) == 1
    for p in (dev, dev2):
        dd.check_status(dd.lib.ddla_c_free(h.ptr, p))


def test_communication_selection():
    h = dd.Handle(dd.Backend.GPU)
    try:
        mode = ctypes.c_int(-1)
        assert dd.lib.ddla_c_get_communication(h.ptr, ctypes.byref(mode)) == 0 and mode.value == 2
        for m in (0, 1, 2):
            assert dd.lib.ddla_c_communication_available(h.ptr, m) == 1
            assert dd.lib.ddla_c_set_communication(h.ptr, m) == 0
            assert dd.lib.ddla_c_get_communication(h.ptr, ctypes.byref(mode)) == 0 and mode.value == m
        assert dd.lib.ddla_c_set_communication(h.ptr, 7) == 2
        assert dd.lib.ddla_c_get_communication(h.ptr, ctypes.byref(mode)) == 0 and mode.value == 2
        assert dd.lib.ddla_c_communication_available(h.ptr, 3) == 0
        assert dd.lib.ddla_c_communication_available(None, 2) == 0
        assert dd.lib.ddla_c_set_communication(None, 2) == 1
        assert dd.lib.ddla_c_get_communication(None, ctypes.byref(mode)) == 1
        assert dd.lib.ddla_c_get_communication(h.ptr, None) == 2
    finally:
        h.destroy()


def test_single_process_communicator_helpers():
    comm = dd.lib.ddla_c_default_comm()
    assert dd.lib.ddla_c_comm_rank(comm) == 0 and dd.lib.ddla_c_comm_size(comm) == 1
    m, n, t = ctypes.c_int(5), ctypes.c_int(6), ctypes.c_int(ord("d"))
    assert dd.lib.ddla_c_bcast_dims(comm, ctypes.byref(m), ctypes.byref(n), ctypes.byref(t), 0) == 0
    assert (m.value, n.value, t.value) == (5, 6, ord("d"))


def test_wait_limit_error_is_reported_and_handle_recovers(handle):
    """The bounded waits of the persistent kernels, forced from the host: with
    a spin limit of 1 poll the solve sweeps of ppotrs (order 4096, 64 blocks)
    and the panel factorization of ppotrf give up on a predecessor that is
    still computing, every workgroup exits, and the call fails with the wait
    error.  With the default restored the same handle solves correctly."""
    import pytest as _pytest
    import torch

    from oracle import oracle as orc

    n, nrhs = 4096, 8
    a = orc.spd_matrix(n, seed=9, delta=0.5)
    b = orc.normal_matrix(9, 1, n, nrhs)
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
    _, info = dd.ppotrf(A, "L")
    assert info == 0
    B = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)
    previous = dd.lib.ddla_c_set_wait_limit(1)
    try:
        assert previous == 1 << 23
        print("case forced wait limit: ppotrs", flush=True)
        with _pytest.raises(RuntimeError, match="wait limit"):
            dd.ppotrs(A, B, "L")
        print("case forced wait limit: ppotrf", flush=True)
        A2 = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
        with _pytest.raises(RuntimeError, match="wait exceeded its limit"):
            dd.ppotrf(A2, "L")
        # the local potrf is asynchronous: the abandoned factorization shows
        # in its device info word (DDLA_POTRF_INFO_WAIT_LIMIT)
        print("case forced wait limit: local potrf", flush=True)
        t = torch.tensor(np.ascontiguousarray(a.T), device="cuda")
        d_info = torch.full((1,), -7, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        assert dd.lib.ddla_d_potrf(handle.ptr, b"L", n, ctypes.c_void_p(t.data_ptr()), n,
                                   ctypes.c_void_p(d_info.data_ptr())) == 0
        handle.synchronize()
        assert int(d_info.item()) == -1000
    finally:
        assert dd.lib.ddla_c_set_wait_limit(0) == 1
    assert dd.lib.ddla_c_set_wait_limit(0) == 1 << 23
    A3 = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
    B3 = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)
    X, info = dd.pposv(A3, B3, "L")
    assert info == 0
    x = X.to_global_matrix()
    assert orc.backward_error(a, x, b) < 1e-14


def test_ppotrs_three_phase_matches_driver(handle):
    """ppotrs_workspace_size + ppotrs_run on caller scratch give the driver's
    result bit for bit; a workspace below the query is rejected."""
    import torch

    from oracle import oracle as orc

    n, nrhs = 700, 9
    a = orc.spd_matrix(n, seed=12, delta=0.3)
    b = orc.normal_matrix(12, 1, n, nrhs)
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=64, nb=64)
    _, info = dd.ppotrf(A, "L")
    assert info == 0
    B1 = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)
    B2 = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)
    dd.ppotrs(A, B1, "L")
    bb, hb = ctypes.c_longlong(), ctypes.c_longlong()
    assert dd.lib.ddla_d_ppotrs_workspace_size(handle.ptr, b"L", n, nrhs, A.desc_ptr, B2.desc_ptr, ctypes.byref(bb),
                                               ctypes.byref(hb)) == 0
    assert bb.value > 0 and hb.value >= 0
    work = torch.empty(bb.value, dtype=torch.uint8, device="cuda")
    host = (ctypes.c_char * max(1, hb.value))()
    torch.cuda.synchronize()
    st = dd.lib.ddla_d_ppotrs_run(handle.ptr, b"L", b"L", b"N", n, nrhs, ctypes.c_void_p(A.data_ptr), A.desc_ptr,
                                  ctypes.c_void_p(B2.data_ptr), B2.desc_ptr, ctypes.c_void_p(work.data_ptr()),
                                  bb.value, ctypes.cast(host, ctypes.c_void_p), hb.value)
    assert st == 0
    handle.synchronize()
    assert B2.to_global_matrix().tobytes() == B1.to_global_matrix().tobytes()
    B3 = dd.DistributedMatrix.from_global_matrix(handle, b, mb=64, nb=64)  # the Python wrapper of the same run
    dd.ppotrs_run(A, B3, "L", work.data_ptr(), bb.value)
    handle.synchronize()
    assert B3.to_global_matrix().tobytes() == B1.to_global_matrix().tobytes()
    st = dd.lib.ddla_d_ppotrs_run(handle.ptr, b"L", b"L", b"N", n, nrhs, ctypes.c_void_p(A.data_ptr), A.desc_ptr,
                                  ctypes.c_void_p(B2.data_ptr), B2.desc_ptr, ctypes.c_void_p(work.data_ptr()),
                                  bb.value - 256, ctypes.cast(host, ctypes.c_void_p), hb.value)
    assert st == 6


def test_set_stream_takes_ownership_and_null_restores_default():
    """ddla_c_set_stream (include/ddla/ddla_handle_t.h:286-300 of the reference):
    the handle takes over a stream created by the caller, runs on it, and
    destroys it when replaced by the null (default) stream."""
    from oracle import oracle as orc

    hip = ctypes.CDLL("libamdhip64.so")
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    h = dd.Handle(dd.Backend.GPU)
    try:
        h.set_stream(stream.value)
        assert h.stream == stream.value
        n, nrhs = 300, 4
        a = orc.spd_matrix(n, seed=13, delta=0.3)
        b = orc.normal_matrix(13, 1, n, nrhs)
        A = dd.DistributedMatrix.from_global_matrix(h, a, mb=64, nb=64)
        B = dd.DistributedMatrix.from_global_matrix(h, b, mb=64, nb=64)
        _, info = dd.pposv(A, B, "L")
        assert info == 0
        assert orc.backward_error(a, B.to_global_matrix(), b) < 1e-14
        A.free()
        B.free()
        h.set_stream(0)  # the handle destroys the stream it owned
        assert h.stream == 0
    finally:
        h.destroy()


def test_device_normals_match_the_oracle(handle):
    """ddla_c_fill_normal (the benchmark's generator) against the oracle's
    restatement of the same Philox4x64-10 + Box-Muller convention, element by
    element (the device and host libm differ by at most a few ulps)."""
    import torch

    from oracle import oracle as orc

    for seed, stream, first, count in ((1, 0, 0, 1001), (2, 1, 12345, 777), (7, 9, (1 << 33) + 3, 64)):
        out = torch.empty(count, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert dd.lib.ddla_c_fill_normal(handle.ptr, seed, stream, first, count, ctypes.c_void_p(out.data_ptr())) == 0
        handle.synchronize()
        want = orc.fill_normal(seed, stream, first, count)
        got = out.cpu().numpy()
        assert np.abs(got - want).max() <= 1e-14 * max(1.0, np.abs(want).max())


def test_launch_profiler_records_the_schedule(handle):
    """The per-launch event profiler behind the benchmark's roofline figures:
    one FP64 pposv at n = 2048 (four 512-column panels, uplo L) records the
    launches the schedule makes, per class: 4 panels, 1 trailing update on the
    handle's stream (columns beyond panel p + 2 exist only for p = 0), 5
    look-ahead updates, 1 solve (both sweeps) and 1 diagonal-block inversion;
    every duration is positive, the trailing update's flops are its
    m (m + 1) kb with m = 512, kb = 512."""
    from oracle import oracle as orc

    n, nrhs = 2048, 4
    a = orc.spd_matrix(n, seed=14, delta=0.3)
    b = orc.normal_matrix(14, 1, n, nrhs)
    A = dd.DistributedMatrix.from_global_matrix(handle, a, mb=512, nb=512)
    B = dd.DistributedMatrix.from_global_matrix(handle, b, mb=512, nb=512)
    handle.profile(True)
    try:
        _, info = dd.pposv(A, B, "L")
        prof = handle.profile_read()
    finally:
        handle.profile(False)
    assert info == 0
    counts = {k: v[0] for k, v in prof.items()}
    assert counts["panel"] == 4 and counts["trailing_update"] == 1 and counts["lookahead_update"] == 5
    assert counts["solve"] == 1 and counts["trtri"] == 1
    for k, (cnt, ms, fl, _) in prof.items():
        if cnt:
            assert ms > 0, k
    assert prof["trailing_update"][2] == 512 * 513 * 512
