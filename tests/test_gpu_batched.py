"""GPU tests of the batched fused potrf + potrs + logdet (ddla_{s,d}_pposv_vbatched).

Schedules: the persistent fused kernel for every member, or the blocked path
(64-column steps with gemmVbatched trailing updates) for the members of more
than split_blocks blocks, or the left-looking persistent kernel for every
member; each test pins the schedule it runs.
Every member is checked against the oracle on its own: the factor against
oracle potrf, the solution against oracle potrs, the log-determinant against
2 * the oracle's diag_logdet of its factor; the unused triangle (NaN-poisoned)
and the padding rows must come back bitwise unchanged.  Tolerances: factor
1e-12 / 2e-5 relative to the largest entry (FP64 / FP32), solutions 5e-9 /
2e-3 relative, logdet 1e-12 / 1e-5 relative to n.  The full config-4 batch
(4096 members, orders uniform in [32, 512], seed 3) is checked through
per-member backward errors of a sample, computed in extended precision.
"""
import numpy as np
import pytest

import libddla_amd as dd
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SIZES = [1, 5, 33, 64, 65, 100, 128, 129, 200, 257, 300, 512, 0, 77]
TOL = {np.float64: (1e-12, 5e-9, 1e-12), np.float32: (2e-5, 2e-3, 1e-5)}


def make_batch(handle, sizes, dtype, uplo, nrhs, pad=3, seed0=11):
    vb = dd.VBatch(handle, sizes, nrhs=nrhs, dtype=dtype, pad=pad)
    mats, rhs, raws = [], [], []
    for i, n in enumerate(sizes):
        a = orc.spd_matrix(n, seed0 + i, 0.5) if n else np.zeros((0, 0))
        b = orc.normal_matrix(seed0 + i, 1, n, nrhs) if n else np.zeros((0, nrhs))
        ld = int(vb.ld[i])
        raw = np.full((ld, n), -313.0, dtype=dtype, order="F")
        mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
        raw[:n, :n] = np.where(mask, a, np.nan)
        vb.set_raw(i, raw)
        vb.set_rhs(i, b)
        mats.append(a)
        rhs.append(b)
        raws.append(raw)
    handle.synchronize()
    return vb, mats, rhs, raws


def check_member(vb, i, a, b, raw, dtype, uplo):
    n = a.shape[0]
    ftol, xtol, ltol = TOL[dtype]
    got = vb.get_raw(i)
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    outside = np.ones(raw.shape, bool)
    outside[:n, :n] = mask
    assert got[~outside].tobytes() == raw[~outside].tobytes()  # unused triangle and padding untouched
    if n == 0:
        return
    l, info = orc.potrf("L", a.astype(np.float64))
    assert info == 0
    want = np.tril(l) if uplo == "L" else np.tril(l).T
    err = np.abs(np.where(mask, got[:n, :n], 0) - np.where(mask, want, 0)).max() / np.abs(want).max()
    assert err < ftol, (i, n, err)
    x = vb.get_rhs(i)
    xo = orc.potrs("L", "L", l, b.astype(np.float64))
    assert np.abs(x - xo).max() / max(1.0, np.abs(xo).max()) < xtol, (i, n)
    mag, _ = orc.diag_logdet(l)
    ld = vb.get_logdet()[i]
    assert abs(ld - 2 * mag) <= ltol * max(1.0, n), (i, n, ld, 2 * mag)


# (workgroups per CU, largest first, split_blocks[, left-looking]): fused only; blocked above 1 and 3
# blocks; the left-looking kernel for every member
SCHEDULES = [(1, True, 16), (2, False, 16), (1, True, 1), (2, True, 3), (1, True, 16, True), (2, False, 16, True),
             (1, True, 1, False, 128), (2, True, 2, False, 128)]


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("nrhs", [1, 17])
def test_batched_members_match_oracle(handle, dtype, uplo, nrhs, schedule):
    print(f"case batched {np.dtype(dtype).name} uplo={uplo} nrhs={nrhs} schedule={schedule}", flush=True)
    dd.set_vbatched_schedule(handle, dtype, *schedule)
    vb, mats, rhs, raws = make_batch(handle, SIZES, dtype, uplo, nrhs)
    dd.pposv_vbatched(vb, uplo)
    info = vb.get_info()
    assert (info == 0).all(), info
    for i in range(len(SIZES)):
        check_member(vb, i, mats[i], rhs[i], raws[i], dtype, uplo)


@pytest.mark.parametrize("schedule", [(1, True, 16), (1, True, 1), (1, True, 16, True), (2, True, 16, True),
                                      (1, True, 1, False, 128)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_failures_are_per_member(handle, dtype, schedule):
    """A non-PD member reports its pivot (B untouched, logdet NaN); an order
    above max_n reports -1 and is not touched; the others are solved.  With
    split 1 the members above 64 take the blocked path (the failing pivot is in
    the third block column); the last two schedules run the left-looking
    kernel."""
    print(f"case batched failures {np.dtype(dtype).name} schedule={schedule}", flush=True)
    dd.set_vbatched_schedule(handle, dtype, *schedule)
    vb, mats, rhs, raws = make_batch(handle, [130, 70, 5, 40], dtype, "L", 2)
    # member 0: final pivot negative
    raw0 = raws[0].copy()
    raw0[129, 129] = -1.0
    vb.set_raw(0, raw0)
    handle.synchronize()
    dd.pposv_vbatched(vb, "L", max_n=512)
    info = vb.get_info()
    assert info[0] == 130 and info[1] == 0 and info[2] == 0 and info[3] == 0
    assert vb.get_rhs(0).tobytes() == rhs[0].astype(dtype).tobytes()
    assert np.isnan(vb.get_logdet()[0])
    check_member(vb, 1, mats[1], rhs[1], raws[1], dtype, "L")
    check_member(vb, 3, mats[3], rhs[3], raws[3], dtype, "L")
    vb2, m2, r2, raw2 = make_batch(handle, [20, 90, 30], dtype, "L", 1)
    dd.pposv_vbatched(vb2, "L", max_n=64)  # member 1 (order 90) is above max_n
    info2 = vb2.get_info()
    assert list(info2) == [0, -1, 0]
    assert vb2.get_raw(1).tobytes() == raw2[1].tobytes()
    assert vb2.get_rhs(1).tobytes() == r2[1].astype(dtype).tobytes()
    check_member(vb2, 0, m2[0], r2[0], raw2[0], dtype, "L")
    check_member(vb2, 2, m2[2], r2[2], raw2[2], dtype, "L")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_run_tune_and_graph_agree_bitwise(handle, dtype):
    """Tuning leaves the inputs unchanged and records its schedule; for that
    schedule the driver, the workspace run and each replay of a captured graph
    give the same results bit for bit (no member's arithmetic depends on timing).
    The graph is replayed three times with eager work and re-uploaded inputs in
    between."""
    sizes = [300, 17, 512, 64, 129, 1, 250, 400]
    vb, mats, rhs, raws = make_batch(handle, sizes, dtype, "L", 2)

    def reload():
        for i, n in enumerate(sizes):
            vb.set_raw(i, raws[i])
            vb.set_rhs(i, rhs[i])
        handle.synchronize()

    occ, ordered = dd.pposv_vbatched_tune(vb, "L")
    assert occ in (1, 2)
    for i in range(len(sizes)):  # inputs unchanged by tuning
        assert vb.get_raw(i).tobytes() == raws[i].tobytes()
    sched = dd.vbatched_schedule(handle, dtype)
    assert sched["tuned"] and sched["workgroups_per_cu"] == occ and sched["largest_first"] == ordered
    assert 1 <= sched["split_blocks"] <= 16
    dd.pposv_vbatched(vb, "L")
    assert (vb.get_info() == 0).all()
    for i in range(len(sizes)):
        check_member(vb, i, mats[i], rhs[i], raws[i], dtype, "L")
    ref_a = [vb.get_raw(i).copy() for i in range(len(sizes))]
    ref_x = [vb.get_rhs(i).copy() for i in range(len(sizes))]
    ref_l = vb.get_logdet().copy()

    def same():
        for i in range(len(sizes)):
            assert vb.get_raw(i).tobytes() == ref_a[i].tobytes()
            assert vb.get_rhs(i).tobytes() == ref_x[i].tobytes()
        assert vb.get_logdet().tobytes() == ref_l.tobytes()

    reload()
    nbytes = dd.pposv_vbatched_workspace_size(vb)
    work = dd.batched._Device(handle, nbytes)
    dd.pposv_vbatched_run(vb, work.ptr.value, nbytes, "L")
    same()
    with pytest.raises(dd.DdlaExceptionError):
        dd.pposv_vbatched_run(vb, work.ptr.value, nbytes - 1, "L")
    g = dd.BatchedGraph(vb, "L")
    for rep in range(3):
        reload()
        dd.pposv_vbatched(vb, "L")  # eager work between replays
        handle.synchronize()
        reload()
        g.launch()
        same()
    g.destroy()


def test_batched_invalid_arguments(handle):
    for bad in ((3, True, 4), (1, True, 0), (1, True, 17)):
        with pytest.raises(dd.DdlaError):
            dd.set_vbatched_schedule(handle, np.float64, *bad)
    with pytest.raises(dd.DdlaError):
        dd.check_status(dd.lib.ddla_c_vbatched_set_kernel(handle.ptr, 8, 3), "ddla_c_vbatched_set_kernel")
    with pytest.raises(ValueError):
        dd.set_vbatched_schedule(handle, np.float64, 1, True, 4, False, 96)
    dd.set_vbatched_schedule(handle, np.float64, 2, False, 5, True)
    sched = dd.vbatched_schedule(handle, np.float64)
    assert sched == {"tuned": True, "workgroups_per_cu": 2, "largest_first": False, "split_blocks": 5,
                     "left_looking": True, "step_columns": 64}
    dd.set_vbatched_schedule(handle, np.float64, 1, True, 3, False, 128)
    sched = dd.vbatched_schedule(handle, np.float64)
    assert sched["step_columns"] == 128 and not sched["left_looking"] and sched["split_blocks"] == 3
    dd.set_vbatched_schedule(handle, np.float64, 1, True, 16)  # back to the right-looking kernels
    assert not dd.vbatched_schedule(handle, np.float64)["left_looking"]
    vb = dd.VBatch(handle, [4, 4], nrhs=1)
    with pytest.raises(dd.DdlaExceptionError):
        dd.pposv_vbatched(vb, "X")
    with pytest.raises(dd.DdlaExceptionError):
        dd.pposv_vbatched(vb, "L", max_n=1025)
    empty = dd.VBatch(handle, [], nrhs=1)
    dd.pposv_vbatched(empty, "L")  # no-op


@pytest.mark.parametrize("kernel", ["blocked", "left_looking", "blocked_128"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_full_batch_4096_members(handle, dtype, kernel):
    """4096 members, orders uniform in [32, 512] (seed 3), A_i = G G^T / n + 0.5 I,
    one right-hand side each: every member succeeds; a sample of 64 members has
    backward error ||b - A x||_inf / (||A||_inf ||x||_inf) <= 64 n eps and a
    log-determinant within 1e-9 n (FP64) / 1e-3 n (FP32) of numpy's slogdet."""
    print(f"case batched full {np.dtype(dtype).name} kernel={kernel}", flush=True)
    # blocked path above order 256 (64- or 128-column steps), or the left-looking kernel for every member
    dd.set_vbatched_schedule(handle, dtype, 1, True, 4, kernel == "left_looking",
                             128 if kernel == "blocked_128" else 64)
    rng = np.random.RandomState(3)
    sizes = rng.randint(32, 513, size=4096)
    vb = dd.VBatch(handle, sizes, nrhs=1, dtype=dtype)
    vb.generate(3, 0.5)
    sample = rng.choice(4096, 64, replace=False)
    a_in = {int(i): vb.get_matrix(int(i)).astype(np.float64) for i in sample}
    b_in = {int(i): vb.get_rhs(int(i)).astype(np.float64) for i in sample}
    dd.pposv_vbatched(vb, "L")
    info = vb.get_info()
    assert (info == 0).all()
    logdet = vb.get_logdet()
    eps = float(np.finfo(dtype).eps)
    for i in sample:
        i = int(i)
        a = np.tril(a_in[i]) + np.tril(a_in[i], -1).T
        x = vb.get_rhs(i).astype(np.longdouble)
        r = b_in[i].astype(np.longdouble) - a.astype(np.longdouble) @ x
        bwd = float(np.abs(r).max() / (np.abs(a).sum(1).max() * np.abs(x).max()))
        assert bwd <= 64 * sizes[i] * eps, (i, bwd)
        sign, ref = np.linalg.slogdet(a)
        assert sign > 0
        assert abs(logdet[i] - ref) <= (1e-9 if dtype == np.float64 else 1e-3) * sizes[i], (i, logdet[i], ref)


@pytest.mark.parametrize("occupancy", [1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_left_looking_large_orders(handle, dtype, uplo, occupancy):
    """The left-looking kernel up to the largest order (1024 = 16 block
    columns), orders off the 64 grid and a short last block, with 17
    right-hand sides (passes of 4 columns and a last one of 1) and with one:
    every member against the oracle, the unused triangle and padding bitwise."""
    print(f"case batched left-looking large {np.dtype(dtype).name} uplo={uplo} occ={occupancy}", flush=True)
    dd.set_vbatched_schedule(handle, dtype, occupancy, True, 16, True)
    sizes = [1024, 1000, 513, 700, 960, 2, 191]
    for nrhs in (17, 1):
        vb, mats, rhs, raws = make_batch(handle, sizes, dtype, uplo, nrhs, seed0=40 + nrhs)
        dd.pposv_vbatched(vb, uplo, max_n=1024)
        assert (vb.get_info() == 0).all(), vb.get_info()
        for i in range(len(sizes)):
            check_member(vb, i, mats[i], rhs[i], raws[i], dtype, uplo)
