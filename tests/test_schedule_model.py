"""CPU model of the blocked factorization's stream schedule.

The library exports the schedule it issues (ddla_c_potrf_schedule: the same
step list potrf_blocked executes, generated without a GPU).  A happens-before
model of HIP streams and events runs over it:
  * kernels on one stream run in issue order;
  * hipEventRecord(e, s) captures everything issued on s so far;
    hipStreamWaitEvent(s2, e) orders every later kernel on s2 after the last
    capture of e issued before the wait (events are re-recorded, so the
    capture that counts is the latest one).
Checked for every panel count from 1 to 40 and for ragged last panels:
  * each panel p is applied exactly once to every column after it;
  * every update that writes panel q's columns happens before panel q is
    factored, and each update happens after the panel it applies;
  * any two launches that write overlapping columns are ordered;
  * with split-product updates (three rotating buffers: panel q's split goes
    to buffer q % 3 as part of its factorization, every update of panel p reads
    buffer p % 3), a buffer is rewritten only after its previous readers;
  * everything is ordered before the main stream's end.
Planted mistakes in the schedule (a wait removed, a look-ahead update moved
to another stream, the last join dropped) must be caught, except for the
waits that are redundant in the schedule as issued (named in the test).
"""
import ctypes

import pytest

from libddla_amd._capi import lib

RESET, FACTOR, UPDATE, RECORD, WAIT = range(5)
MAIN, SIDE, LOOK = range(3)


def schedule(n, nb):
    count = ctypes.c_int()
    assert lib.ddla_c_potrf_schedule(n, nb, None, 0, ctypes.byref(count)) == 0
    buf = (ctypes.c_int * (5 * count.value))()
    assert lib.ddla_c_potrf_schedule(n, nb, buf, count.value, ctypes.byref(count)) == 0
    return [tuple(buf[5 * i:5 * i + 5]) for i in range(count.value)]


def happens_before(steps):
    """Kernel launches [(index, step)] and, per launch, the set of launches
    ordered before it; plus the set ordered before the main stream's end."""
    frontier = {MAIN: frozenset(), SIDE: frozenset(), LOOK: frozenset()}
    event = {}
    launches, before = [], {}
    for i, (op, stream, a, b, c) in enumerate(steps):
        if op == RECORD:
            event[a] = frontier[stream]
        elif op == WAIT:
            frontier[stream] = frontier[stream] | event.get(a, frozenset())
        else:
            before[i] = frontier[stream]
            frontier[stream] = frontier[stream] | {i}
            launches.append(i)
    return launches, before, frontier[MAIN]


def check(steps, n, nb, split=False):
    """Raises AssertionError naming the first violated property."""
    launches, before, end = happens_before(steps)
    np_ = (n + nb - 1) // nb
    col = lambda q: min(n, q * nb)
    kind = {i: steps[i] for i in launches}

    def ordered(x, y):
        return x in before[y] or y in before[x]

    def writes(i):  # [lo, hi) columns written (rows from lo down)
        op, _, a, b, c = kind[i]
        if op == FACTOR:
            return col(a), col(a + 1)
        if op == UPDATE:
            return min(b, n), min(c, n)
        return 0, 0

    resets = [i for i in launches if kind[i][0] == RESET]
    assert len(resets) == 1, "one counter reset"
    factors = {kind[i][2]: i for i in launches if kind[i][0] == FACTOR}
    assert sorted(factors) == list(range(np_)), "every panel factored once"
    # each panel applied exactly once to every later column
    for p in range(np_):
        cover = [0] * n
        for i in launches:
            op, _, a, b, c = kind[i]
            if op == UPDATE and a == p:
                for j in range(min(b, n), min(c, n)):
                    cover[j] += 1
        assert cover[:col(p + 1)] == [0] * col(p + 1), f"panel {p} applied to its own or earlier columns"
        assert cover[col(p + 1):] == [1] * (n - col(p + 1)), f"panel {p} not applied exactly once to later columns"
    for i in launches:
        assert resets[0] == i or resets[0] in before[i], "the reset precedes every launch"
        op, _, a, b, c = kind[i]
        if op == UPDATE and b < min(c, n):
            assert factors[a] in before[i], f"update by panel {a} before the panel is factored"
            for q in range(np_):  # the update writes panel q's columns: it precedes q's factorization
                if b < col(q + 1) and min(c, n) > col(q):
                    assert i in before[factors[q]], f"panel {q} factored before an update of its columns"
    for x in launches:
        for y in launches:
            if x < y:
                lx, hx = writes(x)
                ly, hy = writes(y)
                if max(lx, ly) < min(hx, hy):
                    assert ordered(x, y), f"launches {kind[x]} and {kind[y]} write columns concurrently"
    if split:
        for q in range(3, np_):
            readers = [i for i in launches if kind[i][0] == UPDATE and kind[i][2] == q - 3 and
                       kind[i][3] < min(kind[i][4], n)]
            for r in readers:
                assert r in before[factors[q]], f"split buffer {q % 3} rewritten while panel {q - 3}'s update reads it"
    assert set(launches) <= end | set(), "every launch ordered before the main stream's end"


@pytest.mark.parametrize("panels", list(range(1, 41)))
def test_schedule_orders_every_conflict(panels):
    nb = 64
    n = panels * nb
    steps = schedule(n, nb)
    check(steps, n, nb, split=True)


@pytest.mark.parametrize("n,nb", [(1, 512), (100, 64), (1000, 512), (2000, 768), (32768, 768), (4097, 512)])
def test_schedule_ragged_and_default_widths(n, nb):
    check(schedule(n, nb), n, nb, split=True)


def test_default_width_rule():
    # panel width <= 0 takes the FP64 rule: 768 columns from order 16384, else 512
    assert schedule(16384, 0) == schedule(16384, 768)
    assert schedule(8192, 0) == schedule(8192, 512)


def mutants(steps):
    """Schedules with one planted mistake each."""
    waits = [i for i, s in enumerate(steps) if s[0] == WAIT]
    for i in waits:
        yield f"wait {steps[i]} removed at {i}", steps[:i] + steps[i + 1:]
    # the first two look-ahead updates moved onto the side stream: the first then
    # runs before the side stream has seen the counter reset, the second races
    # with the next step's look-ahead update of the same columns (later moves can
    # leave a valid, only slower, schedule)
    updates = [i for i, s in enumerate(steps) if s[0] == UPDATE and s[1] == LOOK]
    for i in updates[:2]:
        s = list(steps)
        s[i] = (UPDATE, SIDE) + steps[i][2:]
        yield f"look-ahead update {steps[i]} moved to the side stream", s


@pytest.mark.parametrize("panels", [4, 6, 9])
def test_planted_mistakes_are_caught(panels):
    """Every look-ahead update moved onto the side stream, and every removed
    wait but the redundant ones, breaks one of the properties, so the model
    does not pass schedules by accident.  The redundant waits: the main
    stream's wait for the panel event at step 0 (it factored panel 0 itself)
    and at the steps whose main-stream update is empty (p + 3 >= panels), and
    the final wait for the look-ahead event (the side stream's last
    factorization waited for it, and the main stream waits for that)."""
    nb, n = 64, panels * 64
    steps = schedule(n, nb)
    check(steps, n, nb, split=True)
    redundant = set()
    p = -1
    for i, (op, stream, a, b, c) in enumerate(steps[:-3]):  # the last three steps are the final join
        if op == WAIT and stream == LOOK and a == 0:
            p += 1  # each step opens with the look-ahead stream's wait for the panel event
        if op == WAIT and stream == MAIN and a == 0 and p >= 0 and (p == 0 or p + 3 >= panels):
            redundant.add(i)
    final_look_wait = max(i for i, s in enumerate(steps) if s[0] == WAIT and s[1] == MAIN and s[2] == 1)
    redundant.add(final_look_wait)
    caught, passed = 0, set()
    for i, s in enumerate(steps):
        if s[0] == WAIT:
            try:
                check(steps[:i] + steps[i + 1:], n, nb, split=True)
                passed.add(i)
            except AssertionError:
                caught += 1
    assert passed == redundant
    for name, bad in mutants(steps):
        if name.startswith("look-ahead update"):
            with pytest.raises(AssertionError):
                check(bad, n, nb, split=True)
            caught += 1
    assert caught >= 2 * panels - 2
