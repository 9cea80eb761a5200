"""CPU model of the in-kernel publication protocol of the persistent panel
and solve kernels (panel_tile.h, solve_tile.h, stage_tile.h).

Memory, as the protocol uses gfx950's: each XCD has its own L2; a plain
store enters the writer's in-flight queue and later lands in its XCD's L2 as
a dirty line; `s_waitcnt vmcnt(0)` waits until the wave's in-flight stores
have landed; `buffer_wbl2` (agent-scope release) takes the XCD's dirty lines
at issue and writes them back to the coherence point (memory) some time
later, and a following `s_waitcnt vmcnt(0)` waits for that write-back; an
agent-scope atomic store or load is performed at memory (never served from
an L1 or L2 copy); a plain load may be served from the reader's L2.

Writer (one workgroup publishing block t): stores of the block's words,
s_waitcnt vmcnt(0) in every wavefront, barrier, one thread's buffer_wbl2,
s_waitcnt vmcnt(0), atomic store of the counter.  Reader (another XCD):
relaxed atomic polls of the counter until it shows t, then agent-scope atomic
loads of the block.  Every interleaving of the two programs and of the
asynchronous landings and write-backs is explored; the reader must always see
the published words.  Mutants that must fail: the wait after the write-back
dropped (the counter overtook its block on the GPU this way once), the wait
before the barrier dropped, no write-back, and plain loads in the reader (a
stale L2 line from the buffer's previous use).

Tickets: a workgroup takes its ticket from a counter when it starts and
waits only on the workgroup holding the previous ticket, which has therefore
started.  With R resident slots and workgroups dispatched in any order,
every schedule of the model finishes; a mutant that waits on the previous
workgroup index instead (dispatch order is not index order) deadlocks.
"""
import itertools

WORDS = 2       # words in the published block
OLD, NEW = 0, 1  # the block's previous contents (in memory and a stale reader L2 line) and the published ones


def writer_program(mutant):
    prog = [("store", w) for w in range(WORDS)]
    if mutant != "no_wait_before_barrier":
        prog.append(("wait",))
    if mutant != "no_writeback":
        prog.append(("wbl2",))
        if mutant != "no_wait_after_writeback":
            prog.append(("wait",))
    prog.append(("publish",))
    return prog


def explore(mutant=None):
    """All reachable outcomes of the reader's block reads (a set of tuples)."""
    wprog = writer_program(mutant)
    plain_reader = mutant == "plain_loads"
    # state: (wpc, inflight stores (frozenset of words), l2 dirty (frozenset), pending write-backs
    #         (tuple of frozensets), memory (tuple), counter in memory, counter store in flight,
    #         rpc, reader values (tuple))
    start = (0, frozenset(), frozenset(), (), (OLD,) * WORDS, 0, False, 0, ())
    seen, stack, outcomes = {start}, [start], set()
    while stack:
        st = stack.pop()
        wpc, infl, dirty, wbs, mem, ctr, ctr_fly, rpc, vals = st
        nxt = []
        # writer steps
        if wpc < len(wprog):
            op = wprog[wpc]
            if op[0] == "store":
                nxt.append((wpc + 1, infl | {op[1]}, dirty, wbs, mem, ctr, ctr_fly, rpc, vals))
            elif op[0] == "wait":
                if not infl and not wbs:
                    nxt.append((wpc + 1, infl, dirty, wbs, mem, ctr, ctr_fly, rpc, vals))
            elif op[0] == "wbl2":
                nxt.append((wpc + 1, infl, frozenset(), wbs + (dirty,), mem, ctr, ctr_fly, rpc, vals))
            elif op[0] == "publish":
                nxt.append((wpc + 1, infl, dirty, wbs, mem, ctr, True, rpc, vals))
        # asynchronous events: a store lands in the writer's L2, a write-back reaches memory,
        # the counter store reaches memory
        for w in infl:
            nxt.append((wpc, infl - {w}, dirty | {w}, wbs, mem, ctr, ctr_fly, rpc, vals))
        for k, lines in enumerate(wbs):
            m = tuple(NEW if w in lines else mem[w] for w in range(WORDS))
            nxt.append((wpc, infl, dirty, wbs[:k] + wbs[k + 1:], m, ctr, ctr_fly, rpc, vals))
        if ctr_fly:
            nxt.append((wpc, infl, dirty, wbs, mem, 1, False, rpc, vals))
        # reader: poll until the counter shows the block, then read its words
        if rpc == 0:
            if ctr == 1:
                nxt.append((wpc, infl, dirty, wbs, mem, ctr, ctr_fly, 1, vals))
        elif rpc <= WORDS:
            w = rpc - 1
            v = OLD if plain_reader else mem[w]  # a plain load hits the reader's stale L2 line
            nxt.append((wpc, infl, dirty, wbs, mem, ctr, ctr_fly, rpc + 1, vals + (v,)))
        else:
            outcomes.add(vals)
        for s in nxt:
            if s not in seen:
                seen.add(s)
                stack.append(s)
    return outcomes


def test_protocol_publishes_whole_blocks():
    assert explore() == {(NEW,) * WORDS}


def test_protocol_mutants_fail():
    for mutant in ("no_wait_after_writeback", "no_wait_before_barrier", "no_writeback", "plain_loads"):
        outcomes = explore(mutant)
        assert any(v != (NEW,) * WORDS for v in outcomes), mutant


def run_tickets(order, slots, by_index):
    """Workgroups 0..len(order)-1 dispatched in `order` into `slots` resident
    slots; each takes a ticket at start (or uses its index when by_index) and
    waits until the holder of the previous ticket / index has finished.
    Returns True if every workgroup finishes (schedules explored exhaustively
    over which runnable workgroup finishes next)."""
    count = len(order)

    def step(queue, resident, done, next_ticket):
        if len(done) == count:
            return True
        # dispatch while a slot is free (the dispatcher does not wait on running workgroups)
        if queue and len(resident) < slots:
            wg = queue[0]
            key = next_ticket if not by_index else wg
            return step(queue[1:], resident + ((wg, key),), done, next_ticket + 1)
        runnable = [r for r in resident if r[1] == 0 or (r[1] - 1) in done]
        if not runnable:
            return False  # every resident workgroup waits on one that is not done
        for r in runnable:
            rest = tuple(x for x in resident if x != r)
            if not step(queue, rest, done | {r[1]}, next_ticket):
                return False
        return True

    return step(tuple(order), (), frozenset(), 0)


def test_tickets_never_deadlock():
    for count in (3, 4, 5):
        for slots in (1, 2, 3):
            for order in itertools.permutations(range(count)):
                assert run_tickets(order, slots, by_index=False), (order, slots)


def test_waiting_on_indices_can_deadlock():
    # dispatched last-index first into one slot, the first workgroup waits on an index never started
    assert not run_tickets((2, 1, 0), 1, by_index=True)
    assert any(not run_tickets(order, 2, by_index=True) for order in itertools.permutations(range(4)))
