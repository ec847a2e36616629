"""CPU models of two executor protocols (k_persistent in tasktorrent_amd/csrc/kernels.hip), run as
interleavings of their atomic steps chosen at random.

Claim credits: a queue keeps pushed-but-unclaimed entries in `count`; a claimer takes a credit
(fetch-add -1, giving it back when the old value was not above its reserve: 0 for a primary
worker, the number of idle primary workers for a secondary one) and then a slot (fetch-add on
the head); a producer reserves slots on the tail, writes the entries, then adds credits. Checked:
a claimed slot is always below the tail, no slot is claimed twice, every pushed slot is claimed.

Release elision: an agent-scope release writes back every dirty line of one XCD's L2, so for each
task only the item that takes the task's running-item counter of its XCD to zero releases; every
item then decrements the task's remaining count. Stores travel to the L2 asynchronously; each wave
waits for its own before the post-item barrier. Checked: when that count reaches zero, no store of
any item of the task is still held in an L2 or in flight (all are visible to the successors), and
the same model without the store wait does lose stores.
"""
import random

import pytest


def _claims(seed, n_prod=4, per=5, n_cons=6):
    rnd = random.Random(seed)
    st = {"count": 0, "head": 0, "tail": 0}
    written, taken = set(), []

    def producer(n):
        pos = st["tail"]
        st["tail"] += n  # fetch-add on the tail
        yield
        for r in range(n):
            written.add(pos + r)
            yield
        st["count"] += n  # credits after the entries
        yield

    def consumer(primary):
        while True:
            reserve = 0 if primary else rnd.randrange(3)
            if st["count"] <= reserve:
                if len(taken) >= n_prod * per:
                    return
                yield
                continue
            old = st["count"]
            st["count"] -= 1  # fetch-add -1
            yield
            if old <= reserve:
                st["count"] += 1  # give the credit back
                yield
                continue
            slot = st["head"]
            st["head"] += 1
            assert slot < st["tail"], "a claim passed the tail"
            yield
            while slot not in written:  # claimed before the producer wrote it: poll
                yield
            taken.append(slot)
            yield

    actors = [producer(per) for _ in range(n_prod)] + [consumer(c % 2 == 0) for c in range(n_cons)]
    live = list(range(len(actors)))
    for _ in range(100000):
        if not live:
            break
        a = rnd.choice(live)
        try:
            next(actors[a])
        except StopIteration:
            live.remove(a)
    assert not live, "did not finish"
    assert sorted(taken) == list(range(n_prod * per))
    assert st["count"] == 0 and st["head"] == st["tail"] == n_prod * per


@pytest.mark.parametrize("seed", range(40))
def test_claim_credits_never_pass_the_tail(seed):
    _claims(seed)


def _elision(seed, n_items=16, n_xcd=4, wait=True):
    """One task's items on random XCDs. A store is issued into flight first and reaches its XCD's L2
    at a later step of the scheduler's choosing; `wait` models the per-wave store wait before the
    post-item barrier (s_waitcnt vmcnt(0)), after which the item's stores are in the L2 before its
    relaxed decrement. Returns True when every store was written back before the task's count hit 0."""
    rnd = random.Random(seed)
    run = [0] * n_xcd
    dirty = [set() for _ in range(n_xcd)]  # stores held in each XCD's L2
    visible = set()
    left = [n_items]
    done_check = []
    in_flight = {}  # item -> XCD its stores are travelling to

    def deliver(it):
        if it in in_flight:
            dirty[in_flight.pop(it)].add(it)

    def item(it):
        x = rnd.randrange(n_xcd)
        run[x] += 1  # start
        yield
        in_flight[it] = x  # the item's stores are issued
        yield
        if wait:  # every wave waits for its stores to be acknowledged by the L2
            deliver(it)
            yield
        run[x] -= 1
        last_here = run[x] == 0
        yield
        if last_here:  # agent-scope release: write back the whole L2 of this XCD
            visible.update(dirty[x])
            dirty[x].clear()
            yield
        left[0] -= 1
        if left[0] == 0:
            done_check.append(set(range(n_items)) <= visible)
        yield

    def memory():  # stores in flight reach the L2 in any order, at any time
        while True:
            if in_flight:
                deliver(rnd.choice(sorted(in_flight)))
            yield

    actors = [item(i) for i in range(n_items)]
    mem = memory()
    live = list(range(n_items))
    while live:
        if rnd.random() < 0.3:
            next(mem)
            continue
        a = rnd.choice(live)
        try:
            next(actors[a])
        except StopIteration:
            live.remove(a)
    assert len(done_check) == 1
    return done_check[0]


@pytest.mark.parametrize("seed", range(200))
def test_release_elision_publishes_every_item(seed):
    assert _elision(seed, n_items=1 + seed % 16, n_xcd=1 + seed % 8)


def test_release_elision_needs_the_store_wait():
    """Without the per-wave store wait a store can reach the L2 after the last item's write-back:
    the model must find such an interleaving (it checks that the model can fail at all)."""
    fails = sum(not _elision(seed, n_items=2 + seed % 15, n_xcd=1 + seed % 4, wait=False) for seed in range(300))
    assert fails > 0
