"""CPU models of two executor protocols (k_persistent in tasktorrent_amd/csrc/kernels.hip), run as
interleavings of their atomic steps chosen at random.

Claim credits: a queue keeps pushed-but-unclaimed entries in `count`; a claimer takes a credit
(fetch-add -1, giving it back when the old value was not above its reserve: 0 for a primary
worker, the number of idle primary workers for a secondary one) and then a slot (fetch-add on
the head); a producer reserves slots on the tail, writes the entries, then adds credits. Checked:
a claimed slot is always below the tail, no slot is claimed twice, every pushed slot is claimed.

Release elision: an agent-scope release writes back every dirty line of one XCD's L2, so for each
task only the item that takes the task's running-item counter of its XCD to zero releases; every
item then decrements the task's remaining count. Checked: when that count reaches zero, no store
of any item of the task is still held in an L2 (all are visible to the successors).
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


def _elision(seed, n_items=16, n_xcd=4):
    rnd = random.Random(seed)
    run = [0] * n_xcd
    dirty = [set() for _ in range(n_xcd)]  # stores held in each XCD's L2
    visible = set()
    left = [n_items]
    done_check = []

    def item(it):
        x = rnd.randrange(n_xcd)
        run[x] += 1  # start
        yield
        dirty[x].add(it)  # the item's stores (complete before its decrement)
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

    actors = [item(i) for i in range(n_items)]
    live = list(range(n_items))
    while live:
        a = rnd.choice(live)
        try:
            next(actors[a])
        except StopIteration:
            live.remove(a)
    assert done_check == [True]


@pytest.mark.parametrize("seed", range(200))
def test_release_elision_publishes_every_item(seed):
    _elision(seed, n_items=1 + seed % 16, n_xcd=1 + seed % 8)
