"""CPU model of the row-block POTRF schedule (item_potrf_rows in tasktorrent_amd/csrc/tile_ops.hpp).

Item rb owns rows [64 rb, 64 rb + 64) of the tile and waits only on the progress counters of
lower-numbered items: before computing L(rb, j) it needs prog[j] >= j + 1 (panel j factored and
inverted), before updating A(rb, q) with L(q, j) it needs prog[q] >= j + 1. The model runs the
items as generators in random interleavings (an item advances only when its wait is satisfied)
and checks that every interleaving finishes and yields the Cholesky factor, which is what the
counters must guarantee on the device (numpy.linalg.cholesky as the reference result).
"""
import random

import numpy as np
import pytest

DB = 64


def _run(n, seed):
    rng = np.random.default_rng(seed)
    pick = random.Random(seed)
    B = rng.standard_normal((n, n))
    A0 = B @ B.T / n + np.eye(n)
    A = np.tril(A0).copy()
    nrb = (n + DB - 1) // DB
    prog = [0] * nrb
    xinv = {}

    def blk(r):
        return slice(r * DB, min(n, r * DB + DB))

    def item(rb):
        pr = blk(rb)
        acc_d = -np.tril(A[pr, pr])
        for j in range(rb):
            yield j, j + 1
            L = A[pr, blk(j)] @ xinv[j].T
            A[pr, blk(j)] = L
            prog[rb] = j + 1
            acc_d += L @ L.T
            for q in range(j + 1, rb):
                yield q, j + 1
                A[pr, blk(q)] -= L @ A[blk(q), blk(j)].T
        D = np.tril(-acc_d)
        Lc = np.linalg.cholesky(D + np.tril(D, -1).T)
        A[pr, pr] = Lc
        xinv[rb] = np.linalg.inv(Lc)
        prog[rb] = rb + 1

    gens = {rb: item(rb) for rb in range(nrb)}
    waits = {rb: None for rb in range(nrb)}
    live = set(range(nrb))
    steps = 0
    while live:
        ready = [rb for rb in sorted(live) if waits[rb] is None or prog[waits[rb][0]] >= waits[rb][1]]
        assert ready, "every live item waits: the schedule deadlocked"
        rb = pick.choice(ready)
        try:
            waits[rb] = next(gens[rb])
        except StopIteration:
            live.discard(rb)
        steps += 1
    return np.abs(np.tril(A) - np.linalg.cholesky(A0)).max()


@pytest.mark.parametrize("n", [1, 64, 100, 320, 500, 512])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_row_block_schedule_any_interleaving(n, seed):
    assert _run(n, seed + n) < 1e-13
