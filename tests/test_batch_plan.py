"""CPU tests of the merge plan of vare_flip_step_batched (vare_debug_batch_plan): how the operations the
tanks recorded between two sync points are issued.  Every operation is issued exactly once, each tank's
operations keep their order, a merged launch holds one kernel and at most one operation per tank (and at
most 32 tanks), identical sequences merge completely, and a tank with extra or missing operations falls
back into step with the others."""
import numpy as np
import pytest

from vare_amd import _lib as L

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(L.LIB_PATH), reason="library not built")

MAX_TANKS = 32


def check_plan(seqs):
    t, p, it = L.batch_plan(seqs)
    n = sum(len(q) for q in seqs)
    assert len(t) == n
    seen = set()
    last = [-1] * len(seqs)
    for tank, pos in zip(t, p):
        assert (tank, pos) not in seen                      # issued once
        seen.add((tank, pos))
        assert pos == last[tank] + 1                        # in the tank's own order
        last[tank] = pos
    assert last == [len(q) - 1 for q in seqs]               # all issued
    items = {}
    for tank, pos, i in zip(t, p, it):
        items.setdefault(i, []).append((tank, seqs[tank][pos]))
    for i, ops in items.items():
        keys = {k for _, k in ops}
        tanks = [a for a, _ in ops]
        assert len(keys) == 1, ops                           # one kernel per launch
        assert len(set(tanks)) == len(tanks)                # one operation per tank
        k = keys.pop()
        assert len(ops) <= (MAX_TANKS if k >= 0 else 1)     # copies and events are issued one by one
    assert list(it) == sorted(it)                           # items are consecutive in issue order
    return t, p, it, items


def kernel_items(items):
    return [ops for ops in items.values() if ops[0][1] >= 0]


def test_identical_sequences_merge_completely():
    seq = [0, 1, 2, 3, 1, 2, 3, 4, -1]
    seqs = [seq] * 12
    _, _, _, items = check_plan(seqs)
    ks = kernel_items(items)
    assert len(ks) == 8 and all(len(ops) == 12 for ops in ks)


def test_more_tanks_than_one_launch_holds():
    seqs = [[5, 6, 7, -1]] * 40
    _, _, _, items = check_plan(seqs)
    ks = kernel_items(items)
    assert sorted(len(ops) for ops in ks) == sorted([32, 8] * 3)


def test_extra_operation_falls_back_in_step():
    # one tank sorts a long bin (an extra kernel 9) in the middle of its binning
    base = [0, 1, 2, 3, 4, 5, -1]
    extra = [0, 1, 2, 9, 3, 4, 5, -1]
    seqs = [base] * 5 + [extra] + [base] * 2
    _, _, _, items = check_plan(seqs)
    ks = kernel_items(items)
    assert len(ks) == 7                                      # six shared kernels + the lone extra one
    assert sorted(len(ops) for ops in ks) == [1] + [8] * 6


def test_uneven_solver_chunks():
    # the first tank enqueued 6 iterations of (a, b, c), the others 4: the first 4 are shared
    it = [10, 11, 12]
    seqs = [it * 6 + [-1]] + [it * 4 + [-1]] * 3
    _, _, _, items = check_plan(seqs)
    ks = kernel_items(items)
    assert sorted(len(ops) for ops in ks) == [1] * 6 + [4] * 12


def test_random_sequences_keep_every_guarantee():
    rng = np.random.default_rng(11)
    for trial in range(200):
        ntank = int(rng.integers(1, 48))
        base = list(rng.integers(0, 6, size=int(rng.integers(1, 30))))
        seqs = []
        for _ in range(ntank):
            q = list(base)
            for _ in range(int(rng.integers(0, 4))):         # insert, drop or change a few operations
                op = int(rng.integers(0, 3))
                pos = int(rng.integers(0, len(q) + 1))
                if op == 0:
                    q.insert(pos, int(rng.integers(-2, 8)))
                elif op == 1 and q:
                    q.pop(min(pos, len(q) - 1))
                elif q:
                    q[min(pos, len(q) - 1)] = int(rng.integers(-2, 8))
            seqs.append(q)
        if sum(len(q) for q in seqs) == 0:
            continue
        check_plan(seqs)


def test_empty_sequences():
    t, p, it = L.batch_plan([[], [3, -1], []])
    assert list(t) == [1, 1] and list(p) == [0, 1]
