"""The work-queue protocol of the persistent batched SIRT launch
(batch_sirt_persistent in csrc/batch2d.hip), modelled on the host and run
under random interleavings of workgroups: items are dealt to 8 queues by
decreasing work, each queue holds one iteration's items stage after stage
(slice copies, FP, BP) and repeats them per iteration, and an item waits
until the previous stage of its slice is complete.  The model checks that
every schedule drains (no deadlock, whatever the number of workgroups,
down to one) and that every stage of a slice runs only after the stage it
depends on, as the kernel's counters demand.  A planted mistake -- an FP
item that does not wait for its slice copies -- is caught."""
import random

import pytest

K_PAD = 2


def units(h, nd, stage, db, py):
    """persist_units(): items of one stage of a slice."""
    if stage == 0:
        return (h + 2 * K_PAD + 63) // 64
    if stage == 1:
        return (nd + db - 1) // db
    return (h + 4 * py - 1) // (4 * py)


def build(slices, db, py):
    """build_persistent(): queues of one iteration's items."""
    work = [na * (nd * max(w, h) + w * h) for (w, h, na, nd) in slices]
    order = sorted(range(len(slices)), key=lambda i: -work[i])
    qs, load = [[] for _ in range(8)], [0] * 8
    for i in order:
        x = min(range(8), key=lambda k: load[k])
        qs[x].append(i)
        load[x] += work[i]
    items, qoff = [], []
    for x in range(8):
        qoff.append(len(items))
        for st in range(3):
            for i in qs[x]:
                w, h, na, nd = slices[i]
                items += [(i, st, u) for u in range(units(h, nd, st, db, py))]
    qoff.append(len(items))
    return items, qoff


def need(slices, s, st, it, db, py, fp_waits_for_copy=True):
    """Completed items of the previous stage that item (s, st, it) waits for."""
    w, h, na, nd = slices[s]
    prev = 2 if st == 0 else st - 1
    if st == 1 and not fp_waits_for_copy:
        return prev, 0
    return prev, (it if st == 0 else it + 1) * units(h, nd, prev, db, py)


def run(slices, iters, n_wg, db, py, seed, fp_waits_for_copy=True):
    rnd = random.Random(seed)
    items, qoff = build(slices, db, py)
    heads, done = [0] * 8, [[0] * len(slices) for _ in range(3)]
    wgs = [dict(q=0, home=b % 8, state="deq", item=None) for b in range(n_wg)]
    active, finished, steps = set(range(n_wg)), 0, 0
    total = (qoff[8] - qoff[0]) * iters
    while active:
        steps += 1
        assert steps < 10 * total * n_wg + 10000, "no progress (deadlock)"
        b = rnd.choice(sorted(active))
        g = wgs[b]
        if g["state"] == "deq":
            got = None
            while g["q"] < 8:
                x = (g["home"] + g["q"]) % 8
                n = qoff[x + 1] - qoff[x]
                i = heads[x]
                heads[x] += 1
                if i < n * iters:
                    s, st, u = items[qoff[x] + i % n]
                    got = (s, st, u, i // n)
                    break
                g["q"] += 1
            if got is None:
                active.discard(b)
                continue
            g["item"], g["state"] = got, "wait"
        elif g["state"] == "wait":
            s, st, u, it = g["item"]
            prev, k = need(slices, s, st, it, db, py, fp_waits_for_copy)
            if done[prev][s] >= k:
                g["state"] = "run"
        else:
            s, st, u, it = g["item"]
            w, h, na, nd = slices[s]
            # the stage this item reads is complete for its slice
            if st == 0:
                assert done[2][s] == it * units(h, nd, 2, db, py)
            else:
                assert done[st - 1][s] == (it + 1) * units(h, nd, st - 1, db, py), (s, st, it)
            done[st][s] += 1
            finished += 1
            g["state"] = "deq"
    assert finished == total
    for s, (w, h, na, nd) in enumerate(slices):
        for st in range(3):
            assert done[st][s] == iters * units(h, nd, st, db, py)


def _slices(rnd, n):
    out = []
    for _ in range(n):
        size = rnd.randint(1, 300)
        out.append((size, size, rnd.randint(1, 200), rnd.randint(1, 450)))
    return out


@pytest.mark.parametrize("seed", range(12))
def test_protocol_drains_in_dependency_order(seed):
    rnd = random.Random(seed)
    slices = _slices(rnd, rnd.randint(1, 12))
    run(slices, rnd.randint(1, 4), rnd.choice([1, 2, 3, 8, 9, 40]), rnd.choice([32, 64, 128]),
        rnd.choice([2, 4, 8]), seed)


def test_single_workgroup_never_waits():
    """One workgroup takes the items in queue order: every dependency is
    already complete when it dequeues an item."""
    rnd = random.Random(5)
    run(_slices(rnd, 6), 3, 1, 32, 4, 5)


def test_planted_missing_wait_is_caught():
    rnd = random.Random(1)
    slices = _slices(rnd, 6)
    with pytest.raises(AssertionError):
        for seed in range(20):
            run(slices, 2, 16, 32, 4, seed, fp_waits_for_copy=False)
