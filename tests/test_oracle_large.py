"""CPU tests of the host model's large region (requests above 3072 B): the
class of a request, first-fit search order from the ticket's start word,
disjoint aligned blocks, no-op double and interior frees, NULL only when no
aligned run of free units is left, and units reused across classes.

The reference hands these requests to the CUDA device heap (src/halloc.cu:
185-187, 233-235), whose algorithm is not in its tree; its tests pin that
boundary only by non-NULL and disjoint blocks (tst/corr/run-all-tests.pl:23-46
checkptr at 3080 B .. 32 KiB).  The checks here pin the region's own scheme
against brute force over the dumped state."""
import numpy as np
import pytest

from oracle import model as M
from tests.helpers import (LARGE_SB, LARGE_UNIT, REC_NULL, large_class_of, large_state, large_units,
                           live_large_blocks, unit_used)

MEM = 64 << 20          # 16 MiB large region: 4096 units, 64 words
LUNITS = large_units(MEM)


@pytest.fixture(scope="module", autouse=True)
def _built():
    M.build()


def opts():
    return M.make_opts(memory=MEM, sb_sz_sh=21, nheads=1)


def one_lane(ops_list):
    ops = np.zeros((len(ops_list), 1, 2), dtype=np.uint32)
    for i, op in enumerate(ops_list):
        ops[i, 0] = op
    return ops


def fibonacci_start(ticket, c, n):
    """The word (or run) a group of class c with this ticket looks at first."""
    h = ((ticket * 0x9E3779B1) & 0xFFFFFFFF) ^ ((c * 0x85EBCA77) & 0xFFFFFFFF)
    return (h * n) >> 32


def expected_unit(lbits, lunits, ticket, c):
    """First fit from the start word: the lowest free aligned field of the
    first word (cyclically from the start) that has one; above 64 units, the
    first aligned run of all-free words (cyclically from the start run)."""
    nw = len(lbits)
    k = 1 << c
    used = lambda u: u >= lunits or unit_used(lbits, u)
    if k <= 64:
        w0 = fibonacci_start(ticket, c, nw)
        for i in range(nw):
            w = (w0 + i) % nw
            for p in range(0, 64, k):
                if not any(used(w * 64 + p + j) for j in range(k)):
                    return w * 64 + p
        return None
    m = k // 64
    ngr = nw // m
    if ngr == 0:
        return None
    g0 = fibonacci_start(ticket, c, ngr)
    for i in range(ngr):
        g = (g0 + i) % ngr
        if not any(used(g * k + j) for j in range(k)):
            return g * k
    return None


def test_large_class_and_record_of_each_request():
    rng = np.random.default_rng(1)
    sizes = np.concatenate([[3073, 4096, 4097, 8192, 8193, 65536, 65537, 1 << 20],
                            rng.integers(3073, 300000, 60)])
    h = M.ModelHeap(opts())
    recs, ptrs = h.replay(one_lane([(1, int(s)) for s in sizes]))
    for s, r, p in zip(sizes, recs[:, 0], ptrs[:, 0]):
        c = large_class_of(s)
        assert (1 << c) * LARGE_UNIT >= s and (c == 0 or (1 << (c - 1)) * LARGE_UNIT < s)
        if r[0] == REC_NULL:
            continue
        assert r[0] == 16 + c and r[1] == LARGE_SB
        assert r[2] % (1 << c) == 0 and r[2] + (1 << c) <= LUNITS       # aligned to its size, inside
        assert int(p) - 1 == h.nsbs * (1 << 21) + int(r[2]) * LARGE_UNIT  # after the slabs
    assert M.check_dump(opts(), h.dump())[0] == 0


def malloc1(h, n):
    recs, ptrs = h.wave([1], [n])
    return recs[0], int(ptrs[0])


def free1(h, p):
    recs, _ = h.wave([2], [p])
    return recs[0]


def test_large_search_is_first_fit_from_the_ticket_start():
    rng = np.random.default_rng(2)
    h = M.ModelHeap(opts())
    live = []
    hits = 0
    for step in range(1500):
        if live and rng.random() < 0.45:
            free1(h, live.pop(int(rng.integers(len(live)))))
            continue
        n = int(rng.choice([4096, 8192, 20000, 65536, 300000, 600000, 2 << 20]))
        c = large_class_of(n)
        tick, lbits, _ = large_state(h.dump(), LUNITS)
        want = expected_unit(lbits, LUNITS, int(tick[c]), c)
        r, p = malloc1(h, n)
        got = None if r[0] == REC_NULL else int(r[2])
        assert got == want, (step, n, got, want)
        if got is not None:
            live.append(p)
            hits += 1
    assert hits > 300
    assert M.check_dump(opts(), h.dump())[0] == 0


def test_large_random_program_keeps_blocks_disjoint():
    from tests.helpers import random_ops
    rng = np.random.default_rng(3)
    for lanes in (1, 7, 64):
        h = M.ModelHeap(opts())
        ops = random_ops(rng, 120, lanes, [4000, 9000, 17000, 40000, 70000, 150000], p_free=0.35)
        recs, _ = h.replay(ops)
        blocks = sorted(live_large_blocks(ops, recs))
        for (u0, n0), (u1, _) in zip(blocks, blocks[1:]):
            assert u0 + n0 <= u1
        tot, det = M.check_dump(opts(), h.dump())
        assert tot == 0, det
        _, lbits, lstart = large_state(h.dump(), LUNITS)
        assert sum(n for _, n in blocks) == sum(bin(int(w)).count("1") for w in lbits) - (len(lbits) * 64 - LUNITS)


def test_large_double_and_interior_frees_change_nothing():
    o = opts()
    h = M.ModelHeap(o)
    blocks = [malloc1(h, n) for n in (65536, 8192, 16384, 4096)]
    assert all(r[1] == LARGE_SB for r, _ in blocks)
    (_, p0), (_, p1), (_, p2), (_, p3) = blocks
    before = h.dump()
    # interior pointers (one 4 KiB unit in), and pointers off a unit boundary
    for p in (p0 + LARGE_UNIT, p0 + 3 * LARGE_UNIT, p2 + LARGE_UNIT, p1 + 16, p3 + 2048):
        r = free1(h, p)
        assert list(r[:3]) == [0, REC_NULL, REC_NULL]
    assert (h.dump() == before).all()
    r = free1(h, p1)
    assert r[0] == 2 and r[1] == LARGE_SB and r[2] == blocks[1][0][2]
    after = h.dump()
    assert list(free1(h, p1)[:3]) == [0, REC_NULL, REC_NULL]        # a double free
    assert list(free1(h, p1 + LARGE_UNIT)[:3]) == [0, REC_NULL, REC_NULL]
    assert (h.dump() == after).all()
    # later blocks never overlap the live ones
    for n in (8192, 4096, 4096, 65536, 16384):
        malloc1(h, n)
    tot, det = M.check_dump(o, h.dump())
    assert tot == 0, det
    # the same through a replay program (op kinds 2 and 3 as the GPU replays them)
    ops = np.zeros((9, 1, 2), dtype=np.uint32)
    ops[:4, 0] = [(1, 65536), (1, 8192), (1, 16384), (1, 4096)]
    ops[4:, 0] = [(3, 0), (3, 2), (2, 1), (2, 1), (3, 1)]
    recs, _ = M.ModelHeap(o).replay(ops)
    assert [list(x[:3]) for x in recs[4:, 0]] == [[0, REC_NULL, REC_NULL]] * 2 + [
        [2, LARGE_SB, recs[1, 0, 2]]] + [[0, REC_NULL, REC_NULL]] * 2


def test_large_units_are_reused_across_classes_and_null_only_when_full():
    o = opts()
    h = M.ModelHeap(o)
    n4 = LUNITS + 64
    recs, _ = h.replay(np.tile(np.array([[[1, 4096]]], dtype=np.uint32), (n4, 1, 1)))
    got = recs[:, 0, 0] != REC_NULL
    assert got.sum() == LUNITS and got[:LUNITS].all()      # every unit, then NULL
    _, lbits, _ = large_state(h.dump(), LUNITS)
    assert all(int(w) == (1 << 64) - 1 for w in lbits)
    # free everything: the units serve 64 KiB blocks (class 4) and 1 MiB blocks (class 8)
    frees = np.zeros((LUNITS, 1, 2), dtype=np.uint32)
    frees[:, 0, 0] = 2
    frees[:, 0, 1] = np.arange(LUNITS)
    ops = np.concatenate([np.tile(np.array([[[1, 4096]]], dtype=np.uint32), (LUNITS, 1, 1)), frees])
    g = M.ModelHeap(o)
    g.replay(ops)
    _, lbits, lstart = large_state(g.dump(), LUNITS)
    assert not lbits.any() and not lstart.any()
    recs, _ = g.replay(np.tile(np.array([[[1, 65536]]], dtype=np.uint32), (LUNITS // 16 + 4, 1, 1)))
    got = recs[:, 0, 0] != REC_NULL
    assert got.sum() == LUNITS // 16 and (recs[got][:, 0, 0] == 16 + 4).all()
    assert M.check_dump(o, g.dump())[0] == 0


def test_large_null_only_when_no_aligned_run_is_free():
    rng = np.random.default_rng(4)
    o = opts()
    h = M.ModelHeap(o)
    live = []
    nulls = 0
    for step in range(2500):
        if live and rng.random() < 0.3:
            free1(h, live.pop(int(rng.integers(len(live)))))
        n = int(rng.choice([4096, 12000, 64000, 250000, 1 << 20, 4 << 20]))
        c = large_class_of(n)
        _, lbits, _ = large_state(h.dump(), LUNITS)
        any_run = expected_unit(lbits, LUNITS, 0, c) is not None
        r, p = malloc1(h, n)
        assert (r[0] != REC_NULL) == any_run, step
        nulls += r[0] == REC_NULL
        if p:
            live.append(p)
    assert nulls > 100
    assert M.check_dump(o, h.dump())[0] == 0


def test_checker_catches_large_region_mistakes():
    o = opts()
    h = M.ModelHeap(o)
    h.replay(one_lane([(1, 65536), (1, 8192), (1, 5000)]))
    d = h.dump()
    assert M.check_dump(o, d)[0] == 0
    n = len(d)
    lwords = (LUNITS + 63) // 64
    lb0 = n - LUNITS - lwords
    starts = np.nonzero(d[n - LUNITS:])[0]
    u = int(starts[0])
    bad = d.copy()                       # a live block's unit marked free
    bad[lb0 + u // 64] &= ~np.uint64(1 << (u % 64))
    assert M.check_dump(o, bad)[1][6] > 0
    bad = d.copy()                       # a stray used unit
    free_u = next(x for x in range(LUNITS) if not unit_used(d[lb0:lb0 + lwords], x))
    bad[lb0 + free_u // 64] |= np.uint64(1 << (free_u % 64))
    assert M.check_dump(o, bad)[1][6] > 0
    bad = d.copy()                       # a start word off its alignment
    bad[n - LUNITS + u + 1] = bad[n - LUNITS + u]
    bad[n - LUNITS + u] = 0
    assert M.check_dump(o, bad)[1][6] > 0
