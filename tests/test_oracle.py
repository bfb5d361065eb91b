"""CPU tests of the host model (oracle/) against closed forms, brute force,
the reference's size table and the reference test suite's invariants."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import model as M
from tests.helpers import (REC_NULL, BLOCK_SIZES, assert_disjoint, class_switch_program, crowded_program, head_words,
                           live_intervals, pressure_program, random_batch_ops, random_ops, wave_refree_ops)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _built():
    M.build()


def primes_below(n):
    if n < 3:
        return []
    sieve = bytearray([1]) * n
    sieve[0:2] = b"\x00\x00"
    for i in range(2, int(n ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(sieve[i * i::i]))
    return [i for i in range(n) if sieve[i]]


def test_size_id_is_smallest_fitting_block():
    # src/halloc.cu:53-67: the class is the smallest 2*unit / 3*unit block holding n (>= 16 B)
    L = M.lib()
    for n in range(0, 3073):
        want = next(i for i, b in enumerate(BLOCK_SIZES) if b >= max(n, 16))
        assert L.hm_size_id(n) == want, n


@pytest.mark.parametrize("sz", [16, 24, 256, 384])
def test_chunk_arithmetic_round_trip(sz):
    # src/size-info.h:50-70: chunk_mul(v) = v*sz, chunk_div(v*sz + r) = v for r < sz
    L = M.lib()
    cs = L.hm_chunk_val(sz)
    for v in list(range(0, 4096)) + [262143, 174759, 16383, 10921]:
        if v * sz >= 1 << 32:
            continue
        assert L.hm_chunk_mul(v, cs) == v * sz
        for r in (0, 1, sz - 1):
            if v * sz + r < 1 << 32:
                assert L.hm_chunk_div(v * sz + r, cs) == v


def test_max_prime_below_matches_sieve():
    # src/utils.cu:8-24: largest prime p < n with n % p and nb % p nonzero
    L = M.lib()
    P = primes_below(70000)
    rng = np.random.default_rng(0)
    for n in list(range(4, 400)) + [int(x) for x in rng.integers(400, 70000, 300)]:
        nb = int(rng.integers(1, 1 << 20))
        cands = [p for p in P if 3 <= p < n and n % p and nb % p]
        want = max(cands) if cands else 0
        assert L.hm_max_prime_below(n, nb) == want, (n, nb)


def test_golden_size_table_is_reproducible():
    # the committed table is what its generator (an independent restatement
    # of src/halloc.cu:363-394 and src/utils.cu:8-24) writes
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(GOLDEN, "gen_size_table.py"), "--check"])
    assert r.returncode == 0


def test_reference_size_table_matches_golden():
    # the reference's own table at 4 MiB slabs (src/halloc.cu:363-394)
    gold = json.load(open(os.path.join(GOLDEN, "size_table_ref_sb22.json")))["rows"]
    t = M.size_table(M.make_opts(hash_mode=0))
    F = {f: i for i, f in enumerate(M.SIZE_TABLE_FIELDS)}
    for row in gold:
        r = t[row["id"]]
        assert r[F["block_sz"]] == row["block_sz"]
        assert r[F["chunk_id"]] == row["chunk_id"]
        assert r[F["chunk_bytes"]] == row["chunk_bytes"]
        assert r[F["ncib"]] == row["ncib"]
        assert r[F["nblocks"]] == row["nblocks"]
        assert r[F["nchunks"]] == row["nchunks"]
        assert r[F["hash_step"]] == row["hash_step"]
        assert r[F["roomy"]] == row["roomy"]
        assert r[F["busy"]] == row["busy"]
        assert r[F["sparse"]] == row["sparse"]


def test_wave_run_tables_need_2mib_slabs():
    # with 64-block runs, 1 MiB slabs leave the 3072 B class 320 blocks and no
    # admissible hash prime below 320/256 + 320/64 = 6 (5 divides 320, 3 divides 6)
    with pytest.raises(ValueError):
        M.size_table(M.make_opts(sb_sz_sh=20, hash_mode=1))


@pytest.mark.parametrize("hash_mode", [0, 1])
@pytest.mark.parametrize("sh", [20, 21, 22, 24, 26])
def test_size_table_properties(hash_mode, sh):
    fr = dict(busy=0.835, roomy=0.6, sparse=0.012)
    if hash_mode == 1 and sh == 20:
        pytest.skip("not a valid wave-run geometry (see test_wave_run_tables_need_2mib_slabs)")
    t = M.size_table(M.make_opts(sb_sz_sh=sh, hash_mode=hash_mode))
    F = {f: i for i, f in enumerate(M.SIZE_TABLE_FIELDS)}
    sb = 1 << sh
    P = primes_below(1 << 17)
    for sid in range(16):
        r = {f: int(t[sid][i]) for f, i in F.items()}
        assert r["block_sz"] == BLOCK_SIZES[sid]
        assert r["chunk_bytes"] in (16, 24, 256, 384)
        assert r["block_sz"] % r["chunk_bytes"] == 0 and r["ncib"] == r["block_sz"] // r["chunk_bytes"]
        assert r["ncib"] in (1, 2, 4, 8)
        tmod = r["tmod"]
        assert tmod == (64 if hash_mode else 1 << max(2 - sid // 2, 0))
        # nblocks: the largest multiple of tmod that fits; with 64-block runs,
        # one run fewer where 11 (the run stride) would divide the run count
        assert r["nblocks"] % tmod == 0
        fit = sb // r["block_sz"] // tmod * tmod
        if hash_mode and (fit // 64) % 11 == 0:
            fit -= 64
        assert r["nblocks"] == fit and r["nblocks"] * r["block_sz"] <= sb
        if hash_mode:
            # first probes start at run (ticket / 64) * 11 mod runs: every run is reached
            runs = r["nblocks"] // 64
            assert len({(j * 11) % runs for j in range(runs)}) == runs
        assert r["nchunks"] == r["nblocks"] * r["ncib"]
        # hash step = ncib * the largest admissible prime
        n, nb = r["nblocks"] // 256 + r["nblocks"] // 64, r["nblocks"]
        p = r["hash_step"] // r["ncib"]
        assert r["hash_step"] % r["ncib"] == 0
        assert p == max(q for q in P if 3 <= q < n and n % q and nb % q)
        # thresholds: fraction * nchunks truncated (double arithmetic)
        for f in ("busy", "roomy", "sparse"):
            assert r[f] == int(fr[f] * r["nchunks"])
            assert abs(Fraction(r[f]) - Fraction(fr[f]) * r["nchunks"]) < 1
        assert r["sparse"] <= r["roomy"] <= r["busy"] <= r["nchunks"]


def test_wave_run_initial_positions():
    # hash_mode 1: 64 consecutive tickets fill the tail of one aligned 64-block
    # run and the head of the next, contiguously, and never collide
    L = M.lib()
    rng = np.random.default_rng(1)
    for sh in (21, 22, 26):
        t = M.size_table(M.make_opts(sb_sz_sh=sh))
        for sid in range(16):
            ncib, nchunks = int(t[sid][3]), int(t[sid][0])
            for old in [0x0B0B0B0B, 0, 1, 63, 64, 0xFFFFFFFF, 0xFFFFFFC0] + [int(x) for x in rng.integers(0, 1 << 32, 10)]:
                pos = [L.hm_ichunk_init((old + l) & 0xFFFFFFFF, sh, sid, 1) for l in range(64)]
                split = 64 - (old & 63)
                a, b = pos[:split], pos[split:]
                assert a[0] % ncib == 0 and (a[0] - (old & 63) * ncib) % (64 * ncib) == 0
                assert a == [a[0] + l * ncib for l in range(split)]
                if b:
                    assert b[0] % (64 * ncib) == 0
                    assert b == [b[0] + l * ncib for l in range(len(b))]
                assert len(set(pos)) == 64 and max(pos) < nchunks


def test_reference_hash_initial_positions_are_block_aligned():
    L = M.lib()
    t = M.size_table(M.make_opts(hash_mode=0))
    for sid in range(16):
        ncib, nchunks = int(t[sid][3]), int(t[sid][0])
        for cv in range(0x0B0B0B0B, 0x0B0B0B0B + 256):
            ich = L.hm_ichunk_init(cv, 22, sid, 0)
            assert ich % ncib == 0 and ich < nchunks


@pytest.mark.parametrize("sid", [0, 7, 15])
def test_probe_sequence_block_aligned_and_spreading(sid):
    # src/slab.cuh:492-495 on a 2 MiB slab: every probe lands on a block start;
    # the probes of one lane keep reaching new blocks (brute-force enumeration)
    L = M.lib()
    t = M.size_table(M.make_opts(sb_sz_sh=21))
    ncib, nchunks, nblocks = int(t[sid][3]), int(t[sid][0]), int(t[sid][9])
    seen = set()
    for itry in range(0, 4 * nblocks):
        ich = L.hm_probe_next(0, itry, 21, sid, 1)
        assert ich % ncib == 0 and ich < nchunks
        seen.add(ich)
    assert len(seen) >= nblocks // 4


# ------------------------------------------------------------ model behaviour
@pytest.mark.parametrize("lanes", [1, 64])
def test_model_replay_invariants(lanes):
    rng = np.random.default_rng(lanes)
    # 96 slabs: room for a head and a cached head for each of the 16 classes
    opts = M.make_opts(memory=256 << 20, sb_sz_sh=21, nheads=1)
    h = M.ModelHeap(opts)
    table = M.size_table(opts)
    ops = random_ops(rng, 300 if lanes == 64 else 4000, lanes, BLOCK_SIZES + [1, 17, 100, 1000, 3000])
    recs, _ = h.replay(ops)
    tot, det = M.check_dump(opts, h.dump())
    assert tot == 0, det
    assert_disjoint(live_intervals(ops, recs, table), 1 << 21)
    assert (recs[ops[..., 0] == 1][:, 0] != M.REC_NULL).all()


def test_model_free_everything_returns_slabs():
    # freeslabs (tst/corr/freeslabs/freeslabs.cu:35-71): after freeing every
    # block, every non-head slab is free and reusable by another class
    opts = M.make_opts(memory=64 << 20, sb_sz_sh=21, nheads=1)
    h = M.ModelHeap(opts)
    rng = np.random.default_rng(5)
    n = 64 * 200
    ops = np.zeros((200 * 2, 64, 2), dtype=np.uint32)
    ops[:200, :, 0] = 1
    ops[:200, :, 1] = rng.choice(BLOCK_SIZES[:10], size=(200, 64))
    ops[200:, :, 0] = 2
    ops[200:, :, 1] = np.arange(n).reshape(200, 64)
    h.replay(ops)
    d = h.dump()
    tot, det = M.check_dump(opts, d)
    assert tot == 0, det
    nsbs = h.nsbs
    ctr = d[:nsbs]
    assert ((ctr >> np.uint64(32)) == 0).all()
    desc = d[nsbs:3 * nsbs].reshape(nsbs, 2)
    heads = (desc[:, 1] >> np.uint64(32)) != 0
    free = (ctr >> np.uint64(8)) & np.uint64(0xF) == 0xF
    assert (heads | free).all()


def test_model_exhaustion_only_when_full():
    # NULL only when no slab of the class has room left by the scheme's rules
    opts = M.make_opts(memory=8 << 20, sb_sz_sh=21, nheads=1)
    h = M.ModelHeap(opts)
    t = M.size_table(opts)
    nsbs = h.nsbs
    busy, nchunks = int(t[0][7]), int(t[0][0])
    cap = nsbs * nchunks
    ops = np.zeros((cap, 1, 2), dtype=np.uint32)
    ops[:, 0] = (1, 16)
    recs, _ = h.replay(ops)
    nulls = np.nonzero(recs[:, 0, 0] == M.REC_NULL)[0]
    assert len(nulls) > 0
    first = nulls[0]
    # replay the prefix up to the first NULL on a fresh heap and inspect it
    h2 = M.ModelHeap(opts)
    h2.replay(ops[:first + 1])
    d = h2.dump()
    ctr = d[:nsbs] >> np.uint64(32)
    desc = d[nsbs:3 * nsbs].reshape(nsbs, 2)
    is_head = (desc[:, 1] >> np.uint64(32)) != 0
    assert (desc[:, 0] & np.uint64(0xFFFFFFFF) == 0).all()   # every slab serves class 0
    assert (ctr[is_head] >= busy).all()
    assert (ctr[~is_head] >= (busy + nchunks) // 2).all()     # the last-resort limit
    assert int(ctr.sum()) == first                              # every success is counted


def test_checker_catches_planted_mistakes():
    opts = M.make_opts(memory=32 << 20, sb_sz_sh=21, nheads=1)
    h = M.ModelHeap(opts)
    ops = np.zeros((50, 64, 2), dtype=np.uint32)
    ops[:, :, 0] = 1
    ops[:, :, 1] = 48
    h.replay(ops)
    d = h.dump()
    assert M.check_dump(opts, d)[0] == 0
    nsbs = h.nsbs
    nsw = (nsbs + 63) // 64
    off_bits = 3 * nsbs + M.NSETS * nsw + 3 * 16
    # 1: a reserved bit cleared
    bad = d.copy()
    w = off_bits + int(np.nonzero(d[off_bits:] != 0)[0][0])
    bad[w] &= bad[w] - np.uint64(1)
    assert M.check_dump(opts, bad)[1][0] > 0
    # 2: a counter off by one
    bad = d.copy()
    b = int(np.nonzero((d[:nsbs] >> np.uint64(32)) > 0)[0][0])
    bad[b] += np.uint64(1 << 32)
    assert M.check_dump(opts, bad)[1][0] > 0
    # 3: a free slab dropped from the free set
    bad = d.copy()
    free_b = int(np.nonzero(((d[:nsbs] >> np.uint64(8)) & np.uint64(0xF)) == 0xF)[0][0])
    bad[3 * nsbs + free_b // 64] &= ~np.uint64(1 << (free_b % 64))
    assert M.check_dump(opts, bad)[1][2] > 0
    # 6: a filed slab (not a head, holding chunks) and the sets it is in;
    # 1 KiB blocks fill and detach a few slabs
    ops = np.zeros((100, 64, 2), dtype=np.uint32)
    ops[:, :, 0] = 1
    ops[:, :, 1] = 1024
    h.replay(ops)
    d = h.dump()
    assert M.check_dump(opts, d)[0] == 0
    desc_hi = d[nsbs:3 * nsbs].reshape(nsbs, 2)[:, 1] >> np.uint64(32)
    filed = [b for b in range(nsbs) if desc_hi[b] == 0 and (d[b] >> np.uint64(32)) > 0]
    assert filed
    b = filed[0]
    size_id = int(d[nsbs + 2 * b]) & 0xFFFFFFFF
    word = lambda s: 3 * nsbs + s * nsw + b // 64
    bit = np.uint64(1 << (b % 64))
    ins = [s for s in range(M.NSETS) if d[word(s)] & bit]
    assert ins and 0 not in ins
    # ... also in the free set, whose pop would re-initialize it
    bad = d.copy()
    bad[word(0)] |= bit
    assert M.check_dump(opts, bad)[1][5] > 0
    # ... moved to the busy set of another size class
    bad = d.copy()
    for s in ins:
        bad[word(s)] &= ~bit
    bad[word(1 + M.NSIZES + (size_id + 1) % M.NSIZES)] |= bit
    assert M.check_dump(opts, bad)[1][5] > 0
    # ... and a stale entry of a class set on a free slab is a hint, not an error
    bad = d.copy()
    bad[3 * nsbs + (1 + M.NSIZES + size_id) * nsw + free_b // 64] |= np.uint64(1 << (free_b % 64))
    assert M.check_dump(opts, bad)[0] == 0


def test_crowded_head_is_replaced_below_its_limit():
    # the crowded-head rule: a probe sequence of 12 failed rounds replaces the
    # head even though its count is below the busy limit; the old head is filed
    # by its count and the live blocks stay disjoint
    opts = M.make_opts(memory=32 << 20, sb_sz_sh=21, nheads=1)
    h = M.ModelHeap(opts)
    ops = crowded_program()
    recs, _ = h.replay(ops)
    assert h.crowded() >= 1
    assert not ((ops[..., 0] == 1) & (recs[..., 0] == REC_NULL)).any()
    d = h.dump()
    assert M.check_dump(opts, d)[0] == 0
    nsbs = h.nsbs
    count = (d[:nsbs] >> np.uint64(32)).astype(np.int64)
    refs = d[nsbs:3 * nsbs].reshape(nsbs, 2)[:, 1] >> np.uint64(32)
    table = M.size_table(opts)
    busy = int(table[8][M.SIZE_TABLE_FIELDS.index("busy")])
    used = [b for b in range(nsbs) if count[b] > 0]
    assert len(used) >= 2
    detached = [b for b in used if refs[b] == 0]
    assert detached and all(count[b] < busy for b in detached)
    assert_disjoint(live_intervals(ops, recs, table), 1 << 21)
    # the same program without the churn never crowds its head
    h2 = M.ModelHeap(opts)
    h2.replay(ops[:100])
    assert h2.crowded() == 0


def test_concurrent_model_and_glibc_baselines_run_clean():
    rate, bad = M.bench_plain(M.make_opts(nheads=8), 8, 1 << 14, 2)
    assert bad == 0 and rate > 0
    rate, bad = M.bench_glibc(4, 1 << 14, 2)
    assert bad == 0 and rate > 0


def test_model_pressure_program_reuses_slabs_across_classes():
    opts = M.make_opts(memory=16 << 20, sb_sz_sh=21, nheads=1)
    h = M.ModelHeap(opts)
    ops = pressure_program()
    recs, _ = h.replay(ops)
    m = ops[..., 0] == 1
    assert (recs[:2500][m[:2500]][:, 0] == M.REC_NULL).sum() > 0      # the pool ran out
    assert (recs[5500:][m[5500:]][:, 0] != M.REC_NULL).all()          # and came back
    first = set(recs[:2500, :, 1][m[:2500]].tolist()) - {M.REC_NULL}
    later = set(recs[5500:, :, 1][m[5500:]].tolist()) - {M.REC_NULL}
    assert later and later <= first
    assert M.check_dump(opts, h.dump())[0] == 0


def freeslabs_sizes():
    """Sizes of the reference's freeslabs sweep (tst/corr/freeslabs/freeslabs.cu:39-69)."""
    out, size, step = [], 16, 8
    while size <= 1024:
        out.append(size)
        size += step
        step = 128 if size >= 256 else 16 if size >= 64 else 8
    return out


def test_model_freeslabs_with_idle_head_stealing():
    # 24 slabs, 8 fixed head sets: the sweep parks idle heads of 13 classes
    # (up to 16 slabs per class); each class must still get 40% of memory
    opts = M.make_opts(memory=64 << 20, sb_sz_sh=21, nheads=8, fixed_sets=1)
    h = M.ModelHeap(opts)
    for size in freeslabs_sizes():
        n = int(0.4 * (64 << 20) / size) // 64 * 64
        steps = n // 64
        ops = np.zeros((2 * steps, 64, 2), dtype=np.uint32)
        ops[:steps, :, 0] = 1
        ops[:steps, :, 1] = size
        ops[steps:, :, 0] = 2
        ops[steps:, :, 1] = np.arange(n).reshape(steps, 64)
        iheads = np.arange(2 * steps, dtype=np.uint32) % 8
        recs, _ = h.replay(ops, iheads)
        assert (recs[:steps, :, 0] != M.REC_NULL).all(), size
    assert M.check_dump(opts, h.dump())[0] == 0


def test_model_mixed_sizes_over_eight_heads_share_instead_of_failing():
    # 16 classes x 8 fixed XCD sets want 128 head slabs from a 96-slab pool:
    # pairs without a slab share another XCD's head of their class, so
    # nothing fails
    opts = M.make_opts(nheads=8, fixed_sets=1)
    h = M.ModelHeap(opts)
    rng = np.random.default_rng(21)
    nw = 1024
    ops = np.zeros((nw, 64, 2), dtype=np.uint32)
    ops[:, :, 0] = 1
    ops[:, :, 1] = rng.integers(1, 3073, (nw, 64))
    recs, _ = h.replay(ops, np.arange(nw, dtype=np.uint32) % 8)
    assert (recs[:, :, 0] != M.REC_NULL).all()
    d = h.dump()
    assert M.check_dump(opts, d)[0] == 0
    refs = (d[:h.nsbs] >> np.uint64(16)) & np.uint64(0xFF)
    assert refs.max() > 1                      # some heads are shared


def test_batch_equals_back_to_back_calls():
    # on a fresh heap, one batched call of R blocks per lane yields exactly the
    # blocks of R back-to-back plain calls (same tickets, same runs)
    for size, R in ((16, 4), (48, 7), (256, 16), (3000, 3)):
        opts = M.make_opts(nheads=1)
        a = M.ModelHeap(opts)
        bops = np.zeros((1, 64, 4), dtype=np.uint32)
        bops[0, :, :3] = (1, size, R)
        brecs = a.replay_batch(bops)
        b = M.ModelHeap(opts)
        pops = np.zeros((R, 64, 2), dtype=np.uint32)
        pops[:, :, 0] = 1
        pops[:, :, 1] = size
        precs, _ = b.replay(pops)
        assert (brecs[0, :, :R, :3].transpose(1, 0, 2) == precs[:, :, :3]).all(), (size, R)
        # the batch touched the ticket once, the counter once: same end state
        assert M.check_dump(opts, a.dump())[0] == 0


def test_model_refree_frees_nothing():
    # the wave/refree program: the first free of each block records its
    # length, the second records nothing, and the end state equals that of
    # the same program without the second free
    lanes, sizes = 64, (16, 48, 96, 192)
    ops = wave_refree_ops(lanes, sizes)
    opts = M.make_opts(memory=128 << 20, sb_sz_sh=21, nheads=4)
    a = M.ModelHeap(opts)
    recs = a.replay_batch(ops)
    skip = ops.copy()
    for i in range(len(sizes)):
        skip[7 * i + 3] = 0  # drop the second free
    b = M.ModelHeap(opts)
    b.replay_batch(skip)
    assert (a.dump() == b.dump()).all()
    assert M.check_dump(opts, a.dump())[0] == 0
    # blocks of 16, 48, 96 and 192 B span 1, 2, 4 and 8 chunks (16 B chunks
    # for 2*unit blocks, 24 B chunks for 3*unit blocks)
    for i, (size, k) in enumerate(zip(sizes, (1, 2, 4, 8))):
        first, second = recs[7 * i + 2], recs[7 * i + 3]
        assert (first[:, :, 0] == k).all()
        assert (second[:, :, 0] == 0).all() and (second[:, :, 1] == M.REC_NULL).all()
    # the heap ends empty
    nsbs = a.nsbs
    assert ((a.dump()[:nsbs] >> np.uint64(32)) == 0).all()


@pytest.mark.parametrize("nheads", [1, 8])
def test_model_batch_programs_keep_invariants(nheads):
    opts = M.make_opts(memory=128 << 20, sb_sz_sh=21, nheads=nheads)
    h = M.ModelHeap(opts)
    rng = np.random.default_rng(9)
    ops = random_batch_ops(rng, 150, 64, BLOCK_SIZES + [20, 100, 700])
    recs = h.replay_batch(ops, np.arange(150, dtype=np.uint32) % 8)
    assert M.check_dump(opts, h.dump())[0] == 0
    m = ops[..., 0] == 1
    cnt = ops[..., 2]
    got = recs[..., 0]
    for st, l in zip(*np.nonzero(m)):
        assert (got[st, l, :cnt[st, l]] != M.REC_NULL).all()


def test_serving_generator_shape():
    from halloc_amd.workloads import pages_for, serving_requests
    arr, ln, tag = serving_requests(200000, seed=0)
    gaps = np.diff(arr.astype(np.float64))
    assert (gaps >= 0).all()
    assert abs(gaps.mean() - 500.0) < 5.0          # 2000 requests per ms
    assert abs(np.median(ln) - 512) < 10
    assert ln.min() >= 1 and ln.max() <= 32768
    assert (pages_for([1, 16, 17, 32768]) == [1, 1, 2, 2048]).all()


def test_serve_checksum_helper_by_loop():
    from tests.helpers import _hash3, serve_checksums
    tags = np.array([7, 0xFFFFFFFF, 123456], dtype=np.uint32)
    prod = np.array([5, 0, 40])
    want = []
    for tg, n in zip(tags, prod):
        s = 0
        for t in range(n):
            s += int(_hash3(tg, t, 0x5E7A11C3)[0])
        want.append(s & 0xFFFFFFFF)
    assert list(serve_checksums(tags, prod)) == want


def test_model_threads_race_mixed_sizes():
    """The model under real threads (the GPU's concurrent mode): 2^16 mixed
    sizes over 8 head sets in 96 slabs; no NULL, invariants hold before and
    after the frees, and every count returns to zero."""
    rng = np.random.default_rng(21)
    sizes = rng.integers(1, 3073, 1 << 16).astype(np.uint32)
    o = M.make_opts(nheads=8)
    h = M.ModelHeap(o)
    p = h.malloc_threads(sizes, nthreads=32)
    assert (p != 0).all()
    assert M.check_dump(o, h.dump())[0] == 0
    h.free_threads(p, nthreads=32)
    d = h.dump()
    assert M.check_dump(o, d)[0] == 0
    assert ((d[:h.nsbs] >> np.uint64(32)) == 0).all()


@pytest.mark.parametrize("frac", [0.01, 0.05, 0.11, 0.3])
def test_hash_step_frac_prime_bound(frac):
    """A tuned stride keeps the table's shape: the prime is the largest below
    frac * nblocks that divides neither the bound nor nblocks."""
    t0 = M.size_table(M.make_opts())
    t = M.size_table(M.make_opts(hash_step_frac=frac))
    f = {n: i for i, n in enumerate(M.SIZE_TABLE_FIELDS)}
    same = [n for n in M.SIZE_TABLE_FIELDS if n != "hash_step"]
    assert (t[:, [f[n] for n in same]] == t0[:, [f[n] for n in same]]).all()
    for row in t:
        nb, k, hs = int(row[f["nblocks"]]), int(row[f["ncib"]]), int(row[f["hash_step"]])
        bound = int(frac * nb)
        p = hs // k
        P = set(primes_below(bound))
        assert hs % k == 0 and p in P and bound % p and nb % p
        assert not any(q in P and bound % q and nb % q for q in range(p + 1, bound))


# ------------------------------------------------------- adaptive head sets
def _spread_and_seen(h):
    n = len(h.dump())
    d = h.dump()
    lunits = (h.opts.memory - int(h.opts.halloc_fraction * h.opts.memory)) // 4096
    lwords = (lunits + 63) // 64
    base = n - lunits - lwords - 40 - 2
    return int(d[base]), int(d[base + 1])


def test_adaptive_spread_follows_the_classes_seen():
    # 96 slabs, 16 sets: one class keeps every set; each new class lowers the
    # spread to the largest power of two s with 3 * s * classes <= 96, and the
    # pairs above it hold no head afterwards
    opts = M.make_opts(nheads=16)
    h = M.ModelHeap(opts)
    nw = 64
    ops = np.zeros((nw, 64, 2), dtype=np.uint32)
    ops[:, :, 0] = 1
    ops[:, :, 1] = 16
    h.replay(ops, np.arange(nw, dtype=np.uint32) % 16)
    assert _spread_and_seen(h) == (16, 1)
    # classes seen -> spread: 1-2 -> 16, 3-4 -> 8, 5-8 -> 4, 9-16 -> 2
    expect = [None, 16, 16, 8, 8, 4, 4, 4, 4, 2, 2, 2, 2, 2, 2, 2, 2]
    for k, size in enumerate(BLOCK_SIZES[1:], start=2):
        ops[:, :, 1] = size
        h.replay(ops, np.arange(nw, dtype=np.uint32) % 16)
        sp, seen = _spread_and_seen(h)
        assert seen == (1 << k) - 1
        assert sp == expect[k], (k, sp)
    nsw = (h.nsbs + 63) // 64
    hb = 3 * h.nsbs + M.NSETS * nsw

    def heads():
        return h.dump()[hb:hb + 16 * 16].reshape(16, 16)

    before = heads()
    outside = int((before[2:] != 0xFFFFFFFF).sum())
    assert outside > 0                               # taken before the spread shrank
    # the inner pairs' 16 B cached heads went to the other classes' head
    # searches while free slabs were scarce (a cached head is a slab its
    # class holds idle)
    cached = h.dump()[hb + 16 * 16:hb + 2 * 16 * 16].reshape(16, 16)
    assert (cached[:2, 0] == 0xFFFFFFFF).all()
    # more 16 B blocks than two inner heads hold: each inner pair then adopts
    # 16 B heads of outer pairs (two each in 250 waves) instead of taking a
    # free slab, and no outer pair takes a new head
    ops[:, :, 1] = 16
    for _ in range(250):
        h.replay(ops, np.arange(nw, dtype=np.uint32) % 16)
    after = heads()
    adopted = set(before[2:, 0].tolist()) & set(after[:2, 0].tolist())
    assert len(adopted) == 2 and (after[2:, 0] == 0xFFFFFFFF).sum() == 4
    assert ((after[2:] != 0xFFFFFFFF) <= (before[2:] != 0xFFFFFFFF)).all()
    assert M.check_dump(opts, h.dump())[0] == 0


def test_adaptive_sets_serve_many_classes_on_a_small_heap():
    # 16 classes over 16 sets on 96 slabs, one wave at a time: every request
    # is served until the requested bytes pass 55% of the slab heap
    opts = M.make_opts(nheads=16)
    h = M.ModelHeap(opts)
    rng = np.random.default_rng(0)
    nw = 2600
    ops = np.zeros((nw, 64, 2), dtype=np.uint32)
    ops[:, :, 0] = 1
    ops[:, :, 1] = rng.integers(8, 3073, (nw, 64))
    recs, _ = h.replay(ops, np.arange(nw, dtype=np.uint32) % 16)
    nul = (recs[:, :, 0] == M.REC_NULL).ravel()
    first = int(np.nonzero(nul)[0][0]) if nul.any() else nul.size
    assert ops[..., 1].ravel()[:first].astype(np.int64).sum() > 0.55 * 96 * (4 << 20)
    assert M.check_dump(opts, h.dump())[0] == 0


def test_adaptive_sets_null_only_when_the_class_has_no_room():
    # one lane, 16 sets, 96 slabs, sizes over all 16 classes until the first
    # NULL: then no slab is free, and every slab of the failing class is at
    # the limit the scheme fills it to
    opts = M.make_opts(nheads=16)
    t = M.size_table(opts)
    rng = np.random.default_rng(5)
    n = 400000
    sizes = rng.integers(8, 3073, n).astype(np.uint32)
    ops = np.zeros((n, 1, 2), dtype=np.uint32)
    ops[:, 0, 0] = 1
    ops[:, 0, 1] = sizes
    iheads = np.arange(n, dtype=np.uint32) % 16
    recs, _ = M.ModelHeap(opts).replay(ops, iheads)
    nul = np.nonzero(recs[:, 0, 0] == M.REC_NULL)[0]
    assert len(nul) > 0
    first = int(nul[0])
    h = M.ModelHeap(opts)
    h.replay(ops[:first + 1], iheads[:first + 1])
    d = h.dump()
    nsbs, nsw = h.nsbs, (h.nsbs + 63) // 64
    sets = d[3 * nsbs:3 * nsbs + M.NSETS * nsw]
    assert not sets[:nsw].any()                       # no free slab
    sid = int(M.lib().hm_size_id(int(sizes[first])))
    busy, nch = int(t[sid][7]), int(t[sid][0])
    desc = d[nsbs:3 * nsbs].reshape(nsbs, 2)
    cnt = d[:nsbs] >> np.uint64(32)
    mine = (desc[:, 0] & np.uint64(0xFFFFFFFF)) == sid
    assert mine.any()
    assert (cnt[mine] >= busy).all()                   # every slab of the class is full
    filled = float((cnt.astype(np.float64) * np.array([int(t[int(x) & 0xFFFF][11]) if (int(x) & 0xFFFFFFFF) < 16 else 0
                                                    for x in desc[:, 0]])).sum()) / (nsbs * (4 << 20))
    assert filled > 0.6, filled                        # most of the heap is in use by then


@pytest.mark.timeout(120)
def test_stale_waves_outside_the_spread_are_served_and_exhaust_cleanly():
    # waves that still use a set above the spread (read before it shrank)
    # are served from the pair their set folds onto, which is replaced as it
    # fills; allocating on such a set until the heap is exhausted ends with
    # NULLs, not a loop, and the outer pair takes no head
    opts = M.make_opts(memory=64 << 20, sb_sz_sh=21, nheads=16)
    h = M.ModelHeap(opts)
    ops = np.zeros((16, 64, 2), dtype=np.uint32)
    ops[:, :, 0] = 1
    ops[:, :, 1] = np.array(BLOCK_SIZES)[:, None]
    h.replay(ops, np.zeros(16, dtype=np.uint32))
    sp, _ = _spread_and_seen(h)
    assert sp == 1                                   # 24 slabs, 16 classes
    got = nul = 0
    for _ in range(20000):
        p = h.wave_raw(13, np.full(64, 48, dtype=np.uint32))
        got += int((p != 0).sum())
        nul += int((p == 0).sum())
        if nul:
            break
    # the other 15 classes keep a head (and a cached head) each on this
    # 24-slab heap; the 48 B class fills the rest
    assert nul > 0 and got * 48 > 0.3 * 24 * (2 << 20)
    d = h.dump()
    nsw = (h.nsbs + 63) // 64
    hb = 3 * h.nsbs + M.NSETS * nsw
    assert (d[hb + 13 * 16:hb + 14 * 16] == 0xFFFFFFFF).all()
    assert M.check_dump(opts, d)[0] == 0


def test_adaptive_sets_limited_to_what_one_class_can_hold():
    # 96 slabs hold heads for at most 32 sets (3 * 32 <= 96): asked for 64
    # adaptive sets, the heap runs 32, so a program gives the same records and
    # the same state as at 32 sets. Counting the 32 sets that no wave can use,
    # free slabs looked scarce with 32 of them free, and pairs shared heads.
    # Fixed sets are cut the same way; 24 slabs keep 8 fixed sets.
    ops, ih = class_switch_program()
    runs = {}
    for nh in (32, 64):
        h = M.ModelHeap(M.make_opts(nheads=nh))
        recs, _ = h.replay(ops, ih)
        runs[nh] = (recs, h.dump())
        assert M.check_dump(M.make_opts(nheads=nh), runs[nh][1])[0] == 0
    assert (runs[32][0] == runs[64][0]).all()
    assert len(runs[32][1]) == len(runs[64][1]) and (runs[32][1] == runs[64][1]).all()
    fixed = M.ModelHeap(M.make_opts(nheads=64, fixed_sets=1)).dump()
    assert len(fixed) == len(runs[32][1])
    small = dict(memory=64 << 20, sb_sz_sh=21, fixed_sets=1)   # 24 slabs
    d8 = M.ModelHeap(M.make_opts(nheads=8, **small)).dump()
    d1 = M.ModelHeap(M.make_opts(nheads=1, **small)).dump()
    d16 = M.ModelHeap(M.make_opts(nheads=16, **small)).dump()
    assert len(d8) - len(d1) == 7 * 16 * 3 and len(d16) == len(d8)   # 3 dump words per (set, class)


def test_class_switch_frees_idle_heads_instead_of_sharing():
    # 32 sets on 96 slabs: 64 B blocks on every set leave 32 heads and 32
    # cached heads of that class idle once they are freed.  The 16 B heads
    # that follow (spread 16 for two classes) come from those idle slabs, one
    # slab per pair; sharing one head among the pairs instead (what a scarce
    # free set led to) puts 16 ticket sequences into one slab, whose crowded
    # head is then replaced by sharing again
    opts = M.make_opts(nheads=32)
    h = M.ModelHeap(opts)
    ops, ih = class_switch_program()
    recs, _ = h.replay(ops, ih)
    assert (recs[:, :, 0] != M.REC_NULL).all()
    heads, _ = head_words(h.dump(), h.nsbs, 32)
    col = heads[:, 0][heads[:, 0] != 0xFFFFFFFF] & 0x7FFFFFFF
    assert len(col) == 16 and len(set(col.tolist())) == 16
    assert M.check_dump(opts, h.dump())[0] == 0
