"""GPU parity of the MI355X allocator against the host model (oracle/).

Deterministic mode: one wavefront replays a program of mallocs and frees; the
records (size class, slab, chunk) and the whole allocator state must equal the
model's bit for bit.  Concurrent mode: full-grid workloads must satisfy the
reference test suite's invariants (tst/common/common.cu:412-529,
tst/corr/freeslabs/freeslabs.cu:35-71), checked on the dumped state.
"""
import dataclasses

import numpy as np
import pytest

import halloc_amd as H
from oracle import model as M
from tests.helpers import (class_switch_program, crowded_program, fill_program, head_words, mixed_sizes_tail,
                           BLOCK_SIZES, LARGE_SB, REC_NULL, assert_disjoint, graph_sizes, large_class_of, large_state,
                           large_units, live_intervals, pressure_program, random_batch_ops, random_ops, sweep_sizes,
                           wave_refree_ops)

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def model_opts(o: H.HallocOpts):
    return M.make_opts(memory=o.memory, halloc_fraction=o.halloc_fraction, busy_fraction=o.busy_fraction,
                       roomy_fraction=o.roomy_fraction, sparse_fraction=o.sparse_fraction,
                       sb_sz_sh=o.sb_sz_sh, nheads=o.nheads, hash_mode=1, hash_step_frac=o.hash_step_frac,
                       fixed_sets=o.fixed_sets)


@pytest.fixture
def halloc():
    started = []

    def start(opts):
        H.Halloc.init(opts)
        started.append(opts)
        return opts

    yield start
    if started:
        H.Halloc.shutdown()


def replay_both(opts, ops, iheads=None):
    """Replay ops on the GPU, then on the model with the GPU's head set per step."""
    recs_g, xcc = H.Halloc.replay(ops, iheads)
    heap = M.ModelHeap(model_opts(opts))
    recs_m, _ = heap.replay(ops, iheads=xcc % opts.nheads)
    return recs_g, recs_m, heap


def assert_same_state(heap, opts):
    dg = H.Halloc.dump()
    dm = heap.dump()
    assert dg.shape == dm.shape
    diff = np.nonzero(dg != dm)[0]
    assert diff.size == 0, f"{diff.size} state words differ, first at {diff[:8]}"
    tot, det = M.check_dump(model_opts(opts), dg)
    assert tot == 0, det


@pytest.mark.parametrize("frac", [0.0, 0.05])
def test_geometry_matches_model(halloc, frac):
    with pytest.raises(H.HallocError):
        H.Halloc.init(H.HallocOpts(cached_arena=2))
    opts = halloc(H.HallocOpts(hash_step_frac=frac))
    g = H.Halloc.geometry()
    say(f"geometry nsbs={g['nsbs']} nheads={g['nheads']} hash_step_frac={frac}")
    t = M.size_table(model_opts(opts))
    assert (g["size_table"] == t).all()
    assert g["nsbs"] == int(0.75 * (512 << 20)) // (4 << 20)


def test_replay_wave_tuned_stride(halloc):
    say("replay one wave, hash_step_frac 0.11")
    opts = halloc(H.HallocOpts(memory=512 << 20, sb_sz_sh=22, nheads=8, hash_step_frac=0.11))
    rng = np.random.default_rng(17)
    ops = random_ops(rng, 200, 64, BLOCK_SIZES + [20, 40, 200, 700, 2500])
    recs_g, recs_m, heap = replay_both(opts, ops)
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)


@pytest.mark.parametrize("seed", [0, 1])
def test_replay_single_lane(halloc, seed):
    say(f"replay single lane seed={seed}")
    opts = halloc(H.HallocOpts(memory=256 << 20, sb_sz_sh=21, nheads=1))
    rng = np.random.default_rng(100 + seed)
    ops = random_ops(rng, 3000, 1, BLOCK_SIZES + [0, 1, 17, 100, 1000, 3000])
    recs_g, recs_m, heap = replay_both(opts, ops)
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)


@pytest.mark.parametrize("nheads", [1, 8])
def test_replay_wave(halloc, nheads):
    say(f"replay one wave nheads={nheads}")
    opts = halloc(H.HallocOpts(memory=512 << 20, sb_sz_sh=21, nheads=nheads))
    rng = np.random.default_rng(7 + nheads)
    ops = random_ops(rng, 200, 64, BLOCK_SIZES + [20, 40, 200, 700, 2500])
    recs_g, recs_m, heap = replay_both(opts, ops)
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)
    table = M.size_table(model_opts(opts))
    assert_disjoint(live_intervals(ops, recs_g, table), 1 << 21)


def test_replay_full_waves_of_one_class(halloc):
    # the plain workload's shape: whole waves of 16 B, then whole waves freed
    say("replay full waves 16 B")
    opts = halloc(H.HallocOpts(memory=64 << 20, sb_sz_sh=21, nheads=1))
    n = 400
    ops = np.zeros((2 * n, 64, 2), dtype=np.uint32)
    ops[:n, :, 0] = 1
    ops[:n, :, 1] = 16
    ops[n:, :, 0] = 2
    ops[n:, :, 1] = np.arange(n * 64).reshape(n, 64)
    recs_g, recs_m, heap = replay_both(opts, ops)
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)


def test_replay_pressure_exhaustion_and_reuse(halloc):
    # 6 slabs: exhaustion (NULL), head replacement, busy-set heads, slabs freed
    # and taken over by other classes
    say("replay pressure")
    opts = halloc(H.HallocOpts(memory=16 << 20, sb_sz_sh=21, nheads=1))
    ops = pressure_program()
    recs_g, recs_m, heap = replay_both(opts, ops)
    nnull = int((recs_g[:2500, :, 0][ops[:2500, :, 0] == 1] == REC_NULL).sum())
    say(f"pressure nulls={nnull}")
    assert nnull > 0
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)
    phase1 = set(recs_g[:2500, :, 1][ops[:2500, :, 0] == 1].tolist()) - {REC_NULL}
    phase3 = set(recs_g[5500:, :, 1][ops[5500:, :, 0] == 1].tolist()) - {REC_NULL}
    assert phase3 and phase3 <= phase1


@pytest.mark.parametrize("batch", [False, True])
def test_replay_mixed_sizes_two_run_free(halloc, batch):
    # a slab keeps two 16 B blocks at chunks off the multiples of 8 and is
    # then taken from the sparse set as the 128 B head (8 chunks a block);
    # one wave frees both from two lanes: the two-run free path, whose run
    # starts are then not multiples of the slab's block length, must clear
    # both blocks' lengths (bit-exact state against the model)
    say(f"replay mixed sizes two-run free batch={batch}")
    opts = halloc(H.HallocOpts(memory=32 << 20, sb_sz_sh=21, nheads=1, busy_fraction=0.1, roomy_fraction=0.05))
    fill = fill_program(16, 260)
    heap = M.ModelHeap(model_opts(opts))
    if not batch:
        plan, _ = M.ModelHeap(model_opts(opts)).replay(fill)
        tail, sb, big_step = mixed_sizes_tail(plan)
        ops = np.concatenate([fill, tail])
        recs_g, _ = H.Halloc.replay(ops)
        recs_m, _ = heap.replay(ops)
        big = recs_g[big_step, 0]
    else:
        def as_batch(o):
            b = np.zeros(o.shape[:2] + (4,), dtype=np.uint32)
            b[..., :2] = o
            b[..., 2] = o[..., 0] == 1
            return b
        plan = M.ModelHeap(model_opts(opts)).replay_batch(as_batch(fill))
        tail, sb, big_step = mixed_sizes_tail(plan[:, :, 0])
        ops = as_batch(np.concatenate([fill, tail]))
        recs_g, _ = H.Halloc.replay_batch(ops)
        recs_m = heap.replay_batch(ops)
        big = recs_g[big_step, 0, 0]
    # the 128 B block went to the slab that holds the two 16 B blocks
    assert big[0] == 6 and big[1] == sb
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)


def test_replay_class_switch_takes_idle_slabs(halloc):
    # 32 sets on 96 slabs (the defaults): 64 B blocks on every set, all freed,
    # then 16 B blocks: the 16 B heads come from the slabs the 64 B class holds
    # idle, one per pair (bit-exact against the model)
    say("replay class switch, 32 sets")
    opts = halloc(H.HallocOpts(memory=512 << 20, nheads=32))
    ops, ih = class_switch_program()
    recs_g, recs_m, heap = replay_both(opts, ops, ih)
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)
    heads, _ = head_words(H.Halloc.dump(), H.Halloc.geometry()["nsbs"], 32)
    col = heads[:, 0][heads[:, 0] != 0xFFFFFFFF] & 0x7FFFFFFF
    assert len(col) == 16 and len(set(col.tolist())) == 16


def test_plain_kernel_with_more_sets_than_the_heap_holds(halloc):
    # 64 sets asked for on 96 slabs: the heap runs the 32 that one class can
    # hold heads in, and the 16 B kernel runs as fast as at 32 sets. Counting
    # the 32 sets no wave could use, free slabs looked scarce, pairs shared
    # heads, and launches took 4-6x as long (adaptive sets) or 86-362 ms
    # (fixed sets)
    say("plain 16 B at 64 sets on 512 MiB")
    us = {}
    for nh, fixed in ((32, 0), (64, 0), (64, 1)):
        halloc(H.HallocOpts(memory=512 << 20, nheads=nh, fixed_sets=fixed))
        assert H.Halloc.geometry()["nheads"] == 32
        p = H.Halloc.run_plain(nthreads=1 << 20, bs=256, size=16, iters=1, warmup=10, launches=10)
        assert p["nulls"] == 0 and p["errors"] == 0
        us[(nh, fixed)] = p["total_ms"] / 10 * 1e3
    say(f"us per launch: {us}")
    assert us[(64, 0)] < 2 * us[(32, 0)] and us[(64, 1)] < 2 * us[(32, 0)]


def test_plain_kernel_after_a_class_switch(halloc):
    # the plain 16 B kernel on a heap whose every set first served 64 B
    # blocks (caller kernel, all freed): about 50 us per launch, as on a fresh
    # heap; when the 16 B pairs shared one head instead, launches took 9-27 ms
    say("plain 16 B after 64 B blocks, library defaults")
    opts = halloc(H.HallocOpts(memory=512 << 20))
    t = H.Halloc.run_touch(1 << 18, 8, 4, 1)
    assert t["nulls"] == 0
    p = H.Halloc.run_plain(warmup=2, launches=5)
    say(f"ms per launch: {p['total_ms'] / 5:.3f} timed, {[round(float(x), 3) for x in p['ms']]} one by one")
    assert p["nulls"] == 0 and p["errors"] == 0
    assert p["total_ms"] / 5 < 1.0 and float(np.max(p["ms"])) < 1.0
    # a third class on top, then the whole state: every invariant holds
    t = H.Halloc.run_touch(1 << 17, 32, 2, 1)
    assert t["nulls"] == 0
    assert H.Halloc.free_bytes() == H.Halloc.geometry()["arena_bytes"]
    assert M.check_dump(model_opts(opts), H.Halloc.dump())[0] == 0


def test_replay_crowded_head_replaced_below_its_limit(halloc):
    # a head kept at about 78% fill by churn: a group whose probes need 12
    # failed rounds has its head replaced though the count is below the limit
    say("replay crowded head")
    opts = halloc(H.HallocOpts(memory=32 << 20, sb_sz_sh=21, nheads=1))
    ops = crowded_program()
    mh = M.ModelHeap(model_opts(opts))
    mh.replay(ops)
    assert mh.crowded() >= 1
    recs_g, recs_m, heap = replay_both(opts, ops)
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)
    table = M.size_table(model_opts(opts))
    assert_disjoint(live_intervals(ops, recs_g, table), 1 << 21)


def test_replay_shared_and_stolen_heads(halloc):
    # mallocs spread over 8 head sets in a 96-slab pool: pairs without a slab
    # share another set's head of their class; later size sweeps take over
    # idle heads of old classes
    say("replay shared heads")
    opts = halloc(H.HallocOpts(nheads=8, fixed_sets=1))
    rng = np.random.default_rng(5)
    nw = 600
    ops = np.zeros((2 * nw, 64, 2), dtype=np.uint32)
    ops[:nw, :, 0] = 1
    ops[:nw, :, 1] = rng.integers(1, 3073, (nw, 64))
    ops[nw:, :, 0] = 2
    ops[nw:, :, 1] = rng.permutation(nw * 64).reshape(nw, 64)
    iheads = np.arange(2 * nw, dtype=np.uint32) % 8
    recs_g, recs_m, heap = replay_both(opts, ops, iheads)
    assert (recs_g[:nw, :, 0] != REC_NULL).all()
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)
    d = H.Halloc.dump()
    nsbs = H.Halloc.geometry()["nsbs"]
    assert ((d[:nsbs] >> np.uint64(16)) & np.uint64(0xFF)).max() > 1


def large_program(rng, nsteps, lanes):
    """Mixed program over both paths: mallocs of slab and large sizes (up to
    4 MiB blocks), frees, and frees of interior pointers of live large blocks
    (op kind 3, 4 KiB into the block).  Then, with no malloc in between, every
    live block is freed, freed again and its interior pointer freed: the
    second and third free of each block change nothing."""
    sizes = BLOCK_SIZES[::3] + [3073, 4096, 5000, 9000, 17000, 40000, 65536, 70000, 300000, 1 << 20, 4 << 20]
    ops = np.zeros((nsteps, lanes, 2), dtype=np.uint32)
    live = []
    for st in range(nsteps):
        new = []
        for l in range(lanes):
            u = rng.random()
            if live and u < 0.3:
                ops[st, l] = (2, live.pop(int(rng.integers(len(live)))))
            elif live and u < 0.38:
                big = [j for j in live if ops.reshape(-1, 2)[j][1] > 4096]
                if big:
                    ops[st, l] = (3, big[int(rng.integers(len(big)))])
            else:
                ops[st, l] = (1, int(rng.choice(sizes)))
                new.append(st * lanes + l)
        live.extend(new)
    rounds = -(-len(live) // lanes)
    tail = np.zeros((3 * rounds, lanes, 2), dtype=np.uint32)
    for kind_i, kind in enumerate((2, 2, 3)):
        for j, op in enumerate(live):
            tail[kind_i * rounds + j // lanes, j % lanes] = (kind, op)
    return np.concatenate([ops, tail]), nsteps + rounds


def spread_word(g, d):
    """(sets in use, classes seen) from a dump."""
    lw = (g["large_units"] + 63) // 64
    b = len(d) - g["large_units"] - lw - 40 - 2
    return int(d[b]), int(d[b + 1])


@pytest.mark.parametrize("nheads", [16, 8])
def test_replay_adaptive_sets(halloc, nheads):
    # the wave's own head set (XCD and CU), folded onto the sets in use as
    # size classes appear: one class first, then all of them; records, the
    # spread and the whole state equal the model's
    say(f"replay adaptive sets nheads={nheads}")
    opts = halloc(H.HallocOpts(memory=512 << 20, nheads=nheads))
    rng = np.random.default_rng(70 + nheads)
    one = np.zeros((40, 64, 2), dtype=np.uint32)
    one[:, :, 0] = 1
    one[:, :, 1] = 16
    mixed = random_ops(rng, 200, 64, BLOCK_SIZES + [20, 40, 200, 700, 2500])
    mixed[mixed[..., 0] == 2, 1] += 40 * 64   # frees name ops of the mixed part
    ops = np.concatenate([one, mixed])
    recs_g, recs_m, heap = replay_both(opts, ops)
    g = H.Halloc.geometry()
    sp, seen = spread_word(g, H.Halloc.dump())
    say(f"spread {sp}, classes seen {bin(seen).count('1')}")
    assert sp < nheads and bin(seen).count("1") == 16
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)


def test_adaptive_sets_many_classes_default_options(halloc):
    # the drop-in defaults (32 sets, adaptive) at 512 MiB with requests over
    # all 16 classes, in kernels of 20 000 until one returns NULL: the sets
    # in use have folded to 2, the heap then serves more than half of its
    # slab bytes in requested bytes, and no slab is free. The requests of one
    # kernel run concurrently, so the fill is what the heap holds after the
    # kernel with the first NULL (the requests it served included), not the
    # bytes of whole kernels before it
    say("adaptive sets: mixed classes until exhaustion, default options")
    opts = halloc(H.HallocOpts())
    g = H.Halloc.geometry()
    rng = np.random.default_rng(31)
    held, total = [], 0
    for rnd in range(200):
        sizes = rng.integers(8, 3073, 20000).astype(np.uint32)
        p = H.Halloc.malloc_sizes(sizes, bs=128)
        held.append(p)
        total += int(sizes[p != 0].sum())
        if (p == 0).any():
            break
    say(f"first NULL in kernel {rnd} ({int((p == 0).sum())} of 20000): the heap holds "
        f"{total / g['arena_bytes']:.3f} of its slab bytes in requested bytes")
    assert (p == 0).any() and total > 0.5 * g["arena_bytes"]
    d = H.Halloc.dump()
    nsw = g["nsw"]
    assert not d[3 * g["nsbs"]:3 * g["nsbs"] + nsw].any()          # no free slab left
    assert spread_word(g, d)[0] == 2
    assert M.check_dump(model_opts(opts), d)[0] == 0
    allp = np.concatenate(held)
    H.Halloc.free_ptrs(allp[allp != 0])
    assert H.Halloc.free_bytes() == g["arena_bytes"]
    d = H.Halloc.dump()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()


@pytest.mark.parametrize("lanes", [8, 64])
def test_replay_large_requests(halloc, lanes):
    # requests above 3072 B from the large region, replayed bit-exactly with
    # their (class, unit) records and the whole state, including interior
    # pointers and double frees (both no-ops)
    say(f"replay large lanes={lanes}")
    opts = halloc(H.HallocOpts(memory=256 << 20, sb_sz_sh=21, nheads=1))
    rng = np.random.default_rng(11 + lanes)
    ops, refree = large_program(rng, 300 if lanes == 8 else 60, lanes)
    recs_g, recs_m, heap = replay_both(opts, ops)
    big = (ops[..., 0] == 1) & (recs_g[..., 1] == LARGE_SB)
    say(f"large blocks={int(big.sum())} interior frees={int((ops[..., 0] == 3).sum())}")
    assert big.sum() > 50
    assert (recs_g[ops[..., 0] == 3][:, 0] == 0).all()   # an interior pointer frees nothing
    assert (recs_g[refree:][ops[refree:, :, 0] != 0][:, 0] == 0).all()  # nor a second free
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)
    # everything came back
    g = H.Halloc.geometry()
    d = H.Halloc.dump()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()
    _, lbits, lstart = large_state(d, g["large_units"])
    assert not lstart.any() and sum(bin(int(w)).count("1") for w in lbits) == len(lbits) * 64 - g["large_units"]


def test_replay_large_random_sizes_one_lane(halloc):
    say("replay large one lane")
    opts = halloc(H.HallocOpts(memory=64 << 20, sb_sz_sh=21, nheads=1))
    rng = np.random.default_rng(99)
    ops = random_ops(rng, 3000, 1, BLOCK_SIZES, large_every=3)
    recs_g, recs_m, heap = replay_both(opts, ops)
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)


@pytest.mark.parametrize("cached", [0, 1])
def test_plain_workload_invariants(halloc, cached):
    say(f"plain 2^18 lanes x 2 iters cached_arena={cached}")
    opts = halloc(H.HallocOpts(cached_arena=cached))
    r = H.Halloc.run_plain(nthreads=1 << 18, bs=256, size=16, iters=2, warmup=1, launches=3)
    assert r["errors"] == 0 and r["nulls"] == 0
    d = H.Halloc.dump()
    tot, det = M.check_dump(model_opts(opts), d)
    assert tot == 0, det
    g = H.Halloc.geometry()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()


def test_alloc_check_and_persistence(halloc):
    # check_alloc (tst/common/common.cu:468-529): no NULL, sorted blocks do not
    # overlap; each block keeps its first word across kernel launches
    say("alloc check 2^16 mixed sizes")
    opts = halloc(H.HallocOpts())
    rng = np.random.default_rng(21)
    sizes = rng.integers(1, 3073, 1 << 16).astype(np.uint32)
    ptrs = H.Halloc.malloc_sizes(sizes)
    assert (ptrs != 0).all()
    order = np.argsort(ptrs)
    p, s = ptrs[order], sizes[order]
    assert (p[:-1] + s[:-1] <= p[1:]).all()
    align = np.where(np.isin(np.array(BLOCK_SIZES)[np.searchsorted(BLOCK_SIZES, np.maximum(s, 16))] % 3, [0]), 8, 16)
    assert (p % align.astype(np.uint64) == 0).all()
    assert (H.Halloc.read_words(ptrs) == sizes).all()
    d = H.Halloc.dump()
    tot, det = M.check_dump(model_opts(opts), d)
    assert tot == 0, det
    H.Halloc.free_ptrs(ptrs)
    d = H.Halloc.dump()
    g = H.Halloc.geometry()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()
    assert M.check_dump(model_opts(opts), d)[0] == 0


def test_freeslabs_reuse_across_classes(halloc):
    # tst/corr/freeslabs/freeslabs.cu:35-71 with its 512 MiB pool and default
    # options: 40% of memory at each size 16..1024 in turn, allocated and freed
    say("freeslabs")
    opts = halloc(H.HallocOpts())
    size, step = 16, 8
    while size <= 1024:
        n = int(0.4 * opts.memory / size)
        ptrs = H.Halloc.malloc_sizes(np.full(n, size, dtype=np.uint32))
        assert (ptrs != 0).all(), size
        H.Halloc.free_ptrs(ptrs)
        size += step
        step = 128 if size >= 256 else 16 if size >= 64 else 8
    d = H.Halloc.dump()
    assert M.check_dump(model_opts(opts), d)[0] == 0


def expected_extfrag(g, d, ideal):
    # ha_extfrag (src/statistics.cuh:50-69): 1 - max_alloc / free_mem, from
    # the dump alone.  free_mem: per slab, its bytes less its count x chunk
    # size, plus the whole large region (the reference adds its device heap);
    # max_alloc: the largest 2^c units whose aligned run of unit bits is all
    # clear (+ the bytes of the free slabs when ideal)
    nsbs, sb = g["nsbs"], g["sb_sz"]
    count = (d[:nsbs] >> np.uint64(32)).astype(np.int64)
    cs = (d[nsbs:3 * nsbs].reshape(nsbs, 2)[:, 1] & np.uint64(0xFFFFFFFF)).astype(np.int64)
    chunk = np.where(cs != 0, (1 << (cs & 0xFFFF)) * (cs >> 16), 0)  # chunk size: (multiplier << 16) | shift
    free_mem = int((sb - count * chunk).sum()) + g["large_bytes"]
    free_slabs = int(sb * (cs == 0).sum())
    _, lbits, _ = large_state(d, g["large_units"])
    used = np.unpackbits(np.asarray(lbits, dtype="<u8").view(np.uint8), bitorder="little")[:g["large_units"]]
    best, k = 0, 1
    while k <= g["large_units"]:
        runs = used[:g["large_units"] // k * k].reshape(-1, k)
        if not (~runs.any(axis=1)).any():
            break
        best, k = k, 2 * k
    max_alloc = best * 4096 + (free_slabs if ideal else 0)
    return 1.0 - max_alloc / free_mem


def test_extfrag_and_free_bytes(halloc):
    # ha_extfrag against the same formula evaluated on the dumped state, on an
    # empty heap (0.75 and 0: the large region is a quarter of the pool and the
    # slabs all free), after one 4 KiB block (which halves the largest aligned
    # run: 0.875 and 0.125 exactly), and after 2^16 blocks of 64 B
    say("extfrag")
    halloc(H.HallocOpts())
    g = H.Halloc.geometry()
    assert H.Halloc.free_bytes() == g["arena_bytes"]
    assert abs(H.Halloc.extfrag(False) - 0.75) < 1e-12 and abs(H.Halloc.extfrag(True)) < 1e-12
    big = H.Halloc.malloc_sizes(np.array([4096], dtype=np.uint32))
    assert big[0] != 0
    assert abs(H.Halloc.extfrag(False) - 0.875) < 1e-12 and abs(H.Halloc.extfrag(True) - 0.125) < 1e-12
    ptrs = H.Halloc.malloc_sizes(np.full(1 << 16, 64, dtype=np.uint32))
    assert H.Halloc.free_bytes() == g["arena_bytes"] - 64 * (1 << 16)
    d = H.Halloc.dump()
    for ideal in (False, True):
        assert abs(H.Halloc.extfrag(ideal) - expected_extfrag(g, d, ideal)) < 1e-12
    H.Halloc.free_ptrs(ptrs)
    H.Halloc.free_ptrs(big)
    assert abs(H.Halloc.extfrag(False) - 0.75) < 1e-12


def test_size_sweep_lists_random_free_order(halloc):
    # size-class sweep shape: U{1..8} requests per lane, log-uniform 16 B..3 KiB
    # plus 1% in (3 KiB, 64 KiB] on the device heap; frees in random order
    say("size sweep 2^16 lanes")
    opts = halloc(H.HallocOpts(memory=1 << 30))
    rng = np.random.default_rng(3)
    sizes = sweep_sizes(rng, 1 << 16)
    order = rng.permutation(sizes.size).astype(np.uint32)
    r = H.Halloc.run_lists(sizes, order, want_ptrs=True)
    assert r["nulls"] == 0 and r["errors"] == 0
    ptrs, sz = r["ptrs"].ravel(), sizes.ravel()
    live = sz != 0
    p, s = ptrs[live], sz[live].astype(np.uint64)
    o = np.argsort(p)
    assert (p[o][:-1] + s[o][:-1] <= p[o][1:]).all()
    g = H.Halloc.geometry()
    d = H.Halloc.dump()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()
    assert M.check_dump(model_opts(opts), d)[0] == 0


def test_mixed_classes_served_with_heap_sized_head_sets(halloc):
    # 132 000 requests uniform on 8..3072 B (all 16 classes, about half the
    # slab bytes of a 512 MiB heap): with the head sets sized to the heap
    # (nheads 0 -> 2 sets) every request is served, blocks are disjoint, and
    # freeing them returns the whole heap
    say("mixed classes 8..3072 B at 512 MiB, nheads 0")
    opts = halloc(H.HallocOpts(memory=512 << 20, nheads=0))
    g = H.Halloc.geometry()
    assert g["nsbs"] == 96 and g["nheads"] == 2
    rng = np.random.default_rng(0)
    sizes = rng.integers(8, 3073, 132000).astype(np.uint32)
    ptrs = H.Halloc.malloc_sizes(sizes, bs=128)
    assert (ptrs != 0).all(), int((ptrs == 0).sum())
    o = np.argsort(ptrs)
    p, s = ptrs[o], sizes[o].astype(np.uint64)
    assert (p[:-1] + s[:-1] <= p[1:]).all()
    H.Halloc.free_ptrs(ptrs)
    assert H.Halloc.free_bytes() == g["arena_bytes"]
    d = H.Halloc.dump()
    assert M.check_dump(model_opts(dataclasses.replace(opts, nheads=g["nheads"])), d)[0] == 0


def test_phase_workload_invariants(halloc):
    # 48 slabs of 2 MiB for the 9 classes of 16..256 B
    say("phase 2^16 lanes x 4 slots")
    opts = halloc(H.HallocOpts(memory=128 << 20, sb_sz_sh=21))
    r = H.Halloc.run_phase(nthreads=1 << 16, slots=8, nphases=3, max_steps=60)
    say(f"phase steps={r['phase_steps']} nulls={r['nulls']}")
    assert r["errors"] == 0
    assert r["mallocs"] == r["frees"]
    assert r["trace"][:, 1].max() > 0.5
    g = H.Halloc.geometry()
    d = H.Halloc.dump()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()
    assert M.check_dump(model_opts(opts), d)[0] == 0


@pytest.mark.parametrize("nheads", [1, 8])
def test_replay_batch(halloc, nheads):
    # hamalloc_batch / hafree_batch on one wave, bit-exact against the model
    say(f"replay batch nheads={nheads}")
    opts = halloc(H.HallocOpts(memory=256 << 20, sb_sz_sh=21, nheads=nheads))
    rng = np.random.default_rng(40 + nheads)
    ops = random_batch_ops(rng, 120, 64, BLOCK_SIZES + [20, 100, 700])
    recs_g, xcc = H.Halloc.replay_batch(ops)
    heap = M.ModelHeap(model_opts(opts))
    recs_m = heap.replay_batch(ops, xcc % nheads)
    assert (recs_g == recs_m).all()
    assert_same_state(heap, opts)


@pytest.mark.parametrize("nheads", [1, 8])
def test_replay_batch_runs_and_refree(halloc, nheads):
    # every round of these frees is one slab's runs (the batched run path);
    # the second free of the same blocks finds every length 0 and frees
    # nothing; bit-exact against the model for block lengths of 1, 2, 4 and
    # 8 chunks
    say(f"replay batch runs nheads={nheads}")
    opts = halloc(H.HallocOpts(memory=256 << 20, sb_sz_sh=21, nheads=nheads))
    ops = wave_refree_ops(64, (16, 48, 96, 192, 24, 700))
    recs_g, xcc = H.Halloc.replay_batch(ops)
    heap = M.ModelHeap(model_opts(opts))
    recs_m = heap.replay_batch(ops, xcc % nheads)
    assert (recs_g == recs_m).all()
    # the double frees freed nothing: the heap ends empty and consistent
    assert_same_state(heap, opts)
    g = H.Halloc.geometry()
    d = H.Halloc.dump()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()


@pytest.mark.parametrize("batch", [4, 16])
def test_graph_builds_plain_and_batched(halloc, batch):
    # both forms build, check and free the same graphs; the heap ends empty
    say(f"graphs batch={batch}")
    opts = halloc(H.HallocOpts(memory=1 << 30))
    sizes = graph_sizes(np.random.default_rng(0), 64)
    r = H.Halloc.run_graphs(sizes, batch=batch, waves=256, plain_chunk_nodes=1 << 20)
    say(f"graphs nodes={r['nodes']} plain={r['plain_ms']:.2f} ms batch={r['batch_ms']:.2f} ms")
    assert r["nulls"] == 0 and r["errors"] == 0
    g = H.Halloc.geometry()
    d = H.Halloc.dump()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()
    assert M.check_dump(model_opts(opts), d)[0] == 0


@pytest.mark.parametrize("batch", [4, 16])
def test_graph_forms_check_each_other(halloc, batch):
    # the batched kernel's graphs pass the plain kernel's per-node check and
    # the plain kernel's graphs pass the batched kernel's class-major check:
    # the two forms agree on every node's header and adjacency contents
    say(f"graph cross check batch={batch}")
    halloc(H.HallocOpts(memory=1 << 30))
    nn = graph_sizes(np.random.default_rng(8), 40)
    r = H.Halloc.run_graphs(nn, batch=batch, waves=512, cross=True)
    assert r["nulls"] == 0 and r["errors"] == 0, r
    g = H.Halloc.geometry()
    d = H.Halloc.dump()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all()  # every block came back


def test_autotune_small_space():
    # the tuner's candidates run the batched graph form; every accepted one is
    # violation-free and the best is no slower than the start
    from halloc_amd import autotune
    say("autotune small space")
    nn = graph_sizes(np.random.default_rng(9), 24)
    space = {"batch": [4, 16], "nheads": [1, 8], "hash_step_frac": [0.0, 0.05]}
    best, ms, hist = autotune.tune(nn, memory=1 << 30, sweeps=1, log=say, space=space)
    assert np.isfinite(ms) and ms <= hist[0][1]
    # one sweep tries every value that differs from the configuration it has reached
    assert 1 + 3 <= len(hist) <= 1 + sum(len(v) for v in space.values())
    assert best["batch"] in (4, 8, 16)
    # the plain form is tuned on its own, over allocator options only
    pspace = {"nheads": [1, 8], "hash_step_frac": [0.0, 0.05]}
    pbest, pms, phist = autotune.tune(nn, memory=1 << 30, sweeps=1, log=say, space=pspace, form="plain")
    assert np.isfinite(pms) and pms <= phist[0][1]
    assert 1 + 2 <= len(phist) <= 1 + sum(len(v) for v in pspace.values())


def _large_block(n):
    """Bytes of the large-region block that serves a request of n bytes."""
    return 4096 << large_class_of(n)


def test_large_blocks_reuse_and_exhaustion(halloc):
    # requests above 3 KiB: blocks of 4 KiB << c aligned to their size, from
    # the region outside the slabs; a kernel that only allocates gets NULL
    # only when no aligned run of free units is left; freed units serve any
    # class again
    say("large blocks")
    opts = halloc(H.HallocOpts(memory=64 << 20))
    g = H.Halloc.geometry()
    lunits = large_units(64 << 20)
    assert g["large_bytes"] == lunits * 4096 == (64 << 20) - int(0.75 * (64 << 20))
    base = g["arena_base"] + g["arena_bytes"]
    assert g["large_units"] == lunits
    rng = np.random.default_rng(12)
    sizes = rng.integers(3073, 65537, 160).astype(np.uint32)
    p1 = H.Halloc.malloc_sizes(sizes)
    assert (p1 != 0).all()
    ext = np.array([_large_block(s) for s in sizes], dtype=np.uint64)
    assert ((p1 - np.uint64(base)) % ext == 0).all()            # aligned to the block size
    assert (H.Halloc.read_words(p1) == sizes).all()
    o = np.argsort(p1)
    assert (p1[o][:-1] + ext[o][:-1] <= p1[o][1:]).all()
    H.Halloc.free_ptrs(p1)
    # every unit as a 4 KiB block: exactly lunits succeed
    p4 = H.Halloc.malloc_sizes(np.full(lunits + 300, 4096, dtype=np.uint32))
    assert int((p4 != 0).sum()) == lunits
    H.Halloc.free_ptrs(p4)
    # the same units as 64 KiB blocks (class 4), then 1 MiB blocks (class 8)
    for n, k in ((65536, 16), (1 << 20, 256)):
        p = H.Halloc.malloc_sizes(np.full(lunits // k + 40, n, dtype=np.uint32))
        got = np.sort(p[p != 0])
        assert len(got) == lunits // k, (n, len(got))
        assert (got[:-1] + np.uint64(n) <= got[1:]).all()
        H.Halloc.free_ptrs(p)
    d = H.Halloc.dump()
    _, lbits, lstart = large_state(d, lunits)
    assert not lbits.any() and not lstart.any()
    assert M.check_dump(model_opts(opts), d)[0] == 0


def test_large_double_and_interior_frees_are_noops_under_load(halloc):
    # a grid frees half of its large blocks twice and hands interior and
    # unaligned pointers of the other half to hafree: the other half keeps its
    # contents, later blocks never overlap it, and the region's invariants hold
    say("large double and interior frees, 1 GiB")
    opts = halloc(H.HallocOpts(memory=1 << 30))
    lunits = large_units(1 << 30)
    rng = np.random.default_rng(13)
    sizes = rng.integers(3073, 300000, 500).astype(np.uint32)   # about 110 MiB of the 256 MiB region
    p = H.Halloc.malloc_sizes(sizes)
    ok = p != 0
    assert ok.all(), int((~ok).sum())
    half = np.zeros(len(p), dtype=bool)
    half[::2] = True
    gone, keep = p[half & ok], p[~half & ok]
    H.Halloc.free_ptrs(gone)
    H.Halloc.free_ptrs(gone)                                    # double free
    big = keep[sizes[~half & ok] > 4096]
    H.Halloc.free_ptrs(np.concatenate([big + np.uint64(4096), keep + np.uint64(512)]))
    assert (H.Halloc.read_words(keep) == sizes[~half & ok]).all()
    d = H.Halloc.dump()
    tot, det = M.check_dump(model_opts(opts), d)
    assert tot == 0, det
    _, lbits, lstart = large_state(d, lunits)
    assert int((lstart != 0).sum()) == len(keep)
    more = rng.integers(3073, 300000, 300).astype(np.uint32)
    q = H.Halloc.malloc_sizes(more)
    allp = np.concatenate([keep, q[q != 0]])
    alls = np.concatenate([[_large_block(s) for s in sizes[~half & ok]], [_large_block(s) for s in more[q != 0]]])
    o = np.argsort(allp)
    assert (allp[o][:-1] + alls[o][:-1].astype(np.uint64) <= allp[o][1:]).all()
    assert (H.Halloc.read_words(keep) == sizes[~half & ok]).all()
    H.Halloc.free_ptrs(np.concatenate([keep, q]))
    _, lbits, lstart = large_state(H.Halloc.dump(), lunits)
    assert not lstart.any() and sum(bin(int(w)).count("1") for w in lbits) == len(lbits) * 64 - lunits
