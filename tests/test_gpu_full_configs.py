"""The five benchmark workloads at their full sizes, each checked on the dumped
allocator state, plus the reference's own correctness programs at their own
sizes: tst/corr/test/test.cu test 0 and the whole tst/corr/run-all-tests.pl
sweep (every checkptr size from 16 B to 32 KiB, freeslabs, and the 30
prob-checkptr configurations at their own sizes), one test per entry so that
each stays short."""
import os

import numpy as np
import pytest

import halloc_amd as H
from halloc_amd import harness
from halloc_amd.workloads import graph_sizes, pages_for, serving_requests, sweep_sizes
from oracle import model as M
from tests.helpers import large_state, serve_checksums

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def model_opts(o, nheads=None):
    return M.make_opts(memory=o.memory, halloc_fraction=o.halloc_fraction, busy_fraction=o.busy_fraction,
                       roomy_fraction=o.roomy_fraction, sparse_fraction=o.sparse_fraction,
                       sb_sz_sh=o.sb_sz_sh, nheads=nheads or o.nheads, hash_mode=1,
                       hash_step_frac=o.hash_step_frac)


def assert_heap_empty_and_sound(opts):
    g = H.Halloc.geometry()
    d = H.Halloc.dump()
    assert ((d[:g["nsbs"]] >> np.uint64(32)) == 0).all(), "a slab still counts live chunks"
    assert H.Halloc.free_bytes() == g["arena_bytes"]
    tot, det = M.check_dump(model_opts(opts, g["nheads"]), d)
    assert tot == 0, det
    _, lbits, lstart = large_state(d, g["large_units"])
    assert not lstart.any()
    assert sum(bin(int(w)).count("1") for w in lbits) == len(lbits) * 64 - g["large_units"]


@pytest.fixture
def heap():
    started = []

    def start(opts):
        H.Halloc.init(opts)
        started.append(opts)
        return opts

    yield start
    if started:
        H.Halloc.shutdown()


def test_config1_plain_full(heap):
    # 2^20 lanes of 256, 512 MiB, the library defaults: 10 launches of one
    # malloc/write/verify/free per lane, then one launch of 10 iterations
    say("config 1: plain 16 B, 2^20 lanes x 10")
    opts = heap(H.HallocOpts(memory=512 << 20))
    r = H.Halloc.run_plain(nthreads=1 << 20, bs=256, size=16, iters=1, warmup=1, launches=10)
    assert r["errors"] == 0 and r["nulls"] == 0, r
    r = H.Halloc.run_plain(nthreads=1 << 20, bs=256, size=16, iters=10, warmup=0, launches=1)
    assert r["errors"] == 0 and r["nulls"] == 0, r
    assert_heap_empty_and_sound(opts)


def test_config2_phase_full(heap):
    # the reference's phase-throughput with its experiments' flags
    # (-f0.95 -F0.05 -e0.91 -g5): 2^20 threads, one block of U{16..256} B
    # each, 5 allocation and 5 free phases of 8 chain steps, 512 MiB, the
    # library defaults; every block held is non-NULL and the heap ends empty
    say("config 2: phase-throughput, 2^20 threads, 10 phases")
    opts = heap(H.HallocOpts(memory=512 << 20))
    o = harness.defaults_for("phase-throughput")
    o.parse(["-n1048576", "-s16", "-S256", "-duniform", "-l1", "-i8", "-t10", "-f0.95", "-F0.05", "-e0.91",
             "-g5", f"-m{512 << 20}"])
    r = H.Halloc.run_test(harness.TESTS["phase-throughput"], o.nthreads, o.bs, o.ntries, o.nallocs, o.niters,
                          o.alloc_sz, o.max_alloc_sz, o.max_alloc_sh, harness.DISTRS[o.distr], o.period_mask,
                          o.group_sh, o.probabs, 0)
    say(f"mallocs={r['mallocs']} frees={r['frees']} nulls={r['nulls']}")
    assert r["nulls"] == 0 and r["checks_failed"] == 0, r
    assert r["mallocs"] == r["frees"] > 4 * (1 << 20)
    assert_heap_empty_and_sound(opts)


def test_phase_fill_target_workload_full(heap):
    # the fill-target variant (run_phase): 2^20 lanes x 4 slots, 5 phases
    # between 0.91 and 0.09 of the slab bytes requested; the target lies above
    # what the busy threshold lets slabs hold, so its allocation phases end
    # in NULLs; every block is intact and the heap ends empty and sound
    say("phase fill target, 2^20 lanes, 5 phases")
    opts = heap(H.HallocOpts(memory=512 << 20))
    r = H.Halloc.run_phase(nthreads=1 << 20, slots=4, nphases=5, max_steps=400)
    say(f"phase steps={r['phase_steps']} nulls={r['nulls']} peak fill={r['trace'][:, 1].max():.3f}")
    assert r["errors"] == 0
    assert r["mallocs"] == r["frees"]
    assert len([s for s in r["phase_steps"] if s]) == 5
    assert_heap_empty_and_sound(opts)


def test_config3_size_sweep_full(heap):
    # 2^19 lanes in 8 rounds (each round's live set fits the 1 GiB pool),
    # U{1..8} requests log-uniform on [16 B, 3 KiB] plus 1% on (3 KiB, 64 KiB],
    # frees in random order; every round's blocks disjoint and intact
    say("config 3: size sweep, 2^19 lanes, 1 GiB")
    opts = heap(H.HallocOpts(memory=1 << 30, nheads=0))
    rng = np.random.default_rng(0)
    nlarge = 0
    for rnd in range(8):
        sizes = sweep_sizes(rng, (1 << 19) // 8)
        order = rng.permutation(sizes.size).astype(np.uint32)
        r = H.Halloc.run_lists(sizes, order, want_ptrs=True)
        assert r["nulls"] == 0 and r["errors"] == 0, (rnd, r)
        p, s = r["ptrs"].ravel(), sizes.ravel()
        live = s != 0
        p, s = p[live], s[live].astype(np.uint64)
        ext = np.where(s > 3072, np.uint64(4096) << np.array([max(0, int(-(-int(x) // 4096) - 1).bit_length())
                                                             for x in s], dtype=np.uint64), s)
        o = np.argsort(p)
        assert (p[o][:-1] + ext[o][:-1] <= p[o][1:]).all(), rnd
        nlarge += int((s > 3072).sum())
    say(f"large requests {nlarge}")
    assert nlarge > 20000
    assert_heap_empty_and_sound(opts)


def test_config4_batching_full(heap):
    # 4096 random-graph builds on 4 GiB: both forms build, check and free the
    # same graphs, then each form's graphs are checked and freed by the other
    say("config 4: 4096 graphs, 4 GiB, both forms and the cross check")
    opts = heap(H.HallocOpts(memory=4 << 30))
    nn = graph_sizes(np.random.default_rng(0), 4096)
    r = H.Halloc.run_graphs(nn, batch=16, waves=4096, repeats=1)
    say(f"nodes={r['nodes']} plain={r['plain_ms']:.2f} ms batch={r['batch_ms']:.2f} ms")
    assert r["nulls"] == 0 and r["errors"] == 0, r
    assert r["nodes"] == int(nn.sum())
    # the cross check holds every graph of one form at once: 1024 graphs fit
    c = H.Halloc.run_graphs(nn[:1024], batch=16, waves=4096, cross=True)
    assert c["nulls"] == 0 and c["errors"] == 0, c
    assert_heap_empty_and_sound(opts)


def test_config5_serving_full(heap):
    # 16 GiB pool, 65536 slots, 100 000 requests at 2000 req/ms; every
    # request's checksum equals the value its tag and length fix, and the same
    # requests replayed alone give the same checksums
    say("config 5: serving, 16 GiB, 65536 slots, 100000 requests")
    opts = heap(H.HallocOpts(memory=16 << 30))
    g = H.Halloc.geometry()
    arr, ln, tag = serving_requests(100000, seed=0)
    r = H.Halloc.run_serve(arr, ln, tag, slots=65536, watermark=g["arena_bytes"] // 16, sample_every=4)
    say(f"completed={r['completed']} pages={r['pages']} kernel={r['kernel_ms']:.1f} ms")
    assert r["completed"] == len(arr) and r["nulls"] == 0
    assert (r["produced"] == ln).all() and (r["bad"] == 0).all()
    assert (r["checksum"] == serve_checksums(tag, ln)).all()
    assert r["pages"] == int(pages_for(ln).sum())
    pick = np.unique(np.concatenate([np.argsort(ln)[-2:], np.arange(0, len(arr), 20011)]))
    for i in pick:
        a = H.Halloc.run_serve(np.zeros(1, np.uint64), ln[i:i + 1], tag[i:i + 1], slots=256, watermark=0)
        assert a["completed"] == 1 and a["nulls"] == 0
        assert a["checksum"][0] == r["checksum"][i] and a["produced"][0] == r["produced"][i]
    assert_heap_empty_and_sound(opts)


def test_reference_test0_full(heap):
    # tst/corr/test/test.cu test 0: 2^21 / 8 threads x 8 blocks of 16 B on the
    # 128 MiB heap its main initializes (ha_init(halloc_opts_t(MEMORY)))
    say("tst/corr/test test 0: 262144 threads x 8 x 16 B, 128 MiB")
    opts = heap(H.HallocOpts(memory=4 * 16 * (2 << 20)))
    r = H.Halloc.run_corr_test0()
    assert r["blocks"] == 2 << 20
    assert r["nulls"] == 0 and r["bad_contents"] == 0 and r["bad_gaps"] == 0, r
    assert_heap_empty_and_sound(opts)


RUN_ALL = harness.run_all_sweep()


def _scaled_tries(argv, div):
    return [f"-t{max(2, -(-int(a[2:]) // div))}" if a.startswith("-t") else a for a in argv]


@pytest.mark.parametrize("test,argv", RUN_ALL, ids=[f"{t}:{' '.join(a)}" for t, a in RUN_ALL])
def test_run_all_sweep(test, argv):
    # every entry of tst/corr/run-all-tests.pl:20-75 at its own thread counts
    # and sizes; the try counts are cut 16-fold (the reference's 16384 tries
    # of 2^20 threads per prob-checkptr configuration take about 45 s each);
    # HALLOC_FULL_SWEEP=1 runs the full counts
    full = os.environ.get("HALLOC_FULL_SWEEP") == "1"
    o = harness.defaults_for(test)
    o.parse(argv if full else _scaled_tries(argv, 16))
    lines = []
    ok = harness.run_test(test, o, out=lines.append)
    assert ok, (test, argv, lines[-3:])
