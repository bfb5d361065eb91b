"""The reference test harness on the device (halloc_amd/harness.py over
ha_run_test): the run-all-tests sweep, shortened, passes; the perf tests run
without NULLs; a failing check is reported, not hidden."""
import numpy as np
import pytest

import halloc_amd as H
from halloc_amd import harness

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_run_all_sweep_quick():
    # tst/corr/run-all-tests.pl shortened: every fourth checkptr size, one
    # freeslabs try, the 30 prob-checkptr configurations with an eighth of the
    # threads and two tries each (512 MiB, as the reference runs them)
    say("harness run-all quick")
    lines = []
    n, ok = harness.run_all(quick=True, out=lines.append)
    say("\n".join(l for l in lines if not l.startswith("ok")))
    assert n == len(harness.run_all_sweep(quick=True)) and ok == n


def test_throughput_all_quick():
    # tst/perf/throughput/throughput-all.sh with two tries per size: every
    # size class from 16 to 3072 B at the script's thread counts, no NULLs
    say("harness throughput-all quick")
    lines = []
    n, ok = harness.throughput_all(quick=True, out=lines.append)
    say("\n".join(lines))
    assert n == len(harness.THROUGHPUT_ALL) == 16 and ok == n


def test_exp_replay_quick(tmp_path):
    # tst/exp/scaling and tst/exp/speed at four thread counts and a sixteenth
    # of the launches: every run passes its NULL check and reports a rate
    say("harness exp quick")
    t = harness.exp_scaling(quick=True, out=say)
    t.update(harness.exp_speed(quick=True, out=say))
    assert len(t["thru"]) == 1 + 4 * 3 and len(t["lat"]) == 1 + 3 * 2
    assert len(t["single"]) == 17 and len(t["combi"]) == 5
    for key in ("thru", "single", "combi"):
        for row in t[key][1:]:
            assert all(isinstance(x, float) and x > 0 for x in row[-4:]), (key, row)
    for row in t["lat"][1:]:
        assert 0 < row[3] <= row[4] <= row[5] and 0 < row[6] <= row[7] <= row[8], row
    assert len(harness.write_exp_csvs(t, str(tmp_path))) == 4


@pytest.mark.parametrize("test,argv", [
    ("throughput", ["-n65536", "-s16", "-S3072", "-dexpequal", "-l4", "-t2", "-m268435456"]),
    ("priv-throughput", ["-n262144", "-T256", "-s16", "-l1", "-i4", "-t2", "-m268435456"]),
    ("phase-throughput", ["-n262144", "-f0.95", "-F0.05", "-e0.91", "-s16", "-S256", "-l1", "-t10"]),
    ("phase-alloc-write", ["-n262144", "-f0.95", "-F0.05", "-e0.91", "-s16", "-S256", "-l1", "-t10"]),
])
def test_perf_tests(test, argv):
    say(f"harness {test}")
    o = harness.defaults_for(test)
    o.parse(argv)
    lines = []
    assert harness.run_test(test, o, out=lines.append)
    say("\n".join(lines))
    assert any("pair throughput" in l for l in lines)


@pytest.mark.parametrize("test,argv", [
    ("latency", ["-n65536", "-s16", "-S256", "-duniform", "-l2", "-t3", "-m268435456"]),
    ("phase-latency", ["-n65536", "-f0.95", "-F0.05", "-e0.91", "-s16", "-S256", "-l1", "-i4", "-t4"]),
])
def test_latency_tests(test, argv):
    # tst/perf/latency and tst/perf/phase-latency: the reference's output
    # lines, and every timed operation counted once
    say(f"harness {test}")
    o = harness.defaults_for(test)
    o.parse(argv)
    lines = []
    assert harness.run_test(test, o, out=lines.append)
    say("\n".join(lines))
    assert any(l.startswith("avg pair latency") for l in lines) and len(lines) == 7
    H.Halloc.init(o.halloc_opts())
    try:
        r = H.Halloc.run_test(harness.TESTS[test], o.nthreads, o.bs, o.ntries, o.nallocs,
                              1 if test == "latency" else o.niters, o.alloc_sz, o.max_alloc_sz,
                              o.max_alloc_sh, harness.DISTRS[o.distr], o.period_mask, o.group_sh,
                              o.probabs, seed=5)
        assert H.Halloc.free_bytes() == H.Halloc.geometry()["arena_bytes"]
    finally:
        H.Halloc.shutdown()
    assert r["nulls"] == 0 and r["checks_failed"] == 0, r
    # every malloc timed; in the chain every free of a block held is timed
    assert r["lat_n"][0] == r["mallocs"] > 0
    if test == "latency":
        assert r["lat_n"][0] == r["lat_n"][1] == r["frees"] == o.total_nallocs()
    else:
        assert 0 < r["lat_n"][1] <= r["frees"]
    for k in range(2):
        assert 0 < r["lat_min"][k] <= r["lat_avg"][k] <= r["lat_max"][k] < 1e9, r


def test_phase_extfrag():
    # tst/perf/phase-extfrag: one line per try; the ideal measure never
    # exceeds the plain one, and both stay in [0, 1]
    say("harness phase-extfrag")
    o = harness.defaults_for("phase-extfrag")
    o.parse(["-n262144", "-f0.95", "-F0.05", "-e0.91", "-s16", "-S256", "-l1", "-t6"])
    lines = []
    assert harness.run_test("phase-extfrag", o, out=lines.append)
    say("\n".join(lines))
    assert [l.split()[2] for l in lines] == [str(t) for t in range(6)]
    for l in lines:
        ef, efi = float(l.split()[3]), float(l.split()[4])
        assert 0.0 <= efi <= ef + 0.005 and ef <= 1.0, l


def test_exhaustion_is_reported():
    # more than the pool holds: the check reports NULLs instead of passing
    say("harness exhaustion")
    o = harness.defaults_for("checkptr")
    o.parse(["-n262144", "-s3072", "-l4", "-t1", "-m33554432"])
    lines = []
    assert not harness.run_test("checkptr", o, out=lines.append)
    assert "cannot allocate enough memory" in lines[0]


@pytest.mark.parametrize("sample", ["grid-points", "add-strings", "random-graph"])
def test_samples(sample):
    # the reference samples' mains: every point in its own cell once, every
    # concatenation equal to its two strings, everything freed
    say(f"sample {sample}")
    lines = []
    assert harness.run_sample(sample, out=lines.append)
    say("\n".join(lines))


def test_typed_new_and_delete():
    # new (halloc) T(args) / hadelete (include/halloc.h under HALLOC_CPP): every
    # object constructed once and destroyed once, arrays intact, nothing left
    say("sample objects")
    H.Halloc.init(H.HallocOpts(memory=4 << 30))  # up to 16 classes; the adaptive head sets fold to 16
    try:
        for n, hi, seed in ((1 << 20, 64, 1), (300000, 700, 2), (1000, 1, 3)):
            r = H.Halloc.run_sample("objects", n, max_len=hi, seed=seed)
            assert r["nulls"] == 0 and r["errors"] == 0, (n, hi, r)
            assert H.Halloc.free_bytes() == H.Halloc.geometry()["arena_bytes"]
    finally:
        H.Halloc.shutdown()


def test_sample_string_lengths_vary():
    # strings of 1..600 characters reach 14 size classes: the head sets fold to
    # 16 on 4 GiB (768 slabs), which leaves room for that many heads
    say("sample add-strings 1..600")
    H.Halloc.init(H.HallocOpts(memory=4 << 30))
    try:
        r = H.Halloc.run_sample("add-strings", 100000, min_len=1, max_len=600, seed=3)
        assert r["nulls"] == 0 and r["errors"] == 0, r
        r = H.Halloc.run_sample("grid-points", 200000, ncells=5, seed=4)
        assert r["nulls"] == 0 and r["errors"] == 0, r
        assert H.Halloc.free_bytes() == H.Halloc.geometry()["arena_bytes"]
    finally:
        H.Halloc.shutdown()


def touch_expected(n, words, iters, reps):
    """The word sums ha_run_touch must report: stores leave word w of lane i
    at (iters-1)*words + w + i; a load pass sums the block; each of the
    1 + reps update launches applies x -> x*M + it for it in 0..iters-1."""
    u = np.uint64
    x = ((iters - 1) * words + np.arange(words, dtype=np.uint64)[None, :]
         + np.arange(n, dtype=np.uint64)[:, None]).astype(np.uint64)
    store = int(x.sum(dtype=np.uint64))
    load = (iters * store) % (1 << 64)
    with np.errstate(over="ignore"):
        for _ in range(1 + reps):
            for it in range(iters):
                x = x * u(6364136223846793005) + u(it)
    return store, load, int(x.sum(dtype=np.uint64))


@pytest.mark.parametrize("cached", [0, 1])
def test_caller_block_accesses(cached):
    # a caller's kernel over its blocks in both arena types: stores, loads and
    # updates see exactly their own values, and every block is freed
    say(f"caller block accesses cached_arena={cached}")
    H.Halloc.init(H.HallocOpts(memory=64 << 20, cached_arena=cached))
    try:
        n, words, iters, reps = 5000, 8, 4, 2
        r = H.Halloc.run_touch(n, words, iters, reps)
        store, load, update = touch_expected(n, words, iters, reps)
        assert r["nulls"] == 0
        assert (r["store_sum"], r["load_sum"], r["update_sum"]) == (store, load, update)
        assert H.Halloc.free_bytes() == H.Halloc.geometry()["arena_bytes"]
    finally:
        H.Halloc.shutdown()


def test_long_loop_192B_at_16_sets():
    # tst/perf/priv-throughput at 192 B with 32 alloc/free iterations per
    # launch on 16 head sets: the crowded-head and forced replacements run
    # under full concurrency; no NULL, no check failure, every block freed
    # and every heap invariant holds afterwards
    import dataclasses
    from oracle import model as M
    say("priv-throughput 192 B, 16 sets, -i32")
    o = harness.defaults_for("priv-throughput")
    o.parse(["-n1048576", "-s192", "-l1", "-i32", "-t2", "-m536870912"])
    opts = dataclasses.replace(o.halloc_opts(), nheads=16)
    H.Halloc.init(opts)
    try:
        r = harness._kernel("priv-throughput", o)
        assert r["nulls"] == 0 and r["checks_failed"] == 0, r
        assert H.Halloc.free_bytes() == H.Halloc.geometry()["arena_bytes"]
        mo = M.make_opts(memory=opts.memory, halloc_fraction=opts.halloc_fraction,
                         busy_fraction=opts.busy_fraction, roomy_fraction=opts.roomy_fraction,
                         sparse_fraction=opts.sparse_fraction, sb_sz_sh=opts.sb_sz_sh, nheads=opts.nheads,
                         hash_mode=1, hash_step_frac=opts.hash_step_frac, fixed_sets=opts.fixed_sets)
        assert M.check_dump(mo, H.Halloc.dump())[0] == 0
    finally:
        H.Halloc.shutdown()
