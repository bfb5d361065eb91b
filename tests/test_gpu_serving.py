"""Serving loop (continuous batching) on the GPU.

Requests arrive on a Poisson clock while others are decoding; each grows page
by page through hamalloc and is read back and freed at completion.  Every
request's checksum must equal the value its tag and token count fix, the same
alone as in the mix; the heap must be empty afterwards with its invariants
intact, also when the pool runs dry and pages come back NULL.
"""
import numpy as np
import pytest

import halloc_amd as H
from halloc_amd.workloads import pages_for, serving_requests
from oracle import model as M
from tests.helpers import serve_checksums

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def check_empty_heap(opts):
    g = H.Halloc.geometry()
    assert H.Halloc.free_bytes() == g["arena_bytes"]
    mo = M.make_opts(memory=opts.memory, sb_sz_sh=opts.sb_sz_sh, nheads=opts.nheads,
                     hash_step_frac=opts.hash_step_frac)
    tot, detail = M.check_dump(mo, H.Halloc.dump())
    assert tot == 0, detail


def report_bad(r, tag, ln, arr):
    """Prints what the damaged requests held: whose token sat where theirs should."""
    idx = np.nonzero(r["bad"])[0]
    if len(idx) == 0:
        return
    by_tag = {int(t): i for i, t in enumerate(tag)}
    say(f"{len(idx)} requests with damaged tokens, {int(r['bad'].sum())} tokens")
    for i in idx[:12]:
        x = int(r["bad_first"][i] & 0xFFFFFFFF)
        y = int(r["bad_first"][i] >> 32)
        o = by_tag.get(x, -1)
        say(f"  req {i} len {ln[i]} bad {r['bad'][i]} first ({r['first_ns'][i]}..{r['done_ns'][i]} ns) "
            f"holds tag {x:#x} (req {o}, len {ln[o] if o >= 0 else '-'}) token {y}"
            + (f" ({r['first_ns'][o]}..{r['done_ns'][o]} ns)" if o >= 0 else ""))


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


def test_serve_mixed_and_alone(heap):
    opts = heap(H.HallocOpts(memory=2 << 30, nheads=8))
    arr, ln, tag = serving_requests(20000, seed=1)
    say(f"serve mixed nreq={len(arr)} max_len={ln.max()}")
    r = H.Halloc.run_serve(arr, ln, tag, slots=65536, watermark=64 << 20)
    assert r["completed"] == len(arr) and r["admitted"] == len(arr)
    assert r["nulls"] == 0
    assert (r["produced"] == ln).all()
    report_bad(r, tag, ln, arr)
    assert (r["bad"] == 0).all()
    assert (r["checksum"] == serve_checksums(tag, ln)).all()
    assert r["pages"] == int(pages_for(ln).sum())
    assert (r["first_ns"] + 1000 >= arr).all()  # one 100 MHz tick of rounding
    assert (r["done_ns"] >= r["first_ns"]).all()
    assert len(r["lat_ns"]) > 0
    check_empty_heap(opts)
    # the same requests alone: same checksums and token counts
    pick = np.unique(np.concatenate([np.argsort(ln)[-3:], np.arange(0, len(arr), 4001)]))
    for i in pick:
        say(f"serve alone request={i} len={ln[i]}")
        a = H.Halloc.run_serve(np.zeros(1, np.uint64), ln[i:i + 1], tag[i:i + 1], slots=256, watermark=0)
        assert a["completed"] == 1 and a["nulls"] == 0
        assert a["produced"][0] == r["produced"][i]
        assert a["checksum"][0] == r["checksum"][i]
    check_empty_heap(opts)


@pytest.mark.parametrize("nheads,cached", [(1, 0), (8, 0), (1, 1), (8, 1)])
def test_serve_pool_runs_dry(heap, nheads, cached):
    """With no watermark the heap runs dry; slabs then move between XCDs all
    the time, and no token may come back damaged in either arena mode."""
    opts = heap(H.HallocOpts(memory=64 << 20, sb_sz_sh=21, nheads=nheads, cached_arena=cached))
    arr, ln, tag = serving_requests(4000, rate_per_ms=4000, median=2048, sigma=1.0, seed=2)
    say(f"serve dry nheads={nheads} cached_arena={cached} nreq={len(arr)} total_pages={int(pages_for(ln).sum())}")
    r = H.Halloc.run_serve(arr, ln, tag, slots=16384, watermark=0)
    assert r["completed"] == len(arr)
    assert r["nulls"] > 0, "the workload should exhaust a 48 MiB heap"
    assert (r["produced"] <= ln).all()
    assert ((r["produced"] < ln).sum()) == r["nulls"]
    report_bad(r, tag, ln, arr)
    assert (r["bad"] == 0).all()
    assert (r["checksum"] == serve_checksums(tag, r["produced"])).all()
    assert r["pages"] == int(pages_for(r["produced"]).sum())
    check_empty_heap(opts)


def test_serve_watermark_gates_admission(heap):
    """A watermark at the full heap admits only while nothing is allocated."""
    opts = heap(H.HallocOpts(memory=64 << 20, sb_sz_sh=21, nheads=8))
    g = H.Halloc.geometry()
    arr, ln, tag = serving_requests(500, rate_per_ms=1000, seed=3)
    r = H.Halloc.run_serve(arr, ln, tag, slots=256, watermark=g["arena_bytes"])
    assert r["completed"] == len(arr) and r["nulls"] == 0
    assert (r["checksum"] == serve_checksums(tag, ln)).all()
    assert r["stalls"] > 0
    check_empty_heap(opts)


def test_serve_rejects_bad_input(heap):
    heap(H.HallocOpts(memory=64 << 20, sb_sz_sh=21, nheads=8))
    g = H.Halloc.geometry()
    arr, ln, tag = serving_requests(10, seed=4)
    with pytest.raises(H.HallocError):
        H.Halloc.run_serve(arr[::-1], ln, tag, slots=256)
    bad = ln.copy()
    bad[3] = 0
    with pytest.raises(H.HallocError):
        H.Halloc.run_serve(arr, bad, tag, slots=256)
    with pytest.raises(H.HallocError):
        H.Halloc.run_serve(arr, ln, tag, slots=256, watermark=g["arena_bytes"] + 1)
    with pytest.raises(H.HallocError):
        H.Halloc.run_serve(arr, ln, tag, slots=300)
