"""The reference test harness's option set (halloc_amd/harness.py): parsing,
the alloc/free probabilities and the sweep list, on the CPU."""
import itertools
import math

import numpy as np
import pytest

from halloc_amd import harness
from halloc_amd.harness import CommonOpts, OptionError, run_all_sweep


def parsed(argv):
    o = CommonOpts()
    o.parse(argv)
    return o


def test_defaults_and_size_options():
    o = CommonOpts()
    assert (o.nthreads, o.ntries, o.alloc_sz, o.max_alloc_sz, o.nallocs, o.bs) == (1 << 20, 8, 16, 16, 4, 128)
    assert o.sparse_fraction == 0.0125 and o.sb_sz_sh == 22 and o.memory == 512 << 20
    assert parsed(["-s64"]).max_alloc_sz == 64          # -s alone sets both ends
    assert parsed(["-S96"]).alloc_sz == 96              # -S alone too
    o = parsed(["-s8", "-S3072"])
    assert (o.alloc_sz, o.max_alloc_sz, o.max_alloc_sh) == (8, 3072, 8)  # 8 << 8 = 2048 <= 3072 < 4096
    assert parsed(["-q3"]).period_mask == 7 and parsed(["-q0"]).period_mask == 0
    with pytest.raises(OptionError):
        parsed(["-s64", "-S32"])
    with pytest.raises(OptionError):
        parsed(["-T2048"])
    with pytest.raises(OptionError):
        parsed(["-dgauss"])
    with pytest.raises(OptionError):
        parsed(["-f1.0", "-F1.0", "-e0.5"])  # e + f + F > 2


def test_probabilities_known_answer():
    # -f0.95 -F0.05 -e0.91 (README:48) gives p00 = p11 = 0.9526, p01 = p10 = 0.1
    o = parsed(["-f0.95", "-F0.05", "-e0.91"])
    (p00, p01), (p10, p11) = o.probabs
    assert abs(p00 - 0.9526) < 1e-4 and abs(p11 - 0.9526) < 1e-4
    assert abs(p01 - 0.1) < 1e-9 and abs(p10 - 0.1) < 1e-9


@pytest.mark.parametrize("f,F,e", [(0.95, 0.05, 0.91), (0.5, 0.5, 0.75), (0.8, 0.3, 0.6), (0.7, 0.1, 0.9)])
def test_probabilities_reach_the_fractions(f, F, e):
    # the two-phase chain (phase 0 acts with p00 on free / p01 on holding
    # threads, phase 1 with p10 / p11): starting from F holding, one phase-0
    # step leaves f holding and one phase-1 step brings it back to F, and in
    # both phases the fraction of threads that act is e
    (p00, p01), (p10, p11) = parsed([f"-f{f}", f"-F{F}", f"-e{e}"]).probabs
    a0 = F * (1 - p01) + (1 - F) * p00
    assert abs(a0 - f) < 1e-6
    a1 = f * (1 - p11) + (1 - f) * p10
    assert abs(a1 - F) < 1e-6
    assert abs((1 - F) * p00 + F * p01 - e) < 1e-6
    assert abs((1 - f) * p10 + f * p11 - e) < 1e-6


def test_expected_sizes():
    assert parsed(["-s16"]).expected_sz() == 16
    assert parsed(["-s8", "-S64", "-duniform"]).expected_sz() == 36
    # levels [8,15] [16,31] [32,63] [64,64]: means 11.5 23.5 47.5 64
    assert parsed(["-s8", "-S64", "-dexpuniform"]).expected_sz() == pytest.approx((11.5 + 23.5 + 47.5 + 64) / 4)
    # the reference's weights 1/2, 1/4, 1/8, 1/16
    assert parsed(["-s8", "-S64", "-dexpequal"]).expected_sz() == pytest.approx(
        11.5 / 2 + 23.5 / 4 + 47.5 / 8 + 64 / 16)


def test_active_thread_count():
    # nptrs_cont counts tid with (tid >> g) & period_mask == 0 among the first n
    for q, g, n in itertools.product((0, 1, 2), (0, 3, 5), (1, 100, 1000, 4096)):
        o = parsed([f"-q{q}", f"-g{g}"])
        tid = np.arange(n)
        assert o.nptrs_cont(n) == int(((tid >> g) & o.period_mask == 0).sum()), (q, g, n)


def test_run_all_sweep_list():
    full = run_all_sweep()
    chk = [a for t, a in full if t == "checkptr"]
    sizes = [int(a[2][2:]) for a in chk]
    assert sizes[:8] == [16, 24, 32, 40, 48, 56, 64, 80] and sizes[-1] <= 32 * 1024
    assert all(b > a for a, b in zip(sizes, sizes[1:]))
    # 25% of the memory per size up to 2 KiB, 12.5% above
    for a in chk:
        s, n = int(a[2][2:]), int(a[3][2:])
        assert n == math.floor((0.25 if s <= 2048 else 0.125) * (512 << 20) / s)
    assert sum(t == "freeslabs" for t, _ in full) == 1
    prob = [a for t, a in full if t == "prob-checkptr"]
    assert len(prob) == 3 * 2 * 5
    assert {a[1] for a in prob} == {"-t16384", "-t3277"}


def test_run_all_parts_cover_the_sweep_once(monkeypatch):
    # run-all --part I/N: the N parts together run every entry exactly once
    seen = []
    monkeypatch.setattr(harness, "run_test", lambda test, o, out=None: seen.append((test, o.nthreads)) or True)
    full = [(t, a) for t, a in run_all_sweep(quick=True)]
    for i in range(3):
        n, ok = harness.run_all(quick=True, out=lambda m: None, part=(i, 3))
        assert n == ok == len(range(i, len(full), 3))
    assert len(seen) == len(full)
    assert [t for t, _ in seen] != [t for t, _ in full] or len(full) < 3
    assert sorted(t for t, _ in seen) == sorted(t for t, _ in full)


def test_throughput_all_list():
    # throughput-all.sh: the 16 size classes in order, 4 / 2 / 1 allocations
    # per thread, threads halved every two classes above 256 B
    sizes = [sz for sz, _, _ in harness.THROUGHPUT_ALL]
    assert sizes == [16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072]
    for sz, nallocs, n in harness.THROUGHPUT_ALL:
        assert nallocs == (4 if sz <= 64 else 2 if sz <= 128 else 1)
        assert n is None if sz <= 256 else n == (1 << 20) >> (1 + (sizes.index(sz) - 9) // 2)


def test_exp_tables_follow_the_experiment_scripts(monkeypatch, tmp_path):
    # tst/exp/scaling/exp-run.pl and tst/exp/speed/exp-run.pl: which runs, with
    # which flags, into which rows (the device runs are replaced by a stub)
    calls = []

    def stub(test, args):
        calls.append((test, args))
        return lambda k: 1.5
    monkeypatch.setattr(harness, "_exp_run", stub)
    t = harness.exp_scaling(out=lambda m: None)
    assert len(t["thru"]) == 1 + 16 * 3 and len(t["lat"]) == 1 + 21 * 2
    assert t["thru"][1][:3] == [16, 1, 32] and t["thru"][-1][:3] == [256, 1, 1 << 20]
    assert ("throughput", ["-n1048576", "-l4", "-s16", "-i32", "-t16"]) in calls
    assert ("throughput", ["-n32", "-l1", "-s256", "-i1", "-t512"]) in calls
    assert ("latency", ["-n1", "-l1", "-s16", "-i16", "-t8"]) in calls
    calls.clear()
    t.update(harness.exp_speed(out=lambda m: None))
    assert [r[0] for r in t["single"][1:]][:3] == [16, 24, 32] and len(t["single"]) == 17
    assert [r[0] for r in t["combi"][1:]] == ["8..32", "8..64", "8..256", "8..3072"]
    # threads: 40% of 512 MiB over the bytes per thread, at most 2^20
    assert ("throughput", ["-n69905", "-l1", "-s3072", "-i1", "-t512"]) in calls
    assert ("throughput", ["-n1048576", "-l4", "-s8", "-S32", "-i1", "-t512"]) in calls
    assert ("throughput", ["-n139446", "-l1", "-s8", "-S3072", "-i32", "-t16"]) in calls
    paths = harness.write_exp_csvs(t, str(tmp_path))
    assert sorted(p.rsplit("/", 1)[1] for p in paths) == sorted(harness.EXP_FILES.values())
    lines = open(tmp_path / "exp-log-lat.csv").read().splitlines()
    assert lines[0].startswith("alloc_sz nallocs nthreads malloc_min") and lines[1] == "16 1 1 " + " ".join(["1.50"] * 6)


def test_exp_fields_match_common_pl():
    v = harness._exp_fields(["malloc throughput 1.5 Mmallocs/s", "free throughput 2.5 Mfrees/s",
                             "pair throughput 3.00 Mpairs/s", "malloc speed 4.00 GiB/s", "pair speed 5.00 GiB/s",
                             "min malloc latency 6.00 cycles", "max free latency 7.00 cycles",
                             "avg pair latency 9.00 cycles"])
    assert v == {"thru_malloc": 1.5, "thru_free": 2.5, "thru_pair": 3.0, "speed_malloc": 4.0, "speed_pair": 5.0,
                 "lat_malloc_min": 6.0, "lat_free_max": 7.0}


def test_head_sets_follow_the_classes_spanned():
    # 512 MiB with 4 MiB slabs: 96 slabs; 3 * sets * classes stays within them
    assert parsed(["-s16"]).size_classes() == 1 and parsed(["-s16"]).nheads() == 32
    assert parsed(["-s8", "-S32"]).size_classes() == 3 and parsed(["-s8", "-S32"]).nheads() == 8
    assert parsed(["-s8", "-S64"]).size_classes() == 5 and parsed(["-s8", "-S64"]).nheads() == 4
    assert parsed(["-s8", "-S256"]).size_classes() == 9 and parsed(["-s8", "-S256"]).nheads() == 2
    assert parsed(["-s8", "-S3072"]).size_classes() == 16 and parsed(["-s8", "-S3072"]).nheads() == 2
    assert parsed(["-s3072"]).nheads() == 32 and parsed(["-s100", "-S200"]).size_classes() == 3
    assert parsed(["-s8", "-S3072", f"-m{1 << 30}"]).nheads() == 4
    assert parsed(["-s8", "-S3072", f"-m{4 << 30}"]).nheads() == 16
    assert parsed(["-s16", f"-m{256 << 20}"]).nheads() == 16  # 48 slabs
