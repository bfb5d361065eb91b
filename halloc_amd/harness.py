"""The reference test harness's command line over the MI355X allocator.

The reference builds one CUDA program per test, all parsing the option set of
``CommonOpts`` (tst/common/common.h:118-298, tst/common/common.cu:130-355), and
sweeps them with tst/corr/run-all-tests.pl.  This module keeps that option set,
its derived quantities and its tests, so a reference sweep can be re-run
one-for-one; the kernels are the library's (``ha_run_test``):

    python -m halloc_amd.harness checkptr -l1 -t4 -s64 -n1000000
    python -m halloc_amd.harness prob-checkptr -i5 -t8 -f0.5 -F0.5 -e0.75 -g5 -l4 -s8 -S64 -duniform
    python -m halloc_amd.harness throughput -s16 -S3072 -dexpequal -n262144 -l4
    python -m halloc_amd.harness priv-throughput -n1048576 -T256 -s16 -l1
    python -m halloc_amd.harness phase-throughput -f0.95 -F0.05 -e0.91 -s16 -S256 -l1 -i1 -t10
    python -m halloc_amd.harness latency -n65536 -s16 -l1 -t4
    python -m halloc_amd.harness phase-latency -f0.95 -F0.05 -e0.91 -s16 -S256 -l1 -i4 -t4
    python -m halloc_amd.harness phase-alloc-write -f0.95 -F0.05 -e0.91 -s16 -S256 -l1 -i1 -t10
    python -m halloc_amd.harness phase-extfrag -f0.95 -F0.05 -e0.91 -s16 -S256 -l1 -i1 -t10
    python -m halloc_amd.harness freeslabs
    python -m halloc_amd.harness run-all [--quick] [--part I/N]
    python -m halloc_amd.harness throughput-all [--quick]   (tst/perf/throughput/throughput-all.sh)
    python -m halloc_amd.harness exp scaling|speed|all [--quick] [--out DIR]   (tst/exp/scaling, tst/exp/speed)
    python -m halloc_amd.harness grid-points      (samples/grid-points)
    python -m halloc_amd.harness add-strings      (samples/add-strings)
    python -m halloc_amd.harness random-graph     (samples/random-graph)

Options (same letters and ranges as the reference): -a allocator (halloc),
-m memory, -C halloc fraction, -B busy fraction, -R roomy fraction, -b slab
size shift, -D device, -n threads, -t tries, -T block size, -s / -S smallest /
largest allocation, -l allocations per thread, -i iterations, -q period shift,
-g group shift, -d distribution (uniform, expuniform, expequal), -f / -F / -e
allocation, free and execution fractions.  Random draws come from a
counter-based hash of (seed, thread, draw) instead of the reference's
per-thread state array.
"""
from __future__ import annotations

import dataclasses
import getopt
import math
import sys

DISTRS = {"uniform": 1, "expuniform": 2, "expequal": 3}
TESTS = {"checkptr": 0, "prob-checkptr": 1, "throughput": 2, "priv-throughput": 3, "phase-throughput": 4,
         "latency": 5, "phase-latency": 6, "phase-alloc-write": 7, "phase-extfrag": 8}
_WARMED = ("throughput", "priv-throughput", "phase-throughput", "latency", "phase-latency", "phase-alloc-write",
           "phase-extfrag")
NBYTES_IN_GIB = 1024.0 ** 3


class OptionError(ValueError):
    pass


def parse_int(s: str, lo: int = 0, hi: int | None = None) -> int:
    """parse_int of tst/common/common.cu:66-86: an integer within [lo, hi]."""
    try:
        v = int(s)
    except ValueError:
        raise OptionError(f"{s} is not an integer") from None
    if v < lo or (hi is not None and v > hi):
        raise OptionError(f"{v} is not in [{lo}, {hi if hi is not None else 'inf'}]")
    return v


def parse_double(s: str) -> float:
    try:
        return float(s)
    except ValueError:
        raise OptionError(f"{s} is not a number") from None


@dataclasses.dataclass
class CommonOpts:
    """CommonOpts with the defaults of tst/common/common.h:123-130."""
    allocator: str = "halloc"
    memory: int = 512 * 1024 * 1024
    halloc_fraction: float = 0.75
    busy_fraction: float = 0.835
    roomy_fraction: float = 0.6
    sparse_fraction: float = 0.0125  # the harness's value; the library default is 0.012
    sb_sz_sh: int = 22
    device: int = 0
    nthreads: int = 1024 * 1024
    ntries: int = 8
    alloc_sz: int = 16
    max_alloc_sz: int = 16
    nallocs: int = 4
    niters: int = 1
    bs: int = 128
    period_mask: int = 0
    group_sh: int = 0
    distr: str = "uniform"
    alloc_fraction: float = 1.0
    free_fraction: float = 0.0
    exec_fraction: float = 1.0
    max_alloc_sh: int = 0
    probabs: tuple = ((1.0, 1.0), (1.0, 1.0))

    def __post_init__(self):
        self.recompute_fields()

    # -- parse_cmdline (tst/common/common.cu:130-261)
    def parse(self, argv: list[str], device_memory: int | None = None) -> list[str]:
        """Applies the options in argv; returns the other arguments."""
        opts, rest = getopt.getopt(argv, "ha:m:C:B:R:D:b:n:t:T:s:S:l:i:q:g:d:f:F:e:")
        min_explicit = max_explicit = False
        for c, v in opts:
            if c == "-h":
                raise OptionError(__doc__)
            if c == "-a":
                if v != "halloc":
                    raise OptionError(f"allocator {v} is not supported (only halloc)")
                self.allocator = v
            elif c == "-m":
                self.memory = parse_int(v, 4096)
            elif c == "-C":
                self.halloc_fraction = parse_double(v)
            elif c == "-B":
                self.busy_fraction = parse_double(v)
            elif c == "-R":
                self.roomy_fraction = parse_double(v)
            elif c == "-b":
                self.sb_sz_sh = parse_int(v, 20, 26)
            elif c == "-D":
                self.device = parse_int(v, 0)
            elif c == "-n":
                self.nthreads = parse_int(v, 0)
            elif c == "-t":
                self.ntries = parse_int(v, 1)
            elif c == "-T":
                self.bs = parse_int(v, 1, 1024)
            elif c == "-s":
                min_explicit = True
                self.alloc_sz = parse_int(v, 0)
                if max_explicit:
                    if self.max_alloc_sz < self.alloc_sz:
                        raise OptionError("max allocation size should be >= min allocation size")
                else:
                    self.max_alloc_sz = self.alloc_sz
            elif c == "-S":
                max_explicit = True
                self.max_alloc_sz = parse_int(v, 0)
                if min_explicit:
                    if self.max_alloc_sz < self.alloc_sz:
                        raise OptionError("max allocation size should be >= min allocation size")
                else:
                    self.alloc_sz = self.max_alloc_sz
            elif c == "-l":
                self.nallocs = parse_int(v, 1)
            elif c == "-i":
                self.niters = parse_int(v, 1)
            elif c == "-q":
                sh = parse_int(v, 0, 31)
                self.period_mask = (1 << sh) - 1 if sh > 0 else 0
            elif c == "-g":
                self.group_sh = parse_int(v, 0, 31)
            elif c == "-d":
                if v not in DISTRS:
                    raise OptionError(f"unknown distribution {v}")
                self.distr = v
            elif c == "-f":
                self.alloc_fraction = parse_double(v)
            elif c == "-F":
                self.free_fraction = parse_double(v)
            elif c == "-e":
                self.exec_fraction = parse_double(v)
        if device_memory:  # the pool is capped at 3/4 of the device
            self.memory = min(self.memory, int(0.75 * device_memory))
        self.recompute_fields()
        return rest

    # -- recompute_fields (tst/common/common.cu:294-341)
    def recompute_fields(self):
        sh = 0
        while self.alloc_sz > 0 and self.max_alloc_sz >= self.alloc_sz << (sh + 1):
            sh += 1
        self.max_alloc_sh = sh
        f, F, e = self.alloc_fraction, self.free_fraction, self.exec_fraction
        if e + f + F > 2:
            raise OptionError("too large change fraction")
        if e < abs(f - F):
            raise OptionError("too small change fraction")
        # state 0 = holds nothing (next action: allocate), 1 = holds blocks (next action: free)
        p00 = (e + f - F) / (2 * (1 - F)) if F < 1 else 1.0
        p01 = (e - f + F) / (2 * F) if F > 0 else 1.0
        p10 = (e - f + F) / (2 * (1 - f)) if f < 1 else 1.0
        p11 = (e + f - F) / (2 * f) if f > 0 else 1.0  # the reference leaves it unset for f = 0
        self.probabs = ((p00, p01), (p10, p11))

    def is_random(self) -> bool:
        return self.alloc_sz != self.max_alloc_sz

    def group(self) -> int:
        return 1 << self.group_sh

    def period(self) -> int:
        return self.period_mask + 1

    def nptrs_cont(self, nts: int) -> int:
        """Active threads among the first nts (tst/common/common.h:287-290)."""
        gp = self.group() * self.period()
        return nts // gp * self.group() + min(nts % gp, self.group())

    def expected_sz(self) -> float:
        """Expected allocation size (tst/common/common.cu:263-285).  For
        expequal it keeps the reference's sum, which halves the weight at every
        level including the last (weights sum to 1 - 2^-(max_alloc_sh+1))."""
        a, b, m = self.alloc_sz, self.max_alloc_sz, self.max_alloc_sh
        if a == b:
            return float(a)
        if self.distr == "uniform":
            return (a + b) / 2
        levels = [((a << sh), min((a << (sh + 1)) - 1, b)) for sh in range(m + 1)]
        if self.distr == "expuniform":
            return sum((lo + hi) / 2 for lo, hi in levels) / (m + 1)
        e, p = 0.0, 1.0
        for lo, hi in levels:
            p /= 2
            e += (lo + hi) / 2 * p
        return e

    def total_nallocs(self) -> float:
        return float(self.nptrs_cont(self.nthreads)) * self.nallocs * self.niters * self.ntries

    def total_sz(self) -> float:
        return self.expected_sz() * self.total_nallocs()

    def size_classes(self) -> int:
        """How many of the 16 size classes (16..3072 B, src/halloc.cu:367-393)
        the drawn sizes can fall in; sizes above 3072 B use none."""
        blocks = [16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072]
        lo, hi = max(self.alloc_sz, 1), min(self.max_alloc_sz, 3072)
        return sum(1 for i, b in enumerate(blocks) if b >= lo and (i == 0 or blocks[i - 1] < hi))

    def nheads(self) -> int:
        """Head sets: every (set, class) pair pins a partly filled head slab,
        so the largest power of two up to 32 with 3 * sets * classes <= slabs
        (the classes the sizes span; 32 sets for one class from 96 slabs)."""
        nsbs = int(self.memory * self.halloc_fraction) >> self.sb_sz_sh
        ncls = max(1, self.size_classes())
        n = 1
        while 2 * n <= 32 and 3 * 2 * n * ncls <= nsbs:
            n *= 2
        return n

    def halloc_opts(self):
        """The allocator options, with the head-set count of nheads()."""
        import halloc_amd as H
        return H.HallocOpts(memory=self.memory, halloc_fraction=self.halloc_fraction,
                            busy_fraction=self.busy_fraction, roomy_fraction=self.roomy_fraction,
                            sparse_fraction=self.sparse_fraction, sb_sz_sh=self.sb_sz_sh, nheads=self.nheads())


def _kernel(test: str, o: CommonOpts, seed: int = 0) -> dict:
    import halloc_amd as H
    return H.Halloc.run_test(TESTS[test], o.nthreads, o.bs, o.ntries, o.nallocs, o.niters, o.alloc_sz,
                             o.max_alloc_sz, o.max_alloc_sh, DISTRS[o.distr], o.period_mask, o.group_sh,
                             o.probabs, seed)


def run_test(test: str, o: CommonOpts, out=print, clock: str = "wall") -> bool:
    """One reference test on a fresh heap: the test body of
    tst/corr/checkptr, tst/corr/prob-checkptr, tst/corr/freeslabs,
    tst/perf/throughput, tst/perf/priv-throughput, tst/perf/phase-throughput,
    tst/perf/latency, tst/perf/phase-latency, tst/perf/phase-alloc-write or
    tst/perf/phase-extfrag.  Returns whether it passed (the reference exits
    with -1 where this returns False).  clock "wall" times launch plus
    synchronization on the host, as the reference does (omp_get_wtime);
    "kernel" uses the kernels' own HIP-event time."""
    import halloc_amd as H
    if o.nthreads == 0:
        return True
    H.Halloc.init(o.halloc_opts())
    try:
        if test == "freeslabs":
            return _freeslabs(o, out)
        if test in _WARMED:
            # the reference's warm-up: min(4 * bs, n) threads, one try
            w = dataclasses.replace(o, nthreads=min(4 * o.bs, o.nthreads), ntries=1)
            if test != "priv-throughput":
                w.niters = 1
            _kernel(test, w)
        if test in ("checkptr", "throughput", "latency"):
            o = dataclasses.replace(o, niters=1)
        r = _kernel(test, o)
    finally:
        H.Halloc.shutdown()
    if test in ("checkptr", "prob-checkptr"):
        ok = r["checks_failed"] == 0
        if not ok:
            out(f"cannot allocate enough memory or pointers overlap: {r['checks_failed']} of "
                f"{r['checked'] + r['checks_failed']} pointers, first at try {r['failed_try']}")
        return ok
    if test == "phase-extfrag":
        # tst/perf/phase-extfrag/phase-extfrag.cu:84-93
        if r["checks_failed"]:
            out("cannot allocate enough memory")
            return False
        for t, (ef, efi) in enumerate(r["extfrag"]):
            out(f"external fragmentation {t} {ef:.2f} {efi:.2f}")
        return True
    if test in ("latency", "phase-latency"):
        # tst/perf/latency/latency.cu:117-128, tst/perf/phase-latency/phase-latency.cu:172-183
        if r["checks_failed"]:
            out("cannot allocate enough memory")
            return False
        names = ("malloc", "free")
        for k in range(2):
            out(f"min {names[k]} latency {r['lat_min'][k]:.2f} cycles")
            out(f"avg {names[k]} latency {r['lat_avg'][k]:.2f} cycles")
            out(f"max {names[k]} latency {r['lat_max'][k]:.2f} cycles")
        out(f"avg pair latency {r['lat_avg'][0] + r['lat_avg'][1]:.2f} cycles")
        return True
    n = o.total_nallocs()
    if clock == "wall":
        tm, tf = r["malloc_wall_ms"] / 1e3, r["free_wall_ms"] / 1e3
    else:
        tm, tf = r["malloc_ms"] / 1e3, r["free_ms"] / 1e3
    if test == "throughput":
        out(f"malloc throughput {n / tm * 1e-6:.2f} Mmallocs/s")
        out(f"free throughput {n / tf * 1e-6:.2f} Mfrees/s")
        out(f"pair throughput {n / (tm + tf) * 1e-6:.2f} Mpairs/s")
        out(f"malloc speed {o.total_sz() / tm / NBYTES_IN_GIB:.2f} GiB/s")
        out(f"pair speed {o.total_sz() / (tm + tf) / NBYTES_IN_GIB:.2f} GiB/s")
        return r["checks_failed"] == 0
    if test in ("phase-throughput", "phase-alloc-write"):
        # tst/perf/phase-throughput/phase-throughput.cu:110-135, phase-alloc-write.cu:116-144
        npairs = 0.5 * n * o.exec_fraction
        nm = 0.25 * n * (o.exec_fraction + o.alloc_fraction - o.free_fraction)
        tp = tm + tf
        if o.niters == 1:
            out(f"malloc throughput {nm / tm * 1e-6:.2f} Mmallocs/s")
            out(f"free throughput {nm / tf * 1e-6:.2f} Mfrees/s")
        out(f"pair throughput {npairs / tp * 1e-6:.2f} Mpairs/s")
        if o.niters == 1:
            out(f"malloc speed {nm * o.expected_sz() / tm / NBYTES_IN_GIB:.2f} GiB/s")
        out(f"pair speed {npairs * o.expected_sz() / tp / NBYTES_IN_GIB:.2f} GiB/s")
        return r["checks_failed"] == 0
    out(f"pair throughput {n / tm * 1e-6:.2f} Mpairs/s")
    out(f"pair speed {o.total_sz() / tm / NBYTES_IN_GIB:.2f} GiB/s")
    return r["nulls"] == 0


def _freeslabs(o: CommonOpts, out) -> bool:
    """tst/corr/freeslabs/freeslabs.cu:12-71: per try, sizes 16..1024 (steps
    8, 16 from 64, 128 from 256), each allocating 40% of the memory (at most
    nthreads blocks), checked, then freed: every slab must come back."""
    max_n = o.nthreads
    for _ in range(o.ntries):
        step, sz = 8, 16
        while sz <= 1024:
            n = min(max_n, int(math.floor(0.4 * o.memory / sz)))
            t = dataclasses.replace(o, nthreads=n, nallocs=1, niters=1, ntries=1, period_mask=0, group_sh=0,
                                    alloc_sz=sz, max_alloc_sz=sz)
            r = _kernel("checkptr", t)
            if r["checks_failed"]:
                out(f"allocation size {sz}: {r['checks_failed']} pointers NULL or overlapping")
                return False
            step = 128 if sz >= 256 else 16 if sz >= 64 else 8
            sz += step
    return True


def run_all_sweep(memory: int = 512 * 1024 * 1024, quick: bool = False):
    """The test list of tst/corr/run-all-tests.pl:20-75 as (test, argv) pairs.
    quick: every fourth checkptr size, one freeslabs try, and the
    prob-checkptr configurations with an eighth of the threads and two tries
    (the pool stays as large: the size classes need their slabs)."""
    scale = 1 / 8 if quick else 1.0
    tests = []
    step, sz, k = 8, 16, 0
    while sz <= 32 * 1024:
        fraction = 0.25 if sz <= 2 * 1024 else 0.125
        n = int(math.floor(fraction * memory / sz))
        if n and (not quick or k % 4 == 0):
            tests.append(("checkptr", ["-l1", "-t4", f"-s{sz}", f"-n{n}", f"-m{memory}"]))
        k += 1
        step = (1024 * 1024 if sz >= 1024 * 1024 else 128 * 1024 if sz >= 128 * 1024 else 16 * 1024
                if sz >= 16 * 1024 else 2 * 1024 if sz >= 2 * 1024 else 256 if sz >= 256 else 16
                if sz >= 64 else 8)
        sz += step
    tests.append(("freeslabs", [f"-m{memory}"] + (["-t1"] if quick else [])))
    for group in (0, 5, 10):
        for niters in (1, 5):
            ntries = math.ceil((1024 if group == 1 else 16384) / niters)
            if quick:
                ntries = 2
            fixed = [f"-i{niters}", f"-t{ntries}", "-f0.5", "-F0.5", "-e0.75", f"-g{group}", f"-m{memory}"]
            n = lambda k: f"-n{max(1, int(k * 1024 * scale))}"
            tests.append(("prob-checkptr", fixed + ["-l4", n(1024), "-s8", "-S64", "-duniform"]))
            tests.append(("prob-checkptr", fixed + ["-l1", n(1024), "-s8", "-S256", "-duniform"]))
            tests.append(("prob-checkptr", fixed + ["-l4", n(1024), "-s8", "-S256", "-dexpequal"]))
            tests.append(("prob-checkptr", fixed + ["-l1", n(64), "-s8", "-S3072", "-duniform"]))
            tests.append(("prob-checkptr", fixed + ["-l4", n(128), "-s8", "-S3072", "-dexpequal"]))
    return tests


def defaults_for(test: str) -> CommonOpts:
    """Per-test defaults the reference's mains set before parsing."""
    o = CommonOpts()
    if test == "freeslabs":
        o.ntries = 4
    return o


# tst/perf/throughput/throughput-all.sh:4-19: (size, allocations per thread, threads or None for the default)
THROUGHPUT_ALL = ((16, 4, None), (24, 4, None), (32, 4, None), (48, 4, None), (64, 4, None), (96, 2, None),
                  (128, 2, None), (192, 1, None), (256, 1, None), (384, 1, 512 * 1024), (512, 1, 512 * 1024),
                  (768, 1, 256 * 1024), (1024, 1, 256 * 1024), (1536, 1, 128 * 1024), (2048, 1, 128 * 1024),
                  (3072, 1, 64 * 1024))


def throughput_all(quick: bool = False, out=print) -> tuple[int, int]:
    """The throughput test at every size class, as throughput-all.sh runs it
    (quick: two tries instead of eight).  Prints one line per size and
    returns (runs, runs without NULLs)."""
    nruns = nok = 0
    for sz, nallocs, nthreads in THROUGHPUT_ALL:
        argv = [f"-s{sz}", f"-l{nallocs}"] + ([f"-n{nthreads}"] if nthreads else [])
        o = defaults_for("throughput")
        o.parse(argv + (["-t2"] if quick else []))
        lines = []
        ok = run_test("throughput", o, out=lines.append)
        pair = next((l for l in lines if l.startswith("pair throughput")), "no result")
        out(f"{'ok  ' if ok else 'FAIL'} throughput {' '.join(argv)}: {pair}")
        nruns += 1
        nok += ok
    return nruns, nok


# the flags every experiment of tst/exp/scaling/exp-run.pl and
# tst/exp/speed/exp-run.pl shares
EXP_COMMON = ["-f0.95", "-F0.05", "-e0.91", f"-m{512 * 1024 * 1024}", "-g5"]
EXP_FILES = {"thru": "exp-log-thru.csv", "lat": "exp-log-lat.csv", "single": "exp-log-single.csv",
             "combi": "exp-log-combi.csv"}


def _exp_fields(lines) -> dict:
    """The figures tst/exp/common.pl:runtest reads from a test's output lines
    (-1 where a line is absent)."""
    v = {}
    for line in lines:
        w = line.split()
        if "throughput" in w:
            v["thru_" + ("malloc" if "malloc" in w else "free" if "free" in w else "pair")] = float(w[2])
        elif "speed" in w:
            v["speed_" + ("malloc" if "malloc" in w else "pair")] = float(w[2])
        elif "latency" in w and w[0] in ("min", "avg", "max") and w[1] in ("malloc", "free"):
            v[f"lat_{w[1]}_{w[0]}"] = float(w[3])
    return v


def _exp_run(test: str, args: list[str]):
    """runtest of tst/exp/common.pl: phase-<test> with the shared flags;
    returns a lookup of the extracted figures."""
    test = "phase-" + test
    o = defaults_for(test)
    o.parse(EXP_COMMON + args)
    lines = []
    if not run_test(test, o, out=lines.append):
        raise RuntimeError(f"{test} {' '.join(args)} failed: {'; '.join(lines)}")
    v = _exp_fields(lines)
    return lambda k: v.get(k, -1.0)


def exp_scaling(quick: bool = False, out=print) -> dict:
    """tst/exp/scaling/exp-run.pl: pair throughput per thread count, spree
    (one iteration per launch, 512 launches) and priv (32 iterations per
    launch, 16 launches), then latency over 8 launches of 16 iterations.
    quick: four thread counts and a sixteenth of the launches."""
    total = 32 if quick else 512
    thru = [["alloc_sz", "nallocs", "nthreads", "priv_pair", "spree_pair", "spree_malloc", "spree_free"]]
    nts = [32, 1024, 32768, 1 << 20] if quick else [32 << k for k in range(16)]
    for n in nts:
        for sz in (16, 256):
            for nallocs in (1, 4):
                if nallocs == 4 and sz > 64:
                    continue
                args = [f"-n{n}", f"-l{nallocs}", f"-s{sz}"]
                spree = _exp_run("throughput", args + ["-i1", f"-t{total}"])
                priv = _exp_run("throughput", args + ["-i32", f"-t{total // 32}"])
                thru.append([sz, nallocs, n, priv("thru_pair"), spree("thru_pair"), spree("thru_malloc"),
                             spree("thru_free")])
                out(" ".join(str(x) for x in thru[-1]))
    lat = [["alloc_sz", "nallocs", "nthreads", "malloc_min", "malloc_avg", "malloc_max", "free_min", "free_avg",
            "free_max"]]
    total = 16 if quick else 128
    for n in ([1, 1024, 1 << 20] if quick else [1 << k for k in range(21)]):
        for sz in (16, 256):
            g = _exp_run("latency", [f"-n{n}", "-l1", f"-s{sz}", "-i16", f"-t{total // 16}"])
            lat.append([sz, 1, n] + [g(f"lat_{a}_{b}") for a in ("malloc", "free") for b in ("min", "avg", "max")])
            out(" ".join(str(x) for x in lat[-1]))
    return {"thru": thru, "lat": lat}


def exp_speed(quick: bool = False, out=print) -> dict:
    """tst/exp/speed/exp-run.pl: every size class alone, then the mixed
    ranges 8..32, 8..64, 8..256 and 8..3072 B; threads sized to 40% of the
    512 MiB pool (at most 2^20); spree and priv as in exp_scaling."""
    total = 32 if quick else 512
    cols = ["block", "priv_speed", "priv_thru", "spree_speed", "spree_speed_malloc", "speed_thru",
            "speed_thru_malloc"]

    def speed_row(label, args):
        spree = _exp_run("throughput", args + ["-i1", f"-t{total}"])
        priv = _exp_run("throughput", args + ["-i32", f"-t{total // 32}"])
        row = [label, priv("speed_pair"), priv("thru_pair"), spree("speed_pair"), spree("speed_malloc"),
               spree("thru_pair"), spree("thru_malloc")]
        out(" ".join(str(x) for x in row))
        return row

    memory, max_n = 512 * 1024 * 1024, 1 << 20
    single = [cols]
    for sz in (16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072):
        nallocs = 4 if sz <= 64 else 2 if sz <= 128 else 1
        n = min(int(math.floor(0.4 * memory / (sz * nallocs))), max_n)
        single.append(speed_row(sz, [f"-n{n}", f"-l{nallocs}", f"-s{sz}"]))
    combi = [cols]
    for lo, hi in ((8, 32), (8, 64), (8, 256), (8, 3072)):
        nallocs = 4 if hi <= 64 else 1
        n = min(int(math.floor(0.4 * memory / ((lo + hi) / 2 * nallocs))), max_n)
        combi.append(speed_row(f"{lo}..{hi}", [f"-n{n}", f"-l{nallocs}", f"-s{lo}", f"-S{hi}"]))
    return {"single": single, "combi": combi}


def write_exp_csvs(tables: dict, outdir: str) -> list[str]:
    """The experiment logs under the reference's file names, space separated."""
    import os
    os.makedirs(outdir, exist_ok=True)
    paths = []
    for key, rows in tables.items():
        path = os.path.join(outdir, EXP_FILES[key])
        with open(path, "w") as f:
            for row in rows:
                f.write(" ".join(f"{x:.2f}" if isinstance(x, float) else str(x) for x in row) + "\n")
        paths.append(path)
    return paths


def run_all(quick: bool = False, memory: int = 512 * 1024 * 1024, out=print, part=(0, 1)) -> tuple[int, int]:
    """part (i, n): only the entries whose index is i modulo n, so that the
    full sweep (about half an hour) can run in pieces."""
    ntests = nok = 0
    i, n = part
    for k, (test, argv) in enumerate(run_all_sweep(memory=memory, quick=quick)):
        if k % n != i:
            continue
        o = defaults_for(test)
        o.parse(argv)
        ok = run_test(test, o, out=lambda m: None)
        out(f"{'ok  ' if ok else 'FAIL'} {test} {' '.join(argv)}")
        ntests += 1
        nok += ok
    out(f"tests: {ntests} TOTAL, {nok} SUCCEEDED, {ntests - nok} FAILED")
    return ntests, nok


def run_sample(sample: str, out=print) -> bool:
    """A reference sample's main: a warm-up run, then the measured one
    (samples/grid-points/grid-points.cu:203-216: 10 000 points in 8^3 cells,
    then 10^6 in 32^3, 512 MiB; samples/add-strings/add-strings.cu:410-423:
    strings of 31 characters; samples/random-graph).  Returns whether every
    check passed."""
    import halloc_amd as H
    H.Halloc.init(H.HallocOpts(memory=512 << 20))
    try:
        if sample == "grid-points":
            H.Halloc.run_sample(sample, 10000, ncells=8)
            n = 1000000
            r = H.Halloc.run_sample(sample, n, ncells=32)
            out(f"allocation time {r['alloc_ms']:.2f} ms")
            out(f"free time {r['free_ms']:.2f} ms")
            out(f"allocation performance {n / r['alloc_ms'] * 1e-3:.2f} Mpoints/s")
            out(f"free performance {n / r['free_ms'] * 1e-3:.2f} Mpoints/s")
        elif sample == "random-graph":
            # samples/random-graph/random-graph.cu:238-246: 10^4 vertices of
            # degree <= 4, then 10^6 of degree <= 8
            H.Halloc.run_sample(sample, 10000, max_len=4)
            n, d = 1000000, 8
            r = H.Halloc.run_sample(sample, n, max_len=d)
            out(f"GPU time: {r['alloc_ms']:.3f} ms")
            out(f"GPU performance: {0.5 * (d + 1) * n / r['alloc_ms'] * 1e-3:.3f} Medges/s")
        else:
            H.Halloc.run_sample(sample, 10000)
            n = 1000000
            r = H.Halloc.run_sample(sample, n)
            out(f"GPU allocation time: {r['alloc_ms'] / 2:.2f} ms")
            out(f"GPU allocation performance: {n / (r['alloc_ms'] / 2) * 1e-3:.2f} Mstrings/s")
            out(f"GPU concatenation time: {r['work_ms']:.2f} ms")
            out(f"GPU concatenation performance: {n / r['work_ms'] * 1e-3:.2f} Mstrings/s")
            out(f"GPU freeing time: {r['free_ms'] / 3:.2f} ms")
            out(f"GPU freeing performance: {n / (r['free_ms'] / 3) * 1e-3:.2f} Mstrings/s")
    finally:
        H.Halloc.shutdown()
    if r["nulls"] or r["errors"]:
        out(f"{r['nulls']} NULLs, {r['errors']} failed checks")
    return not r["nulls"] and not r["errors"]


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] in ("-h", "--help"):
        print(__doc__)
        return 0
    test, args = argv[0], argv[1:]
    if test == "run-all":
        # run-all [--quick] [--part I/N]
        part = (0, 1)
        if "--part" in args:
            i, n = args[args.index("--part") + 1].split("/")
            part = (int(i), int(n))
        n, ok = run_all(quick="--quick" in args, part=part)
        return 0 if ok == n else 1
    if test == "exp":
        # exp scaling|speed|all [--quick] [--out DIR]
        which = next((x for x in args if x in ("scaling", "speed", "all")), "all")
        outdir = args[args.index("--out") + 1] if "--out" in args else "."
        quick = "--quick" in args
        tables = {}
        if which in ("scaling", "all"):
            tables.update(exp_scaling(quick))
        if which in ("speed", "all"):
            tables.update(exp_speed(quick))
        for path in write_exp_csvs(tables, outdir):
            print(f"wrote {path}")
        return 0
    if test == "throughput-all":
        n, ok = throughput_all(quick="--quick" in args)
        return 0 if ok == n else 1
    if test in ("grid-points", "add-strings", "random-graph"):
        return 0 if run_sample(test) else 255
    if test not in TESTS and test != "freeslabs":
        print(f"unknown test {test}", file=sys.stderr)
        return 2
    o = defaults_for(test)
    try:
        o.parse(args)
    except OptionError as e:
        print(e, file=sys.stderr)
        return 255
    return 0 if run_test(test, o) else 255


if __name__ == "__main__":
    sys.exit(main())
