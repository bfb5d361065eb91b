"""The reference test harness on the device (halloc_amd/harness.py over
ha_run_test): the run-all-tests sweep, shortened, passes; the perf tests run
without NULLs; a failing check is reported, not hidden."""
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


@pytest.mark.parametrize("test,argv", [
    ("throughput", ["-n65536", "-s16", "-S3072", "-dexpequal", "-l4", "-t2", "-m268435456"]),
    ("priv-throughput", ["-n262144", "-T256", "-s16", "-l1", "-i4", "-t2", "-m268435456"]),
])
def test_perf_tests(test, argv):
    say(f"harness {test}")
    o = harness.defaults_for(test)
    o.parse(argv)
    lines = []
    assert harness.run_test(test, o, out=lines.append)
    say("\n".join(lines))
    assert any("pair throughput" in l for l in lines)


def test_exhaustion_is_reported():
    # more than the pool holds: the check reports NULLs instead of passing
    say("harness exhaustion")
    o = harness.defaults_for("checkptr")
    o.parse(["-n262144", "-s3072", "-l4", "-t1", "-m33554432"])
    lines = []
    assert not harness.run_test("checkptr", o, out=lines.append)
    assert "cannot allocate enough memory" in lines[0]
