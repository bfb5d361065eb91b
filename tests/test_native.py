"""CPU checks of the arithmetic shared with the device code: the wave-run layout and division-free
remainders (halloc_amd/csrc/ha_bits.h) against brute force, and the graph workload's Poisson table
(halloc_amd/csrc/ha_workload.h) against scipy."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_run_union_arithmetic(tmp_path):
    exe = tmp_path / "run_union_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "native", "run_union_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def test_poisson8_table_matches_scipy(tmp_path):
    # threshold k = floor(2^24 P(X <= k)), X ~ Poisson(8); and the inverted
    # degrees of 2^24 evenly spaced uniforms reproduce the distribution
    import numpy as np
    from scipy.stats import poisson
    exe = tmp_path / "workload_table"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(HERE, "native", "workload_table.cpp")],
                   check=True)
    t = np.array([int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.split()], dtype=np.int64)
    k = np.arange(len(t))
    want = np.floor(poisson.cdf(k, 8) * 2**24).astype(np.int64)
    assert (t == want).all(), np.nonzero(t != want)
    # degree = number of thresholds <= u: P(degree = k) = (t[k] - t[k-1]) / 2^24
    u = np.arange(0, 2**24, 97, dtype=np.int64)
    deg = np.searchsorted(t, u, side="right")
    hist = np.bincount(deg, minlength=len(t)) / len(u)
    assert np.abs(hist[:30] - poisson.pmf(np.arange(30), 8)).max() < 1e-4
