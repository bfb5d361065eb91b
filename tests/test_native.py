"""CPU check of the device layout arithmetic (halloc_amd/csrc/ha_bits.h) against brute force."""
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
