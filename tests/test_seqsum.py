"""The lens path's running-sum emulation (focus-stacker_amd/csrc/fg_seqsum.h) against the plain
loop it reproduces, on the CPU: tests/c/seqsum_check.cpp compiled with g++ and run."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_seq_add_matches_the_running_loop(tmp_path):
    exe = str(tmp_path / "seqsum_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra", "-Werror",
                    os.path.join(ROOT, "tests", "c", "seqsum_check.cpp"), "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert " 0 mismatches" in r.stdout
