"""CPU: the bench's byte model and its command line, without a GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


def test_pipeline_bytes_per_frame_pixel(bench):
    """Pass 1 = 3s + 14, pass 2 = 3s + 41, collapse = (16 + 3s)/N: 61 + 19/N B at u8 and
    67 + 22/N B at u16 (two-pass streaming model of the fused path)."""
    assert bench.alg_bytes_per_frame_px(1, 30) == pytest.approx(61 + 19 / 30)
    assert bench.alg_bytes_per_frame_px(2, 143) == pytest.approx(67 + 22 / 143)


def test_dominant_kernel_models_fit_the_pipeline(bench):
    """k_pass1 carries the whole pass-1 term; level-0 k_pass2 carries the pass-2 term less the
    accumulator levels above level 0 (2 x 12 B x 1/3)."""
    for s in (1, 2):
        m = bench.kernel_models(s)
        assert m["k_pass1"][1] == 3 * s + 14
        assert m["k_pass2_level0"][1] == (3 * s + 41) - 8
        for per_launch in bench.TRAFFIC.values():
            assert set(per_launch) == set(m)


def test_job_cores_follows_omp_num_threads(bench, monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "5")
    assert bench.job_cores() == 5
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.job_cores() == len(os.sched_getaffinity(0))


def test_percentile(bench):
    assert bench.pct([5, 1, 3, 2, 4], 50) == 3
    assert bench.pct([5, 1, 3, 2, 4], 99) == 5
    assert bench.pct([], 50) is None


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "--gpus must be 1" in r.stderr
    assert r.stdout == ""
