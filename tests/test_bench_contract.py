"""The bench's bookkeeping on the CPU: the algorithmic bytes of the pair-force
launch, which kernel the default LJ line times (the row pass with the fused
final half step) and that the committed traffic calibration is the one for
that kernel; the multi-GPU guard."""
import os
import subprocess
import sys

import bench
from mddem_amd import hip
from mddem_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pair_force_bytes():
    # 16 B fixed-point position + 4 B per half-list pair + 3 x 4 B force
    assert bench.force_kernel_bytes(256000, 9742717, 4, 16.0) == 256000 * 28.0 + 4.0 * 9742717
    # FP64: 24 B position, 3 x 8 B force
    assert bench.force_kernel_bytes(10, 7, 8) == 10 * 48.0 + 28.0


def test_default_line_times_the_fused_row_pass():
    w = W.lj_fcc(ncell=40)
    assert bench.post_fused(w, hip.MIXED, 2)
    assert bench.kernel_name(2, True, True) == "k_lj_post<2>"
    # not fused: FP64, no thermostat, eight lanes, cell tiles, above 262,144
    assert not bench.post_fused(w, hip.FP64, 2)
    assert not bench.post_fused(W.lj_fcc(ncell=6, thermostat=False), hip.MIXED, 2)
    assert not bench.post_fused(w, hip.MIXED, 3)
    assert not bench.post_fused(w, hip.MIXED, -1)
    assert not bench.post_fused(W.lj_fcc(ncell=41), hip.MIXED, 2)
    assert bench.kernel_name(3, True, False) == "k_lj<float,false,3>"
    assert bench.kernel_name(-1, True) == "k_lj_tile<float,false,true>"


def test_traffic_calibration_covers_the_timed_kernel():
    t = bench.traffic_from_profiles("lj_fcc_256000", "k_lj_post<2>")
    assert t is not None and 90e6 < t < 150e6
    assert bench.traffic_from_profiles("lj_fcc_256000", "k_lj<float,false,2>") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
