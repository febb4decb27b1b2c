"""bench.py's contract pieces that need no GPU.

- The roofline kernel's HBM traffic comes from the committed FETCH_SIZE /
  WRITE_SIZE summary, found by the kernel's name: a renamed or re-instantiated
  step kernel without a new profile row would report `traffic` as null.
- More than one GPU is not run: the command exits with a message saying so.
"""
import os
import subprocess
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_roofline_kernel_has_a_traffic_row():
    kernel = bench.STEP_KERNELS[1][0]
    traffic, row = bench.measured_traffic(kernel)
    assert row == kernel
    # at least the algorithmic 6 n b of one launch (both chains, 64^3 reduced, FP64)
    n = 64 ** 3
    assert traffic is not None and traffic >= 6 * n * 8


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)
