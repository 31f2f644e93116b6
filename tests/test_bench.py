"""The bench's host-side contract: one GPU only, the roofline's byte accounting and the
measured-traffic figure it reads from the committed PMC summary (no GPU needed)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_algorithmic_bytes_closed_form():
    # probe 3 rows x 2 columns; one build table of 5 rows -> B = floorPow2(5) = 4 buckets
    rels = [([np.zeros(3, np.int32)] * 2, [0, 1]), ([np.zeros(5, np.int32)] * 2, [1, 2])]
    count = 7
    want = 4 * 2 * 3 + (4 * 2 * 5 + 8 * 5 + 4 * (4 + 1)) + 4 * 3 * count + 16
    assert bench.algorithmic_bytes(rels, count, n_attrs=3) == want


def test_measured_traffic_is_close_to_the_algorithmic_bytes_of_the_chain():
    # the 4-way chain at s=0: R1 64 Mi rows, R2..R4 16 Mi rows, 4,295,896,600 results
    traffic = bench.measured_traffic()
    assert isinstance(traffic, int) and traffic > 0
    r1, rb, count = 64 << 20, 16 << 20, 4_295_896_600
    alg = 8 * r1 + 3 * (8 * rb + 8 * rb + 4 * ((1 << 24) + 1)) + 20 * count + 16
    # reads of records, prefixes and suffix runs come on top of the compulsory bytes; a
    # re-read of the output or a lost calibration factor would push the ratio far above this
    assert 1.0 <= traffic / alg < 1.25
