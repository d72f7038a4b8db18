"""bench.py's command-line contract on the CPU: one GPU only, and the measured HBM traffic it reports comes from
the committed rocprofv3 summary."""

from __future__ import annotations

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def test_more_than_one_gpu_is_refused(capsys):
    import bench

    assert bench.main(["--gpus", "2"]) == 2
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_traffic_profile_is_readable():
    import bench

    traffic = bench.measured_traffic("deep_k_1m")
    assert isinstance(traffic, int) and traffic > 0
    headline = bench.measured_traffic("deep_sparse_20m")
    assert isinstance(headline, int) and 10e9 < headline < 40e9
    wide = bench.measured_traffic("wide_vocab_2m")
    assert isinstance(wide, int) and 1e9 < wide < 20e9
    assert bench.measured_traffic("plain_60k") is None  # no PMC traffic profile for this workload
