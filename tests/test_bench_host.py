"""CPU checks of bench.py's host logic: the committed traffic summary is matched to the
factor and sweep layout the plain leg runs."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_read_traffic_matches_layout(tmp_path):
    p = tmp_path / "t.md"
    p.write_text("# title\n\ntraffic_bytes_per_iter: 100\ntraffic_config: subtree_budget=0 amalg_ratio=0.9 level_pairs=1\n"
                 "text\ntraffic_bytes_per_iter: 200\ntraffic_config: subtree_budget=128 amalg_ratio=0.7 level_pairs=1\n"
                 "traffic_bytes_per_iter: 300\ntraffic_config: subtree_budget=256 amalg_ratio=0.7\n")
    assert bench.read_traffic(str(p), 0, 0.9, 1) == 100
    assert bench.read_traffic(str(p), 128, 0.7, True) == 200
    assert bench.read_traffic(str(p), 256, 0.7, 0) == 300  # no level_pairs key: unpaired
    assert bench.read_traffic(str(p), 256, 0.7, 1) is None
    assert bench.read_traffic(str(p), 0, 0.7, 1) is None
    assert bench.read_traffic(str(tmp_path / "missing.md"), 0, 0.9, 1) is None


def test_committed_traffic_summary_parses():
    vals = []
    with open(os.path.join(ROOT, "profiles", "traffic_icosphere164k_fp32_refine.md")) as f:
        vals = [ln for ln in f if ln.startswith("traffic_config:")]
    assert vals
