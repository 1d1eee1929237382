"""The bench's command-line contract and byte accounting (CPU only)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr


def test_compulsory_bytes_per_unit():
    # heatbath: 7 half-planes read (6 staple directions' links and the opposite
    # parity of its own), 1 written; overrelaxation also reads the old link
    for prec, link in ((1, 16), (0, 32)):
        assert bench.LINK_BYTES[prec] == link
        assert bench.BYTES_PER_UPDATE["heatbath"][prec] == 8 * link
        assert bench.BYTES_PER_UPDATE["overrelax"][prec] == 9 * link
    # smearing: read and write per smeared link; the temporal step's staples
    # read all four directions (4 VOL links read, VOL written)
    assert bench.APE_TRANSFERS == {0: 2, 1: 2, 2: 5}
    assert bench.APE_SMEARED == {0: 3, 1: 4, 2: 1}


def test_roofline_traffic_comes_from_the_committed_profile():
    for kernel in ("heatbath", "overrelax"):
        t = bench.load_traffic(kernel, "48x48x48x48_fp32")
        assert t is not None and 3.0e8 < t < 6.0e8    # bytes per launch, near the 340/382 MB compulsory
    assert bench.load_traffic("heatbath", "no such workload") is None


def test_cpu_model_names_the_host():
    m = bench.cpu_model()
    assert m is None or len(m) > 0
