"""bench.py's contract pieces that need no GPU: the one-GPU guard and the
HBM traffic figure it reports beside the step kernel's roofline."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout.strip() == ""


def test_traffic_figure_comes_from_the_committed_counter_summary():
    """The summary states FETCH_SIZE and WRITE_SIZE per launch in KiB with
    their scale factors; the figure bench.py reports is their scaled sum,
    and it sits below the launch's canonical (algorithmic) bytes, since the
    kernel reads 16-bit column offsets and no row pointers."""
    t = bench.measured_traffic(1024)
    assert isinstance(t, int) and t > 0
    assert bench.measured_traffic(512) is None  # only the measured batch
    text = open(bench.PMC_SUMMARY).read()
    m = re.search(r"traffic per launch = \(([\d,]+) x ([\d.]+) \+ ([\d,]+)\) KiB", text)
    assert m, "the summary states how the figure is formed"
    fetch, scale, write = (int(m.group(1).replace(",", "")), float(m.group(2)),
                           int(m.group(3).replace(",", "")))
    assert abs((fetch * scale + write) * 1024 - t) / t < 1e-4
    canonical = int(re.search(r"= ([\d,]+) B\.\n", text).group(1).replace(",", ""))
    assert canonical == 4272467252
    assert 0.5 * canonical < t < canonical
