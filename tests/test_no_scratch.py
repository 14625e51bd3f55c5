"""Register budget of the library's kernels (CPU: hipcc cross-compiles for
gfx950): no kernel of ours spills to scratch memory.  A spill turns register
traffic into memory traffic on every use; the library's sort kernels
(rocPRIM, behind hipCUB) are their own and not checked."""
import concurrent.futures as cf
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "adfem.jl_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature",
         "-Xclang", "-packed-fp32-ops", "-x", "hip", "-c",
         "-Rpass-analysis=kernel-resource-usage"]


def _usage(src, out_dir):
    obj = os.path.join(out_dir, os.path.basename(src) + ".o")
    r = subprocess.run([HIPCC] + FLAGS + [src, "-o", obj], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    found, name = [], None
    for ln in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", ln)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", ln)
        if m and name:
            found.append((name, int(m.group(1))))
            name = None
    return found


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_no_kernel_of_ours_spills():
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with tempfile.TemporaryDirectory() as d, cf.ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda s: (s, _usage(s, d)), srcs))
    spills, ours = [], 0
    for src, kernels in results:
        for name, scratch in kernels:
            if "admfem" not in name:  # rocPRIM / hipCUB
                continue
            ours += 1
            if scratch:
                spills.append((os.path.basename(src), name, scratch))
    assert ours > 50, "kernel resource report not parsed"
    assert not spills, spills
