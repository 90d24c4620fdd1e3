"""Every device kernel compiles for gfx950 without private (scratch) memory.

A kernel with scratch reads and writes per-lane memory through the cache
hierarchy: a grouped kernel that indexes its by-value argument group with the
proof index can end up copying the whole group (kilobytes) into scratch for
every lane, which made the grinding kernels several times slower.  The
compiler's resource remarks name each kernel's scratch size; all must be 0."""
import concurrent.futures
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "whir-p3_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["k_lde.hip", "k_merkle.hip", "k_poly.hip", "prover.hip"]

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="needs hipcc")


def remarks(src, tmp):
    out = os.path.join(tmp, os.path.splitext(src)[0] + ".dev.o")
    r = subprocess.run([HIPCC, "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", os.path.join(CSRC, src), "-o", out,
                        "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    kernels, cur = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and cur:
            kernels[cur] = int(m.group(1))
    return kernels


def test_no_kernel_uses_scratch(tmp_path):
    with concurrent.futures.ThreadPoolExecutor(len(SOURCES)) as ex:
        found = list(ex.map(lambda s: remarks(s, str(tmp_path)), SOURCES))
    kernels = {k: v for d in found for k, v in d.items()}
    assert len(kernels) > 30
    assert any("k_grind" in k for k in kernels)
    bad = {k: v for k, v in kernels.items() if v}
    assert not bad, bad
