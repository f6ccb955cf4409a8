"""Structure of the persistent grouped-contraction kernels in the gfx950
code object (compiled here, CPU only).

The kernels loop: one thread takes the next work unit from a device queue,
a barrier publishes it, the workgroup processes it.  If the optimiser treats
"thread 0" as loop-invariant it can thread the loop latch on it, giving a
second cycle that re-enters unit processing without a new queue fetch for
every lane but thread 0 (a workgroup that never drains).  The checks: each
grouped kernel has exactly one outermost loop, every nested loop is an
innermost one (K and chunk-sum loops), and the queue fetch sits in the
outermost loop's header block before the single read of the fetched index.
"""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tenet.jl_amd", "csrc", "k_gemm.hip")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def isa(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("isa") / "k_gemm.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "--cuda-device-only", "-S", SRC, "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, cwd=os.path.dirname(SRC))
    return out.read_text()


def _kernel_body(text, name):
    start = text.index(f"\n{name}:")
    end = text.index(".Lfunc_end", start)
    return text[start:end]


@pytest.mark.parametrize("kernel", ["grouped_gemm_c64_kernel", "grouped_gemm_c128_kernel"])
def test_persistent_loop_has_one_cycle_through_the_queue_fetch(isa, kernel):
    name = next(m for m in re.findall(r"^(_Z\S+):", isa, flags=re.M) if kernel in m)
    body = _kernel_body(isa, name)
    assert len(re.findall(r"This Loop Header: Depth=1", body)) == 1
    # a nested loop that is not innermost would be a second path around the unit
    assert "This Loop Header: Depth=2" not in body
    assert "This Loop Header: Depth=3" not in body
    hdr = body.index("This Loop Header: Depth=1")
    fetch = body.index("global_atomic_add", hdr)
    reads = [m.start() for m in re.finditer(r"ds_read_b32 v\d+, v\d+ offset:(17408|34816)\b", body)]
    assert len(reads) == 1 and fetch < reads[0]
