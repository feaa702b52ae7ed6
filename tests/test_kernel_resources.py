"""Resource budget of the step kernel, read from the gfx950 code-object
metadata that hipcc emits: no private (scratch) memory, at most 96 VGPRs
(five waves per SIMD) and LDS for five workgroups per CU.  A kernel that
breaks one of these still computes the right bits but several times slower,
so the budget is checked at build time, without a GPU."""
import os
import re
import subprocess

import pytest

from simplewinderosion_amd import build as B

LDS_PER_CU = 160 * 1024


@pytest.fixture(scope="module")
def metadata(tmp_path_factory):
    try:
        cc = B.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("isa") / "swe_step.s")
    cmd = [cc] + B.HIP_FLAGS + ["--cuda-device-only", "-S", "-o", out, os.path.join(B.CSRC, "swe_step.hip")]
    subprocess.run(cmd, check=True, capture_output=True)
    text = open(out).read()
    kernels = {}
    for block in text.split("  - .agpr_count:")[1:]:
        m = re.search(r"\.name:\s+(\S+)", block)
        if not m:
            continue
        fields = dict(re.findall(r"\.(\w+):\s+(\d+)\s*$", block, re.M))
        kernels[m.group(1)] = {k: int(v) for k, v in fields.items()}
    return kernels


def test_step_kernel_budget(metadata):
    steps = {k: v for k, v in metadata.items() if "k_step" in k}
    assert len(steps) == 32  # periodic x track x layers x deposit x relax-only
    for name, f in steps.items():
        assert f["private_segment_fixed_size"] == 0, name
        assert f["vgpr_count"] <= 96, (name, f["vgpr_count"])
        assert f.get("vgpr_spill_count", 0) == 0, name
        assert 5 * f["group_segment_fixed_size"] <= LDS_PER_CU, (name, f["group_segment_fixed_size"])
