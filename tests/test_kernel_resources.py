"""Resource budget of the step kernel, read from the gfx950 code-object
metadata that hipcc emits, for every compiled tile shape: no private
(scratch) memory, the VGPRs of the shape's waves per SIMD (96 for five, 128
for four) and LDS for the workgroups those waves make per CU.  A kernel that
breaks one of these still computes the right bits but several times slower,
so the budget is checked at build time, without a GPU."""
import os
import re
import subprocess

import pytest

from simplewinderosion_amd import build as B

LDS_PER_CU = 160 * 1024


def _kernels(text):
    kernels = {}
    for block in text.split("  - .agpr_count:")[1:]:
        m = re.search(r"\.name:\s+(\S+)", block)
        if not m:
            continue
        fields = dict(re.findall(r"\.(\w+):\s+(\d+)\s*$", block, re.M))
        kernels[m.group(1)] = {k: int(v) for k, v in fields.items()}
    return kernels


# the shape table of csrc/swe_step.h: (cells x, cells y, threads, waves per SIMD)
SHAPES = [(16, 32, 256, 5), (32, 16, 256, 5), (16, 16, 128, 5), (16, 32, 256, 4), (32, 32, 512, 4)]


def test_shape_table_matches_the_source():
    text = open(os.path.join(B.CSRC, "swe_step.h")).read()
    table = text[text.index("kStepShape[kStepShapes] = {"):]
    got = [tuple(int(v) for v in m) for m in re.findall(r"\{(\d+), (\d+), (\d+), (\d+)\}", table)][:len(SHAPES)]
    assert got == SHAPES and B.N_SHAPES == len(SHAPES)


@pytest.fixture(scope="module")
def metadata(tmp_path_factory):
    try:
        cc = B.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("isa")
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        out = str(d / ("shape%d.s" % i))
        cmd = [cc] + B.HIP_FLAGS + ["-DSWE_STEP_SHAPE=%d" % i, "--cuda-device-only", "-S", "-o", out,
                                    os.path.join(B.CSRC, B.SHAPE_SOURCE)]
        subprocess.run(cmd, check=True, capture_output=True)
        return _kernels(open(out).read())

    with ThreadPoolExecutor(max_workers=len(SHAPES)) as ex:
        return list(ex.map(one, range(len(SHAPES))))


@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_step_kernel_budget(metadata, shape):
    """Every instance of a shape: no scratch, the VGPRs of its waves per SIMD,
    LDS for the workgroups those waves make per CU."""
    sx, sy, threads, waves = SHAPES[shape]
    steps = {k: v for k, v in metadata[shape].items() if "k_step" in k}
    assert len(steps) == 48  # periodic x track x layers x deposit x mode (step + relax, relax only, step only)
    vgpr_max = (512 // waves) // 8 * 8
    groups_per_cu = 4 * waves // (threads // 64)
    for name, f in steps.items():
        assert f["private_segment_fixed_size"] == 0, name
        assert f["vgpr_count"] <= vgpr_max, (name, f["vgpr_count"])
        assert f.get("vgpr_spill_count", 0) == 0, name
        assert groups_per_cu * f["group_segment_fixed_size"] <= LDS_PER_CU, (name, f["group_segment_fixed_size"])
