"""Build step of the package: compiles libswe.so for gfx950 in-tree with hipcc.

The floating-point flags are part of the library's contract: no contraction
(-ffp-contract=off), correctly rounded float division and square root, IEEE
denormals (the HIP default), and no packed-FP32 instruction selection.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libswe.so")

SOURCES = ["swe_kernels.hip", "swe_step.hip", "swe_capi.hip", "swe_terrain.cpp", "swe_state.cpp", "swe_export.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build libswe.so")
    return exe


# the step kernel is compiled once per tile shape (SWE_STEP_SHAPE = index
# into kStepShape in csrc/swe_step.h)
SHAPE_SOURCE = "swe_step_shapes.hip"
N_SHAPES = 5
OBJ_DIR = os.path.join(ROOT, "_build", "obj")


def _compile(args):
    src, obj, defines = args
    cmd = [hipcc()] + HIP_FLAGS + defines + ["-c", "-o", obj, src]
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    return cmd, res


def build(verbose=False, jobs=None):
    """Compile every HIP source into simplewinderosion_amd/libswe.so: one
    object per source (the step kernel once per tile shape), compiled in
    parallel, then linked."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJ_DIR, exist_ok=True)
    units = [(os.path.join(CSRC, s), os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o"), []) for s in SOURCES]
    units += [(os.path.join(CSRC, SHAPE_SOURCE), os.path.join(OBJ_DIR, "swe_step_shape%d.o" % i),
               ["-DSWE_STEP_SHAPE=%d" % i]) for i in range(N_SHAPES)]
    jobs = jobs or max(1, min(len(units), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, res in ex.map(_compile, units):
            if verbose:
                print(" ".join(cmd), flush=True)
            if res.returncode != 0:
                sys.stderr.write(res.stdout + res.stderr)
                raise RuntimeError("hipcc failed compiling %s" % cmd[-1])
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + ["-shared", "-o", tmp] + [u[1] for u in units]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libswe.so")
    os.replace(tmp, LIB)
    build_nn(verbose)
    build_demo(verbose)
    return LIB


NN_LIB = os.path.join(HERE, "libswe_nn.so")
NN_SOURCES = ["swe_nn.hip"]


def build_nn(verbose=False):
    """The surrogate's kernels (include/swe_nn.h) as libswe_nn.so."""
    tmp = NN_LIB + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + ["-shared", "-o", tmp] + [os.path.join(CSRC, s) for s in NN_SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libswe_nn.so")
    os.replace(tmp, NN_LIB)
    return NN_LIB


DEMO_SRC = os.path.join(HERE, "examples", "erode_demo.cpp")
DEMO = os.path.join(HERE, "examples", "erode_demo")


def build_demo(verbose=False):
    """The C++ example caller of include/swe_world.hpp (host code only)."""
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), DEMO_SRC, "-o", DEMO,
           "-L", HERE, "-lswe", "-Wl,-rpath," + HERE]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building the C++ example")
    return DEMO


if __name__ == "__main__":
    print(build(verbose=True))
