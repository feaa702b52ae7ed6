"""The C-ABI library loads and exports every symbol include/swe.h declares
(no compute: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from simplewinderosion_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "swe.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(swe_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_bound_api():
    names = declared_functions()
    assert "swe_erode" in names and "swe_create" in names
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(swe_[a-z_0-9]+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_params_layout_matches_c(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "swe.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(swe_params), offsetof(swe_params, seed),'
        ' offsetof(swe_params, window), sizeof(swe_audit_t), sizeof(swe_stats_t));return 0;}\n')
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert [int(v) for v in got] == [ctypes.sizeof(capi.SweParams), capi.SweParams.seed.offset,
                                     capi.SweParams.window.offset, ctypes.sizeof(capi.SweAudit),
                                     ctypes.sizeof(capi.SweStats)]
    import oracle as O
    assert O.load().or_params_size() == ctypes.sizeof(capi.SweParams)
    for name, _ in O.OrParams._fields_:
        assert getattr(O.OrParams, name).offset == getattr(capi.SweParams, name).offset


def test_defaults_are_the_reference_statics():
    p = capi.default_params()
    assert p.max_age == 512                       # wind.hpp:23
    assert p.boundarylayer == 2.0                 # wind.hpp:24
    assert abs(p.suspension - 0.05) < 1e-9        # wind.hpp:25
    assert abs(p.gravity - 0.1) < 1e-8            # wind.hpp:26
    assert list(p.pspeed) == [1.0, 0.0, 0.0]      # wind.hpp:13
    assert abs(p.lrate - 0.1) < 1e-8              # world.hpp:30
    assert abs(p.maxdiff - 0.005) < 1e-9          # world.hpp:31
    assert p.settling == 1.0                      # world.hpp:32
    assert p.mapscale == 80.0                     # cellpool.h:184
    assert p.periodic == 0 and p.layers == 1 and p.track == 1


def test_terrain_generator_host_side():
    import numpy as np
    from simplewinderosion_amd import generate_terrain

    a = generate_terrain(64, 48, 7)
    b = generate_terrain(64, 48, 7)
    c = generate_terrain(64, 48, 8)
    assert a.shape == (64, 48) and np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() == 0.0 and a.max() == 1.0
    # periodic: wrap-around differences are as small as interior ones
    interior = np.abs(np.diff(a, axis=0)).max()
    assert np.abs(a[0] - a[-1]).max() <= 1.5 * interior
    p = generate_terrain(256, 256, 3, n_pyramids=8, pyr_h=(0.3, 0.6), pyr_base=(40, 120))
    assert p.max() > 1.0 + 0.25


def test_gpu_entry_points_fail_cleanly_without_device():
    lib = capi.load()
    ctx = ctypes.c_void_p()
    rc = lib.swe_create(ctypes.byref(ctx), 0, 4, 0)
    assert rc == capi.SWE_EINVAL
    assert lib.swe_erode(None, 1, 1) == capi.SWE_EINVAL


def test_cpp_header_mirror_compiles_and_fails_cleanly_without_gpu(tmp_path):
    """include/swe_world.hpp builds with g++ -std=c++17 and reports errors as exceptions."""
    from simplewinderosion_amd.build import DEMO, build_demo

    if not os.path.exists(DEMO):
        build_demo()
    out = subprocess.run([DEMO, "8", "8", "1", "4", "1"], capture_output=True, text=True, timeout=60)
    if out.returncode == 0:
        pytest.skip("a GPU is present: the GPU test covers the run")
    assert "swe_create failed" in out.stderr


def test_nn_library_exports_every_declared_symbol():
    """libswe_nn.so exports, and simplewinderosion_amd/nn.py binds, every
    function include/swe_nn.h declares."""
    from simplewinderosion_amd import nn

    text = open(os.path.join(ROOT, "include", "swe_nn.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(swe_nn_[a-z_0-9]+)\s*\(", text))
    assert names == set(nn.SIGNATURES), names ^ set(nn.SIGNATURES)
    lib = nn.load()
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", nn.LIB_PATH], capture_output=True, text=True).stdout
    assert names <= set(re.findall(r"\sT\s(swe_nn_[a-z_0-9]+)$", out, flags=re.M))
