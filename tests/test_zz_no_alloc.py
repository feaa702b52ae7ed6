"""swe_normals_noalloc (no device allocation) against swe_normals through the
C-ABI: two calls in a row on two contexts, the second map smaller, and the
size limit of its staging array."""
import ctypes

import numpy as np
import pytest

from simplewinderosion_amd import capi

pytestmark = pytest.mark.gpu


def _ctx(lib, nx, ny, seed, periodic):
    ctx = ctypes.c_void_p()
    assert lib.swe_create(ctypes.byref(ctx), nx, ny, 0) == 0
    p = capi.default_params()
    p.periodic = periodic
    assert lib.swe_set_params(ctx, ctypes.byref(p)) == 0
    h = np.random.default_rng(seed).random((nx, ny), dtype=np.float32) * 0.5
    h = np.ascontiguousarray(h)
    assert lib.swe_upload_layers(ctx, h.ctypes.data_as(capi._FP), None) == 0
    return ctx


def _normals(lib, fn, ctx, n):
    out = np.full(3 * n, np.nan, dtype=np.float32)
    rc = getattr(lib, fn)(ctx, out.ctypes.data_as(capi._FP))
    assert rc == 0, (fn, rc, lib.swe_last_error(ctx))
    return out


@pytest.mark.parametrize("periodic", [1, 0])
def test_noalloc_normals_equal_allocating_call(periodic):
    lib = capi.load()
    # the first map is the larger one: the second call's staging still holds
    # the first call's values beyond the smaller map
    for nx, ny, seed in [(300, 211, 1), (96, 40, 2)]:
        print("noalloc normals: %dx%d periodic=%d" % (nx, ny, periodic), flush=True)
        ctx = _ctx(lib, nx, ny, seed, periodic)
        try:
            ref = _normals(lib, "swe_normals", ctx, nx * ny)
            got = _normals(lib, "swe_normals_noalloc", ctx, nx * ny)
            assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))
            again = _normals(lib, "swe_normals_noalloc", ctx, nx * ny)
            assert np.array_equal(ref.view(np.uint32), again.view(np.uint32))
        finally:
            lib.swe_destroy(ctx)


def test_noalloc_normals_rejects_maps_beyond_staging():
    lib = capi.load()
    ctx = ctypes.c_void_p()
    assert lib.swe_create(ctypes.byref(ctx), 1025, 1024, 0) == 0
    try:
        out = np.zeros(3 * 1025 * 1024, dtype=np.float32)
        rc = lib.swe_normals_noalloc(ctx, out.ctypes.data_as(capi._FP))
        assert rc == capi.SWE_EINVAL
        assert b"exceed" in lib.swe_last_error(ctx)
        assert not out.any()
    finally:
        lib.swe_destroy(ctx)
