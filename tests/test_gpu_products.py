"""Flow-field products written to files (npz, PNG) and batch checkpoints,
through the C-ABI on the GPU."""
import struct
import zlib

import numpy as np
import pytest

import oracle as O
from simplewinderosion_amd import Batch, World, Wind, glm_normalize
from simplewinderosion_amd.world import build_params

pytestmark = pytest.mark.gpu


def read_png_rgba8(path):
    """Minimal PNG reader for the 8-bit RGBA, filter-0 images the library writes."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF, kind
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype = hdr[:4]
    assert depth == 8 and ctype == 6
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert not rows[:, 0].any()  # filter type 0 on every row
    return rows[:, 1:].reshape(h, w, 4)


def _eroded_world(nx, ny, periodic):
    Wind.suspension = 1e-2
    World.settling = 0.05
    World.SEED = 5
    rng = np.random.default_rng(11)
    h = (rng.random((nx, ny)) * 0.4).astype(np.float32)
    w = World(nx, ny, periodic=periodic, layers=2, abrasion=1e-3)
    w.upload(h, np.full((nx, ny), 0.004, np.float32))
    w.erode(nx * ny // 2, n_cycles=2)
    w.synchronize()
    return w


@pytest.mark.parametrize("periodic", [True, False])
def test_npz_export_holds_the_downloaded_fields(statics, tmp_path, periodic):
    print("case npz export periodic=%s" % periodic, flush=True)
    w = _eroded_world(48, 80, periodic)
    path = tmp_path / "fields.npz"
    w.export(path, "npz")
    d = w.download()
    zm, sm = w.master()
    with np.load(path) as z:
        assert sorted(z.files) == sorted(["height", "sediment", "massflow", "momentum", "windmap", "normals",
                                          "height_fx", "sediment_fx"])
        for k in ("height", "sediment", "massflow", "momentum"):
            assert z[k].dtype == np.float32 and np.array_equal(z[k], d[k]), k
        assert np.array_equal(z["windmap"], w.windmap())
        assert np.array_equal(z["normals"], w.normals())
        assert z["height_fx"].dtype == np.int64 and np.array_equal(z["height_fx"], zm)
        assert np.array_equal(z["sediment_fx"], sm)
        assert z["windmap"].shape == (48, 80, 4) and z["momentum"].shape == (3, 48, 80)
    w.close()


def test_png_export_is_the_wind_map(statics, tmp_path):
    print("case png export", flush=True)
    w = _eroded_world(40, 56, True)
    path = tmp_path / "wind.png"
    w.export(path, "png")
    img = read_png_rgba8(path)  # rows = y, columns = x
    wm = w.windmap()            # [x, y, rgba]
    assert img.shape == (56, 40, 4)
    want = np.clip(np.rint(255.0 * np.clip(wm.astype(np.float64), 0, 1)), 0, 255).astype(np.uint8)
    assert np.array_equal(img, want.transpose(1, 0, 2))
    assert (img[..., 3] == 255).all()
    with pytest.raises(ValueError):
        w.export(tmp_path / "x.bin", "tiff")
    w.close()


def test_batch_checkpoint_resume_matches_uninterrupted_run_and_oracle(statics, tmp_path):
    """A three-terrain batch saved after 2 cycles and resumed in a fresh batch
    equals the uninterrupted 4-cycle run on every terrain; terrain 1 also
    equals the oracle's 4 cycles."""
    sizes = [(64, 96), (40, 72), (33, 50)]
    print("case batch checkpoint", sizes, flush=True)
    Wind.suspension = 1e-2
    World.settling = 0.05
    rng = np.random.default_rng(2)
    winds = [glm_normalize((np.cos(a), 0.5, np.sin(a))) for a in rng.uniform(0, 6.28, len(sizes))]
    heights = [(rng.random(s) * 0.4).astype(np.float32) for s in sizes]
    seds = [np.full(s, 0.003, np.float32) for s in sizes]
    parts = [a * b // 2 for a, b in sizes]

    def fresh():
        b = Batch(sizes, periodic=True, layers=2, abrasion=1e-3)
        for k in range(len(sizes)):
            b.set_terrain(k, 100 + k, winds[k])
            b.upload(k, heights[k], seds[k])
        return b

    b = fresh()
    b.erode(parts, n_cycles=2)
    path = tmp_path / "batch.swe"
    b.save(path)
    b.erode(parts, n_cycles=2)
    ref = [(b.master(k), b.download(k), b.audit(k)) for k in range(len(sizes))]
    b.close()

    World.settling = 0.6  # the checkpoint's params must win
    b2 = Batch(sizes, periodic=True, layers=1)
    b2.load(path)
    for k in range(len(sizes)):
        nx, ny, seed, ps = b2.terrain(k)
        assert (nx, ny, seed) == (sizes[k][0], sizes[k][1], 100 + k)
        assert np.array_equal(np.float32(ps), np.float32(winds[k]))
    b2.erode(parts, n_cycles=2)
    for k in range(len(sizes)):
        (zr, sr), dr, ar = ref[k]
        zg, sg = b2.master(k)
        dg = b2.download(k)
        assert np.array_equal(zr, zg) and np.array_equal(sr, sg), k
        for f in ("massflow", "momentum", "height", "sediment"):
            assert np.array_equal(dr[f], dg[f]), (k, f)
        assert ar == b2.audit(k), k
    k = 1
    World.SEED = 100 + k
    Wind.pspeed = winds[k]
    p = build_params(abrasion=1e-3, periodic=True, layers=2)
    st = O.OracleState(*sizes[k], heights[k], sediment=seds[k], params=O.params_from(p))
    st.erode(parts[k], 4)
    zg, sg = b2.master(k)
    assert np.array_equal(zg.ravel(), st.zm) and np.array_equal(sg.ravel(), st.sm)
    b2.close()


def test_checkpoint_refuses_a_batch_of_other_sizes(statics, tmp_path):
    print("case checkpoint size mismatch", flush=True)
    b = Batch([(32, 32), (16, 48)], periodic=True)
    b.save(tmp_path / "b.swe")
    c = Batch([(32, 32), (48, 16)], periodic=True)
    with pytest.raises(Exception):
        c.load(tmp_path / "b.swe")
    w = World(32, 32, periodic=True)
    with pytest.raises(Exception):
        w.load(tmp_path / "b.swe")
    b.close()
    c.close()
    w.close()
