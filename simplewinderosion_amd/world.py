"""Python mirror of the reference's simulation interface over libswe.so.

``Wind`` carries the parameter statics of ``struct Wind`` (wind.hpp:13,15-26)
and ``World`` those of ``class World`` (world.hpp:12-32) plus
``World.erode(cycles)`` (world.hpp:21,36-77).  As in the reference the
parameters are class-level statics read at each ``erode`` call; the map and
the extensions (periodic boundary, sediment layer, snapshot window) belong to
one ``World`` instance, which owns one device context.
"""
import ctypes

import numpy as np

from . import capi


class Wind:
    """Statics of struct Wind (wind.hpp:15-26); pspeed is the member default (wind.hpp:13)."""

    maxAge = 512
    boundarylayer = 2.0
    suspension = 0.05
    gravity = 0.1
    pspeed = (1.0, 0.0, 0.0)  # glm::normalize(vec3(1,0,0))


def glm_normalize(v):
    """glm::normalize in float32 with glm's operation order: v * (1/sqrt(dot(v,v)))."""
    v = np.asarray(v, dtype=np.float32)
    d = np.float32(np.float32(v[0] * v[0]) + np.float32(v[1] * v[1])) + np.float32(v[2] * v[2])
    r = np.float32(1.0) / np.sqrt(np.float32(d))
    return tuple(float(np.float32(c * r)) for c in v)


class SweError(RuntimeError):
    pass


def build_params(abrasion=0.0, mapscale=80.0, window=1, periodic=False, track=True, layers=1,
                 deposit_on_death=False, spawn=0, cascade_order=-1):
    """swe_params from the Wind/World statics plus per-map options (no device needed)."""
    p = capi.SweParams()
    capi.load().swe_params_default(ctypes.byref(p))
    p.update(
        max_age=int(Wind.maxAge), boundarylayer=float(Wind.boundarylayer),
        suspension=float(Wind.suspension), gravity=float(Wind.gravity), pspeed=Wind.pspeed,
        lrate=float(World.lrate), maxdiff=float(World.maxdiff), settling=float(World.settling),
        abrasion=float(abrasion), mapscale=float(mapscale), window=int(window), periodic=int(bool(periodic)),
        track=int(bool(track)), layers=int(layers), deposit_on_death=int(bool(deposit_on_death)),
        spawn=int(spawn), seed=int(World.SEED) & 0xFFFFFFFFFFFFFFFF, cascade_order=int(cascade_order),
    )
    return p


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)) if a is not None else None


class World:
    """class World (world.hpp:9-24) bound to one libswe context on one GPU."""

    SEED = 0
    lrate = 0.1
    maxdiff = 0.005
    settling = 1.0

    def __init__(self, nx=512, ny=512, device=0, periodic=False, track=True, layers=1, window=1,
                 abrasion=0.0, deposit_on_death=False, mapscale=80.0, spawn=0, cascade_order=-1):
        self.lib = capi.load()
        self.cascade_order = int(cascade_order)
        self.nx, self.ny = int(nx), int(ny)
        self.periodic = bool(periodic)
        self.track = bool(track)
        self.layers = int(layers)
        self.window = int(window)
        self.abrasion = float(abrasion)
        self.deposit_on_death = bool(deposit_on_death)
        self.mapscale = float(mapscale)
        self.spawn = int(spawn)
        self.ctx = ctypes.c_void_p()
        rc = self.lib.swe_create(ctypes.byref(self.ctx), self.nx, self.ny, int(device))
        if rc != 0:
            raise SweError("swe_create(%d, %d, device=%d) failed with %d" % (nx, ny, device, rc))
        self._push()

    # -- parameters ---------------------------------------------------------
    def options(self):
        return dict(abrasion=self.abrasion, mapscale=self.mapscale, window=self.window,
                    periodic=self.periodic, track=self.track, layers=self.layers,
                    deposit_on_death=self.deposit_on_death, spawn=self.spawn, cascade_order=self.cascade_order)

    def params(self):
        return build_params(**self.options())

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.swe_last_error(self.ctx)
            raise SweError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def _push(self):
        p = self.params()
        self._check(self.lib.swe_set_params(self.ctx, ctypes.byref(p)), "swe_set_params")

    # -- World::erode --------------------------------------------------------
    def erode(self, cycles, n_cycles=1):
        """World::erode(cycles): `cycles` particles per map-cycle (world.hpp:36-77)."""
        self._push()
        self._check(self.lib.swe_erode(self.ctx, int(cycles), int(n_cycles)), "swe_erode")

    def set_path(self, path):
        """0 automatic, 1 global-memory march, 2 tile-owned step kernel (window 1)."""
        self._check(self.lib.swe_set_path(self.ctx, int(path)), "swe_set_path")

    def synchronize(self):
        self._check(self.lib.swe_synchronize(self.ctx), "swe_synchronize")

    # -- map transfer ----------------------------------------------------------
    def upload(self, height, sediment=None):
        h = np.ascontiguousarray(height, dtype=np.float32).reshape(self.nx * self.ny)
        s = None if sediment is None else np.ascontiguousarray(sediment, dtype=np.float32).reshape(self.nx * self.ny)
        self._check(self.lib.swe_upload_layers(self.ctx, _fp(h), _fp(s)), "swe_upload_layers")

    def _flat(self, a, dtype, count, what):
        """a as a contiguous 1-D array of exactly `count` elements: the C side
        copies that many from the pointer."""
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.size != count:
            raise ValueError("%s: %d values for a %dx%d map, expected %d" % (what, a.size, self.nx, self.ny, count))
        return a.reshape(count)

    def upload_master(self, height_fx, sediment_fx=None):
        n = self.nx * self.ny
        h = self._flat(height_fx, np.int64, n, "height_fx")
        s = None if sediment_fx is None else self._flat(sediment_fx, np.int64, n, "sediment_fx")
        self._check(self.lib.swe_upload_master(self.ctx, _ip(h), _ip(s)), "swe_upload_master")

    def upload_fields(self, massflow=None, momentum=None):
        n = self.nx * self.ny
        m = None if massflow is None else self._flat(massflow, np.float32, n, "massflow")
        mo = None if momentum is None else self._flat(momentum, np.float32, 3 * n, "momentum")
        self._check(self.lib.swe_upload_fields(self.ctx, _fp(m), _fp(mo)), "swe_upload_fields")

    def download(self):
        n = self.nx * self.ny
        h = np.empty(n, np.float32)
        s = np.empty(n, np.float32)
        mf = np.empty(n, np.float32)
        mo = np.empty(3 * n, np.float32)
        self._check(self.lib.swe_download_layers(self.ctx, _fp(h), _fp(s), _fp(mf), _fp(mo)), "swe_download_layers")
        shp = (self.nx, self.ny)
        return {"height": h.reshape(shp), "sediment": s.reshape(shp), "massflow": mf.reshape(shp),
                "momentum": mo.reshape((3,) + shp)}

    def master(self):
        n = self.nx * self.ny
        h = np.empty(n, np.int64)
        s = np.empty(n, np.int64)
        self._check(self.lib.swe_download_master(self.ctx, _ip(h), _ip(s)), "swe_download_master")
        return h.reshape(self.nx, self.ny), s.reshape(self.nx, self.ny)

    def windmap(self):
        """RGBA wind map of the EMA momentum (SimpleWindErosion.cpp:224-230), [nx, ny, 4]."""
        out = np.empty(4 * self.nx * self.ny, np.float32)
        self._check(self.lib.swe_windmap(self.ctx, _fp(out)), "swe_windmap")
        return out.reshape(self.nx, self.ny, 4)

    def mesh(self):
        """updatenode's vertex attributes (cellpool.h:302-333): [nx, ny, 12] = P, normal, T-P, B-P."""
        out = np.empty(12 * self.nx * self.ny, np.float32)
        self._check(self.lib.swe_mesh(self.ctx, _fp(out)), "swe_mesh")
        return out.reshape(self.nx, self.ny, 12)

    def normals(self):
        out = np.empty(3 * self.nx * self.ny, np.float32)
        self._check(self.lib.swe_normals(self.ctx, _fp(out)), "swe_normals")
        return out.reshape(3, self.nx, self.ny)

    # -- known-answer entry points -------------------------------------------
    def debug_step(self, state, age, key):
        """One fly() call (wind.hpp:28-117) of each particle against the
        current snapshot: state [n, 7] (pos xyz, speed xyz, sediment), age [n]
        calls made, key [n] RNG keys.  Returns (state, age, alive, pickup)."""
        st = np.ascontiguousarray(state, np.float32).reshape(-1, 7).copy()
        n = st.shape[0]
        ag = np.ascontiguousarray(age, np.int32).reshape(-1).copy()
        ky = np.ascontiguousarray(key, np.uint64).reshape(-1)
        if ag.size != n or ky.size != n:
            raise ValueError("state, age and key describe different particle counts")
        alive = np.zeros(n, np.int32)
        pick = np.zeros(n, np.int64)
        i32 = ctypes.POINTER(ctypes.c_int32)
        self._check(self.lib.swe_debug_step(self.ctx, n, _fp(st), ag.ctypes.data_as(i32),
                                            ky.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                            alive.ctypes.data_as(i32), _ip(pick)), "swe_debug_step")
        return st, ag, alive, pick

    def debug_cascade(self, x, y, events=1):
        """One World::cascade (world.hpp:79-153) at (x, y) with `events`
        events against the current snapshot: [(neighbour cell, fixed-point
        amount received)] for the in-map neighbours in offset order."""
        cnt = ctypes.c_int32()
        cell = np.zeros(8, np.int64)
        fx = np.zeros(8, np.int64)
        self._check(self.lib.swe_debug_cascade(self.ctx, int(x), int(y), int(events), ctypes.byref(cnt),
                                               _ip(cell), _ip(fx)), "swe_debug_cascade")
        return [(int(cell[i]), int(fx[i])) for i in range(cnt.value)]

    def set_rng(self, seed):
        """World::SEED of the particles' counter RNG (swe_set_rng).  SEED is
        a static of World, as in the reference, and erode() pushes the
        statics: the class attribute is set too."""
        World.SEED = int(seed)
        self._check(self.lib.swe_set_rng(self.ctx, int(seed)), "swe_set_rng")

    # -- audit, counters, timing -------------------------------------------------
    def audit(self):
        a = capi.SweAudit()
        self._check(self.lib.swe_audit(self.ctx, ctypes.byref(a)), "swe_audit")
        return {"height_sum": a.height_sum, "sediment_sum": a.sediment_sum, "lost": a.lost, "cycle": a.cycle}

    @property
    def cycle(self):
        c = ctypes.c_uint64()
        self._check(self.lib.swe_get_cycle(self.ctx, ctypes.byref(c)), "swe_get_cycle")
        return c.value

    @cycle.setter
    def cycle(self, v):
        self._check(self.lib.swe_set_cycle(self.ctx, int(v)), "swe_set_cycle")

    def set_lost(self, v):
        self._check(self.lib.swe_set_lost(self.ctx, int(v)), "swe_set_lost")

    def save(self, path):
        """Checkpoint the exact state (masters, EMA fields, params, counters) to `path`."""
        self._check(self.lib.swe_save_state(self.ctx, str(path).encode()), "swe_save_state")

    def load(self, path):
        """Restore a checkpoint written by save(); the map size must match.  The
        statics are set from the file so later erode() calls keep its params."""
        self._check(self.lib.swe_load_state(self.ctx, str(path).encode()), "swe_load_state")
        p = capi.SweParams()
        self._check(self.lib.swe_get_params(self.ctx, ctypes.byref(p)), "swe_get_params")
        _statics_from(p)
        self.abrasion, self.mapscale, self.window = p.abrasion, p.mapscale, p.window
        self.periodic, self.track, self.layers = bool(p.periodic), bool(p.track), p.layers
        self.deposit_on_death = bool(p.deposit_on_death)
        self.spawn = p.spawn
        self.cascade_order = p.cascade_order

    def export(self, path, fmt="npz"):
        """Write the flow-field and map products: fmt "npz" (height, sediment,
        massflow, momentum, windmap, normals and the exact masters) or "png"
        (the wind map of SimpleWindErosion.cpp:224-230 as RGBA8)."""
        self._check(self.lib.swe_export(self.ctx, str(path).encode(), _export_format(fmt)), "swe_export")

    def set_timing(self, on):
        """False/0: off; True/1: time every kernel launch; n > 1: every n-th."""
        self._check(self.lib.swe_set_timing(self.ctx, int(on)), "swe_set_timing")

    def set_tune(self, knob, value):
        """Launch tuning knob (capi.SWE_TUNE_*); changes speed, never results."""
        self._check(self.lib.swe_set_tune(self.ctx, int(knob), int(value)), "swe_set_tune")

    def get_tune(self, knob):
        v = ctypes.c_int64()
        self._check(self.lib.swe_get_tune(self.ctx, int(knob), ctypes.byref(v)), "swe_get_tune")
        return v.value

    def autotune(self, particles, trial_cycles=2, cache=None, retune=False):
        """Pick the fastest launch setting for this workload (map state is
        kept); see simplewinderosion_amd.tuning."""
        from . import tuning

        return tuning.autotune(self, particles, trial_cycles=trial_cycles,
                               cache=tuning.DEFAULT_CACHE if cache is None else cache, retune=retune)

    def stats(self):
        s = capi.SweStats()
        self._check(self.lib.swe_stats(self.ctx, ctypes.byref(s)), "swe_stats")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def reset_stats(self):
        self._check(self.lib.swe_reset_stats(self.ctx), "swe_reset_stats")

    def close(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self.lib.swe_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def generate_terrain(nx, ny, seed, octaves=6, base_cycles=4.0, n_pyramids=0, pyr_h=(0.3, 0.6),
                     pyr_base=(40.0, 120.0)):
    """Seeded periodic fBm terrain in [0,1] plus optional square pyramids (host-side)."""
    lib = capi.load()
    out = np.empty(int(nx) * int(ny), np.float32)
    rc = lib.swe_generate_terrain(int(nx), int(ny), int(seed) & 0xFFFFFFFFFFFFFFFF, int(octaves),
                                  float(base_cycles), int(n_pyramids), float(pyr_h[0]), float(pyr_h[1]),
                                  float(pyr_base[0]), float(pyr_base[1]), _fp(out))
    if rc != 0:
        raise SweError("swe_generate_terrain failed (%d)" % rc)
    return out.reshape(int(nx), int(ny))


def _statics_from(p):
    """Set the Wind / World statics from a swe_params (after a checkpoint load)."""
    Wind.maxAge, Wind.boundarylayer, Wind.suspension, Wind.gravity = (
        p.max_age, p.boundarylayer, p.suspension, p.gravity)
    Wind.pspeed = tuple(p.pspeed)
    World.lrate, World.maxdiff, World.settling, World.SEED = p.lrate, p.maxdiff, p.settling, p.seed


def _export_format(fmt):
    formats = {"npz": capi.SWE_EXPORT_NPZ, "png": capi.SWE_EXPORT_PNG}
    if fmt not in formats:
        raise ValueError("export format must be 'npz' or 'png', not %r" % (fmt,))
    return formats[fmt]


class Batch:
    """Many terrains of different sizes eroded together: one launch per
    snapshot window covers every terrain (window 1).  Shared parameters come
    from the Wind / World statics and the options; each terrain has its own
    seed and prevailing wind."""

    def __init__(self, sizes, device=0, periodic=True, track=True, layers=1, abrasion=0.0,
                 deposit_on_death=False, mapscale=80.0, spawn=0, cascade_order=-1):
        self.lib = capi.load()
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        n = len(self.sizes)
        nx = (ctypes.c_int32 * n)(*[a for a, _ in self.sizes])
        ny = (ctypes.c_int32 * n)(*[b for _, b in self.sizes])
        self.opts = dict(abrasion=abrasion, mapscale=mapscale, window=1, periodic=periodic, track=track,
                         layers=layers, deposit_on_death=deposit_on_death, spawn=spawn,
                         cascade_order=cascade_order)
        self.ctx = ctypes.c_void_p()
        rc = self.lib.swe_create_batch(ctypes.byref(self.ctx), n, nx, ny, int(device))
        if rc != 0:
            raise SweError("swe_create_batch(%d terrains) failed with %d" % (n, rc))
        self._push()

    _check = World._check

    def _push(self):
        p = build_params(**self.opts)
        self._check(self.lib.swe_set_params(self.ctx, ctypes.byref(p)), "swe_set_params")

    def params(self):
        return build_params(**self.opts)

    def set_terrain(self, k, seed, pspeed=None):
        ps = None if pspeed is None else (ctypes.c_float * 3)(*[float(v) for v in pspeed])
        self._check(self.lib.swe_set_terrain(self.ctx, int(k), int(seed) & 0xFFFFFFFFFFFFFFFF, ps),
                    "swe_set_terrain")

    def upload(self, k, height, sediment=None):
        nx, ny = self.sizes[k]
        h = np.ascontiguousarray(height, dtype=np.float32).reshape(nx * ny)
        s = None if sediment is None else np.ascontiguousarray(sediment, dtype=np.float32).reshape(nx * ny)
        self._check(self.lib.swe_upload_terrain(self.ctx, int(k), _fp(h), _fp(s)), "swe_upload_terrain")

    def erode(self, particles, n_cycles=1):
        """particles: per-terrain counts (one World::erode(N_k) each per cycle)."""
        self._push()
        arr = (ctypes.c_int64 * len(self.sizes))(*[int(v) for v in particles])
        self._check(self.lib.swe_erode_batch(self.ctx, arr, int(n_cycles)), "swe_erode_batch")

    def synchronize(self):
        self._check(self.lib.swe_synchronize(self.ctx), "swe_synchronize")

    def set_timing(self, on):
        """False/0: off; True/1: time every kernel launch; n > 1: every n-th."""
        self._check(self.lib.swe_set_timing(self.ctx, int(on)), "swe_set_timing")

    def set_tune(self, knob, value):
        """Launch tuning knob (capi.SWE_TUNE_*); changes speed, never results."""
        self._check(self.lib.swe_set_tune(self.ctx, int(knob), int(value)), "swe_set_tune")

    def get_tune(self, knob):
        v = ctypes.c_int64()
        self._check(self.lib.swe_get_tune(self.ctx, int(knob), ctypes.byref(v)), "swe_get_tune")
        return v.value

    def autotune(self, particles, trial_cycles=2, cache=None, retune=False):
        """Pick the fastest launch setting for this workload (map state is
        kept); see simplewinderosion_amd.tuning."""
        from . import tuning

        return tuning.autotune(self, particles, trial_cycles=trial_cycles,
                               cache=tuning.DEFAULT_CACHE if cache is None else cache, retune=retune)

    def stats(self):
        s = capi.SweStats()
        self._check(self.lib.swe_stats(self.ctx, ctypes.byref(s)), "swe_stats")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def reset_stats(self):
        self._check(self.lib.swe_reset_stats(self.ctx), "swe_reset_stats")

    def download(self, k):
        nx, ny = self.sizes[k]
        n = nx * ny
        h, s, mf, mo = (np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.float32),
                        np.empty(3 * n, np.float32))
        self._check(self.lib.swe_download_terrain(self.ctx, int(k), _fp(h), _fp(s), _fp(mf), _fp(mo)),
                    "swe_download_terrain")
        return {"height": h.reshape(nx, ny), "sediment": s.reshape(nx, ny), "massflow": mf.reshape(nx, ny),
                "momentum": mo.reshape(3, nx, ny)}

    def master(self, k):
        nx, ny = self.sizes[k]
        h, s = np.empty(nx * ny, np.int64), np.empty(nx * ny, np.int64)
        self._check(self.lib.swe_download_terrain_master(self.ctx, int(k), _ip(h), _ip(s)),
                    "swe_download_terrain_master")
        return h.reshape(nx, ny), s.reshape(nx, ny)

    def audit(self, k):
        a = capi.SweAudit()
        self._check(self.lib.swe_audit_terrain(self.ctx, int(k), ctypes.byref(a)), "swe_audit_terrain")
        return {"height_sum": a.height_sum, "sediment_sum": a.sediment_sum, "lost": a.lost, "cycle": a.cycle}

    def windmap(self, k):
        nx, ny = self.sizes[k]
        out = np.empty(4 * nx * ny, np.float32)
        self._check(self.lib.swe_windmap_terrain(self.ctx, int(k), _fp(out)), "swe_windmap_terrain")
        return out.reshape(nx, ny, 4)

    def normals(self, k):
        nx, ny = self.sizes[k]
        out = np.empty(3 * nx * ny, np.float32)
        self._check(self.lib.swe_normals_terrain(self.ctx, int(k), _fp(out)), "swe_normals_terrain")
        return out.reshape(3, nx, ny)

    def terrain(self, k):
        """Terrain k's (nx, ny, seed, pspeed) as the context has them."""
        nx, ny, ov = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        seed = ctypes.c_uint64()
        ps = (ctypes.c_float * 3)()
        self._check(self.lib.swe_get_terrain(self.ctx, int(k), ctypes.byref(nx), ctypes.byref(ny),
                                             ctypes.byref(seed), ps, ctypes.byref(ov)), "swe_get_terrain")
        return nx.value, ny.value, seed.value, tuple(ps)

    def save(self, path):
        """Checkpoint every terrain's exact state, its seed and wind, the shared
        params and the counters to `path`."""
        self._check(self.lib.swe_save_state(self.ctx, str(path).encode()), "swe_save_state")

    def load(self, path):
        """Restore a checkpoint written by save() into a batch of the same
        terrain sizes; the statics and options follow the file."""
        self._check(self.lib.swe_load_state(self.ctx, str(path).encode()), "swe_load_state")
        p = capi.SweParams()
        self._check(self.lib.swe_get_params(self.ctx, ctypes.byref(p)), "swe_get_params")
        _statics_from(p)
        self.opts = dict(abrasion=p.abrasion, mapscale=p.mapscale, window=1, periodic=bool(p.periodic),
                         track=bool(p.track), layers=p.layers, deposit_on_death=bool(p.deposit_on_death),
                         spawn=p.spawn, cascade_order=p.cascade_order)

    def export(self, k, path, fmt="npz"):
        """World.export for terrain k."""
        self._check(self.lib.swe_export_terrain(self.ctx, int(k), str(path).encode(), _export_format(fmt)),
                    "swe_export_terrain")

    def close(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self.lib.swe_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
