"""ctypes binding of libswe.so, the C-ABI declared in include/swe.h.

The library is the product: this module only loads it and describes its
structs.  It fails loudly when the library has not been built; there is no
CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswe.so")

SWE_OK = 0
SWE_EINVAL = -1
SWE_EHIP = -2
SWE_ENOMEM = -3
SWE_EAUDIT = -4
SWE_ESTATE = -5

SWE_TUNE_TILE_ORDER = 1
SWE_TUNE_GRAPH = 2
SWE_TUNE_HEAVY_FIRST = 3
SWE_TUNE_SHAPE = 4

SWE_EXPORT_NPZ = 0
SWE_EXPORT_PNG = 1


class SweParams(ctypes.Structure):
    """swe_params (include/swe.h)."""

    _fields_ = [
        ("max_age", ctypes.c_int32),
        ("boundarylayer", ctypes.c_float),
        ("suspension", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("pspeed", ctypes.c_float * 3),
        ("lrate", ctypes.c_float),
        ("maxdiff", ctypes.c_float),
        ("settling", ctypes.c_float),
        ("abrasion", ctypes.c_float),
        ("mapscale", ctypes.c_float),
        ("window", ctypes.c_int32),
        ("periodic", ctypes.c_int32),
        ("track", ctypes.c_int32),
        ("layers", ctypes.c_int32),
        ("deposit_on_death", ctypes.c_int32),
        ("spawn", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("cascade_order", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "pspeed" else v
        return d

    def update(self, **kw):
        for k, v in kw.items():
            if k == "pspeed":
                for i in range(3):
                    self.pspeed[i] = float(v[i])
            else:
                if not hasattr(self, k):
                    raise AttributeError("swe_params has no field %r" % k)
                setattr(self, k, v)
        return self


class SweAudit(ctypes.Structure):
    _fields_ = [
        ("height_sum_lo", ctypes.c_int64),
        ("height_sum_hi", ctypes.c_int64),
        ("sediment_sum_lo", ctypes.c_int64),
        ("sediment_sum_hi", ctypes.c_int64),
        ("lost", ctypes.c_int64),
        ("cycle", ctypes.c_uint64),
    ]

    @property
    def height_sum(self):
        return (self.height_sum_hi << 64) + (self.height_sum_lo & 0xFFFFFFFFFFFFFFFF)

    @property
    def sediment_sum(self):
        return (self.sediment_sum_hi << 64) + (self.sediment_sum_lo & 0xFFFFFFFFFFFFFFFF)


class SweStats(ctypes.Structure):
    _fields_ = [
        ("cycles", ctypes.c_uint64),
        ("particle_steps", ctypes.c_uint64),
        ("march_ms", ctypes.c_double),
        ("march_launches", ctypes.c_uint64),
        ("relax_ms", ctypes.c_double),
        ("relax_launches", ctypes.c_uint64),
        ("max_tile_particles", ctypes.c_uint64),
        ("crowded_tile_windows", ctypes.c_uint64),
        ("far_particles", ctypes.c_uint64),
        ("heavy_tile_windows", ctypes.c_uint64),
        ("light_tile_windows", ctypes.c_uint64),
    ]


class SweTuneResult(ctypes.Structure):
    _fields_ = [
        ("tile_order", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("candidates", ctypes.c_int32),
        ("heavy_first", ctypes.c_int32),
        ("ms_per_cycle", ctypes.c_double),
        ("shape", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


# name -> (restype, argtypes); every symbol include/swe.h declares
_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_I64P = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = {
    "swe_params_default": (None, [ctypes.POINTER(SweParams)]),
    "swe_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "swe_destroy": (ctypes.c_int, [_P]),
    "swe_last_error": (ctypes.c_char_p, [_P]),
    "swe_set_params": (ctypes.c_int, [_P, ctypes.POINTER(SweParams)]),
    "swe_get_params": (ctypes.c_int, [_P, ctypes.POINTER(SweParams)]),
    "swe_upload_layers": (ctypes.c_int, [_P, _FP, _FP]),
    "swe_download_layers": (ctypes.c_int, [_P, _FP, _FP, _FP, _FP]),
    "swe_download_master": (ctypes.c_int, [_P, _I64P, _I64P]),
    "swe_upload_master": (ctypes.c_int, [_P, _I64P, _I64P]),
    "swe_upload_fields": (ctypes.c_int, [_P, _FP, _FP]),
    "swe_erode": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32]),
    "swe_synchronize": (ctypes.c_int, [_P]),
    "swe_set_path": (ctypes.c_int, [_P, ctypes.c_int32]),
    "swe_audit": (ctypes.c_int, [_P, ctypes.POINTER(SweAudit)]),
    "swe_get_cycle": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "swe_set_cycle": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "swe_set_lost": (ctypes.c_int, [_P, ctypes.c_int64]),
    "swe_get_size": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "swe_save_state": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "swe_load_state": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "swe_set_timing": (ctypes.c_int, [_P, ctypes.c_int32]),
    "swe_set_tune": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int64]),
    "swe_autotune": (ctypes.c_int, [_P, _I64P, ctypes.c_int32, ctypes.POINTER(SweTuneResult)]),
    "swe_device_arch": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int32]),
    "swe_get_tune": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "swe_stats": (ctypes.c_int, [_P, ctypes.POINTER(SweStats)]),
    "swe_reset_stats": (ctypes.c_int, [_P]),
    "swe_normals": (ctypes.c_int, [_P, _FP]),
    "swe_normals_noalloc": (ctypes.c_int, [_P, _FP]),
    "swe_debug_step": (ctypes.c_int, [_P, ctypes.c_int32, _FP, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32), _I64P]),
    "swe_debug_cascade": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_int32), _I64P, _I64P]),
    "swe_set_rng": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "swe_windmap": (ctypes.c_int, [_P, _FP]),
    "swe_mesh": (ctypes.c_int, [_P, _FP]),
    "swe_generate_terrain": (
        ctypes.c_int,
        [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float, ctypes.c_int32,
         ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _FP],
    ),
    "swe_version": (ctypes.c_char_p, []),
    "swe_create_batch": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "swe_get_terrains": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "swe_set_terrain": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_uint64, _FP]),
    "swe_upload_terrain": (ctypes.c_int, [_P, ctypes.c_int32, _FP, _FP]),
    "swe_download_terrain": (ctypes.c_int, [_P, ctypes.c_int32, _FP, _FP, _FP, _FP]),
    "swe_download_terrain_master": (ctypes.c_int, [_P, ctypes.c_int32, _I64P, _I64P]),
    "swe_audit_terrain": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(SweAudit)]),
    "swe_erode_batch": (ctypes.c_int, [_P, _I64P, ctypes.c_int32]),
    "swe_get_terrain": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint64), _FP,
                                       ctypes.POINTER(ctypes.c_int32)]),
    "swe_upload_terrain_master": (ctypes.c_int, [_P, ctypes.c_int32, _I64P, _I64P]),
    "swe_upload_terrain_fields": (ctypes.c_int, [_P, ctypes.c_int32, _FP, _FP]),
    "swe_set_terrain_lost": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int64]),
    "swe_windmap_terrain": (ctypes.c_int, [_P, ctypes.c_int32, _FP]),
    "swe_normals_terrain": (ctypes.c_int, [_P, ctypes.c_int32, _FP]),
    "swe_export": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int32]),
    "swe_export_terrain": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32]),
}

_lib = None


def load():
    """Load libswe.so (built in-tree by simplewinderosion_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libswe.so is not built (%s); run `python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def default_params():
    p = SweParams()
    load().swe_params_default(ctypes.byref(p))
    return p
