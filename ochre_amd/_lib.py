"""ctypes binding of libochre_hip.so (the C ABI in include/ochre.h)."""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libochre_hip.so")

OCHRE_OK = 0
STATUS_NAMES = {
    0: "OCHRE_OK",
    -1: "OCHRE_ERR_INVALID_ARG",
    -2: "OCHRE_ERR_OUT_OF_MEMORY",
    -3: "OCHRE_ERR_DEVICE",
    -4: "OCHRE_ERR_CAPACITY",
    -5: "OCHRE_ERR_FLATTEN_LIMIT",
    -6: "OCHRE_ERR_NO_DEVICE",
    -7: "OCHRE_ERR_TOO_LARGE",
}

# Every symbol include/ochre.h declares.
EXPORTS = [
    "ochre_rasterizer_new", "ochre_rasterizer_move_to", "ochre_rasterizer_line_to",
    "ochre_rasterizer_command", "ochre_rasterizer_fill", "ochre_rasterizer_finish",
    "ochre_rasterizer_free", "ochre_context_create", "ochre_context_destroy",
    "ochre_rasterize_plain", "ochre_context_set_timing", "ochre_context_timing",
    "ochre_build_info", "ochre_rasterize_batched", "ochre_context_set_batched_params",
    "ochre_context_get_batched_params", "ochre_autotune_batched", "ochre_coverage_adjoint",
    "ochre_rasterizer_stroke", "ochre_stroke_plain", "ochre_stroke_batched",
    "ochre_context_set_dense_params", "ochre_context_get_dense_params", "ochre_build_atlas",
    "ochre_context_set_dense_small", "ochre_context_get_dense_small", "ochre_context_set_pass_limits",
    "ochre_context_set_dense_large", "ochre_context_get_dense_large",
    "ochre_draw_plain", "ochre_draw_batched",
    "ochre_svg_parse", "ochre_svg_info", "ochre_svg_verbs", "ochre_svg_verb_offsets", "ochre_svg_points",
    "ochre_svg_point_offsets", "ochre_svg_draws", "ochre_svg_free", "ochre_rasterize_batched_noalloc",
    "ochre_autotune_batched_cached", "ochre_rasterizer_finish_many",
]


class OchreError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}")


class Transform(ctypes.Structure):
    _fields_ = [("m", ctypes.c_float * 4), ("offset", ctypes.c_float * 2)]


class TileBuilderC(ctypes.Structure):
    _fields_ = [("user", ctypes.c_void_p), ("tile", ctypes.c_void_p), ("span", ctypes.c_void_p)]


class PathBatchC(ctypes.Structure):
    _fields_ = [("n_paths", ctypes.c_int64), ("verbs", ctypes.c_void_p), ("verb_off", ctypes.c_void_p),
                ("points", ctypes.c_void_p), ("point_off", ctypes.c_void_p), ("weights", ctypes.c_void_p),
                ("weight_off", ctypes.c_void_p), ("transforms", ctypes.c_void_p),
                ("n_verbs", ctypes.c_int64), ("n_points", ctypes.c_int64), ("n_weights", ctypes.c_int64)]


class BatchOutputC(ctypes.Structure):
    _fields_ = [("tile_xy", ctypes.c_void_p), ("tile_alpha", ctypes.c_void_p), ("tile_capacity", ctypes.c_int64),
                ("span_xy", ctypes.c_void_p), ("span_width", ctypes.c_void_p), ("span_capacity", ctypes.c_int64),
                ("path_tile_start", ctypes.c_void_p), ("path_tile_count", ctypes.c_void_p),
                ("path_span_start", ctypes.c_void_p), ("path_span_count", ctypes.c_void_p),
                ("n_tiles", ctypes.c_int64), ("n_spans", ctypes.c_int64), ("tile_coverage", ctypes.c_void_p)]


class CoverageTargetC(ctypes.Structure):
    _fields_ = [("tile_start", ctypes.c_void_p), ("tile_count", ctypes.c_void_p), ("tile_xy", ctypes.c_void_p),
                ("tile_coverage", ctypes.c_void_p), ("span_start", ctypes.c_void_p), ("span_count", ctypes.c_void_p),
                ("span_xy", ctypes.c_void_p), ("span_width", ctypes.c_void_p)]


class BatchedTuningC(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int), ("paths_per_pop", ctypes.c_int), ("tiles_per_chunk", ctypes.c_int),
                ("ms", ctypes.c_float), ("dense_groups_per_cu", ctypes.c_int),
                ("dense_waves_per_group", ctypes.c_int), ("dense_small_groups_per_cu", ctypes.c_int),
                ("dense_large", ctypes.c_int)]


TILE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int16, ctypes.c_int16, ctypes.POINTER(ctypes.c_uint8))
SPAN_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int16, ctypes.c_int16, ctypes.c_uint16)

_lib = None


def load() -> ctypes.CDLL:
    """Loads the HIP library.  Raises if it has not been built: there is no
    CPU fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m ochre_amd.build`")
    # torch's HIP runtime is the one that owns the device memory handed to the
    # library: load it first, so that the library binds to that runtime rather
    # than to a second copy from the system ROCm
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    P, i64, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
    lib.ochre_rasterizer_new.restype = P
    lib.ochre_rasterizer_new.argtypes = []
    lib.ochre_rasterizer_move_to.argtypes = [P, f32, f32]
    lib.ochre_rasterizer_line_to.argtypes = [P, f32, f32]
    lib.ochre_rasterizer_command.argtypes = [P, ctypes.c_uint8, P, f32]
    lib.ochre_rasterizer_fill.argtypes = [P, P, ctypes.c_size_t, P, ctypes.c_size_t, P, ctypes.c_size_t, P]
    lib.ochre_rasterizer_finish.argtypes = [P, P]
    if hasattr(lib, "ochre_rasterizer_finish_many"):
        lib.ochre_rasterizer_finish_many.argtypes = [ctypes.POINTER(P), ctypes.c_size_t, P]
    lib.ochre_rasterizer_free.argtypes = [P]
    lib.ochre_rasterizer_free.restype = None
    lib.ochre_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(P)]
    lib.ochre_context_destroy.argtypes = [P]
    lib.ochre_context_destroy.restype = None
    lib.ochre_rasterize_plain.argtypes = [P, ctypes.POINTER(PathBatchC), ctypes.POINTER(BatchOutputC), P]
    lib.ochre_context_set_timing.argtypes = [P, ctypes.c_int]
    lib.ochre_context_timing.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(i64)]
    lib.ochre_build_info.restype = ctypes.c_char_p
    _optional(lib)
    _lib = lib
    return lib


def _optional(lib):
    """Binds entry points added after the plain pipeline, when present."""
    P = ctypes.c_void_p
    if hasattr(lib, "ochre_rasterize_batched"):
        lib.ochre_rasterize_batched.argtypes = [P, ctypes.POINTER(PathBatchC), ctypes.POINTER(BatchOutputC), P]
        I = ctypes.c_int
        lib.ochre_context_set_batched_params.argtypes = [P, I, I, I, I]
        lib.ochre_context_get_batched_params.argtypes = [P] + [ctypes.POINTER(I)] * 4
        lib.ochre_coverage_adjoint.argtypes = [P, ctypes.POINTER(PathBatchC), ctypes.POINTER(CoverageTargetC), P, P,
                                               P, P]
        lib.ochre_autotune_batched.argtypes = [P, ctypes.POINTER(PathBatchC), ctypes.POINTER(BatchOutputC), P, I,
                                               ctypes.POINTER(BatchedTuningC), I, ctypes.POINTER(I)]
    if hasattr(lib, "ochre_autotune_batched_cached"):
        lib.ochre_autotune_batched_cached.argtypes = [P, ctypes.POINTER(PathBatchC), ctypes.POINTER(BatchOutputC),
                                                      P, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
    if hasattr(lib, "ochre_context_set_dense_params"):
        I = ctypes.c_int
        lib.ochre_context_set_dense_params.argtypes = [P, I, I, I]
        lib.ochre_context_get_dense_params.argtypes = [P] + [ctypes.POINTER(I)] * 3
        lib.ochre_context_set_dense_small.argtypes = [P, I]
        lib.ochre_context_get_dense_small.argtypes = [P, ctypes.POINTER(I)]
    if hasattr(lib, "ochre_context_set_dense_large"):
        lib.ochre_context_set_dense_large.argtypes = [P, ctypes.c_int]
        lib.ochre_context_get_dense_large.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    if hasattr(lib, "ochre_context_set_pass_limits"):
        lib.ochre_context_set_pass_limits.argtypes = [P, ctypes.c_int64, ctypes.c_int64]
    if hasattr(lib, "ochre_build_atlas"):
        i64 = ctypes.c_int64
        lib.ochre_build_atlas.argtypes = [P, ctypes.POINTER(BatchOutputC), i64, P, P, ctypes.c_int, P, i64, P, i64,
                                          ctypes.POINTER(i64), ctypes.POINTER(i64), P]
    if hasattr(lib, "ochre_rasterizer_stroke"):
        lib.ochre_rasterizer_stroke.argtypes = [P, P, ctypes.c_size_t, P, ctypes.c_size_t, P, ctypes.c_size_t,
                                                ctypes.c_float, P]
        for name in ("ochre_stroke_plain", "ochre_stroke_batched"):
            getattr(lib, name).argtypes = [P, ctypes.POINTER(PathBatchC), P, ctypes.POINTER(BatchOutputC), P]
    if hasattr(lib, "ochre_rasterize_batched_noalloc"):
        lib.ochre_rasterize_batched_noalloc.argtypes = [P, ctypes.POINTER(PathBatchC), ctypes.POINTER(BatchOutputC),
                                                        P]
    if hasattr(lib, "ochre_draw_plain"):
        for name in ("ochre_draw_plain", "ochre_draw_batched"):
            getattr(lib, name).argtypes = [P, ctypes.POINTER(PathBatchC), P, P, ctypes.POINTER(BatchOutputC), P]
    if hasattr(lib, "ochre_svg_parse"):
        lib.ochre_svg_parse.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(P)]
        lib.ochre_svg_info.argtypes = [P, P]
        for name in ("ochre_svg_verbs", "ochre_svg_verb_offsets", "ochre_svg_points", "ochre_svg_point_offsets",
                     "ochre_svg_draws"):
            getattr(lib, name).argtypes = [P]
            getattr(lib, name).restype = P
        lib.ochre_svg_free.argtypes = [P]
        lib.ochre_svg_free.restype = None


def check(status: int, what: str):
    if status != OCHRE_OK:
        raise OchreError(status, what)
