"""ctypes binding of libsubpx_hip.so (declarations: include/subpx_hip.h).

The library is built in-tree by ``python -m imageprocessor_new_amd._build``. Nothing here
falls back to another implementation: a missing or unloadable library raises.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsubpx_hip.so"

SPX_OK = 0
SPX_E_ARG = -1
SPX_E_NOMEM = -2
SPX_E_HIP = -3
SPX_E_UNSUPPORTED = -4
SPX_E_CAPACITY = -5

FIT_CONVERGED = 0
FIT_MAX_ITER = 1
FIT_SINGULAR = 2
FIT_DIVERGED = 3
FIT_TOO_FEW = 4
FIT_OUTSIDE = 5
FIT_BAD_START = 6
FIT_STATUS_NAMES = {
    FIT_CONVERGED: "converged",
    FIT_MAX_ITER: "max_iter",
    FIT_SINGULAR: "singular",
    FIT_DIVERGED: "diverged",
    FIT_TOO_FEW: "too_few_pixels",
    FIT_OUTSIDE: "center_outside_patch",
    FIT_BAD_START: "bad_start",
}

MODEL_ERF_DISK6 = 0
MODEL_ERF_ELLIPSE7 = 1


class SpxFitOptions(ctypes.Structure):
    _fields_ = [
        ("max_iter", ctypes.c_int32),
        ("step_tol", ctypes.c_float),
        ("lambda0", ctypes.c_float),
        ("lambda_up", ctypes.c_float),
        ("lambda_down", ctypes.c_float),
        ("init_sigma", ctypes.c_float),
        ("tile_waves", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
        ("bucket_mask", ctypes.c_int32),
        ("learned_iters", ctypes.c_int32),
        ("bucket_waves", ctypes.c_int32),
    ]


class SpxEdgeErfOptions(ctypes.Structure):
    _fields_ = [
        ("band_rad", ctypes.c_int32),
        ("edge_rad", ctypes.c_int32),
        ("smooth_passes", ctypes.c_int32),
        ("loc_rad", ctypes.c_int32),
        ("grad_power", ctypes.c_float),
        ("iters", ctypes.c_int32),
        ("erf_fit_rad", ctypes.c_int32),
        ("erf_gn_iters", ctypes.c_int32),
        ("erf_init_sigma", ctypes.c_float),
        ("erf_sigma_min", ctypes.c_float),
        ("erf_sigma_max", ctypes.c_float),
        ("erf_max_shift", ctypes.c_float),
        ("erf_damp", ctypes.c_float),
        ("erf_cond_max", ctypes.c_float),
    ]


DTYPE_U16 = 0
DTYPE_F32 = 1


class SpxLogquadWindowOptions(ctypes.Structure):
    _fields_ = [
        ("win_rad", ctypes.c_int32),
        ("bg_win_pad", ctypes.c_int32),
        ("bg_dilate", ctypes.c_int32),
        ("fit_dilate", ctypes.c_int32),
        ("min_bg_pts", ctypes.c_int32),
        ("min_plane_pts", ctypes.c_int32),
        ("plane_bg", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("core_frac", ctypes.c_double),
        ("weight_power", ctypes.c_double),
        ("bg_tail_frac", ctypes.c_double),
        ("max_cond", ctypes.c_double),
        ("negdef_eps", ctypes.c_double),
        ("eps", ctypes.c_double),
    ]


class SpxHoleStepOptions(ctypes.Structure):
    _fields_ = [
        ("nv0", ctypes.c_double),
        ("nv1", ctypes.c_double),
        ("rscale", ctypes.c_double),
        ("max_move", ctypes.c_double),
        ("max_delta", ctypes.c_double),
        ("a_med", ctypes.c_double),
        ("relax", ctypes.c_double),
        ("size", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class SpxTileSpec(ctypes.Structure):
    _fields_ = [
        ("layout", ctypes.c_int32),
        ("cols", ctypes.c_int32),
        ("pitch", ctypes.c_float),
        ("jitter", ctypes.c_float),
        ("r_lo", ctypes.c_float),
        ("r_hi", ctypes.c_float),
        ("s_lo", ctypes.c_float),
        ("s_hi", ctypes.c_float),
        ("ellipse_frac", ctypes.c_float),
        ("amplitude", ctypes.c_float),
        ("background", ctypes.c_float),
        ("read_sigma", ctypes.c_float),
    ]


class SpxLmStage(ctypes.Structure):
    """One stage of the one-pass page fit's device Levenberg-Marquardt (all pointers device)."""

    _fields_ = [(name, ctypes.c_int32) for name in ("B", "P", "nf", "nshape", "nb", "npr")] + [
        (name, ctypes.c_void_p) for name in ("fixed", "tm", "lo", "hi", "PL", "pm", "pw", "pj", "pa", "ev", "q", "qn",
                                             "theta", "blur", "cost", "scale", "cst", "g", "A", "lam", "nu", "done",
                                             "alldone")] + [("ftol", ctypes.c_double), ("xtol", ctypes.c_double)]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64

# name -> (restype, argtypes); every symbol include/subpx_hip.h declares.
SIGNATURES = {
    "spx_init": (_I32, [ctypes.c_int]),
    "spx_release": (None, []),
    "spx_last_error": (ctypes.c_char_p, []),
    "spx_version": (ctypes.c_char_p, []),
    "spx_otsu_u16": (_I32, [_P, _I64, _I64, _I64, _I32, _P, _P]),
    "spx_threshold_cc": (_I32, [_P, _I64, _I64, _I64, _I32, _I32, _I32, _P, _P, _P, _I32, _P, _P]),
    "spx_refine_logquad": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _P, _P, _P]),
    "spx_fit_stats": (_I32, [_P, _P]),
    "spx_fit_counters": (_I32, [_P, _I32, _P]),
    "spx_cnn_param_count": (_I64, []),
    "spx_gemm_bf16": (_I32, [_P, _I64, _I64, _P, _I64, _I64, _P, _I32, _I32, _I32, _P]),
    "spx_cnn_prepare": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _P, _P, _P]),
    "spx_cnn_encode": (_I32, [_P, _P, _I32, _P, _P]),
    "spx_cnn_propose": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _P, _P, _P, _P]),
    "spx_cnn_forward": (_I32, [_P, _I32, _P, _P, _P, _P, _P]),
    "spx_cnn_train_step": (_I32, [_P, _P, _P, _I32, _P, _P, _P, _P, _I32, ctypes.c_float, _P, _P]),
    "spx_cnn_to_bf16": (_I32, [_P, _I64, _P, _P]),
    "spx_cnn_grad": (_I32, [_P, _P, _P, _I32, _P, _P, _P, _P, _P]),
    "spx_fit_batch_cells": (_I32, [_P, _I64, _I64, _I64, _P, _P, _P, _P, ctypes.POINTER(SpxFitOptions), _I32, _P,
                                   _I32, _P, _P, _P, _P, _P, _P]),
    "spx_fit_batch_noalloc": (_I32, [_P, _I64, _I64, _I64, _P, _P, _P, _P, ctypes.POINTER(SpxFitOptions), _I32, _P,
                                     _I32, _P, _P, _P, _P, _P, _P]),
    "spx_fit_graph_create": (_I32, [_P, _I64, _I64, _I64, _P, _P, _P, _P, ctypes.POINTER(SpxFitOptions), _I32, _P,
                                    _I32, _P, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_void_p)]),
    "spx_fit_graph_launch": (_I32, [_P, _P]),
    "spx_fit_graph_destroy": (_I32, [_P]),
    "spx_voronoi_labels": (_I32, [_P, _I32, _I64, _I64, _P, _P, _P, _P]),
    "spx_deblend_init": (_I32, [_P, _I64, _I64, _P, _P, _P, _I32, _I32, _I32, _P, _P, _P]),
    "spx_gauss_fit": (_I32, [_P, _P, _I32, _I32, _I32, _P, _I32, _P, _P]),
    "spx_gauss_render": (_I32, [_P, _P, _I32, _I64, _I64, _P, _P]),
    "spx_deblend_clean": (_I32, [_P, _P, _P, _P, _P, _P, _P, _I32, _I32, _I32, _I64, _I64, _P, _P]),
    "spx_tile_medians": (_I32, [_P, _P, _P, _I64, _I64, _I32, ctypes.c_double, ctypes.c_double, _P, _P, _P]),
    "spx_tile_map": (_I32, [_P, _P, _P, _I32, _I32, _I64, _I64, _P, _P]),
    "spx_cell_refine": (_I32, [_P, _P, _I32, _I32, _I32, _I32, _P, ctypes.c_double, _P, _P]),
    "spx_deblend_window": (_I32, [_P, _P, _P, _P, _I32, _I32, _I32, _I32, _P, _P, _P]),
    "spx_voronoi_rois": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _I32, _I32, _P, _P, _P, _P]),
    "spx_radial_symmetry": (_I32, [_P, _P, _I32, _I32, _I32, _I32, _I32, _P, _P, _P]),
    "spx_isophote_curvature": (_I32, [_P, _P, _I32, _I32, _I32, _I32, _I32, ctypes.c_double, _P, _P, _P]),
    "spx_tiff_open": (_I32, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "spx_tiff_page_count": (_I32, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "spx_tiff_page_info": (_I32, [_P, _I32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(ctypes.c_int32)]),
    "spx_tiff_read_page": (_I32, [_P, _I32, _P, _I64]),
    "spx_tiff_close": (_I32, [_P]),
    "spx_onepass_smooth": (_I32, [_P, _I64, _I64, _P, _I32, _P, _P, _P]),
    "spx_onepass_lm_trial": (_I32, [ctypes.POINTER(SpxLmStage), _I32, _P]),
    "spx_onepass_lm_geometry": (_I32, [_P, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _P]),
    "spx_onepass_lm_init": (_I32, [ctypes.POINTER(SpxLmStage), _P]),
    "spx_onepass_lm_update": (_I32, [ctypes.POINTER(SpxLmStage), _P]),
    "spx_onepass_peaks": (_I32, [_P, _I64, _I64, _I32, ctypes.c_double, _P, _P, _P]),
    "spx_onepass_peak_list": (_I32, [_P, _P, _P, _I64, _I64, ctypes.c_double, _P, _P]),
    "spx_onepass_count_above": (_I32, [_P, _I64, _I64, _I32, ctypes.c_double, _P, _P]),
    "spx_onepass_nearest": (_I32, [_P, _I32, _P, _P, _I32, _I32, ctypes.c_double, _P, _P]),
    "spx_onepass_stack": (_I32, [_P, _I64, _I64, _P, _I32, _P, _I32, _I32, _I32, _P, _P, _P]),
    "spx_onepass_versus": (_I32, [_P, _I64, _I64, _P, _I32, _P, _I32, _I32, _I32, ctypes.c_double, _P, _P]),
    "spx_onepass_pair_hist": (_I32, [_P, _P, ctypes.c_double, _P, _I32, _I32, _P, _P, _I32, _I32, ctypes.c_double,
                                     ctypes.c_double, _I32, _I32, _P, _P, _P]),
    "spx_onepass_pair_crop": (_I32, [_P, _I32, _I32, _I32, _I32, _I32, _P, _P]),
    "spx_onepass_outline_hat": (_I32, [_P, _P, _P, _I32, _I32, _I32, ctypes.c_double, _P, _P, _I32, _P, _P]),
    "spx_onepass_coeff": (_I32, [_P, _I32, _P, _I32, _I32, _P, _I32, _P, _P, _P, _I32, _I32, _P, _P]),
    "spx_onepass_grid_coeff": (_I32, [_P, _I32, _P, _I32, _I32, _I32, _P, _P, _P, _I32, _I32, _P, _P]),
    "spx_onepass_synth": (_I32, [_P, _I32, _I32, _I32, _P, _P, _P, _I32, _P, _P, _P, _I32, _I32, _I32, _P, _P]),
    "spx_onepass_project": (_I32, [_P, _P, _P, _I32, _I32, _I32, _P, _P]),
    "spx_onepass_grid_costs": (_I32, [_P, _P, _P, _P, _I32, _I32, _I32, _I32, _P, _P]),
    "spx_onepass_band": (_I32, [_P, _I32, _I32, _P, _I32, _I32, _I32, _I32, _I32, ctypes.c_double, _P, _P]),
    "spx_seed_neighbours": (_I32, [_P, _I32, _I64, _I64, ctypes.c_float, _I32, _P, _P]),
    "spx_triangle_u8": (_I32, [_P, _I64, _I64, _I64, _I32, ctypes.POINTER(ctypes.c_int32), _P]),
    "spx_threshold_mask": (_I32, [_P, _I64, _I64, _I64, _I32, _I32, _P, _P]),
    "spx_morph_cross": (_I32, [_P, _I64, _I64, _I64, _I32, _P, _P]),
    "spx_refine_logquad_window": (_I32, [_P, _I32, _I64, _I64, _I64, _P, _P, _P, _I32, _I32,
                                         ctypes.POINTER(SpxLogquadWindowOptions), _P, _P, _P]),
    "spx_refine_weighted": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _P, _P, _P]),
    "spx_edge_moment": (_I32, [_P, _I32, _I64, _I64, _I64, _P, _I32, ctypes.POINTER(SpxEdgeErfOptions), _P, _P, _P]),
    "spx_edge_erf": (_I32, [_P, _I32, _I64, _I64, _I64, _P, _I32, ctypes.POINTER(SpxEdgeErfOptions), _P, _P, _P]),
    "spx_edge_profile_erf": (_I32, [_P, _I32, _I32, _P, _P, ctypes.POINTER(SpxEdgeErfOptions), _P, _P, _P]),
    "spx_render_page": (_I32, [ctypes.c_uint64, _P, _I32, _I64, _I64, ctypes.c_float, ctypes.c_float, _P, _I64,
                               _P]),
    "spx_hole_templates": (_I32, [ctypes.c_double, ctypes.c_double, _I32, _I32, ctypes.c_double, ctypes.c_double,
                                  _I32, _P, _P, _P]),
    "spx_hole_step": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _I32, _P, _I32, _P, _P, _P, _P, _P,
                             ctypes.POINTER(SpxHoleStepOptions), _P, _P]),
    "spx_hole_render": (_I32, [_P, _I64, _I64, _P, _P, _I32, _I32, _I32, ctypes.c_double, _P, _I32, _P, _P, _P,
                               _P, _P]),
    "spx_hole_sweep": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _I32, _I32, ctypes.c_double, _P, _P, _I32,
                              _P, _P, _P, _P, _P, _P, ctypes.POINTER(SpxHoleStepOptions), _P, _P]),
    "spx_select_kth": (_I32, [_P, _I64, _I64, _P, _P, _I32, _P, _P]),
    "spx_far_splat": (_I32, [_P, _P, _P, ctypes.c_double, _I32, _I32, _I32, _I32, _I32, _P, _P]),
    "spx_outline_coverage": (_I32, [_P, _P, _P, _I32, _I32, ctypes.c_double, _I32, _P, _P]),
    "spx_outline_windows": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _I32, _P, _I32, _P, _P, _P, _P,
                                   ctypes.c_double, _P, _P, _P, _P, _P, _P, _P]),
    "spx_outline_far_splat": (_I32, [_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _P, _P]),
    "spx_outline_stack": (_I32, [_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _I32, _P, _I32, _P, _P, _P, _P, _P,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, _I32, _P, _P, _P]),
    "spx_render_tiles": (_I32, [ctypes.c_uint64, _I32, _I32, ctypes.POINTER(SpxTileSpec), _P, _I64, _P, _P]),
    "spx_fit_batch": (_I32, [_P, _I64, _I64, _I64, _P, _P, _P, _P, ctypes.POINTER(SpxFitOptions),
                             _I32, _P, _P, _P, _P, _P, _P]),
}

_lock = threading.Lock()
_lib = None


class SpxError(RuntimeError):
    """A libsubpx_hip call failed; ``code`` is the SPX_E_* value."""

    def __init__(self, code: int, name: str, text: str):
        super().__init__(f"{name} failed ({code}): {text}")
        self.code = code


def load() -> ctypes.CDLL:
    """Load the library once and declare every exported signature."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m imageprocessor_new_amd._build` first"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error() -> str:
    return (load().spx_last_error() or b"").decode()


def call(name: str, *args) -> int:
    """Call ``name``; raise ValueError for argument errors and SpxError otherwise."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc == SPX_OK:
        return rc
    text = last_error()
    if rc == SPX_E_ARG:
        raise ValueError(f"{name}: {text}")
    raise SpxError(rc, name, text)


def fit_options(*, max_iter=30, step_tol=1e-6, lambda0=1e-3, lambda_up=10.0, lambda_down=0.1,
                init_sigma=1.0, tile_waves=0, persistent=1, bucket_mask=0, learned_iters=0,
                bucket_waves=0) -> SpxFitOptions:
    return SpxFitOptions(int(max_iter), float(step_tol), float(lambda0), float(lambda_up),
                         float(lambda_down), float(init_sigma), int(tile_waves), int(persistent),
                         int(bucket_mask), int(learned_iters), int(bucket_waves))


def bucket_mask_of(half, model) -> int:
    """Bucket mask (SpxFitOptions.bucket_mask) covering every bucket the features of host arrays
    of half sides and models can fall in: a feature's fitted pixels (inside the image and its
    cell) number at most its (2 half + 1)^2 patch pixels, so its bucket's patch class is at
    most that of the whole patch."""
    import numpy as np

    h = np.asarray(half, dtype=np.int64).ravel()
    m = np.asarray(model, dtype=np.int64).ravel()
    ok = (h >= 2) & (h <= 31) & ((m == MODEL_ERF_DISK6) | (m == MODEL_ERF_ELLIPSE7))
    mask = 0
    for md in (MODEL_ERF_DISK6, MODEL_ERF_ELLIPSE7):
        sel = ok & (m == md)
        if not sel.any():
            continue
        top = patch_class(int(((2 * h[sel].max() + 1) ** 2)))
        for c in range(top + 1):
            mask |= 1 << (2 * c + (md == MODEL_ERF_ELLIPSE7))
    return mask


def patch_class(npix: int) -> int:
    """Patch class of a pixel count: 0..4 for at most 4, 8, 16, 32, 64 pixels per lane."""
    npl = (int(npix) + 63) // 64
    return 0 if npl <= 4 else 1 if npl <= 8 else 2 if npl <= 16 else 3 if npl <= 32 else 4
