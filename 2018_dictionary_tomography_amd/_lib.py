"""ctypes binding of libgausstomo.so (the C ABI declared in include/gausstomo.h).

The library is the only compute path: if it is missing this module raises,
there is no fallback.  torch is imported first so that the library binds to
the HIP runtime torch already loaded.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime before the library)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgausstomo.so")

GT_MODE_LINEAR, GT_MODE_L2, GT_MODE_RESID = 0, 1, 2
GT_REG_NULL, GT_REG_JOUBERT, GT_REG_MASS, GT_REG_VOLUME, GT_REG_RADIUS, GT_REG_ISO = 0, 1, 2, 3, 4, 5
GT_REG_NPARAM = 5
GT_KL_CLAMP, GT_KL_LAGRANGE = 0, 1
GT_PSF_EXACT, GT_PSF_SAMPLED = 0, 1


class GTError(RuntimeError):
    """Error reported by the HIP library (argument, workspace or HIP error)."""


class gt_atoms(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n", ctypes.c_int), ("iso", ctypes.c_int),
                ("I", ctypes.c_void_p), ("x", ctypes.c_void_p), ("r", ctypes.c_void_p)]


class gt_geom(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("lenT", ctypes.c_int), ("nA", ctypes.c_int), ("n0", ctypes.c_int),
                ("n1", ctypes.c_int), ("sorted", ctypes.c_int), ("T", ctypes.c_void_p), ("w0", ctypes.c_void_p),
                ("w1", ctypes.c_void_p), ("p0", ctypes.c_void_p), ("p1", ctypes.c_void_p)]


class gt_vol(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("m0", ctypes.c_int), ("m1", ctypes.c_int), ("m2", ctypes.c_int),
                ("sorted", ctypes.c_int), ("y0", ctypes.c_void_p), ("y1", ctypes.c_void_p), ("y2", ctypes.c_void_p)]


class gt_kgrid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("m0", ctypes.c_int), ("m1", ctypes.c_int), ("m2", ctypes.c_int),
                ("sorted", ctypes.c_int), ("k0", ctypes.c_void_p), ("k1", ctypes.c_void_p), ("k2", ctypes.c_void_p)]


class gt_kplanes(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int), ("nA", ctypes.c_int), ("m0", ctypes.c_int), ("m1", ctypes.c_int),
                ("sorted", ctypes.c_int), ("k0", ctypes.c_void_p), ("k1", ctypes.c_void_p), ("w0", ctypes.c_void_p),
                ("w1", ctypes.c_void_p)]


class gt_kplanes64(ctypes.Structure):
    _fields_ = gt_kplanes._fields_


class gt_atoms64(ctypes.Structure):
    _fields_ = gt_atoms._fields_
GT_KP_3D, GT_KP_2P5D, GT_KP_2D_LINE, GT_KP_2D_GRID = 0, 1, 2, 3

_P = ctypes.POINTER
_vp, _sz, _ci = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int

# name -> (restype, argtypes); every symbol declared in include/gausstomo.h
SIGNATURES = {
    "gt_last_error": (ctypes.c_char_p, []),
    "gt_version": (ctypes.c_char_p, []),
    "gt_forward_workspace": (_sz, [_P(gt_atoms), _P(gt_geom)]),
    "gt_forward": (_ci, [_P(gt_atoms), _P(gt_geom), _vp, _vp, _sz, _vp]),
    "gt_derivs_workspace": (_sz, [_P(gt_atoms), _P(gt_geom), _ci]),
    "gt_derivs": (_ci, [_P(gt_atoms), _P(gt_geom), _vp, _ci, _ci, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gt_forward_derivs": (_ci, [_P(gt_atoms), _P(gt_geom), _vp, _ci, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gt_noalloc_scratch_bytes": (_sz, []),
    "gt_forward_derivs_noalloc": (_ci, [_P(gt_atoms), _P(gt_geom), _vp, _ci, _vp, _vp, _vp, _vp, _vp]),
    "gt_volume_workspace": (_sz, [_P(gt_atoms), _P(gt_vol)]),
    "gt_volume": (_ci, [_P(gt_atoms), _P(gt_vol), _vp, _vp, _sz, _vp]),
    "gt_ft_workspace": (_sz, [_P(gt_atoms), _P(gt_kgrid)]),
    "gt_ft_eval": (_ci, [_P(gt_atoms), _P(gt_kgrid), _vp, _vp, _sz, _vp]),
    "gt_ft_planes_workspace": (_sz, [_ci, _P(gt_kplanes)]),
    "gt_ft_planes_eval": (_ci, [_P(gt_atoms), _P(gt_kplanes), _vp, _vp, _sz, _vp]),
    "gt_ft_planes_workspace_f64": (_sz, [_ci, _P(gt_kplanes64)]),
    "gt_ft_planes_eval_f64": (_ci, [_P(gt_atoms64), _P(gt_kplanes64), _vp, _vp, _sz, _vp]),
    "gt_ft_derivs_workspace": (_sz, [_P(gt_atoms), _P(gt_kplanes)]),
    "gt_ft_derivs": (_ci, [_P(gt_atoms), _P(gt_kplanes), _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gt_psf_forward_workspace": (_sz, [_P(gt_atoms), _P(gt_geom)]),
    "gt_psf_forward": (_ci, [_P(gt_atoms), _P(gt_geom), _vp, _vp, _sz, _vp]),
    "gt_psf_derivs_workspace": (_sz, [_P(gt_atoms), _P(gt_geom)]),
    "gt_psf_derivs": (_ci, [_P(gt_atoms), _P(gt_geom), _vp, _ci, _ci, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gt_psf_volume_workspace": (_sz, [_P(gt_atoms), _P(gt_vol)]),
    "gt_psf_volume": (_ci, [_P(gt_atoms), _P(gt_vol), _vp, _vp, _sz, _vp]),
    "gt_psf_forward_rule": (_ci, [_P(gt_atoms), _P(gt_geom), _ci, _vp, _vp, _sz, _vp]),
    "gt_psf_derivs_rule": (_ci, [_P(gt_atoms), _P(gt_geom), _vp, _ci, _ci, _ci, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gt_psf_volume_rule": (_ci, [_P(gt_atoms), _P(gt_vol), _ci, _vp, _vp, _sz, _vp]),
    "gt_sino_moments_workspace": (_sz, [_P(gt_geom)]),
    "gt_sino_moments": (_ci, [_P(gt_geom), _vp, _vp, _ci, ctypes.c_double, ctypes.c_double, _vp, _vp, _sz, _vp]),
    "gt_reg_eval": (_ci, [_P(gt_atoms), _ci, _P(ctypes.c_double), _vp, _vp, _vp, _vp]),
    "gt_fit_run": (_ci, [_vp, _ci, _ci, _ci, _vp]),
    "gt_fit_scratch_bytes": (_sz, [_ci]),
    "gt_fit_max_blocks": (_ci, []),
    "gt_fit_run_blocks": (_ci, [_vp, _ci, _vp, _ci, _ci, _ci, _vp]),
    "gt_profile_enable": (_ci, [_ci]),
    "gt_profile_read": (_ci, [_ci, _P(ctypes.c_double), _P(ctypes.c_long)]),
    "gt_profile_select": (_ci, [ctypes.c_uint]),
}

GT_K_SETUP, GT_K_FORWARD, GT_K_MOMENTS, GT_K_FINISH, GT_K_VOLUME, GT_K_FT, GT_K_FIT = range(7)


def profile_enable(on=True):
    load().gt_profile_enable(1 if on else 0)


def profile_select(kernel_ids=None):
    """Bracket only these kernel ids with timing events (None: all)."""
    mask = 0xFFFFFFFF if kernel_ids is None else sum(1 << int(k) for k in kernel_ids)
    load().gt_profile_select(mask)


def profile_read(kernel_id):
    """(total_ms, count) of the launches of one kernel id since the last read."""
    tot, cnt = ctypes.c_double(0), ctypes.c_long(0)
    check(load().gt_profile_read(kernel_id, ctypes.byref(tot), ctypes.byref(cnt)), "gt_profile_read")
    return tot.value, cnt.value

_lib = None


def load(path: str = LIB_PATH):
    """Load libgausstomo.so and declare every exported entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().gt_last_error().decode(errors="replace")
        raise GTError(f"{what} failed (rc={rc}): {msg}")


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
