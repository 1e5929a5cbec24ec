"""ctypes binding of ``libgisapr_hip.so`` (declared in include/gisapr_hip.h).

The library is loaded from the package directory.  There is no fallback: if
it is missing, or a call is made on a non-HIP device, the call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libgisapr_hip.so"
_lib = None

_vp, _i, _d, _sz, _fp, _ip, _dp = C.c_void_p, C.c_int, C.c_double, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p
_f = C.c_float

# name -> (restype, argtypes); mirrors include/gisapr_hip.h
SIGNATURES = {
    "gis_last_error": (C.c_char_p, []),
    "gis_version": (_i, []),
    "gis_init": (_i, [_i]),
    "gis_splat": (_i, [_vp, _vp, _i, _i, _i, _d, _d, _d, _vp, _vp]),
    "gis_move_chain": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp]),
    "gis_volume_dft_bytes": (_sz, [_i]),
    "gis_volume_dft_workspace": (_sz, [_i]),
    "gis_volume_dft_bytes_ex": (_sz, [_i, _i]),
    "gis_volume_dft_workspace_ex": (_sz, [_i, _i]),
    "gis_volume_dft_ex": (_i, [_vp, _i, _i, _vp, _vp, _sz, _vp]),
    "gis_build_templates_workspace_ex": (_sz, [_i, _i, _i]),
    "gis_build_templates_resident": (_i, [_i, _i]),
    "gis_project_slices": (_i, [_vp, _i, _i, _vp, _i, _i, _i, _f, _vp, _vp, _sz, _vp]),
    "gis_whiten_images": (_i, [_vp, _i, _i, _i, _i, _f, _vp, _vp, _sz, _vp]),
    "gis_extract_slices_workspace": (_sz, [_i, _i]),
    "gis_extract_slices": (_i, [_vp, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "gis_sinc2_filter": (_i, [_vp, _i, _i, _i, _vp]),
    "gis_slice_sample_grid": (_i, [_vp, _i, _vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp]),
    "gis_sample_dft_3d": (_i, [_vp, _i, _i, _i, _vp, _sz, _vp, _vp]),
    "gis_volume_dft": (_i, [_vp, _i, _vp, _vp, _sz, _vp]),
    "gis_build_templates_workspace": (_sz, [_i, _i]),
    "gis_build_templates": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _sz, _vp]),
    "gis_build_template_coefficients": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _sz,
                                             _vp]),
    "gis_build_templates_batched": (_i, [_vp, _sz, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _sz, _vp]),
    "gis_build_templates_batched_ld": (_i, [_vp, _sz, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _vp, _i, _vp, _sz,
                                            _vp]),
    "gis_build_template_coefficients_batched": (_i, [_vp, _sz, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _i, _vp, _vp,
                                                     _vp, _sz, _vp]),
    "gis_prepare_particles_workspace": (_sz, [_i, _i]),
    "gis_prepare_particles_workspace_ex": (_sz, [_i, _i, _i]),
    "gis_prepare_particles": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gis_prepare_particles_ld": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _sz, _vp]),
    "gis_ctf_packed": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    "gis_correlate_workspace": (_sz, [_i, _i, _i]),
    "gis_correlate": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i, _i, _vp, _vp,
                           _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gis_correlate_split_workspace": (_sz, [_i, _i, _i]),
    "gis_fragment_plane": (_sz, [_i]),
    "gis_pack_fragments": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "gis_correlate_split": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i, _i, _vp,
                                 _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gis_correlate_split_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i,
                                         _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gis_pack_fragments_f16": (_i, [_vp, _vp, _i, _i, _f, _vp, _vp]),
    "gis_correlate_split_f16_workspace": (_sz, [_i, _i, _i, C.c_long]),
    "gis_correlate_split_f16": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i,
                                     _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "gis_threshold_mask": (_i, [_vp, _sz, _f, _vp, _vp]),
    "gis_edt_squared": (_i, [_vp, _i, _vp, _vp]),
    "gis_restraint_eval": (_i, [_vp, _f, _vp, _vp, _sz, _vp, _vp, _vp, _vp]),
    "gis_score_sum_segments": (_i, [_vp, _vp, _vp, _i, _vp, _vp]),
    "gis_score_sum": (_i, [_vp, _vp, _i, _vp, _vp]),
    "gis_gather_coefficients": (_i, [_vp, _sz, _i, _vp, _i, _i, _vp, _vp, _vp]),
    "gis_correlate_mixed": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _i, _i, _vp,
                                 _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gis_correlate_mixed_ex": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _i, _i, _vp,
                                    _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]),
}


class GisError(RuntimeError):
    """A failing libgisapr_hip call (launch or runtime error)."""


def lib():
    """Load (once) and return the ctypes library; raise if it is not built."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise ImportError(
                f"{_LIB_PATH} is not built; run `python -m gisapr_amd.build` "
                "(gisapr_amd has no CPU fallback)")
        handle = C.CDLL(os.fspath(_LIB_PATH), mode=C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().gis_last_error().decode(errors="replace")
        if rc < 0:
            raise ValueError(f"{what}: {msg}")
        raise GisError(f"{what}: hip error {rc}: {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def stream_of(device):
    import torch
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_hip(t, name="tensor"):
    import torch
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ValueError(f"{name} must be a tensor on a HIP device (gisapr_amd has no CPU path)")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
