"""Traced hillslope lengths (HSL) on the MI355X.

Drop-in for python/streamlines/lengths.py:33-92 (``hsl``): seeds are the
mid-slope (or ridge) pixels picked from the mapping flags, each is traced
downstream to a thin-channel pixel by slm_hillslope_lengths (replacing kernel
hillslope_lengths), and mid-slope lengths are doubled to whole-hillslope
lengths.
"""
from __future__ import annotations

import numpy as np

from . import hip
from .params import kernel_params
from .useful import pick_seeds, vprint

__all__ = ["hsl"]


def hsl(info, data, do_use_ridges=False, verbose=False, fp_contract=False):
    """Fills data.subsegment_hsl_array (float32, one entry per seed pixel).
    Returns False for an empty seed list, as the reference does."""
    import torch
    vprint(verbose, "Computing lengths...", end="")
    flag = info.is_ridge if do_use_ridges else info.is_midslope
    seeds = pick_seeds(mask=data.mask_array, map=data.mapping_array, flag=flag, pad=info.pad_width)
    if getattr(data, "subsegment_hsl_array", None) is None or data.subsegment_hsl_array.shape[0] != seeds.shape[0]:
        if getattr(data, "subsegment_hsl_array", None) is not None:
            raise ValueError("Mismatched midslope/ridge point arrays: seed pts={0} traj len={1}".format(
                seeds.shape, data.subsegment_hsl_array.shape))
        data.subsegment_hsl_array = np.zeros(seeds.shape[0], dtype=np.float32)
    info.n_seed_points = np.uint32(seeds.shape[0])
    if seeds.shape[0] == 0:
        return False
    dev = torch.device("cuda", torch.cuda.current_device())
    p = kernel_params(info, 1, "HILLSLOPE_LENGTHS", fp_contract)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt).reshape(-1)).to(dev)
    length = t(data.subsegment_hsl_array, np.float32)
    hip.hillslope_lengths(p, torch.from_numpy(seeds).to(dev), t(data.mask_array, np.uint8),
                          t(data.uv_array, np.float32), t(data.mapping_array.view(np.int32), np.int32),
                          int(info.is_thinchannel), int(flag), length, sync=True)
    data.subsegment_hsl_array = length.cpu().numpy().reshape(data.subsegment_hsl_array.shape)
    if not do_use_ridges:
        data.subsegment_hsl_array *= np.float32(2.0)
    vprint(verbose, "...done")
    return True
