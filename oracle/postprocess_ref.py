"""numpy restatement of the fields post-processing (fields.py:141-164,
:296-309) with the dtypes numpy 1.x gave the reference.

TEST INFRASTRUCTURE ONLY: the checker for slm_fields_finish_raw /
slm_fields_finish_scaled.  The product finishes the fields on the GPU.
"""
from __future__ import annotations

import numpy as np


def postprocess_raw(slc_raw: np.ndarray, slt_raw: np.ndarray, density: int):
    """fields.py:296-309 on the raw per-direction sums.

    slc_raw uint32 [...,2]; slt_raw float32 [...,2] (the uint32 sums cast to
    float32, fields.py:288).  numpy-1.x dtypes are reproduced explicitly:
    sqrt(2)*slt is a float32 product, divided by slc in float64, stored
    float32; slt/d^2 is a float32 division.  Returns (slc, slt, sla)."""
    slc = slc_raw.astype(np.uint32)
    slt = slt_raw.astype(np.float32).copy()
    sla = np.zeros(slt.shape, dtype=np.float32)
    pos = slc > 0
    prod = (np.float32(np.sqrt(2.0)) * slt[pos]).astype(np.float32)
    sla[pos] = (prod.astype(np.float64) / slc[pos].astype(np.float64)).astype(np.float32)
    slt[pos] = (slt[pos] / np.float32(int(density) ** 2)).astype(np.float32)
    return slc, slt, sla


def postprocess_scaled(slc: np.ndarray, slt: np.ndarray, density: int, dds, uds, pixel_size):
    """fields.py:141-164: slt scaled by mean ds per direction, sqrt'd;
    slc -> sqrt(2)*(slc/d^2)^(2/3) in float64."""
    ps = np.float32(pixel_size)
    slt = slt.astype(np.float32).copy()
    slt[:, :, 0] = (slt[:, :, 0] * np.float32(np.float32(dds) / ps)).astype(np.float32)
    slt[:, :, 1] = (slt[:, :, 1] * np.float32(np.float32(uds) / ps)).astype(np.float32)
    slt = np.sqrt(slt).astype(np.float32)
    slc_f = np.sqrt(2.0) * np.power(slc.astype(np.float64) / float(int(density) ** 2), 2.0 / 3.0)
    return slc_f, slt
