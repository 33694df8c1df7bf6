"""Streamline count / length / area fields (slc, slt, sla) on the MI355X.

Drop-in for python/streamlines/fields.py: same class name, constructor
signature and in-place ``data`` contract.  ``Fields.integrate`` creates the
seeds (useful.create_seeds), traces both directions on the GPU through
libslm_hip.so (slm_trace_fields, replacing kernel integrate_fields launched at
fields.py:233-274), and post-processes exactly as fields.py:141-164 and
:296-309 do.  The OpenCL platform/device arguments are accepted and ignored:
the current HIP device (torch.cuda.current_device()) is used.
"""
from __future__ import annotations

import numpy as np

from . import hip
from .params import kernel_params
from .useful import create_seeds, vprint

__all__ = ["Fields", "postprocess_raw", "postprocess_scaled"]


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


class Fields:
    def __init__(self, which_cl_platform=0, which_cl_device=0, cl_src_path=None, info=None,
                 data=None, verbose=False, gpu_verbose=False, fp_contract=True):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("Fields needs a HIP device (torch.cuda.is_available() is False)")
        hip.load()
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.cl_src_path = cl_src_path
        self.info = info
        self.data = data
        self.verbose = verbose
        self.gpu_verbose = gpu_verbose
        self.vprogress = verbose
        self.fp_contract = fp_contract
        self.steps = [0, 0]

    def integrate(self):
        """fields.py:69-167: seeds, both directions on the GPU, post-processing."""
        vprint(self.verbose, "Integrating streamline fields...")
        info = self.info
        (self.data.seed_point_array, info.n_seed_points, info.n_padded_seed_points) = create_seeds(
            self.data.mask_array, info.pad_width, info.n_work_items, do_shuffle=info.do_shuffle,
            rng_seed=info.shuffle_rng_seed, verbose=self.verbose)
        self.gpu_integrate()
        dds = self.data.traj_stats_df["ds"]["downstream", "mean"]
        uds = self.data.traj_stats_df["ds"]["upstream", "mean"]
        self.data.slc_array, self.data.slt_array = postprocess_scaled(
            self.data.slc_array, self.data.slt_array, info.subpixel_seed_point_density, dds, uds,
            info.pixel_size)
        vprint(self.verbose, "...done")

    def gpu_integrate(self):
        """fields.py:169-311 with both directions traced by slm_trace_fields."""
        import torch
        info = self.info
        dev = self.device
        mask = self.data.mask_array
        nxp, nyp = mask.shape
        seeds = torch.from_numpy(np.ascontiguousarray(self.data.seed_point_array, dtype=np.float32)).to(dev)
        mask_d = torch.from_numpy(np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)).to(dev)
        uv_d = torch.from_numpy(np.ascontiguousarray(self.data.uv_array, dtype=np.float32).reshape(-1)).to(dev)
        slc_d = torch.empty((2, nxp * nyp), dtype=torch.int32, device=dev)
        slt_d = torch.empty((2, nxp * nyp), dtype=torch.int32, device=dev)
        steps_d = torch.zeros(2, dtype=torch.int64, device=dev)
        ws = None
        for idx, sign in ((0, +1.0), (1, -1.0)):
            info.downup_sign = sign
            p = kernel_params(info, sign, "INTEGRATE_FIELDS", self.fp_contract)
            ws = hip.trace_fields(p, seeds, mask_d, uv_d, slc_d[idx], slt_d[idx], steps_d[idx:idx + 1],
                                  workspace=ws, keep_order=(idx == 1))
        torch.cuda.synchronize(dev)
        self.steps = [int(v) for v in steps_d.cpu()]
        slc_raw = hip.as_u32(slc_d).reshape(2, nxp, nyp).transpose(1, 2, 0)
        slt_raw = hip.as_u32(slt_d).reshape(2, nxp, nyp).transpose(1, 2, 0).astype(np.float32)
        self.raw_slc = np.ascontiguousarray(slc_raw)
        self.raw_slt = np.ascontiguousarray(hip.as_u32(slt_d).reshape(2, nxp, nyp).transpose(1, 2, 0))
        self.data.slc_array, self.data.slt_array, self.data.sla_array = postprocess_raw(
            slc_raw, slt_raw, info.subpixel_seed_point_density)
