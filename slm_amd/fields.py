"""Streamline count / length / area fields (slc, slt, sla) on the MI355X.

Drop-in for python/streamlines/fields.py: same class name, constructor
signature and in-place ``data`` contract.  ``Fields.integrate`` creates the
seeds (useful.create_seeds), traces both directions on the GPU through
libslm_hip.so (one slm_trace_fields_batch launch, replacing the two
integrate_fields launches of fields.py:233-274), and finishes the fields on the GPU (slm_fields_finish_raw /
slm_fields_finish_scaled) exactly as fields.py:296-309 and :141-164 do.  The OpenCL platform/device arguments are accepted and ignored:
the current HIP device (torch.cuda.current_device()) is used.
"""
from __future__ import annotations

import numpy as np

from . import hip
from .params import kernel_params
from .useful import create_seeds, vprint

__all__ = ["Fields"]


class Fields:
    def __init__(self, which_cl_platform=0, which_cl_device=0, cl_src_path=None, info=None,
                 data=None, verbose=False, gpu_verbose=False, fp_contract=False):
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
        ps = np.float32(info.pixel_size)
        sd = np.float32(np.float32(self.data.traj_stats_df["ds"]["downstream", "mean"]) / ps)
        su = np.float32(np.float32(self.data.traj_stats_df["ds"]["upstream", "mean"]) / ps)
        nxp, nyp = self.data.mask_array.shape
        slc = hip.fields_finish_scaled(self._slc_raw, self._slt, info.subpixel_seed_point_density, sd, su)
        self.data.slc_array = slc.cpu().numpy().reshape(nxp, nyp, 2)
        self.data.slt_array = self._slt.cpu().numpy().reshape(nxp, nyp, 2)
        vprint(self.verbose, "...done")

    def gpu_integrate(self):
        """fields.py:169-311 with both directions traced by one slm_trace_fields_batch launch."""
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
        steps_d = torch.zeros(1, dtype=torch.int64, device=dev)
        jobs = []
        for idx, sign in ((0, +1.0), (1, -1.0)):
            info.downup_sign = sign
            p = kernel_params(info, sign, "INTEGRATE_FIELDS", self.fp_contract)
            jobs.append((p, seeds, mask_d, uv_d, slc_d[idx], slt_d[idx]))
        # both directions in one persistent launch; each direction's slc/slt
        # are bit-identical to a one-direction slm_trace_fields call
        hip.trace_fields_batch(hip.fields_jobs(jobs), steps_d)
        slt_f, sla_f = hip.fields_finish_raw(slc_d, slt_d, info.subpixel_seed_point_density)
        torch.cuda.synchronize(dev)
        self.steps = int(steps_d.item())
        self._slc_raw, self._slt = slc_d, slt_f
        self.raw_slc = np.ascontiguousarray(hip.as_u32(slc_d).reshape(2, nxp, nyp).transpose(1, 2, 0))
        self.raw_slt = np.ascontiguousarray(hip.as_u32(slt_d).reshape(2, nxp, nyp).transpose(1, 2, 0))
        self.data.slc_array = self.raw_slc.copy()
        self.data.slt_array = slt_f.cpu().numpy().reshape(nxp, nyp, 2)
        self.data.sla_array = sla_f.cpu().numpy().reshape(nxp, nyp, 2)
