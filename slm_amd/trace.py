"""Trace stage: trajectories then fields, re-embedded into full padded grids.

Restates Trace.compute_trajectories / compute_fields of
python/streamlines/trace.py:97-166.  ``state`` must provide
``merge_active_masks()`` and the JSON ``state`` keys; ``geodata`` provides
``pad_width``, ``roi_nx``, ``roi_ny``, ``roi_pixel_size``; ``preprocess``
provides ``uv_array``; ``trace_params`` carries the JSON ``trace`` keys.
"""
from __future__ import annotations

import numpy as np

from .fields import Fields
from .trajectories import Trajectories
from .useful import Data, Info, get_bbox

__all__ = ["Trace"]


class Trace:
    def __init__(self, state, trace_params, geodata, preprocess, fp_contract=False):
        self.state = state
        self.params = trace_params
        for k, v in vars(trace_params).items():
            setattr(self, k, v)
        self.geodata = geodata
        self.preprocess = preprocess
        self.fp_contract = fp_contract
        self.mapping_array = None
        self.seed_point_array = None

    def do(self):
        self.compute_trajectories()
        self.compute_fields()

    def _setup(self):
        mask_array = self.state.merge_active_masks()
        bbox, bnx, bny = get_bbox(~mask_array)
        pad = self.geodata.pad_width
        nxp = self.geodata.roi_nx + pad * 2
        nyp = self.geodata.roi_ny + pad * 2
        info = Info(self.state, self.params, self.geodata.roi_pixel_size)
        info.set_xy(bnx, bny, pad)
        return mask_array, bbox, pad, nxp, nyp, info

    def compute_trajectories(self):
        mask_array, bbox, pad, nxp, nyp, info = self._setup()
        mapping_array = np.zeros((nxp, nyp), dtype=np.uint32)
        data = Data(info=info, bbox=bbox, pad=pad, mapping_array=mapping_array, mask_array=mask_array,
                    uv_array=self.preprocess.uv_array)
        tr = Trajectories(0, 0, info=info, data=data,
                          do_trace_downstream=getattr(self.params, "do_trace_downstream", True),
                          do_trace_upstream=getattr(self.params, "do_trace_upstream", True),
                          verbose=getattr(self.state, "verbose", False), fp_contract=self.fp_contract)
        tr.integrate()
        offset_xy = np.array((bbox[0] - pad, bbox[2] - pad))
        self.seed_point_array = tr.data.seed_point_array + offset_xy
        self.streamline_arrays_list = tr.data.streamline_arrays_list
        self.traj_stats_df = tr.data.traj_stats_df
        self.traj_nsteps_array = tr.data.traj_nsteps_array
        self.traj_length_array = tr.data.traj_length_array

    def compute_fields(self):
        mask_array, bbox, pad, nxp, nyp, info = self._setup()
        mapping_array = np.zeros((nxp, nyp), dtype=np.uint32)
        data = Data(info=info, bbox=bbox, pad=pad, mask_array=mask_array, uv_array=self.preprocess.uv_array,
                    mapping_array=mapping_array, traj_stats_df=self.traj_stats_df)
        fields = Fields(0, 0, info=info, data=data, verbose=getattr(self.state, "verbose", False),
                        fp_contract=self.fp_contract)
        fields.integrate()
        self.fields = fields
        self.slc_array = np.zeros((nxp, nyp, 2), dtype=np.uint32)
        self.slt_array = np.zeros((nxp, nyp, 2), dtype=np.float32)
        self.sla_array = np.zeros((nxp, nyp, 2), dtype=np.float32)
        bounds = data.bounds_slx
        self.slc_array[bounds] = fields.data.slc_array
        self.slt_array[bounds] = fields.data.slt_array
        self.sla_array[bounds] = fields.data.sla_array
