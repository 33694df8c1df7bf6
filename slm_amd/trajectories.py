"""Un-jittered streamline trajectories on the MI355X.

Drop-in for python/streamlines/trajectories.py: same class name, constructor
signature and outputs (data.streamline_arrays_list, traj_nsteps_array,
traj_length_array, traj_stats_df).  Each direction is one slm_trace_paths
launch (replacing kernel integrate_trajectory, trajectories.py:257-333); the
two directions run on two streams, so their launches overlap; no
memory chunking is needed on a 288 GB device, and since results are per seed
they do not depend on chunking anyway.
"""
from __future__ import annotations

import numpy as np

from . import hip
from .params import kernel_params
from .useful import compute_stats, create_seeds, vprint

__all__ = ["Trajectories"]


class Trajectories:
    def __init__(self, which_cl_platform=0, which_cl_device=0, cl_src_path=None, info=None,
                 data=None, do_trace_downstream=True, do_trace_upstream=True, verbose=False,
                 gpu_verbose=False, fp_contract=False):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("Trajectories needs a HIP device (torch.cuda.is_available() is False)")
        hip.load()
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.cl_src_path = cl_src_path
        self.info = info
        self.data = data
        self.do_trace_downstream = do_trace_downstream
        self.do_trace_upstream = do_trace_upstream
        self.verbose = verbose
        self.gpu_verbose = gpu_verbose
        self.fp_contract = fp_contract

    def integrate(self):
        """trajectories.py:78-164"""
        vprint(self.verbose, "Integrating trajectories...")
        info = self.info
        self.data.seed_point_array, n_seed_points, n_padded_seed_points = create_seeds(
            self.data.mask_array, info.pad_width, info.n_work_items,
            n_seed_points=info.n_trajectory_seed_points, do_shuffle=info.do_shuffle,
            rng_seed=info.shuffle_rng_seed, verbose=self.verbose)
        info.n_seed_points = n_seed_points
        info.n_padded_seed_points = n_padded_seed_points
        if self.data.mapping_array is None:
            self.data.mapping_array = np.zeros_like(self.data.mask_array, dtype=np.uint32)
        self.gpu_compute_trajectories()
        self.data.traj_stats_df = compute_stats(self.data.traj_length_array, self.data.traj_nsteps_array,
                                                info.pixel_size, self.verbose)
        vprint(self.verbose, "...done")

    def gpu_compute_trajectories(self):
        """trajectories.py:196-346, one launch per direction."""
        import torch
        info = self.info
        dev = self.device
        ns = int(info.n_seed_points)
        nl = int(info.max_n_steps)
        mask = self.data.mask_array
        seeds = torch.from_numpy(np.ascontiguousarray(self.data.seed_point_array, dtype=np.float32)).to(dev)
        mask_d = torch.from_numpy(np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)).to(dev)
        uv_d = torch.from_numpy(np.ascontiguousarray(self.data.uv_array, dtype=np.float32).reshape(-1)).to(dev)
        traj_nsteps_array = np.zeros([ns, 2], dtype=np.uint16)
        traj_length_array = np.zeros([ns, 2], dtype=np.float32)
        streamline_arrays_list = [[], []]
        todo = [(self.do_trace_downstream, 0, np.float32(+1.0)), (self.do_trace_upstream, 1, np.float32(-1.0))]
        # both directions launched before either is copied back, on two
        # streams so that the two launches overlap
        main = torch.cuda.current_stream(dev)
        side = hip.side_stream(dev)
        side.wait_stream(main)
        outs = {}
        for do, idx, sign in todo:
            if not do:
                continue
            info.downup_sign = sign
            info.seeds_chunk_offset = np.uint32(0)
            p = kernel_params(info, sign, "INTEGRATE_TRAJECTORY", self.fp_contract)
            traj_d = torch.zeros((ns, nl, 2), dtype=torch.int8, device=dev)
            nsteps_d = torch.zeros(ns, dtype=torch.int16, device=dev)
            length_d = torch.zeros(ns, dtype=torch.float32, device=dev)
            stream = main if idx == 0 else side
            if idx == 1:
                side.wait_stream(main)   # the zero fills above ran on main
            hip.trace_paths(p, seeds, mask_d, uv_d, traj_d, nsteps_d, length_d, stream=stream)
            outs[idx] = (traj_d, nsteps_d, length_d)
        torch.cuda.synchronize(dev)
        for idx, (traj_d, nsteps_d, length_d) in outs.items():
            # only the used steps leave the device: packed end to end there
            # (trajectories.py:320-325 slices each seed's row on the host)
            offsets_d, packed_d = hip.paths_pack(traj_d, nsteps_d, nl)
            offsets = offsets_d.cpu().numpy()
            packed = packed_d.cpu().numpy()
            traj_nsteps_array[:, idx] = nsteps_d.cpu().numpy().view(np.uint16)
            traj_length_array[:, idx] = length_d.cpu().numpy()
            streamline_arrays_list[idx] = np.split(packed, offsets[1:-1])
        self.data.streamline_arrays_list = streamline_arrays_list
        self.data.traj_nsteps_array = traj_nsteps_array
        self.data.traj_length_array = traj_length_array
