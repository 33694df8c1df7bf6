"""Host helpers of the tracing path, with the reference's names and contracts.

Restates from python/streamlines/useful.py:
  Data            :37-115   bbox crop (+-pad) of the grids handed to the tracer
  Info            :117-268  every derived kernel parameter
  get_bbox        :270-289
  create_seeds    :470-517  argwhere(~mask)-pad, legacy MT19937 shuffle
  pick_seeds      :519-538
  compute_stats   :540-571  per-direction min/mean/max of l, n, ds
"""
from __future__ import annotations

import sys

import numpy as np

__all__ = ["Data", "Info", "get_bbox", "create_seeds", "pick_seeds", "compute_stats", "vprint"]

FLAG_NAMES = [
    "is_channel", "is_thinchannel", "is_interchannel", "is_channelhead", "is_channeltail",
    "is_majorconfluence", "is_minorconfluence", "is_majorinflow", "is_minorinflow",
    "is_leftflank", "is_rightflank", "is_midslope", "is_ridge", "was_channelhead",
    "is_subsegmenthead", "is_loop", "is_blockage",
]


def vprint(verbose, *args, **kwargs):
    if verbose:
        print(*args, **kwargs, flush=True)
        sys.stdout.flush()


class Data:
    """Grids cropped to ``bbox`` widened by ``pad`` (useful.py:48-114)."""

    def __init__(self, info=None, bbox=None, pad=0, mask_array=None, uv_array=None,
                 mapping_array=None, traj_stats_df=None, sla_array=None, slc_array=None,
                 slt_array=None, hsl_array=None):
        if bbox is not None:
            bbox = [bbox[0] - pad, bbox[1] + pad, bbox[2] - pad, bbox[3] + pad]
            bounds_grid = np.index_exp[bbox[0]:(bbox[1] + 1), bbox[2]:(bbox[3] + 1)]
            bounds_slx = np.index_exp[bbox[0]:(bbox[1] + 1), bbox[2]:(bbox[3] + 1), :]
        else:
            bounds_grid = np.index_exp[:, :]
            bounds_slx = np.index_exp[:, :, :]
        self.mask_array = mask_array[bounds_grid].copy()
        self.mapping_array = mapping_array[bounds_grid].copy() if mapping_array is not None else None
        self.uv_array = uv_array[bounds_slx].copy() if uv_array is not None else None
        self.sla_array = sla_array[bounds_slx].copy() if sla_array is not None else None
        self.slc_array = slc_array[bounds_slx].copy() if slc_array is not None else None
        self.slt_array = slt_array[bounds_slx].copy() if slt_array is not None else None
        self.hsl_array = hsl_array[bounds_slx].copy() if hsl_array is not None else None
        self.traj_stats_df = traj_stats_df
        self.bounds_grid = bounds_grid
        self.bounds_slx = bounds_slx
        self.count_array = None
        self.link_array = None
        self.label_array = None
        self.traj_length_array = None


class Info:
    """Derived kernel parameters (useful.py:128-220); attribute names as there.

    ``state`` and ``trace`` are any objects carrying the JSON keys of the
    ``state`` and ``trace`` sections as attributes."""

    def __init__(self, state, trace, pixel_size, mapping=None, n_seed_points=None,
                 coarse_label=None):
        self.state = state
        self.trace = trace
        self.mapping = mapping
        self.do_measure_hsl_from_ridges = bool(mapping.do_measure_hsl_from_ridges) if mapping is not None else False
        self.segmentation_threshold = np.uint32(0)
        self.channel_threshold = np.uint32(0)
        if coarse_label is not None:
            self.coarse_label = coarse_label
        if np.float32(trace.max_length) == np.float32(0.0):
            max_length = np.finfo(np.float32).max
        else:
            max_length = trace.max_length
        max_n_steps = np.uint32(trace.max_length / trace.integrator_step_factor)
        if trace.interchannel_max_n_steps == 0:
            interchannel_max_n_steps = max_n_steps
        else:
            interchannel_max_n_steps = trace.interchannel_max_n_steps
        self.max_n_steps = np.uint32(max_n_steps)

        # numpy-1.x promotion: 1.0/np.float32 is float64 (useful.py:166-168)
        seed_point_density = np.float32(trace.subpixel_seed_point_density)
        subpixel_seed_span = 1 - 1.0 / float(seed_point_density)
        subpixel_seed_step = subpixel_seed_span / max(float(seed_point_density), 1.0)

        self.debug = bool(getattr(state, "debug", False))
        self.verbose = bool(getattr(state, "gpu_verbose", False))
        self.n_trajectory_seed_points = np.uint32(trace.n_trajectory_seed_points)
        self.n_seed_points = np.uint32(0)
        self.n_padded_seed_points = np.uint32(0)
        self.do_shuffle = bool(trace.do_shuffle_seed_points)
        self.shuffle_rng_seed = np.uint32(trace.shuffle_rng_seed)
        self.downup_sign = np.float32(np.nan)
        self.gpu_memory_limit_pc = np.uint32(getattr(state, "gpu_memory_limit_pc", 90))
        self.n_work_items = np.uint32(getattr(state, "n_work_items", 32))
        self.chunk_size_factor = np.uint32(getattr(state, "chunk_size_factor", 20))
        self.max_time_per_kernel = np.float32(getattr(state, "max_time_per_kernel", 5.0))
        self.integrator_step_factor = np.float32(trace.integrator_step_factor)
        self.max_integration_step_error = np.float32(trace.max_integration_step_error)
        self.adjusted_max_error = 0.85 * np.sqrt(np.float64(trace.max_integration_step_error))
        self.integration_halt_threshold = np.float32(trace.integration_halt_threshold)
        self.max_length = np.float32(max_length / pixel_size)
        self.pixel_size = np.float32(pixel_size)
        self.trajectory_resolution = np.uint32(trace.trajectory_resolution)
        self.seeds_chunk_offset = np.uint32(0)
        self.subpixel_seed_point_density = np.uint32(trace.subpixel_seed_point_density)
        self.subpixel_seed_halfspan = np.float32(subpixel_seed_span / 2.0)
        self.subpixel_seed_step = np.float32(subpixel_seed_step)
        self.jitter_magnitude = np.float32(trace.jitter_magnitude)
        self.interchannel_max_n_steps = np.uint32(interchannel_max_n_steps)
        self.left_flank_addition = 2147483648
        self.set_flags(self)

    def set_xy(self, nx, ny, pad, bbox=None):
        self.bbox = bbox
        self.pad_width = pad
        self.nx = nx
        self.ny = ny
        self.nx_padded = self.nx + self.pad_width * 2
        self.ny_padded = self.ny + self.pad_width * 2

    def set_thresholds(self, segmentation_threshold=None, channel_threshold=None):
        if segmentation_threshold is not None:
            self.segmentation_threshold = segmentation_threshold
        if channel_threshold is not None:
            self.channel_threshold = channel_threshold

    @staticmethod
    def set_flags(obj):
        for idx, flag in enumerate(FLAG_NAMES):
            setattr(obj, flag, np.uint(2 ** idx))


def get_bbox(array):
    """Index span of True elements (useful.py:270-289)."""
    cols = np.any(array, axis=0)
    rows = np.any(array, axis=1)
    x_min, x_max = np.where(rows)[0][[0, -1]]
    y_min, y_max = np.where(cols)[0][[0, -1]]
    return (x_min, x_max, y_min, y_max), (x_max + 1 - x_min), (y_max + 1 - y_min)


def create_seeds(mask, pad, n_work_items, n_seed_points=None, do_shuffle=True, rng_seed=1,
                 verbose=False):
    """Seed list of useful.py:470-517: pixel centres of unmasked pixels minus pad,
    shuffled by the legacy global MT19937 stream seeded with ``rng_seed``."""
    vprint(verbose, "Generating seed points...", end="")
    seed_point_array = np.argwhere(~mask).astype(np.float32) - pad
    seed_point_array = seed_point_array.astype(np.float32)
    if do_shuffle:
        np.random.seed(int(rng_seed))
        np.random.shuffle(seed_point_array)
    if n_seed_points is not None and n_seed_points > 0:
        # the reference keeps n_seed_points even when fewer seeds exist, and its
        # kernels then read past the list; here the count is capped instead
        seed_point_array = seed_point_array[:n_seed_points].astype(np.float32)
        n_seed_points = min(int(n_seed_points), seed_point_array.shape[0])
    else:
        n_seed_points = seed_point_array.shape[0]
    pad_length = (np.uint32(np.round(n_seed_points / n_work_items + 0.5)) * n_work_items - n_seed_points)
    n_padded_seed_points = n_seed_points + pad_length
    vprint(verbose, "done")
    return seed_point_array.copy(), n_seed_points, n_padded_seed_points


def pick_seeds(mask=None, map=None, flag=None, pad=None):
    """useful.py:519-538"""
    if mask is None and map is not None:
        seed_point_array = np.argwhere((map & flag) > 0).astype(np.float32) - pad
    elif mask is not None and map is None:
        seed_point_array = np.argwhere(~mask).astype(np.float32) - pad
    else:
        seed_point_array = np.argwhere(~mask & ((map & flag) > 0)).astype(np.float32) - pad
    return seed_point_array.astype(np.float32).copy()


def compute_stats(traj_length_array, traj_nsteps_array, pixel_size, verbose=False):
    """Per-direction min/mean/max of (l*pixel_size, n, l*pixel_size/n)
    (useful.py:540-571).  Arithmetic is float32 as in the reference: the mean
    is a float32 row-order sum divided by the count."""
    import pandas as pd

    vprint(verbose, "Computing streamlines statistics")
    pixel_size = np.float32(pixel_size)
    stats = []
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in (0, 1):
            l = (traj_length_array[:, k].astype(np.float32) * pixel_size).astype(np.float32)
            n = traj_nsteps_array[:, k].astype(np.float32)
            ds = (l / n).astype(np.float32)
            lnds = np.stack([l, n, ds], axis=1).astype(np.float32)
            mean = np.zeros(3, dtype=np.float32)
            for row in lnds:
                mean = (mean + row).astype(np.float32)
            mean = (mean / np.float32(lnds.shape[0])).astype(np.float32)
            stats += [lnds.min(axis=0), mean, lnds.max(axis=0)]
    arr = np.array(stats, dtype=np.float32)
    index = [np.array(["downstream"] * 3 + ["upstream"] * 3),
             np.array(["min", "mean", "max", "min", "mean", "max"])]
    df = pd.DataFrame(data=arr, columns=["l", "n", "ds"], index=index)
    vprint(verbose, df.T)
    return df
