"""ctypes binding of libslm_hip.so (include/slm_hip.h) and torch-tensor wrappers.

The library is the only compute path: if it is missing, importing this module's
functions raises; nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import SlmFieldsJob, SlmBatchOpts, SlmKdeParams, SlmNnFlip, SlmParams, SlmTileOpts, ABI_VERSION

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libslm_hip.so")

SLM_FLAG_SYNC = 1
SLM_FLAG_NO_ZERO = 2
SLM_FLAG_KEEP_ORDER = 4
SLM_FLAG_FORCE_FALLBACK = 8
SLM_FLAG_FMM_LANES = 32

_ERR_NAMES = {-1: "SLM_E_INVALID_ARG", -2: "SLM_E_HIP", -3: "SLM_E_OOM", -4: "SLM_E_ABI",
              -5: "SLM_E_WORKSPACE"}

# Exported symbols and their prototypes: (restype, argtypes)
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I = ctypes.c_int
_PROTOS = {
    "slm_abi_version": (_U32, []),
    "slm_last_error": (ctypes.c_char_p, []),
    "slm_device_count": (ctypes.c_int, []),
    "slm_rng_stream": (ctypes.c_int, [_U32, _U32, _U32, _P, _P, _P, _U32]),
    "slm_check_rounding": (ctypes.c_int, [_P, _U32, _P, _U32]),
    "slm_calibrate_valu": (ctypes.c_int, [_U32, _U32, _U32, _P, _P, _U32]),
    "slm_trace_fields_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(SlmParams)]),
    "slm_trace_fields": (ctypes.c_int, [ctypes.POINTER(SlmParams), _P, _P, _P, _P, _P, _P, _P,
                                        ctypes.c_size_t, _P, _U32]),
    "slm_trace_fields_noalloc_max_bytes": (ctypes.c_size_t, []),
    "slm_trace_fields_noalloc": (ctypes.c_int, [ctypes.POINTER(SlmParams), _P, _P, _P, _P, _P, _P, _P, _U32]),
    "slm_trace_fields_tiled_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(SlmParams),
                                                                ctypes.POINTER(SlmTileOpts)]),
    "slm_trace_fields_tiled": (ctypes.c_int, [ctypes.POINTER(SlmParams), ctypes.POINTER(SlmTileOpts), _P, _P, _P,
                                              _P, _P, _P, _P, ctypes.c_size_t, _P, _U32]),
    "slm_trace_paths": (ctypes.c_int, [ctypes.POINTER(SlmParams), _P, _P, _P, _P, _P, _P, _P, _U32]),
    "slm_paths_offsets": (ctypes.c_int, [_P, _U32, _U32, _P, _P, _U32]),
    "slm_paths_pack": (ctypes.c_int, [_P, _P, _P, _U32, _U32, _P, _P, _U32]),
    "slm_trace_fields_record": (ctypes.c_int, [ctypes.POINTER(SlmParams), _P, _P, _P, _P, _P, _U32,
                                               _U32, _P, _P, _P, _U32]),
    "slm_hillslope_lengths": (ctypes.c_int, [ctypes.POINTER(SlmParams), _P, _P, _P, _P, _U32, _U32, _P, _P,
                                             _U32]),
    "slm_kde_histogram": (ctypes.c_int, [ctypes.POINTER(SlmKdeParams), _P, _P, _P, _U32]),
    "slm_kde_pdf": (ctypes.c_int, [ctypes.POINTER(SlmKdeParams), _P, _P, _P, _U32]),
    "slm_map_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(SlmParams)]),
    "slm_nn_conv": (ctypes.c_int, [_P, _I, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _P, _U32]),
    "slm_nn_conv_up": (ctypes.c_int, [_P, _I, _P, _I, _I, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _P, _U32]),
    "slm_nn_wflip": (ctypes.c_int, [_P, _P, _I, _I, _I, _P, _U32]),
    "slm_nn_conv_wgrad": (ctypes.c_int, [_P, _I, _P, _I, _P, _P, _I, _I, _I, _I, _I, _I, _P, _U32]),
    "slm_nn_relu_bias_grad": (ctypes.c_int, [_P, _I, _P, _I, _P, _I, _I, _P, _I, _P, _U32]),
    "slm_nn_relu_bwd": (ctypes.c_int, [_P, _I, _P, _I, _I, _I, _P, _U32]),
    "slm_nn_pool": (ctypes.c_int, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _U32]),
    "slm_nn_pool_bwd": (ctypes.c_int, [_P, _I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _P, _U32]),
    "slm_nn_up": (ctypes.c_int, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _U32]),
    "slm_nn_up_bwd": (ctypes.c_int, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _U32]),
    "slm_nn_add": (ctypes.c_int, [_P, _I, _P, _I, _I, _I, _P, _U32]),
    "slm_nn_loss": (ctypes.c_int, [_P, _I, _P, _P, _P, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P, _P,
                                   _U32]),
    "slm_nn_adam": (ctypes.c_int, [_P, _P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, _I, _P, _U32]),
    "slm_nn_adam_flip": (ctypes.c_int, [_P, _P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, _I, ctypes.POINTER(SlmNnFlip), _I, _P, _U32]),
    "slm_nn_tile_inputs": (ctypes.c_int, [_P, _P, _P, _I, _I, _P, _I, _I, ctypes.c_float, _P, _P, _P, _U32]),
    "slm_nn_tile_targets": (ctypes.c_int, [_P, _P, _P, _I, _I, _P, _I, _I, _I, ctypes.c_float, _P, _P, _P, _U32]),
    "slm_nn_tile_stitch": (ctypes.c_int, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _P, _U32]),
    "slm_nn_fields_log": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, _I, ctypes.c_float, _P, _P, _U32]),
    "slm_trace_fields_batch_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(SlmFieldsJob), _U32]),
    "slm_trace_fields_batch": (ctypes.c_int, [ctypes.POINTER(SlmFieldsJob), _U32, ctypes.POINTER(SlmBatchOpts), _P,
                                              _P, ctypes.c_size_t, _P, _U32]),
    "slm_gradient_uv": (ctypes.c_int, [_P, _U32, _U32, _U32, ctypes.c_float, ctypes.c_int, _P, _P, _P, _U32]),
    "slm_condition_workspace_size": (ctypes.c_size_t, [_U32, _U32]),
    "slm_slopes_workspace_size": (ctypes.c_size_t, [_U32, _U32]),
    "slm_gaussian_filter": (ctypes.c_int, [_P, _P, _P, _U32, _U32, _P, ctypes.c_int, _P, _U32]),
    "slm_map_midslopes_ridges": (ctypes.c_int, [_P, _P, _U32, _U32, _P, ctypes.c_int, ctypes.c_float, _P, ctypes.c_int,
                                                ctypes.c_float, _U32, _U32, _P, _P, ctypes.c_size_t, _P, _U32]),
    "slm_fmm_pairs": (ctypes.c_int, [_P, _U32, _U32, _P, _U32, _P, _U32, _P, _P, _U32]),
    "slm_fmm_workspace_size": (ctypes.c_size_t, [_U32, _U32]),
    "slm_fmm_workspace_size_k": (ctypes.c_size_t, [_U32, _U32, _U32]),
    "slm_fmm_hsl": (ctypes.c_int, [_P, _P, _U32, _P, _U32, _U32, _U32, _U32, ctypes.c_double, _U32, _P, _P, _P, _P,
                                   _P, _P, ctypes.c_size_t, _P, _U32]),
    "slm_hsl_remap": (ctypes.c_int, [_P, _U32, _P, _P, _U32, _P, _P, _U32]),
    "slm_map_hsl_workspace_size": (ctypes.c_size_t, [_U32, _U32]),
    "slm_map_hsl": (ctypes.c_int, [_P, _P, _U32, _U32, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_size_t, _P, _U32]),
    "slm_map_aspect_workspace_size": (ctypes.c_size_t, [_U32, _U32]),
    "slm_map_aspect": (ctypes.c_int, [_P, _P, _P, _U32, _U32, _U32, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, _P, _P, _P, ctypes.c_size_t, _P, _U32]),
    "slm_slope_deg": (ctypes.c_int, [_P, _U32, _U32, _U32, ctypes.c_float, _P, _P, _U32]),
    "slm_condition_uv": (ctypes.c_int, [_P, _U32, _U32, _U32, ctypes.c_float, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, _P, _P, _P, _P, ctypes.c_size_t, _P, _U32]),
    "slm_fields_finish_raw": (ctypes.c_int, [_P, _P, _U32, _U32, _P, _P, _P, _U32]),
    "slm_fields_finish_scaled": (ctypes.c_int, [_P, _U32, _U32, ctypes.c_float, ctypes.c_float, _P, _P, _P,
                                                _U32]),
    "slm_thin_channels": (ctypes.c_int, [ctypes.POINTER(SlmParams), _P, _P, _U32, _P, ctypes.c_size_t,
                                         ctypes.POINTER(_U32), _P, _U32]),
}

# Mapping entry points: C name -> (device array arguments in C order, takes a workspace)
MAP_ARGS = {
    "connect_channels": (("seeds", "mask", "uv", "mapping"), False),
    "map_channel_heads": (("seeds", "mask", "uv", "mapping"), False),
    "count_downchannels": (("seeds", "mask", "uv", "mapping", "count", "link"), True),
    "flag_downchannels": (("seeds", "mask", "mapping", "count", "link"), False),
    "label_confluences": (("seeds", "mask", "slt", "mapping", "count", "link"), False),
    "segment_downchannels": (("seeds", "mask", "mapping", "count", "link", "label"), True),
    "link_hillslopes": (("seeds", "mask", "uv", "link"), False),
    "segment_hillslopes": (("seeds", "mask", "uv", "mapping", "label"), False),
    "subsegment_channel_edges": (("seeds", "mask", "mapping", "link", "label"), False),
    "subsegment_flanks": (("seeds", "mask", "uv", "mapping", "label"), False),
    "segment_hillslopes_flanks": (("seeds", "mask", "uv", "mapping", "label"), False),
    "fix_right_flanks": (("seeds", "mask", "uv", "mapping", "label"), True),
    "fix_left_flanks": (("seeds", "mask", "uv", "mapping", "label"), True),
}
for _name, (_args, _ws) in MAP_ARGS.items():
    _PROTOS["slm_" + _name] = (ctypes.c_int, [ctypes.POINTER(SlmParams)] + [_P] * len(_args) +
                               ([_P, ctypes.c_size_t] if _ws else []) + [_P, _U32])

_lib = None


class SlmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("{} ({}): {}".format(_ERR_NAMES.get(code, "SLM_E?"), code, msg))
        self.code = code


def load(path: str = LIB_PATH):
    """Load and bind libslm_hip.so; raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libslm_hip.so not found at {} -- build it with "
                          "__graft_entry__.build()".format(path))
    lib = ctypes.CDLL(path)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.slm_abi_version() != ABI_VERSION:
        raise ImportError("libslm_hip.so ABI {} != {}".format(lib.slm_abi_version(), ABI_VERSION))
    _lib = lib
    return lib


def exported_symbols():
    return list(_PROTOS)


def check(rc: int):
    if rc != 0:
        raise SlmError(rc, load().slm_last_error().decode(errors="replace"))


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


_SIDE_STREAMS = {}


def side_stream(device):
    """A second stream on device (default priority), created once per device
    and reused: stream creation costs far more than a short launch."""
    import torch
    d = torch.device(device)
    idx = d.index if d.index is not None else torch.cuda.current_device()
    if idx not in _SIDE_STREAMS:
        _SIDE_STREAMS[idx] = torch.cuda.Stream(device=idx)
    return _SIDE_STREAMS[idx]


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _require_cuda(*tensors):
    for t in tensors:
        if t is not None and (not t.is_cuda or not t.is_contiguous()):
            raise ValueError("tensors must be contiguous CUDA (HIP) tensors")


def rng_stream(idx0: int, n: int, n_draws: int, rng0=None, draws=None, stream=None, sync=False):
    _require_cuda(rng0, draws)
    check(load().slm_rng_stream(idx0, n, n_draws, _ptr(rng0), _ptr(draws), _stream(stream),
                                SLM_FLAG_SYNC if sync else 0))


def check_rounding(n_numerators: int = 8, stream=None):
    """slm_check_rounding: returns the 8 counts as Python ints (entries 0, 2,
    4, 6, 7 count mismatches of the short sqrt/divide forms)."""
    import torch
    counts = torch.zeros(8, dtype=torch.int64, device=torch.device("cuda", torch.cuda.current_device()))
    check(load().slm_check_rounding(_ptr(counts), n_numerators, _stream(stream), SLM_FLAG_SYNC))
    return [int(v) for v in counts.cpu()]


def calibrate_valu(n_blocks: int, n_iter: int, out, kind: int = 0, stream=None, sync=False):
    """slm_calibrate_valu: n_blocks blocks of 256 threads, each wave issuing
    16*n_iter independent instructions of one kind (0 v_fma_f32, 1 v_add_f32,
    2 v_add_u32, 3 v_xor_b32, 4 v_mul_f32, 5 fma/add_u32 alternating); out: float32
    [n_blocks]."""
    _require_cuda(out)
    check(load().slm_calibrate_valu(kind, n_blocks, n_iter, _ptr(out), _stream(stream),
                                    SLM_FLAG_SYNC if sync else 0))


def fields_workspace_size(p: SlmParams) -> int:
    n = load().slm_trace_fields_workspace_size(ctypes.byref(p))
    if n == 0:
        check(-1)
    return int(n)


def trace_fields(p: SlmParams, seeds, mask, uv, slc, slt, steps=None, workspace=None,
                 stream=None, accumulate=False, sync=False, keep_order=False, force_fallback=False):
    """One direction of integrate_fields on device tensors.
    seeds float32 [N,2]; mask uint8 [nxp*nyp]; uv float32 [nxp*nyp*2];
    slc/slt int32 [nxp*nyp] (uint32 bit patterns); steps int64 [1] or None."""
    import torch
    _require_cuda(seeds, mask, uv, slc, slt, steps, workspace)
    if workspace is None:
        workspace = torch.empty(fields_workspace_size(p), dtype=torch.uint8, device=slc.device)
    flags = (SLM_FLAG_NO_ZERO if accumulate else 0) | (SLM_FLAG_SYNC if sync else 0) | \
        (SLM_FLAG_KEEP_ORDER if keep_order else 0) | (SLM_FLAG_FORCE_FALLBACK if force_fallback else 0)
    check(load().slm_trace_fields(ctypes.byref(p), _ptr(seeds), _ptr(mask), _ptr(uv), _ptr(slc),
                                  _ptr(slt), _ptr(steps), _ptr(workspace), workspace.numel(),
                                  _stream(stream), flags))
    return workspace


def trace_fields_noalloc(p: SlmParams, seeds, mask, uv, slc, slt, steps=None, stream=None, accumulate=False,
                         sync=False, keep_order=False):
    """slm_trace_fields_noalloc: trace_fields with the library's fixed device
    scratch array as workspace (no allocation by the library or here)."""
    _require_cuda(seeds, mask, uv, slc, slt, steps)
    flags = (SLM_FLAG_NO_ZERO if accumulate else 0) | (SLM_FLAG_SYNC if sync else 0) | \
        (SLM_FLAG_KEEP_ORDER if keep_order else 0)
    check(load().slm_trace_fields_noalloc(ctypes.byref(p), _ptr(seeds), _ptr(mask), _ptr(uv), _ptr(slc), _ptr(slt),
                                          _ptr(steps), _stream(stream), flags))


def noalloc_max_bytes() -> int:
    return int(load().slm_trace_fields_noalloc_max_bytes())


def tile_opts(tile=0, block=0, unit=0, cap_steps=0, halo=0) -> SlmTileOpts:
    o = SlmTileOpts()
    o.tile, o.block, o.unit, o.cap_steps, o.halo = tile, block, unit, cap_steps, halo
    return o


def tiled_workspace_size(p: SlmParams, opts: SlmTileOpts = None) -> int:
    opts = opts if opts is not None else tile_opts()
    n = load().slm_trace_fields_tiled_workspace_size(ctypes.byref(p), ctypes.byref(opts))
    if n == 0:
        check(-1)
    return int(n)


def trace_fields_tiled(p: SlmParams, seeds, mask, uv, slc, slt, steps=None, workspace=None, opts=None,
                       stream=None, accumulate=False, sync=False, force_fallback=False):
    """slm_trace_fields_tiled: same result as trace_fields, LDS-owned tiles."""
    import torch
    _require_cuda(seeds, mask, uv, slc, slt, steps, workspace)
    opts = opts if opts is not None else tile_opts()
    need = tiled_workspace_size(p, opts)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=slc.device)
    flags = (SLM_FLAG_NO_ZERO if accumulate else 0) | (SLM_FLAG_SYNC if sync else 0) | \
        (SLM_FLAG_FORCE_FALLBACK if force_fallback else 0)
    check(load().slm_trace_fields_tiled(ctypes.byref(p), ctypes.byref(opts), _ptr(seeds), _ptr(mask), _ptr(uv),
                                        _ptr(slc),
                                        _ptr(slt), _ptr(steps), _ptr(workspace), workspace.numel(),
                                        _stream(stream), flags))
    return workspace


def trace_paths(p: SlmParams, seeds, mask, uv, traj, nsteps, length, stream=None, sync=False):
    """One direction of integrate_trajectory: traj int8 [n,max_n_steps,2],
    nsteps int16 [n] (uint16 bits), length float32 [n]."""
    _require_cuda(seeds, mask, uv, traj, nsteps, length)
    check(load().slm_trace_paths(ctypes.byref(p), _ptr(seeds), _ptr(mask), _ptr(uv), _ptr(traj),
                                 _ptr(nsteps), _ptr(length), _stream(stream),
                                 SLM_FLAG_SYNC if sync else 0))


def paths_pack(traj, nsteps, max_n_steps, stream=None, sync=False):
    """Recorded trajectories packed end to end on the device
    (slm_paths_offsets, then slm_paths_pack): traj int8 [n, max_n_steps, 2],
    nsteps int16 [n] (uint16 bits).  Returns (offsets int64 [n + 1],
    packed int8 [offsets[n], 2]); reading the total waits for the stream."""
    import torch
    _require_cuda(traj, nsteps)
    n = nsteps.numel()
    if traj.numel() != n * max_n_steps * 2:
        raise ValueError("traj holds %d bytes, not n * max_n_steps * 2 = %d" % (traj.numel(), n * max_n_steps * 2))
    offsets = torch.empty(n + 1, dtype=torch.int64, device=nsteps.device)
    check(load().slm_paths_offsets(_ptr(nsteps), n, int(max_n_steps), _ptr(offsets), _stream(stream), 0))
    if stream is not None:
        stream.synchronize()
    total = int(offsets[n].item())
    packed = torch.empty((total, 2), dtype=torch.int8, device=nsteps.device)
    check(load().slm_paths_pack(_ptr(traj), _ptr(nsteps), _ptr(offsets), n, int(max_n_steps), _ptr(packed),
                                _stream(stream), SLM_FLAG_SYNC if sync else 0))
    return offsets, packed


def trace_fields_record(p: SlmParams, seeds, mask, uv, seed_sel, sub_sel, max_rec, xy, nrec,
                        stream=None, sync=False):
    _require_cuda(seeds, mask, uv, seed_sel, sub_sel, xy, nrec)
    check(load().slm_trace_fields_record(ctypes.byref(p), _ptr(seeds), _ptr(mask), _ptr(uv),
                                         _ptr(seed_sel), _ptr(sub_sel), seed_sel.numel(), max_rec,
                                         _ptr(xy), _ptr(nrec), _stream(stream),
                                         SLM_FLAG_SYNC if sync else 0))


def hillslope_lengths(p: SlmParams, seeds, mask, uv, mapping, thin_flag, from_flag, length, stream=None,
                      sync=False):
    """lengths.cl hillslope_lengths on device tensors: mapping int32 [nxy]
    (uint32 bits), length float32 [n_seed_points] updated in place."""
    _require_cuda(seeds, mask, uv, mapping, length)
    check(load().slm_hillslope_lengths(ctypes.byref(p), _ptr(seeds), _ptr(mask), _ptr(uv), _ptr(mapping),
                                       int(thin_flag), int(from_flag), _ptr(length), _stream(stream),
                                       SLM_FLAG_SYNC if sync else 0))


def batch_opts(chunk=0, blocks_per_cu=0, queue_slices=0, lds_slots=0, refill_batch=0) -> SlmBatchOpts:
    o = SlmBatchOpts()
    o.chunk, o.blocks_per_cu, o.queue_slices, o.lds_slots = chunk, blocks_per_cu, queue_slices, lds_slots
    o.refill_batch = refill_batch
    return o


def fields_jobs(jobs) -> "ctypes.Array":
    """ctypes SlmFieldsJob array from (params, seeds, mask, uv, slc, slt) tuples
    of an SlmParams and device tensors."""
    arr = (SlmFieldsJob * len(jobs))()
    for i, (p, seeds, mask, uv, slc, slt) in enumerate(jobs):
        _require_cuda(seeds, mask, uv, slc, slt)
        arr[i].params = p
        arr[i].seeds_xy = seeds.data_ptr()
        arr[i].mask = mask.data_ptr()
        arr[i].uv = uv.data_ptr()
        arr[i].slc = slc.data_ptr()
        arr[i].slt = slt.data_ptr()
    return arr


def batch_workspace_size(job_array) -> int:
    n = load().slm_trace_fields_batch_workspace_size(job_array, len(job_array))
    if n == 0:
        check(-1)
    return int(n)


def trace_fields_batch(job_array, steps=None, workspace=None, opts=None, stream=None, accumulate=False,
                       keep_order=False, force_fallback=False, sync=False):
    """slm_trace_fields_batch over a fields_jobs() array; returns the workspace."""
    import torch
    _require_cuda(steps, workspace)
    need = batch_workspace_size(job_array)
    if workspace is None or workspace.numel() < need:
        dev = torch.device("cuda", torch.cuda.current_device())
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    flags = (SLM_FLAG_NO_ZERO if accumulate else 0) | (SLM_FLAG_SYNC if sync else 0) | \
        (SLM_FLAG_KEEP_ORDER if keep_order else 0) | (SLM_FLAG_FORCE_FALLBACK if force_fallback else 0)
    check(load().slm_trace_fields_batch(job_array, len(job_array), ctypes.byref(opts) if opts is not None else None,
                                        _ptr(steps), _ptr(workspace), workspace.numel(), _stream(stream), flags))
    return workspace


def map_workspace_size(p: SlmParams) -> int:
    n = load().slm_map_workspace_size(ctypes.byref(p))
    if n == 0:
        check(-1)
    return int(n)


def map_call(name: str, p: SlmParams, workspace=None, stream=None, sync=False, **arrays):
    """Call mapping entry point slm_<name> with device tensors passed by
    keyword (names of MAP_ARGS: seeds float32 [N,2], mask uint8, uv float32,
    slt float32, mapping/count/link/label int32 holding uint32 bits, all
    [nxy_padded]); p.n_seed_points must equal the seed count."""
    import torch
    args, takes_ws = MAP_ARGS[name]
    ts = [arrays[a] for a in args]
    _require_cuda(*ts)
    extra = []
    if takes_ws:
        need = map_workspace_size(p)
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=ts[-1].device)
        extra = [_ptr(workspace), workspace.numel()]
    check(getattr(load(), "slm_" + name)(ctypes.byref(p), *[_ptr(t) for t in ts], *extra, _stream(stream),
                                         SLM_FLAG_SYNC if sync else 0))
    return workspace


CHANNEL_BITS = 0x1 | 0x4   # IS_CHANNEL | IS_INTERCHANNEL, the pixels thin_channels thins


def medial_axis_tiebreak(n_set: int):
    """scikit-image medial_axis's tie-breaker for n_set pixels: numpy's legacy
    RandomState(0).permutation (uint32)."""
    return np.random.RandomState(0).permutation(np.arange(n_set)).astype(np.uint32)


def thin_channels(p: SlmParams, mapping, workspace=None, stream=None, medial_axis=True) -> int:
    """skeletonize(medial_axis(.)) of the channel pixels (medial_axis=False:
    the thinning alone); returns the thinning's iterations."""
    import torch
    _require_cuda(mapping, workspace)
    need = map_workspace_size(p)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=mapping.device)
    tie, n_set = None, 0
    if medial_axis:
        n_set = int(((mapping & CHANNEL_BITS) != 0).sum().item())
        tie = torch.from_numpy(medial_axis_tiebreak(n_set).view(np.int32)).to(mapping.device)
    it = _U32(0)
    check(load().slm_thin_channels(ctypes.byref(p), _ptr(mapping), _ptr(tie), n_set, _ptr(workspace),
                                   workspace.numel(), ctypes.byref(it), _stream(stream), 0))
    return int(it.value)


def gradient_uv(roi, pad: int, pixel_size: float, do_normalize_speed=True, stream=None, sync=False):
    """slm_gradient_uv: roi float32 [nx, ny] device tensor -> (uv float32
    [nx+2pad, ny+2pad, 2], mask bool [nx+2pad, ny+2pad])."""
    import torch
    _require_cuda(roi)
    nx, ny = roi.shape
    nxp, nyp = nx + 2 * pad, ny + 2 * pad
    uv = torch.empty((nxp, nyp, 2), dtype=torch.float32, device=roi.device)
    mask = torch.empty((nxp, nyp), dtype=torch.uint8, device=roi.device)
    check(load().slm_gradient_uv(_ptr(roi), nx, ny, pad, float(np.float32(pixel_size)),
                                 1 if do_normalize_speed else 0, _ptr(uv), _ptr(mask), _stream(stream),
                                 SLM_FLAG_SYNC if sync else 0))
    return uv, mask.to(torch.bool)


def condition_uv(roi, pad: int, pixel_size: float, do_normalize_speed=True, vecsum_threshold=0.95,
                 divergence_threshold=-0.5, curl_threshold=0.0, maps=False, workspace=None, stream=None, sync=False):
    """slm_condition_uv: roi float32 [nx, ny] device tensor -> (uv float32
    [nx+2pad, ny+2pad, 2], mask bool) after the loop and blockage fixes;
    with maps=True also (blockages, blocked_neighbors, loop cells) uint8 [nx, ny]."""
    import torch
    _require_cuda(roi, workspace)
    nx, ny = roi.shape
    nxp, nyp = nx + 2 * pad, ny + 2 * pad
    uv = torch.empty((nxp, nyp, 2), dtype=torch.float32, device=roi.device)
    mask = torch.empty((nxp, nyp), dtype=torch.uint8, device=roi.device)
    m = torch.empty((3, nx, ny), dtype=torch.uint8, device=roi.device) if maps else None
    need = int(load().slm_condition_workspace_size(nx, ny))
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=roi.device)
    check(load().slm_condition_uv(_ptr(roi), nx, ny, pad, float(np.float32(pixel_size)), 1 if do_normalize_speed else 0,
                                  float(vecsum_threshold), float(divergence_threshold), float(curl_threshold),
                                  _ptr(uv), _ptr(mask), _ptr(m), _ptr(workspace), workspace.numel(), _stream(stream),
                                  SLM_FLAG_SYNC if sync else 0))
    if maps:
        return uv, mask.to(torch.bool), m
    return uv, mask.to(torch.bool)


def gaussian_weights(sigma: float, truncate: float = 4.0):
    """The weights scipy.ndimage.gaussian_filter uses (float64 numpy, radius):
    exp(-x^2 / (2 sigma^2)) over |x| <= int(truncate*sigma + 0.5), normalised."""
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    w = np.exp(-0.5 / (float(sigma) ** 2) * x ** 2)
    return w / w.sum(), r


def gaussian_filter(x, sigma: float, truncate: float = 4.0, stream=None, sync=False):
    """slm_gaussian_filter: scipy.ndimage.gaussian_filter of a float32 [nx, ny]
    device tensor ('reflect' borders), bit for bit."""
    import torch
    _require_cuda(x)
    w, r = gaussian_weights(sigma, truncate)
    wd = torch.from_numpy(w).to(x.device)
    out = torch.empty_like(x)
    tmp = torch.empty_like(x)
    check(load().slm_gaussian_filter(_ptr(x), _ptr(out), _ptr(tmp), x.shape[0], x.shape[1], _ptr(wd), r,
                                     _stream(stream), SLM_FLAG_SYNC if sync else 0))
    return out


def map_midslopes_ridges(sla, mask, mapping, midslope_flag, ridge_flag, midslope_filter_sigma=5.0,
                         midslope_threshold=0.03, ridge_filter_sigma=2.0, ridge_threshold=0.8, stream=None,
                         sync=False):
    """slm_map_midslopes_ridges (mapping.py:779-817): sla float32 [nx, ny, 2]
    (or [nx*ny, 2]), mask bool/uint8 [nx, ny], mapping int32 [nx, ny] updated
    in place with the mid-slope and ridge skeleton flags."""
    import torch
    _require_cuda(sla, mask, mapping)
    nx, ny = mapping.shape
    wm, rm = gaussian_weights(midslope_filter_sigma)
    wr, rr = gaussian_weights(ridge_filter_sigma)
    wm_d, wr_d = torch.from_numpy(wm).to(sla.device), torch.from_numpy(wr).to(sla.device)
    m8 = mask.to(torch.uint8).contiguous()
    ws = torch.empty(int(load().slm_slopes_workspace_size(nx, ny)), dtype=torch.uint8, device=sla.device)
    level = float(np.float32(-(np.pi / 4) * ridge_threshold))
    sla_c = sla.contiguous()                  # held while the kernels run
    check(load().slm_map_midslopes_ridges(_ptr(sla_c), _ptr(m8), nx, ny, _ptr(wm_d), rm,
                                          float(np.float32(midslope_threshold)), _ptr(wr_d), rr, level,
                                          int(midslope_flag), int(ridge_flag), _ptr(mapping), _ptr(ws), ws.numel(),
                                          _stream(stream), SLM_FLAG_SYNC if sync else 0))
    return mapping


def fields_finish_raw(slc_raw, slt_raw, density: int, stream=None, sync=False):
    """slm_fields_finish_raw: slc_raw / slt_raw int32 [2, nxy] (uint32 bits)
    -> (slt, sla) float32 device tensors [nxy, 2]."""
    import torch
    _require_cuda(slc_raw, slt_raw)
    nxy = slc_raw.shape[1]
    slt = torch.empty((nxy, 2), dtype=torch.float32, device=slc_raw.device)
    sla = torch.empty_like(slt)
    check(load().slm_fields_finish_raw(_ptr(slc_raw), _ptr(slt_raw), nxy, int(density), _ptr(slt), _ptr(sla),
                                       _stream(stream), SLM_FLAG_SYNC if sync else 0))
    return slt, sla


def fields_finish_scaled(slc_raw, slt, density: int, ds_scale_down, ds_scale_up, stream=None, sync=False):
    """slm_fields_finish_scaled: slt float32 [nxy, 2] scaled and square-rooted
    in place; returns slc float64 [nxy, 2]."""
    import torch
    _require_cuda(slc_raw, slt)
    nxy = slc_raw.shape[1]
    slc = torch.empty((nxy, 2), dtype=torch.float64, device=slc_raw.device)
    check(load().slm_fields_finish_scaled(_ptr(slc_raw), nxy, int(density), float(np.float32(ds_scale_down)),
                                          float(np.float32(ds_scale_up)), _ptr(slt), _ptr(slc), _stream(stream),
                                          SLM_FLAG_SYNC if sync else 0))
    return slc


def as_u32(t) -> np.ndarray:
    """Device int32 tensor holding uint32 bits -> host uint32 array."""
    return t.detach().cpu().numpy().view(np.uint32)
