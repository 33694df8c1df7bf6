"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

Functions (each restates the reference kernel named in oracle/slm_oracle.c):
  lehmer, rng_seed, rng_stream         rng.cl:37-67, integratefields.cl:93-99
  trace_fields                         integratefields.cl:60-114 + jittertrajectory.cl
  trace_fields_record                  one jittered streamline, positions recorded
  trace_paths                          integratetraj.cl:54-97 + trajectory.cl
  hillslope_lengths                    lengths.cl:37-116
  kde_histogram, kde_pdf               kde.cl:383-394, :426-526
  connect_channels ... fix_left_flanks channel mapping and segmentation
                                       (oracle/slm_oracle_map.c), see map_call
The tracing functions have an FP64 twin (f64=True) built from the same source.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32


# name -> number of pointer arguments (SlmParams* first), in C order
_MAP_FUNCS = {
    "connect_channels": 5,          # p, seeds, mask, uv, mapping
    "thin_channels": 2,             # p, mapping
    "map_channel_heads": 5,         # p, seeds, mask, uv, mapping
    "count_downchannels": 7,        # p, seeds, mask, uv, mapping, count, link
    "flag_downchannels": 6,         # p, seeds, mask, mapping, count, link
    "label_confluences": 7,         # p, seeds, mask, slt, mapping, count, link
    "segment_downchannels": 7,      # p, seeds, mask, mapping, count, link, label
    "link_hillslopes": 5,           # p, seeds, mask, uv, link
    "segment_hillslopes": 6,        # p, seeds, mask, uv, mapping, label
    "subsegment_channel_edges": 6,  # p, seeds, mask, mapping, link, label
    "subsegment_flanks": 6,         # p, seeds, mask, uv, mapping, label
    "fix_right_flanks": 6,          # p, seeds, mask, uv, mapping, label
    "fix_left_flanks": 6,           # p, seeds, mask, uv, mapping, label
}

_MAP_DTYPES = {"seeds": np.float32, "mask": np.uint8, "uv": np.float32, "slt": np.float32,
               "mapping": np.uint32, "count": np.uint32, "link": np.uint32, "label": np.uint32}
_MAP_ARGS = {
    "connect_channels": ("seeds", "mask", "uv", "mapping"),
    "thin_channels": ("mapping",),
    "map_channel_heads": ("seeds", "mask", "uv", "mapping"),
    "count_downchannels": ("seeds", "mask", "uv", "mapping", "count", "link"),
    "flag_downchannels": ("seeds", "mask", "mapping", "count", "link"),
    "label_confluences": ("seeds", "mask", "slt", "mapping", "count", "link"),
    "segment_downchannels": ("seeds", "mask", "mapping", "count", "link", "label"),
    "link_hillslopes": ("seeds", "mask", "uv", "link"),
    "segment_hillslopes": ("seeds", "mask", "uv", "mapping", "label"),
    "subsegment_channel_edges": ("seeds", "mask", "mapping", "link", "label"),
    "subsegment_flanks": ("seeds", "mask", "uv", "mapping", "label"),
    "fix_right_flanks": ("seeds", "mask", "uv", "mapping", "label"),
    "fix_left_flanks": ("seeds", "mask", "uv", "mapping", "label"),
}


def map_call(name, p, **arrays):
    """Run oracle mapping function ``name`` (slm_oracle_map.c) in place.

    Arrays are passed by keyword (names as in _MAP_ARGS); the uint32 grids
    (mapping/count/link/label) must be C-contiguous numpy arrays of the right
    dtype, since they are updated in place.  Returns the C return value
    (thin_channels: iteration count) or None."""
    keep = []
    ptrs = []
    for a in _MAP_ARGS[name]:
        arr = arrays[a]
        want = _MAP_DTYPES[a]
        if a in ("mapping", "count", "link", "label"):
            if arr.dtype != want or not arr.flags.c_contiguous:
                raise TypeError("%s must be a C-contiguous %s array" % (a, np.dtype(want)))
        else:
            arr = np.ascontiguousarray(arr, dtype=want)
        keep.append(arr)
        ptrs.append(ctypes.c_void_p(arr.ctypes.data))
    return getattr(load(), "slmo_" + name)(ctypes.byref(p), *ptrs)


def map_set_fast_forward(on: bool):
    """Cycle fast-forward of the mapping oracle (exact); off runs every step."""
    load().slmo_map_set_fast_forward(1 if on else 0)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("oracle/liboracle.so is not built (make -C oracle)")
        lib = ctypes.CDLL(LIB_PATH)
        for pre in ("slmo_", "slmo64_"):
            getattr(lib, pre + "lehmer").restype = _U32
            getattr(lib, pre + "lehmer").argtypes = [ctypes.POINTER(_U32)]
            getattr(lib, pre + "rng_seed").restype = _U32
            getattr(lib, pre + "rng_seed").argtypes = [_U32]
            getattr(lib, pre + "rng_stream").restype = None
            getattr(lib, pre + "rng_stream").argtypes = [_U32, _U32, _U32, _P, _P]
            getattr(lib, pre + "trace_fields").restype = ctypes.c_uint64
            getattr(lib, pre + "trace_fields").argtypes = [_P, _P, _P, _P, _P, _P, ctypes.c_int]
            getattr(lib, pre + "trace_fields_record").restype = _U32
            getattr(lib, pre + "trace_fields_record").argtypes = [_P, _P, _P, _P, _U32, _U32, _P, _U32, _P]
            getattr(lib, pre + "hillslope_lengths").restype = None
            getattr(lib, pre + "hillslope_lengths").argtypes = [_P, _P, _P, _P, _P, _U32, _U32, _P]
            getattr(lib, pre + "trace_paths").restype = None
            getattr(lib, pre + "trace_paths").argtypes = [_P, _P, _P, _P, _P, _P, _P, ctypes.c_int]
        for name, nargs in _MAP_FUNCS.items():
            f = getattr(lib, "slmo_" + name)
            f.restype = _U32 if name == "thin_channels" else None
            f.argtypes = [_P] * nargs
        lib.slmo_map_set_fast_forward.restype = None
        lib.slmo_map_set_fast_forward.argtypes = [ctypes.c_int]
        lib.slmo_kde_histogram.restype = None
        lib.slmo_kde_histogram.argtypes = [_P, _P, _P]
        lib.slmo_kde_pdf.restype = None
        lib.slmo_kde_pdf.argtypes = [_P, _P, _P]
        _lib = lib
    return _lib


def _pre(f64):
    return "slmo64_" if f64 else "slmo_"


def _c(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, ctypes.c_void_p(a.ctypes.data)


def lehmer(state: int, f64=False):
    s = _U32(state)
    v = getattr(load(), _pre(f64) + "lehmer")(ctypes.byref(s))
    return int(v), int(s.value)


def rng_seed(idx: int, f64=False) -> int:
    return int(getattr(load(), _pre(f64) + "rng_seed")(idx))


def rng_stream(idx0: int, n: int, n_draws: int):
    rng0 = np.zeros(n, dtype=np.uint32)
    draws = np.zeros((n, max(n_draws, 1)), dtype=np.uint32)
    load().slmo_rng_stream(idx0, n, n_draws, rng0.ctypes.data, draws.ctypes.data if n_draws else None)
    return rng0, draws[:, :n_draws]


def trace_fields(p, seeds, mask, uv, nthreads=8, f64=False, slc=None, slt=None):
    """Returns (slc uint32 [nxy], slt uint32 [nxy], accepted steps)."""
    nxy = p.nxy_padded
    slc = np.zeros(nxy, dtype=np.uint32) if slc is None else slc
    slt = np.zeros(nxy, dtype=np.uint32) if slt is None else slt
    seeds, ps = _c(seeds, np.float32)
    mask, pm = _c(mask, np.uint8)
    uv, pu = _c(uv, np.float32)
    steps = getattr(load(), _pre(f64) + "trace_fields")(ctypes.byref(p), ps, pm, pu, slc.ctypes.data,
                                                       slt.ctypes.data, int(nthreads))
    return slc, slt, int(steps)


def trace_fields_record(p, seeds, mask, uv, seed_k, sub, max_rec=4096, f64=False):
    """Positions (float32 [n,2]) of one jittered fields streamline and its step count."""
    xy = np.zeros((max_rec, 2), dtype=np.float32)
    nrec = _U32(0)
    seeds, ps = _c(seeds, np.float32)
    mask, pm = _c(mask, np.uint8)
    uv, pu = _c(uv, np.float32)
    n = getattr(load(), _pre(f64) + "trace_fields_record")(ctypes.byref(p), ps, pm, pu, seed_k, sub,
                                                          xy.ctypes.data, max_rec, ctypes.byref(nrec))
    return xy[:min(nrec.value, max_rec)].copy(), int(n)


def trace_paths(p, seeds, mask, uv, nthreads=8, f64=False):
    """Returns (traj int8 [n,max_n_steps,2], nsteps uint16 [n], length float32 [n])."""
    n = p.n_seed_points - p.seeds_chunk_offset
    traj = np.zeros((n, p.max_n_steps, 2), dtype=np.int8)
    nsteps = np.zeros(n, dtype=np.uint16)
    length = np.zeros(n, dtype=np.float32)
    seeds, ps = _c(seeds, np.float32)
    mask, pm = _c(mask, np.uint8)
    uv, pu = _c(uv, np.float32)
    getattr(load(), _pre(f64) + "trace_paths")(ctypes.byref(p), ps, pm, pu, traj.ctypes.data,
                                               nsteps.ctypes.data, length.ctypes.data, int(nthreads))
    return traj, nsteps, length


def kde_histogram(k, logx):
    """histogram_univariate (kde.cl:383-394): uint32 [n_hist_bins]."""
    logx, px = _c(logx, np.float32)
    hist = np.zeros(k.n_hist_bins, dtype=np.uint32)
    load().slmo_kde_histogram(ctypes.byref(k), px, hist.ctypes.data)
    return hist


def kde_pdf(k, hist):
    """pdf_univariate with filter (kde.cl:426-526): float32 [n_pdf_points]."""
    hist, ph = _c(hist, np.uint32)
    pdf = np.zeros(k.n_pdf_points, dtype=np.float32)
    load().slmo_kde_pdf(ctypes.byref(k), ph, pdf.ctypes.data)
    return pdf


def hillslope_lengths(p, seeds, mask, uv, mapping, thin_flag, from_flag, length=None, f64=False):
    """hillslope_lengths (lengths.cl:37-116): float32 [n_seed_points], untouched
    entries keep their initial value (zeros by default)."""
    n = p.n_seed_points
    length = np.zeros(n, dtype=np.float32) if length is None else np.ascontiguousarray(length, np.float32)
    seeds, ps = _c(seeds, np.float32)
    mask, pm = _c(mask, np.uint8)
    uv, pu = _c(uv, np.float32)
    mapping, pmap = _c(mapping, np.uint32)
    getattr(load(), _pre(f64) + "hillslope_lengths")(ctypes.byref(p), ps, pm, pu, pmap, thin_flag, from_flag,
                                                     length.ctypes.data)
    return length
