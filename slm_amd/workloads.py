"""Named tracing workloads built from seeded synthetic DTMs.

A Job holds what Trace.compute_fields hands the tracer (trace.py:131-157):
the bbox-cropped padded mask and uv, the Info object, and the shuffled seed
list, so tests and benchmarks can call the C ABI directly.

  small_256   256x256 at 1 m, relief 20 m, 5% tilt, 2-px NaN border, d=1
  tile_2048   2048x2048 at 2 m, relief 60 m, three V-valleys, d=2
"""
from __future__ import annotations

from dataclasses import dataclass
from types import SimpleNamespace

import numpy as np

from . import preprocess, synth
from .parameters import load_parameters, section
from .params import SlmParams, kernel_params
from .useful import Data, Info, create_seeds, get_bbox


@dataclass
class Job:
    name: str
    roi: np.ndarray            # float32 [nx, ny] DTM (NaN = no data)
    pixel_size: float
    mask: np.ndarray           # bool [nxp, nyp], bbox-cropped, padded
    uv: np.ndarray             # float32 [nxp, nyp, 2]
    info: Info
    seeds: np.ndarray          # float32 [N, 2], shuffled
    trace: SimpleNamespace
    state: SimpleNamespace
    origin: tuple = (0, 0)     # full padded index of cropped index (0, 0)

    @property
    def shape(self):
        return self.mask.shape

    def params(self, sign=1, kernel="INTEGRATE_FIELDS", fp_contract=False, n_seed_points=None,
               jitter=None) -> SlmParams:
        info = self.info
        info.n_seed_points = np.uint32(self.seeds.shape[0] if n_seed_points is None else n_seed_points)
        p = kernel_params(info, sign, kernel, fp_contract)
        if jitter is not None:
            p.jitter_magnitude = float(np.float32(jitter))
        return p

    def n_streamlines(self) -> int:
        return int(self.seeds.shape[0]) * int(self.info.subpixel_seed_point_density) ** 2


def job_from_dtm(name, roi, pixel_size, trace_overrides=None, state_overrides=None, pad_width=2):
    params = load_parameters(overrides={"trace": trace_overrides or {}, "state": state_overrides or {}})
    trace = section(params, "trace")
    state = section(params, "state")
    mask_full, uv_full = preprocess.prepare(roi, pixel_size, pad_width)
    bbox, bnx, bny = get_bbox(~mask_full)
    info = Info(state, trace, pixel_size)
    info.set_xy(int(bnx), int(bny), pad_width)
    data = Data(info=info, bbox=bbox, pad=pad_width, mask_array=mask_full, uv_array=uv_full,
                mapping_array=np.zeros(mask_full.shape, dtype=np.uint32))
    seeds, n, npad = create_seeds(data.mask_array, pad_width, int(info.n_work_items),
                                  do_shuffle=info.do_shuffle, rng_seed=info.shuffle_rng_seed)
    info.n_seed_points = np.uint32(n)
    info.n_padded_seed_points = np.uint32(npad)
    origin = (int(bbox[0]) - pad_width, int(bbox[2]) - pad_width)
    return Job(name, roi, pixel_size, data.mask_array, data.uv_array, info, seeds, trace, state, origin)

    # note: a DTM pixel (i, j) sits at cropped padded index (i + pad - origin[0], j + pad - origin[1])


def small_256(jitter=1.0, seed=0, n=256, d=1, max_error=None, resolution=None, max_length=1000.0):
    """max_error overrides max_integration_step_error (small values make the
    adaptive RK2 step actually adapt); resolution overrides
    trajectory_resolution (128 by default)."""
    roi = synth.make_dtm(n, pixel_size=1.0, relief=20.0, tilt=0.05, seed=seed, nan_border=2)
    tr = {"subpixel_seed_point_density": d, "jitter_magnitude": float(jitter),
          "max_length": float(max_length), "integrator_step_factor": 0.5}
    if max_error is not None:
        tr["max_integration_step_error"] = float(max_error)
    if resolution is not None:
        tr["trajectory_resolution"] = int(resolution)
    return job_from_dtm("small_256" if n == 256 else "small_%d" % n, roi, 1.0, tr)


def tile_2048(jitter=1.0, seed=0, n=2048, d=2):
    roi = synth.make_dtm(n, pixel_size=2.0, relief=60.0, tilt=0.05, seed=seed, nan_border=2, n_valleys=3)
    return job_from_dtm("tile_2048" if n == 2048 else "tile_%d" % n, roi, 2.0,
                        {"subpixel_seed_point_density": d, "jitter_magnitude": float(jitter),
                         "max_length": 1000.0, "integrator_step_factor": 0.5})


def planar(n=64, pixel_size=1.0, slope=(0.3, 0.1), d=1, jitter=0.0, max_length=1000.0):
    """Planar ramp h = -(sx*x + sy*y)*pixel_size with a NaN border: uv is uniform."""
    x = np.arange(n, dtype=np.float64)[:, None]
    y = np.arange(n, dtype=np.float64)[None, :]
    roi = (-(slope[0] * x + slope[1] * y) * pixel_size).astype(np.float32)
    roi[:2, :] = np.nan
    roi[-2:, :] = np.nan
    roi[:, :2] = np.nan
    roi[:, -2:] = np.nan
    return job_from_dtm("planar_%d" % n, roi, pixel_size,
                        {"subpixel_seed_point_density": d, "jitter_magnitude": float(jitter),
                         "max_length": max_length, "integrator_step_factor": 0.5})


def v_valley(n=96, pixel_size=1.0, side_slope=0.2, down_slope=0.05, midslope_offset=12):
    """Planar V-valley along y at x = n//2: h = s*|x-c| - t*y, NaN border.
    Returns (job, mapping): mapping (uint32, cropped padded grid) flags the
    valley axis thin-channel (2) and the lines x = c +- midslope_offset
    mid-slope (2048), the flag values of useful.Info."""
    c = n // 2
    x = np.arange(n, dtype=np.float64)[:, None]
    y = np.arange(n, dtype=np.float64)[None, :]
    roi = ((side_slope * np.abs(x - c) - down_slope * y) * pixel_size).astype(np.float32)
    roi[:2, :] = np.nan
    roi[-2:, :] = np.nan
    roi[:, :2] = np.nan
    roi[:, -2:] = np.nan
    job = job_from_dtm("v_valley_%d" % n, roi, pixel_size, {"subpixel_seed_point_density": 1,
                                                            "jitter_magnitude": 0.0})
    mapping = np.zeros(job.mask.shape, dtype=np.uint32)
    ox, oy = job.origin
    col = lambda i: i + 2 - ox
    mapping[col(c), :] |= np.uint32(job.info.is_thinchannel)
    for xm in (c - midslope_offset, c + midslope_offset):
        mapping[col(xm), :] |= np.uint32(job.info.is_midslope)
    mapping[job.mask] = 0
    return job, mapping


def y_valley(n=64, pixel_size=1.0, side_slope=0.2, down_slope=0.05, junction=0.6, skew=0.6,
             channel_width=1):
    """A main V-valley along y at x = n//2 joined by a tributary V-valley whose
    axis x = c + skew*(yj - y) meets it at y = yj = junction*n; h is the lower
    of the two, so the tributary drains into the main valley.  NaN border.
    Returns (job, mapping): mapping (uint32, cropped padded grid) flags
    IS_CHANNEL on the pixels within channel_width//2 of either axis (the
    tributary only upstream of the junction), as Mapping.map_channels would
    on a well-resolved slt field."""
    c = n // 2
    yj = junction * n
    x = np.arange(n, dtype=np.float64)[:, None]
    y = np.arange(n, dtype=np.float64)[None, :]
    h_main = side_slope * np.abs(x - c) - down_slope * y
    xt = c + skew * (yj - y)
    h_trib = np.where(y <= yj, side_slope * np.abs(x - xt) - down_slope * y, np.inf)
    roi = (np.minimum(h_main, h_trib) * pixel_size).astype(np.float32)
    roi[:2, :] = np.nan
    roi[-2:, :] = np.nan
    roi[:, :2] = np.nan
    roi[:, -2:] = np.nan
    job = job_from_dtm("y_valley_%d" % n, roi, pixel_size, {"subpixel_seed_point_density": 1,
                                                            "jitter_magnitude": 0.0})
    mapping = np.zeros(job.mask.shape, dtype=np.uint32)
    ox, oy = job.origin
    half = channel_width // 2
    on_main = np.abs(x - c) <= half
    on_trib = (np.abs(x - np.round(xt)) <= half) & (y <= yj)
    chan = np.zeros((n, n), dtype=bool)
    chan |= np.broadcast_to(on_main, (n, n))
    chan |= on_trib
    full = np.zeros((n + 4, n + 4), dtype=bool)
    full[2:-2, 2:-2] = chan
    nxp, nyp = job.mask.shape
    mapping[full[ox:ox + nxp, oy:oy + nyp]] = np.uint32(job.info.is_channel)
    mapping[job.mask] = 0
    return job, mapping


def batch_c3(n_tiles=64, seed=1, n_cones=16, d=1, jitter=1.0, max_length=2500.0, side_range=(256, 1536),
             relief=40.0):
    """Batched uneven tiles: n_tiles DTMs whose sides are drawn log-uniformly
    from side_range (x and y independently) with RNG seed ``seed``, pixel size
    1 m or 2 m; n_cones of them are alluvial-fan-like cones with pits left
    unfilled, so trajectories range from a few steps to the step cap.
    max_length 2500 m gives MAX_N_STEPS = 5000 at the 0.5 step factor.
    Returns a list of Jobs."""
    rs = np.random.default_rng(seed)
    lo, hi = np.log(side_range[0]), np.log(side_range[1])
    sides = np.exp(rs.uniform(lo, hi, size=(n_tiles, 2))).astype(int)
    pixel = rs.choice([1.0, 2.0], size=n_tiles)
    cones = set(int(i) for i in rs.choice(n_tiles, size=min(n_cones, n_tiles), replace=False))
    jobs = []
    for i in range(n_tiles):
        nx, ny = int(sides[i, 0]), int(sides[i, 1])
        ps = float(pixel[i])
        cone = i in cones
        roi = synth.make_dtm((nx, ny), pixel_size=ps, relief=relief, tilt=0.05, seed=seed * 1000 + i,
                             nan_border=2, n_valleys=0 if cone else 1, cone=cone,
                             n_pits=max(4, nx * ny // 20000) if cone else 0, fill=not cone)
        jobs.append(job_from_dtm("c3_%02d_%dx%d_%gm%s" % (i, nx, ny, ps, "_cone" if cone else ""), roi, ps,
                                 {"subpixel_seed_point_density": d, "jitter_magnitude": float(jitter),
                                  "max_length": float(max_length), "integrator_step_factor": 0.5}))
    return jobs
