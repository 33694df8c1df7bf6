"""Seeded synthetic DTMs for the tracing workloads.

Surfaces: spectral fractal noise (amplitude spectrum ~ k^-beta, so power
~ k^-2beta; beta=2 gives terrain that is smooth at the pixel scale), scaled to a relief,
plus a planar tilt, optional V-valley incisions and pitted cones, sinks filled
by priority-flood with an epsilon gradient, and a NaN no-data border.  Arrays
are float32 [nx, ny] in the reference's ROI orientation (x = first axis).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_HOSTLIB = None


def _hostlib():
    global _HOSTLIB
    if _HOSTLIB is None:
        path = os.path.join(_HERE, "libslm_host.so")
        if not os.path.exists(path):
            raise ImportError("slm_amd/libslm_host.so is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        lib.slm_fill_sinks.restype = ctypes.c_int
        lib.slm_fill_sinks.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double]
        lib.slm_v_valley_segments.restype = ctypes.c_int
        lib.slm_v_valley_segments.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int] + \
            [ctypes.c_void_p] * 4 + [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _HOSTLIB = lib
    return _HOSTLIB


def fill_sinks(h: np.ndarray, eps: float = 1e-3) -> np.ndarray:
    """Priority-flood + epsilon on a float array (NaN = no data); float64 work."""
    a = np.ascontiguousarray(h, dtype=np.float64).copy()
    rc = _hostlib().slm_fill_sinks(a.ctypes.data, a.shape[0], a.shape[1], float(eps))
    if rc != 0:
        raise MemoryError("slm_fill_sinks failed")
    return a


def spectral_noise(nx: int, ny: int, beta: float, rng: np.random.Generator, stationary: bool = False) -> np.ndarray:
    """Gaussian random field whose amplitude spectrum falls as |k|^-beta
    (power spectrum |k|^-2beta), rescaled to the range [0, 1].  stationary:
    not rescaled; the amplitude at each wavenumber (cycles per pixel) is
    then the same for every grid size, so windows of a large grid have the
    statistics of a small grid's (the large grid adds longer wavelengths)."""
    white = rng.standard_normal((nx, ny))
    f = np.fft.rfft2(white)
    kx = np.fft.fftfreq(nx)[:, None]
    ky = np.fft.rfftfreq(ny)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    f *= k ** (-beta)
    f[0, 0] = 0.0
    z = np.fft.irfft2(f, s=(nx, ny))
    if stationary:
        return z
    z -= z.min()
    z /= max(z.max(), 1e-30)
    return z


def v_valleys(nx: int, ny: int, pixel_size: float, n: int, depth: float, half_width_px: float,
              rng: np.random.Generator, half_length_px: float = 0.0, centre_range=(0.2, 0.8)) -> np.ndarray:
    """Sum of n straight V-shaped incisions (depth in m) across the grid:
    -depth * clip(1 - dist/half_width_px, 0, 1) per valley, with the line's
    angle and centre drawn from rng (angle, then centre x, then centre y,
    centres uniform over centre_range of each side).  half_length_px > 0
    makes each valley a segment of that half length whose depth fades out
    over half_width_px at its ends."""
    ang, cx, cy = np.zeros(n), np.zeros(n), np.zeros(n)
    lo, hi = centre_range
    for k in range(n):
        ang[k] = rng.uniform(0.0, np.pi)
        cx[k], cy[k] = rng.uniform(lo, hi) * nx, rng.uniform(lo, hi) * ny
    s, c = np.sin(ang), np.cos(ang)
    out = np.zeros((nx, ny))
    rc = _hostlib().slm_v_valley_segments(out.ctypes.data, nx, ny, n, cx.ctypes.data, cy.ctypes.data,
                                          s.ctypes.data, c.ctypes.data, float(depth), float(half_width_px),
                                          float(half_length_px))
    if rc != 0:
        raise ValueError("slm_v_valleys failed")
    return out


def cones_with_pits(nx: int, ny: int, pixel_size: float, relief: float, n_pits: int,
                    rng: np.random.Generator) -> np.ndarray:
    """Alluvial-fan-like cone (apex near one edge) with shallow pits."""
    x = np.arange(nx)[:, None].astype(np.float64)
    y = np.arange(ny)[None, :].astype(np.float64)
    ax, ay = rng.uniform(0.0, 0.2) * nx, rng.uniform(0.3, 0.7) * ny
    r = np.hypot(x - ax, y - ay)
    out = relief * (1.0 - r / r.max())
    for _ in range(n_pits):
        px, py = rng.uniform(0, nx), rng.uniform(0, ny)
        rad = rng.uniform(3.0, 12.0)
        out -= 0.05 * relief * np.exp(-((x - px) ** 2 + (y - py) ** 2) / (2 * rad * rad))
    return out


def make_dtm(n: int | tuple, pixel_size: float = 1.0, relief: float = 20.0, tilt: float = 0.05,
             beta: float = 2.0, seed: int = 0, nan_border: int = 2, n_valleys: int = 0,
             valley_depth: float | None = None, cone: bool = False, n_pits: int = 0,
             fill: bool = True, valley_half_width_px: float | None = None,
             noise_scale: float | None = None, valley_half_length_px: float = 0.0,
             valley_centre_range=(0.2, 0.8)) -> np.ndarray:
    """Seeded synthetic DTM (float32 [nx, ny], metres, NaN border).  With
    noise_scale the noise is stationary (spectral_noise) times noise_scale
    metres instead of rescaled to relief."""
    nx, ny = (n, n) if np.isscalar(n) else n
    rng = np.random.default_rng(seed)
    if noise_scale is not None:
        h = noise_scale * spectral_noise(nx, ny, beta, rng, stationary=True)
    else:
        h = relief * spectral_noise(nx, ny, beta, rng)
    xs = np.arange(nx)[:, None] * pixel_size
    ys = np.arange(ny)[None, :] * pixel_size
    h = h + tilt * (0.8 * xs + 0.6 * ys)
    if n_valleys:
        depth = valley_depth if valley_depth is not None else 0.5 * relief
        hw = valley_half_width_px if valley_half_width_px is not None else max(8.0, 0.03 * max(nx, ny))
        h = h + v_valleys(nx, ny, pixel_size, n_valleys, depth, hw, rng, valley_half_length_px, valley_centre_range)
    if cone:
        h = 0.3 * h + cones_with_pits(nx, ny, pixel_size, relief, n_pits, rng)
    if nan_border:
        b = nan_border
        h[:b, :] = np.nan
        h[-b:, :] = np.nan
        h[:, :b] = np.nan
        h[:, -b:] = np.nan
    if fill:
        h = fill_sinks(h, eps=1e-3 * pixel_size)
    return h.astype(np.float32)
