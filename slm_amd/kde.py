"""Univariate kernel-density estimation on the MI355X.

Drop-in for the 'opencl' method of python/streamlines/kde.py and
analysis.Univariate_distribution (analysis.py:659-852): the histogram and the
FP64-accumulated kernel filter run through libslm_hip.so (slm_kde_histogram,
slm_kde_pdf, replacing kernels histogram_univariate and pdf_univariate); the
host keeps the reference's parameter derivation (kde.py:137-157), its 'kde'
macro rendering (pocl.py:195-221), normalisation and mode/threshold search.
numpy-1.x scalar promotion is reproduced where it decides a value: a numpy
float32 scalar combined with a Python scalar gives float64.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import hip
from .params import ABI_VERSION, SlmKdeParams, c_float_literal, c_uint_literal

__all__ = ["SlmKdeParams", "Univariate_distribution", "estimate_univariate_pdf", "kde_params", "prepare_filter",
           "render_kde_macros", "KERNELS"]

KERNELS = {"tophat": 0, "triangle": 1, "epanechnikov": 2, "cosine": 3, "gaussian": 4}


_KDE_MACROS = [("bandwidth", "f"), ("n_data", "u"), ("n_hist_bins", "u"), ("n_pdf_points", "u"),
               ("logx_min", "f"), ("logx_max", "f"), ("logx_range", "f"), ("bin_dx", "f"),
               ("pdf_dx", "f"), ("kdf_width_x", "f"), ("n_kdf_part_points_x", "u"),
               ("logy_min", "f"), ("logy_max", "f"), ("logy_range", "f"), ("bin_dy", "f"),
               ("pdf_dy", "f"), ("kdf_width_y", "f"), ("n_kdf_part_points_y", "u")]


def _s(v) -> str:
    if isinstance(v, np.float32):
        return str(v)
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    return str(int(v))


def render_kde_macros(distbn, kernel_fn: str) -> list:
    """pocl.set_compile_options(..., job_type='kde') (pocl.py:195-221)."""
    out = ["-D", "KERNEL_{}".format(kernel_fn.upper()), "-D", "KDF_IS_{}".format(distbn.kernel.upper())]
    for name, suffix in _KDE_MACROS:
        out += ["-D", "{0}={1}{2}".format(name.upper(), _s(getattr(distbn, name)), suffix)]
    return out


def kde_params(distbn, fp_contract=False) -> SlmKdeParams:
    """SlmKdeParams parsed from the rendered macro literals."""
    m = {}
    for opt in render_kde_macros(distbn, "pdf_univariate")[1::2]:
        if "=" in opt:
            k, v = opt.split("=", 1)
            m[k] = c_uint_literal(v[:-1]) if v.endswith("u") else c_float_literal(v[:-1])
    k = SlmKdeParams()
    k.abi_version = ABI_VERSION
    k.kernel = KERNELS[distbn.kernel]
    k.n_data = m["N_DATA"]
    k.n_hist_bins = m["N_HIST_BINS"]
    k.n_pdf_points = m["N_PDF_POINTS"]
    k.n_kdf_part_points_x = m["N_KDF_PART_POINTS_X"]
    for f in ("logx_min", "logx_range", "bin_dx", "pdf_dx", "kdf_width_x"):
        setattr(k, f, float(m[f.upper()]))
    k.fp_contract = 1 if fp_contract else 0
    return k


def _sum_f64(a) -> np.float64:
    """Python's sum() over float32 elements under numpy 1.x: float64 running sum."""
    s = np.float64(0.0)
    for v in np.asarray(a, dtype=np.float32).ravel():
        s = np.float64(s + np.float64(v))
    return s


def prepare_filter(distbn):
    """Filter width and half-width in pdf points (kde.py:148-157): the
    'hacked Silverman' width 1.06*std*n^-0.2*bandwidth."""
    stddev_x = np.float32(np.std(distbn.logx_data))
    distbn.kdf_width_x = 1.06 * float(stddev_x) * float(np.power(float(distbn.n_data), -0.2)) * \
        float(distbn.bandwidth)
    distbn.kdf_width_y = 0.0
    distbn.n_kdf_part_points_x = int(np.uint32(np.floor(distbn.kdf_width_x / distbn.pdf_dx))) // 2
    distbn.n_kdf_part_points_y = 0


def estimate_univariate_pdf(distbn, fp_contract=False):
    """kde.py:116-194 ('opencl' path, no detrending): histogram + filter on the
    GPU.  Returns (pdf, histogram) normalised as the reference does."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    sl = np.ascontiguousarray(distbn.logx_data, dtype=np.float32).reshape(-1)
    prepare_filter(distbn)
    k = kde_params(distbn, fp_contract)
    lib = hip.load()
    x = torch.from_numpy(sl).to(dev)
    hist = torch.empty(int(distbn.n_hist_bins), dtype=torch.int32, device=dev)
    pdf = torch.empty(int(distbn.n_pdf_points), dtype=torch.float32, device=dev)
    s = hip._stream(None)
    hip.check(lib.slm_kde_histogram(ctypes.byref(k), hip._ptr(x), hip._ptr(hist), s, 0))
    hip.check(lib.slm_kde_pdf(ctypes.byref(k), hip._ptr(hist), hip._ptr(pdf), s, hip.SLM_FLAG_SYNC))
    histogram_array = hip.as_u32(hist).reshape(-1, 1).copy()
    pdf_array = pdf.cpu().numpy().reshape(-1, 1).copy()
    distbn.raw_pdf = pdf_array.copy()
    bin_dx = np.float64(distbn.bin_dx)
    pdf_n = (pdf_array / np.float32(_sum_f64(pdf_array) * bin_dx)).astype(np.float32)
    hist_n = histogram_array / (float(histogram_array.sum(dtype=np.uint64)) * bin_dx)
    return pdf_n, hist_n


class Univariate_distribution:
    """analysis.py:659-852, 'opencl' method."""

    def __init__(self, logx_array=None, logy_array=None, pixel_size=None, method="opencl",
                 n_hist_bins=2048, n_pdf_points=256, search_cdf_min=0.95, search_cdf_max=0.99,
                 logx_min=None, logy_min=None, logx_max=None, logy_max=None, cl_src_path=None,
                 cl_platform=None, cl_device=None, debug=False, verbose=False, gpu_verbose=False):
        fmin = np.finfo(np.float32).min
        if logy_array is None:
            logy_array = logx_array
        if logx_min is None:
            logx_min = logx_array[logx_array > fmin].min()
        if logx_max is None:
            logx_max = logx_array[logx_array > fmin].max()
        if logy_min is None:
            logy_min = logy_array[logy_array > fmin].min()
        if logy_max is None:
            logy_max = logy_array[logy_array > fmin].max()
        sel = (logx_array >= logx_min) & (logx_array <= logx_max) & (logy_array >= logy_min) & \
            (logy_array <= logy_max)
        self.logx_data = logx_array[sel].astype(np.float32).reshape(-1, 1)
        self.logx_min = np.min(self.logx_data)
        self.logx_max = np.max(self.logx_data)
        self.logx_range = self.logx_max - self.logx_min
        self.logy_min = self.logy_max = self.logy_range = 0.0
        self.n_data = self.logx_data.shape[0]
        self.n_hist_bins = n_hist_bins
        self.n_pdf_points = n_pdf_points
        self.bin_dx = float(self.logx_range) / n_hist_bins
        self.pdf_dx = float(self.logx_range) / n_pdf_points
        self.bin_dy = self.pdf_dy = 0.0
        self.logx_vec = np.linspace(float(self.logx_min), float(self.logx_max), n_pdf_points).reshape(-1, 1)
        self.logx_vec_histogram = np.linspace(float(self.logx_min), float(self.logx_max),
                                              n_hist_bins).reshape(-1, 1)
        self.x_vec = np.exp(self.logx_vec)
        self.dlogx = self.logx_vec[1] - self.logx_vec[0]
        self.search_cdf_min = search_cdf_min
        self.search_cdf_max = search_cdf_max
        self.pixel_size = pixel_size
        self.verbose = verbose

    def print(self, *args, **kwargs):
        if self.verbose:
            print(*args, **kwargs)

    def compute_kde_opencl(self, kernel="epanechnikov", bandwidth=0.15, fp_contract=False):
        self.kernel = kernel
        self.bandwidth = bandwidth
        if self.kernel == "gaussian":
            self.bandwidth /= 2.0
        if self.kernel not in KERNELS:
            raise ValueError('PDF kernel "{}" is not among those available: {}'.format(self.kernel, list(KERNELS)))
        self.pdf, self.histogram = estimate_univariate_pdf(self, fp_contract)
        self.cdf = (np.cumsum(self.pdf, dtype=np.float32) * np.float32(self.bin_dx)).astype(np.float32)
        cdf_max = self.cdf[-1]
        self.pdf = (self.pdf / cdf_max).astype(np.float32)
        self.cdf = (self.cdf / cdf_max).astype(np.float32)

    def statistics(self):
        logx, pdf = self.logx_vec, self.pdf
        mean = np.sum(logx * pdf) / np.sum(pdf)
        variance = np.sum((logx - mean) ** 2 * pdf) / np.sum(pdf)
        self.mean = np.exp(mean)
        self.stddev = np.exp(np.sqrt(variance))
        self.var = np.exp(variance)
        self.raw_mean = np.exp(self.logx_data.mean())
        self.raw_stddev = np.exp(self.logx_data.std())
        self.raw_var = np.exp(self.logx_data.var())

    def find_modes(self):
        from scipy.signal import argrelextrema
        pdf = self.pdf
        approx_mode = np.round(self.x_vec[pdf == pdf.max()][0], 2)
        peaks = argrelextrema(np.reshape(pdf, pdf.shape[0]), np.greater, order=3)[0]
        peaks = [pk for pk in list(peaks) if self.x_vec[pk] >= approx_mode * 0.5]
        if peaks:
            self.mode_i = peaks[0]
            self.mode_x = self.x_vec[peaks[0]][0]
        else:
            self.mode_i = self.pdf.shape[0] // 2
            self.mode_x = 100.0

    def locate_threshold(self):
        from scipy.signal import argrelextrema
        from scipy.stats import norm
        x_vec, pdf, cdf = self.x_vec, self.pdf, self.cdf
        detrended = pdf / norm.pdf(np.log(x_vec), np.log(self.mean), np.log(self.stddev))
        n_bins = pdf.shape[0]
        ext = argrelextrema(np.reshape(detrended, n_bins), np.less, order=3)[0]
        ext = np.concatenate((ext, np.array([n_bins - 1])))
        search_cdf_min = self.search_cdf_min
        extrema_i = []
        while search_cdf_min >= 0.70:
            extrema_i = [e for e in ext if x_vec[e] >= self.mode_x
                         and (cdf[e] >= search_cdf_min and cdf[e] <= self.search_cdf_max)]
            if len(extrema_i) >= 1:
                break
            search_cdf_min -= 0.01
        if extrema_i:
            self.channel_threshold_i = extrema_i[0]
            self.channel_threshold_x = x_vec[extrema_i[0]][0]
        else:
            self.channel_threshold_i = None
            self.channel_threshold_x = None
