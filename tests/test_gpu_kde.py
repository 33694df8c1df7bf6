"""GPU parity of the univariate KDE (slm_kde_histogram / slm_kde_pdf) against
the CPU oracle: histograms bit-identical; pdf bit-identical for the kernels
without transcendentals, within 2 float32 ulps for gaussian and cosine (the
device exp/cos and the host libm may differ in the last double ulp)."""
import ctypes

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def _case(kernel, n_hist, n_pdf, n, bandwidth, seed=0):
    from slm_amd.kde import Univariate_distribution, prepare_filter
    rs = np.random.default_rng(seed)
    x = np.concatenate([rs.lognormal(3.0, 0.6, n // 2), rs.lognormal(4.5, 0.3, n - n // 2)]).astype(np.float32)
    lx = np.log(x).astype(np.float32)
    d = Univariate_distribution(logx_array=lx, logy_array=lx, n_hist_bins=n_hist, n_pdf_points=n_pdf)
    d.kernel, d.bandwidth = kernel, bandwidth
    prepare_filter(d)
    return d


@pytest.mark.parametrize("kernel,contract,n_hist,n_pdf,n", [
    ("epanechnikov", True, 2048, 256, 200000),
    ("epanechnikov", False, 512, 128, 50000),
    ("triangle", True, 512, 512, 50000),
    ("tophat", True, 16384, 256, 100000),
    ("gaussian", True, 512, 128, 1000000),
    ("cosine", False, 512, 64, 50000),
])
def test_kde_bit_exact(gpu, kernel, contract, n_hist, n_pdf, n):
    import torch
    from slm_amd import hip
    from slm_amd.kde import kde_params
    d = _case(kernel, n_hist, n_pdf, n, 1.0)
    k = kde_params(d, contract)
    lib = hip.load()
    x = torch.from_numpy(np.ascontiguousarray(d.logx_data.reshape(-1))).to(gpu)
    hist = torch.empty(n_hist, dtype=torch.int32, device=gpu)
    pdf = torch.empty(n_pdf, dtype=torch.float32, device=gpu)
    s = hip._stream(None)
    hip.check(lib.slm_kde_histogram(ctypes.byref(k), hip._ptr(x), hip._ptr(hist), s, 0))
    hip.check(lib.slm_kde_pdf(ctypes.byref(k), hip._ptr(hist), hip._ptr(pdf), s, hip.SLM_FLAG_SYNC))
    oh = oracle.kde_histogram(k, d.logx_data.reshape(-1))
    assert np.array_equal(hip.as_u32(hist), oh)
    op = oracle.kde_pdf(k, oh)
    gp = pdf.cpu().numpy()
    if kernel in ("gaussian", "cosine"):
        ulp = np.abs(gp.view(np.int32).astype(np.int64) - op.view(np.int32).astype(np.int64))
        assert ulp.max() <= 2
    else:
        assert np.array_equal(gp, op)


def test_univariate_distribution_threshold(gpu):
    """Full 'opencl' path of Univariate_distribution: pdf normalised, cdf ends
    at 1, a mode and a threshold in the searched cdf band."""
    from slm_amd.kde import Univariate_distribution
    rs = np.random.default_rng(3)
    x = np.concatenate([rs.lognormal(3.0, 0.5, 80000), rs.lognormal(6.0, 0.4, 20000)]).astype(np.float32)
    lx = np.log(x).astype(np.float32)
    d = Univariate_distribution(logx_array=lx, logy_array=lx, n_hist_bins=2048, n_pdf_points=256,
                                search_cdf_min=0.7, search_cdf_max=0.99)
    d.compute_kde_opencl(kernel="epanechnikov", bandwidth=1.0)
    assert abs(float(d.cdf[-1]) - 1.0) < 1e-6
    d.find_modes()
    d.statistics()
    d.locate_threshold()
    assert 10.0 < d.mode_x < 60.0
    assert d.channel_threshold_x is not None and d.channel_threshold_x > d.mode_x
