"""Pins the CPU oracle (oracle/slm_oracle.c) without a GPU.

* RNG: against exact big-integer arithmetic (brute force over many seeds and
  draws, plus the edge states where the 32-bit +1 wraps and states >= modulus),
  and against the committed golden vectors.
* Kernels: against oracle/pyref.py, a separate pure-Python restatement of the
  OpenCL C (float32 scalars, exact-rational FMA), bit for bit on tiny grids,
  jitter on and off, both FP-contraction modes, both directions, d = 1, 2, 3.
* Invariants the mathematics fixes: sum(slc) = seed writes + accepted steps,
  lattice-exact positions for d in {1, 2}, FP64 twin close to FP32.
"""
import os

import numpy as np
import pytest

from oracle import oracle, pyref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M = 0xfffffffb


def _lehmer_np(s):
    """Vectorised exact Lehmer step in uint64 (products < 2^61)."""
    s = s.astype(np.uint64)
    return (((s + np.uint64(1)) & np.uint64(0xffffffff)) * np.uint64(279470273)) % np.uint64(M)


def test_lehmer_edge_states():
    for s in [0, 1, 2, M - 1, M, M + 1, 0xfffffffe, 0xffffffff, 0x7fffffff, 0x80000000, 123456789]:
        v, s2 = oracle.lehmer(s)
        exp = ((s + 1) % (1 << 32)) * 279470273 % M
        assert v == exp and s2 == exp, s
    # +1 wraps in 32 bits before widening: state 2^32-1 maps to 0
    assert oracle.lehmer(0xffffffff)[0] == 0


def test_rng_stream_brute_force():
    n, nd = 1 << 16, 32
    rng0, draws = oracle.rng_stream(0, n, nd)
    idx = np.arange(n, dtype=np.uint64)
    bs = lambda v: ((v >> np.uint64(24)) & np.uint64(0xff)) | ((v >> np.uint64(8)) & np.uint64(0xff00)) | \
        ((v << np.uint64(8)) & np.uint64(0xff0000)) | ((v << np.uint64(24)) & np.uint64(0xff000000))
    s = bs(_lehmer_np(bs(idx)))
    assert np.array_equal(rng0, s.astype(np.uint32))
    for k in range(nd):
        s = _lehmer_np(s)
        assert np.array_equal(draws[:, k], s.astype(np.uint32))


def test_rng_golden_vectors():
    g = np.load(os.path.join(GOLDEN, "rng_vectors.npz"))
    rng0, draws = oracle.rng_stream(int(g["idx0"]), int(g["rng0"].shape[0]), int(g["draws"].shape[1]))
    assert np.array_equal(rng0, g["rng0"])
    assert np.array_equal(draws, g["draws"])
    for i, si in enumerate(g["seed_idx"]):
        assert oracle.rng_seed(int(si)) == int(g["seed_rng0"][i])


def test_rng_variates_in_range():
    _, draws = oracle.rng_stream(0, 4096, 8)
    v = draws.astype(np.float32) / np.float32(float(0xfffffffb)) - np.float32(0.5)
    assert v.min() >= -0.5 and v.max() <= 0.5
    assert abs(float(v.mean())) < 0.01


def _tiny_job(n, d, jitter, seed=3, max_error=None):
    from slm_amd import workloads
    return workloads.small_256(jitter=jitter, n=n, d=d, seed=seed, max_error=max_error)


PYREF_CASES = [
    # (grid, d, jitter, contract, sign, n_seeds, max_integration_step_error)
    (24, 1, 1.0, True, 1, 40, None),
    (24, 1, 0.0, True, -1, 40, None),
    (24, 2, 1.0, False, 1, 20, None),
    (28, 3, 1.0, True, -1, 10, None),
    (28, 3, 0.0, False, 1, 10, None),
    # adaptive step actually adapting (dt leaves DT_MAX)
    (32, 1, 1.0, True, 1, 30, 1e-6),
    (32, 2, 0.0, False, -1, 15, 1e-5),
]


@pytest.mark.parametrize("grid,d,jitter,contract,sign,ns,max_error", PYREF_CASES)
def test_fields_match_pyref(grid, d, jitter, contract, sign, ns, max_error):
    job = _tiny_job(grid, d, jitter, max_error=max_error)
    ns = min(ns, job.seeds.shape[0])
    p = job.params(sign, fp_contract=contract, n_seed_points=ns)
    p.max_n_steps = min(p.max_n_steps, 300)
    m = job.mask.astype(np.uint8).reshape(-1)
    oslc, oslt, osteps = oracle.trace_fields(p, job.seeds, m, job.uv.reshape(-1), nthreads=1)
    rslc, rslt, rsteps = pyref.fields(p, job.seeds, job.mask, job.uv)
    assert osteps == rsteps
    assert np.array_equal(oslc, rslc)
    assert np.array_equal(oslt, rslt)
    # sum(slc) = one write at each seed + one per accepted step
    assert int(oslc.sum()) == ns * d * d + osteps


@pytest.mark.parametrize("grid,contract,sign,max_error", [(24, True, 1, None), (24, False, -1, None),
                                                          (28, True, -1, None), (32, True, 1, 1e-6)])
def test_paths_match_pyref(grid, contract, sign, max_error):
    job = _tiny_job(grid, 1, 0.0, max_error=max_error)
    ns = min(30, job.seeds.shape[0])
    p = job.params(sign, kernel="INTEGRATE_TRAJECTORY", fp_contract=contract, n_seed_points=ns)
    p.max_n_steps = min(p.max_n_steps, 400)
    m = job.mask.astype(np.uint8).reshape(-1)
    otraj, onst, oln = oracle.trace_paths(p, job.seeds, m, job.uv.reshape(-1), nthreads=1)
    rtraj, rnst, rln = pyref.paths(p, job.seeds, job.mask, job.uv)
    assert np.array_equal(onst, rnst)
    assert np.array_equal(oln.view(np.uint32), rln.view(np.uint32))
    assert np.array_equal(otraj, rtraj)


def test_oracle_threads_invariant():
    """Integer fields do not depend on the order of accumulation."""
    job = _tiny_job(96, 2, 1.0)
    p = job.params(1)
    m = job.mask.astype(np.uint8).reshape(-1)
    a = oracle.trace_fields(p, job.seeds, m, job.uv.reshape(-1), nthreads=1)
    b = oracle.trace_fields(p, job.seeds, m, job.uv.reshape(-1), nthreads=8)
    assert a[2] == b[2] and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("d", [1, 2])
def test_positions_on_lattice(d):
    """For d in {1, 2} every recorded position is a multiple of 1/128 (the
    sub-pixel offsets 0 and -1/4 are, and every step is approximate()d)."""
    job = _tiny_job(96, d, 1.0)
    p = job.params(1)
    m = job.mask.astype(np.uint8).reshape(-1)
    for k in range(0, 40):
        xy, _ = oracle.trace_fields_record(p, job.seeds, m, job.uv.reshape(-1), k, k % (d * d), 4096)
        q = xy.astype(np.float64) * 128.0
        assert np.array_equal(q, np.round(q))


def test_fp64_twin_close():
    """FP64 trace of the same algorithm stays within the 1/128 lattice of the
    FP32 trace for the first steps (it diverges only after a truncation lands
    differently)."""
    job = _tiny_job(96, 1, 1.0)
    p = job.params(1)
    m = job.mask.astype(np.uint8).reshape(-1)
    agree = 0
    for k in range(40):
        a, _ = oracle.trace_fields_record(p, job.seeds, m, job.uv.reshape(-1), k, 0, 64)
        b, _ = oracle.trace_fields_record(p, job.seeds, m, job.uv.reshape(-1), k, 0, 64, f64=True)
        kk = min(len(a), len(b), 5)
        agree += int(np.abs(a[:kk].astype(np.float64) - b[:kk]).max() <= 1.0 / 128 + 1e-9)
    assert agree >= 36


# ---------------------------------------------------------------- KDE

def _kde_case(kernel, n_hist=512, n_pdf=128, n=20000, bandwidth=1.0, seed=0):
    from slm_amd.kde import Univariate_distribution, kde_params, prepare_filter
    rs = np.random.default_rng(seed)
    x = np.concatenate([rs.lognormal(3.0, 0.6, n // 2), rs.lognormal(4.5, 0.3, n - n // 2)]).astype(np.float32)
    lx = np.log(x).astype(np.float32)
    d = Univariate_distribution(logx_array=lx, logy_array=lx, n_hist_bins=n_hist, n_pdf_points=n_pdf)
    d.kernel, d.bandwidth = kernel, bandwidth
    prepare_filter(d)
    return d, lx


@pytest.mark.parametrize("contract", [True, False])
def test_kde_histogram_matches_numpy(contract):
    from slm_amd.kde import kde_params
    d, lx = _kde_case("epanechnikov")
    k = kde_params(d, contract)
    h = oracle.kde_histogram(k, d.logx_data.reshape(-1))
    bx = ((d.logx_data.reshape(-1) - np.float32(k.logx_min)) / np.float32(k.logx_range)).astype(np.float32)
    t = (bx * np.float32(k.n_hist_bins)).astype(np.float32)
    idx = np.minimum(np.trunc(np.maximum(t, 0)).astype(np.int64), k.n_hist_bins - 1)
    assert np.array_equal(h, np.bincount(idx, minlength=k.n_hist_bins).astype(np.uint32))
    assert int(h.sum()) == d.n_data


def _pdf_pure_python(k, hist):
    """filter/pdf_univariate (kde.cl:426-526) in plain Python doubles."""
    import math
    nbp = k.n_hist_bins // k.n_pdf_points
    w = float(np.float32(k.kdf_width_x))
    out = np.zeros(k.n_pdf_points, dtype=np.float32)
    fma = lambda a, b, c: float(np.float64(a) * np.float64(b) + np.float64(c)) if not k.fp_contract else None

    def ks(x):
        if k.kernel == 0:
            return 1.0 if abs(x) < w else 0.0
        if k.kernel == 1:
            return max(1.0 - abs(x / w), 0.0)
        if k.kernel == 2:
            q = x / w
            if k.fp_contract:
                from fractions import Fraction
                return max(float(Fraction(1) - Fraction(q) * Fraction(q)), 0.0)
            return max(1.0 - q * q, 0.0)
        if k.kernel == 3:
            return max(math.cos(1.57079632679489661923 * x), 0.0)
        q = x / (w * float(np.float32(0.4)))
        return math.exp(-0.5 * q * q)

    for p in range(k.n_pdf_points):
        acc = kw = 0.0
        for kk in range(k.n_kdf_part_points_x + 1):
            edge = kk * float(np.float32(k.pdf_dx)) - float(np.float32(np.float32(k.pdf_dx) / np.float32(2.0)))
            cr, cl = p + kk, p - kk + 1
            for h in range(nbp):
                hh = h + 0.5
                if k.fp_contract:
                    from fractions import Fraction
                    x = float(Fraction(float(np.float32(k.bin_dx))) * Fraction(hh) + Fraction(edge))
                else:
                    x = edge + float(np.float32(k.bin_dx)) * hh
                v = ks(x)
                okr, okl = cr < k.n_pdf_points, (kk != 0 and cl > 0)
                acc += float(hist[cr * nbp + h]) * v if okr else 0.0
                acc += float(hist[cl * nbp - h - 1]) * v if okl else 0.0
                kw += v if okr else 0.0
                kw += v if okl else 0.0
        out[p] = np.float32(float(np.float32(acc)) / kw)
    return out


@pytest.mark.parametrize("kernel,contract", [("epanechnikov", True), ("epanechnikov", False),
                                             ("triangle", True), ("tophat", False), ("gaussian", True)])
def test_kde_pdf_matches_pure_python(kernel, contract):
    from slm_amd.kde import kde_params
    d, lx = _kde_case(kernel, n_hist=256, n_pdf=64, n=5000)
    k = kde_params(d, contract)
    h = oracle.kde_histogram(k, d.logx_data.reshape(-1))
    got = oracle.kde_pdf(k, h)
    exp = _pdf_pure_python(k, h)
    assert np.array_equal(got, exp)
    assert k.n_kdf_part_points_x >= 1


def test_kde_pdf_delta_is_kernel_shape():
    """A histogram with all mass in one pdf bin gives back the (normalised)
    filter shape around it: symmetric, peaked, zero beyond the half-width."""
    from slm_amd.kde import kde_params
    d, lx = _kde_case("epanechnikov", n_hist=512, n_pdf=128, n=20000, bandwidth=3.0)
    k = kde_params(d, True)
    nbp = k.n_hist_bins // k.n_pdf_points
    c = 60
    hist = np.zeros(k.n_hist_bins, dtype=np.uint32)
    hist[c * nbp:(c + 1) * nbp] = 100
    pdf = oracle.kde_pdf(k, hist)
    K = k.n_kdf_part_points_x
    assert pdf[c] == pdf.max() and pdf[c] > 0
    assert np.all(pdf[:c - K - 1] == 0) and np.all(pdf[c + K + 1:] == 0)
    for j in range(1, K):
        assert abs(float(pdf[c + j]) - float(pdf[c - j])) <= 0.05 * float(pdf[c])


# ---------------------------------------------------------------- hillslope lengths

def _hsl_case(contract=True, offset=12):
    from slm_amd import workloads
    from slm_amd.params import kernel_params
    from slm_amd.useful import pick_seeds
    job, mapping = workloads.v_valley(midslope_offset=offset)
    seeds = pick_seeds(mask=job.mask, map=mapping, flag=int(job.info.is_midslope), pad=2)
    job.info.n_seed_points = np.uint32(seeds.shape[0])
    p = kernel_params(job.info, 1, "HILLSLOPE_LENGTHS", contract)
    return job, mapping, seeds, p


def test_dt_adapts_with_small_error():
    """With a tiny step-error bound the step size really adapts: steps are
    shorter and more numerous than with the default bound."""
    a = _tiny_job(32, 1, 0.0)
    b = _tiny_job(32, 1, 0.0, max_error=1e-6)
    m = a.mask.astype(np.uint8).reshape(-1)
    sa = oracle.trace_fields(a.params(1), a.seeds, m, a.uv.reshape(-1))[2]
    sb = oracle.trace_fields(b.params(1), b.seeds, m, b.uv.reshape(-1))[2]
    assert sb > 1.5 * sa


@pytest.mark.parametrize("contract", [True, False])
def test_hsl_matches_pyref(contract):
    job, mapping, seeds, p = _hsl_case(contract)
    thin, mid = int(job.info.is_thinchannel), int(job.info.is_midslope)
    p.n_seed_points = 60
    o = oracle.hillslope_lengths(p, seeds, job.mask.astype(np.uint8).reshape(-1), job.uv.reshape(-1),
                                 mapping.reshape(-1), thin, mid)
    r = pyref.hillslope_lengths(p, seeds, job.mask, job.uv, mapping, thin, mid)
    assert np.array_equal(o.view(np.uint32), r.view(np.uint32))


@pytest.mark.parametrize("offset", [8, 12, 20])
def test_hsl_analytic_v_valley(offset):
    """On a planar V-valley (side slope s, down-valley slope t) the streamline
    from x = c +- offset reaches the axis after offset*sqrt(1+(t/s)^2) pixels."""
    job, mapping, seeds, p = _hsl_case(True, offset)
    L = oracle.hillslope_lengths(p, seeds, job.mask.astype(np.uint8).reshape(-1), job.uv.reshape(-1),
                                 mapping.reshape(-1), int(job.info.is_thinchannel), int(job.info.is_midslope))
    exp = offset * np.sqrt(1 + (0.05 / 0.2) ** 2)
    reached = L[L > 0]
    assert reached.size >= 0.9 * L.size
    assert abs(float(np.median(reached)) - exp) < 0.5


@pytest.mark.parametrize("jitter,min_equal", [(0.0, 0.995), (1.0, 0.98)])
def test_fp32_against_fp64_oracle(jitter, min_equal):
    """Errors stated against the FP64 trace (the same algorithm in double
    precision; the GPU is bit-identical to the FP32 oracle).  Positions stay
    on the 1/128 lattice in both, so a streamline's FP32 and FP64 traces are
    identical until an approximate() truncation decides differently (first
    after ~45 steps on the 256^2 workload); from there they are different
    streamlines.  So: the leading 32 positions agree exactly for at least 98%
    of 200 sampled streamlines (all 200 with jitter 1.0, 199 without: RMS 0
    and 2.8e-4 px); over whole maps, slc agrees on at least 98% (jitter 1.0)
    / 99.5% (no jitter) of the unmasked cells, and the total step count
    within 0.1%."""
    from slm_amd import workloads
    job = workloads.small_256(jitter=jitter, d=1)
    m = job.mask.astype(np.uint8).reshape(-1)
    uv = job.uv.reshape(-1)
    cells = ~job.mask.reshape(-1)
    for sign in (1, -1):
        p = job.params(sign)
        c32, _, s32 = oracle.trace_fields(p, job.seeds, m, uv)
        c64, _, s64 = oracle.trace_fields(p, job.seeds, m, uv, f64=True)
        assert (c32 == c64)[cells].mean() >= min_equal
        assert abs(s32 - s64) <= 1e-3 * s64
    p = job.params(1)
    ks = np.random.default_rng(5).integers(0, job.seeds.shape[0], 200)
    same = 0
    for k in ks:
        a, _ = oracle.trace_fields_record(p, job.seeds, m, uv, int(k), 0, 64)
        b, _ = oracle.trace_fields_record(p, job.seeds, m, uv, int(k), 0, 64, f64=True)
        kk = min(len(a), len(b), 32)
        same += int(np.array_equal(a[:kk].astype(np.float64), b[:kk]))
    assert same >= 0.98 * len(ks)
