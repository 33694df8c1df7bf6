"""Host logic of the path (no GPU): parameter derivation and macro rendering,
seed lists, JSON merging, post-processing dtypes, and the C ABI exports."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _info(nx=250, ny=180, d=2, pixel_size=2.0, max_length=1000.0):
    from types import SimpleNamespace
    from slm_amd.parameters import load_parameters, section
    from slm_amd.useful import Info
    prm = load_parameters(overrides={"trace": {"subpixel_seed_point_density": d, "max_length": max_length}})
    info = Info(section(prm, "state"), section(prm, "trace"), pixel_size)
    info.set_xy(nx, ny, 2)
    info.n_seed_points = np.uint32(1234)
    return info


def test_macro_values_closed_form():
    from slm_amd.params import kernel_params, render_macros
    info = _info()
    opts = render_macros(info, "INTEGRATE_FIELDS", -1)
    d = dict(o.split("=", 1) for o in opts[1::2] if "=" in o)
    assert opts[:2] == ["-D", "KERNEL_INTEGRATE_FIELDS"]
    assert d["DOWNUP_SIGN"] == "-1u"                    # np.int8(-1) rendered with 'u'
    assert d["NX_PADDED"] == "254u" and d["NY_PADDED"] == "184u" and d["NXY_PADDED"] == "46736u"
    assert d["PAD_WIDTH_PP5"] == "2.5f"
    assert d["MAX_N_STEPS"] == "2000u"                  # 1000 m / 0.5, metres not pixels
    assert d["MAX_LENGTH"] == "500.0f"                  # 1000 m / 2 m per pixel
    assert d["SUBPIXEL_SEED_HALFSPAN"] == "0.25f" and d["SUBPIXEL_SEED_STEP"] == "0.25f"
    assert d["ADJUSTED_MAX_ERROR"] == repr(float(0.85 * np.sqrt(0.03))) + "f"
    p = kernel_params(info, -1)
    gs = np.float32(np.sqrt(np.float32(np.float32(250) * np.float32(180))))
    assert np.float32(p.grid_scale) == gs
    assert np.float32(p.combo_factor) == -np.float32(gs * np.float32(0.5))
    assert np.float32(p.dt_max) == np.float32(1.0 / 250.0)  # float64 1/nx, then rounded
    assert p.downup_sign == 4294967296.0                  # (float)(-1u)
    assert kernel_params(info, 1).downup_sign == 1.0
    assert p.n_seed_points == 1234 and p.max_n_steps == 2000


def test_macro_flag_duplicates_previous():
    """A false flag re-appends the previous option (pocl.py:283-289 loop)."""
    from slm_amd.params import render_macros
    opts = render_macros(_info(), "INTEGRATE_TRAJECTORY", 1)
    tail = opts[-8:]
    assert tail[0:2] == tail[2:4] == tail[4:6] == tail[6:8]


def test_params_struct_layout_matches_header():
    from slm_amd.params import SlmParams
    hdr = open(os.path.join(ROOT, "include", "slm_hip.h")).read()
    body = hdr[hdr.index("typedef struct SlmParams {"):hdr.index("} SlmParams;")]
    names = re.findall(r"^\s+(?:uint32_t|float)\s+(\w+)", body, re.M)
    assert names == [n for n, _ in SlmParams._fields_]
    assert ctypes.sizeof(SlmParams) == 128


def test_d3_offsets():
    info = _info(d=3)
    assert np.float32(info.subpixel_seed_halfspan) == np.float32((1 - 1.0 / 3.0) / 2.0)
    assert np.float32(info.subpixel_seed_step) == np.float32((1 - 1.0 / 3.0) / 3.0)  # span/d, not span/(d-1)


def test_pad_seed_count_bankers_rounding():
    from slm_amd.params import pad_seed_count
    assert pad_seed_count(64, 32) == 64      # round(2.5) = 2
    assert pad_seed_count(96, 32) == 128     # round(3.5) = 4
    assert pad_seed_count(65, 32) == 96


def test_seeds_match_golden():
    from slm_amd import workloads
    g = np.load(os.path.join(ROOT, "tests", "golden", "seeds_small_256.npz"))
    job = workloads.small_256()
    assert job.seeds.shape[0] == int(g["n_seeds"])
    assert np.array_equal(job.seeds[:1000], g["seeds_head"])


def test_create_seeds_contract():
    from slm_amd.useful import create_seeds
    mask = np.ones((10, 12), dtype=bool)
    mask[3:6, 4:9] = False
    s, n, npad = create_seeds(mask, 2, 32, do_shuffle=False)
    assert n == 15 and npad == 32 and s.dtype == np.float32
    assert np.array_equal(s[0], [1.0, 2.0])   # argwhere - pad
    s2, n2, _ = create_seeds(mask, 2, 32, n_seed_points=4, rng_seed=1)
    assert n2 == 4 and s2.shape == (4, 2)


def test_json_merge_by_key(tmp_path):
    from slm_amd.parameters import load_parameters
    job = tmp_path / "job.json"
    job.write_text('{"trace": {"subpixel_seed_point_density": 2}, "state": {"verbose": true}}')
    prm = load_parameters(str(job))
    assert prm["trace"]["subpixel_seed_point_density"] == 2
    assert prm["trace"]["max_length"] == 1000.0       # untouched keys survive
    assert prm["state"]["verbose"] is True


def test_postprocess_dtypes():
    from oracle.postprocess_ref import postprocess_raw, postprocess_scaled
    slc = np.array([[[0, 3], [7, 1]]], dtype=np.uint32)
    slt_u = np.array([[[0, 10], [16777217, 1]]], dtype=np.uint32)   # 2^24+1 rounds in float32
    slc2, slt2, sla = postprocess_raw(slc, slt_u.astype(np.float32), 2)
    assert slt2.dtype == np.float32 and sla.dtype == np.float32
    assert sla[0, 0, 0] == 0.0
    exp = np.float32(np.float64(np.float32(np.float32(np.sqrt(2.0)) * np.float32(10.0))) / 3.0)
    assert sla[0, 0, 1] == exp
    assert slt2[0, 1, 0] == np.float32(np.float32(16777216.0) / np.float32(4))
    slc3, slt3 = postprocess_scaled(slc2, slt2, 2, np.float32(0.45), np.float32(0.5), 1.0)
    assert slc3.dtype == np.float64
    assert slc3[0, 1, 0] == np.sqrt(2.0) * np.power(7.0 / 4.0, 2.0 / 3.0)
    assert slt3.dtype == np.float32


def test_library_exports_header_symbols():
    """libslm_hip.so loads (no GPU needed) and exports every function the
    header declares; the ctypes binding covers exactly that set."""
    from slm_amd import hip
    hdr = open(os.path.join(ROOT, "include", "slm_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|size_t|uint32_t|const char\*)\s+(slm_\w+)\s*\(", hdr, re.M))
    lib = hip.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(hip.exported_symbols())
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(r"\bT %s$" % name, out, re.M), name
    assert lib.slm_abi_version() == 2


def test_library_rejects_bad_params_without_gpu():
    """Argument checks run on the host before any device call."""
    from slm_amd import hip
    from slm_amd.params import ABI_VERSION, SlmParams
    lib = hip.load()
    p = SlmParams()
    p.abi_version = 99
    assert lib.slm_trace_fields(ctypes.byref(p), None, None, None, None, None, None, None, 0, None, 0) == -4
    p.abi_version = ABI_VERSION
    p.nx_padded, p.ny_padded, p.nxy_padded = 10, 10, 99
    assert lib.slm_trace_paths(ctypes.byref(p), None, None, None, None, None, None, None, 0) == -1
    assert b"nxy_padded" in lib.slm_last_error()
    assert lib.slm_trace_fields_workspace_size(ctypes.byref(p)) == 0


def test_postprocess_known_answers():
    """fields.py:296-309 and :141-164 by hand: slc 4, slt 8 (d=2) gives
    sla = sqrt(2)*8/4, slt = 8/4 = 2, then slt*(0.5/1) = 1 -> sqrt 1 and
    slc = sqrt(2)*(4/4)^(2/3) = sqrt(2); empty pixels stay 0."""
    from oracle.postprocess_ref import postprocess_raw, postprocess_scaled
    slc = np.array([[[4, 0]]], dtype=np.uint32)
    slt = np.array([[[8.0, 0.0]]], dtype=np.float32)
    c, t, a = postprocess_raw(slc, slt, 2)
    assert a[0, 0, 0] == np.float32(np.float64(np.float32(np.sqrt(2.0)) * np.float32(8.0)) / 4.0)
    assert t[0, 0, 0] == 2.0 and a[0, 0, 1] == 0.0 and t[0, 0, 1] == 0.0
    c2, t2 = postprocess_scaled(c, t, 2, np.float32(0.5), np.float32(0.5), 1.0)
    assert t2[0, 0, 0] == 1.0 and t2[0, 0, 1] == 0.0
    assert c2[0, 0, 0] == np.sqrt(2.0) and c2[0, 0, 1] == 0.0


def test_gradient_uv_planar_closed_form():
    """A planar ramp h = -(a x + b y) p has grad = -(a, b) everywhere, so the
    unit flow vector is (a, b)/|(a, b)| at every pixel, edges included."""
    from slm_amd import preprocess
    a, b, p = 0.75, 0.25, 2.0
    x = np.arange(20, dtype=np.float32)[:, None]
    y = np.arange(30, dtype=np.float32)[None, :]
    roi = (-(a * x + b * y) * p).astype(np.float32)
    uv, m = preprocess.gradient_uv(roi, p, 2)
    assert not m.any()
    n = np.hypot(a, b)
    assert np.allclose(uv[..., 0], a / n, atol=1e-6) and np.allclose(uv[..., 1], b / n, atol=1e-6)


def test_v_valleys_c_matches_numpy_expression():
    """The banded C valley loop gives the bits of the whole-grid numpy
    expression -depth*clip(1 - |(x-cx) sin - (y-cy) cos| / hw, 0, 1)."""
    from slm_amd import synth
    for nx, ny, n, hw in [(120, 77, 4, 9.0), (64, 300, 6, 30.5), (257, 255, 3, 61.44)]:
        rng = np.random.default_rng(nx + ny)
        got = synth.v_valleys(nx, ny, 1.0, n, 30.0, hw, np.random.default_rng(nx + ny))
        x = np.arange(nx)[:, None].astype(np.float64)
        y = np.arange(ny)[None, :].astype(np.float64)
        want = np.zeros((nx, ny))
        for _ in range(n):
            ang = rng.uniform(0.0, np.pi)
            cx, cy = rng.uniform(0.2, 0.8) * nx, rng.uniform(0.2, 0.8) * ny
            d = np.abs((x - cx) * np.sin(ang) - (y - cy) * np.cos(ang))
            want -= 30.0 * np.clip(1.0 - d / hw, 0.0, 1.0)
        assert np.array_equal(got, want)
        assert (got < 0).any()


def test_surrogate_tile_cores_partition_grid():
    from slm_amd import gated
    for nx, ny, size, halo in [(100, 37, 32, 4), (256, 256, 64, 8), (8190, 8190, 256, 16)]:
        o = gated.tile_origins(nx, ny, size, halo)
        c = size - 2 * halo
        if nx * ny <= 70000:
            cover = np.zeros((nx, ny), np.int32)
            for i0, j0 in o:
                cover[max(0, i0 + halo):i0 + halo + c, max(0, j0 + halo):j0 + halo + c] += 1
            assert (cover == 1).all()
        assert len(o) == int(np.ceil(nx / c)) * int(np.ceil(ny / c))
        assert o.min() == -halo and (o + halo).max() < max(nx, ny)
