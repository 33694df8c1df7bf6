"""GPU parity of slm_condition_uv (the reference's loop and blockage
conditioning, preprocess.py:531-600) with the CPU oracle
(oracle/condition_ref.py): uv bits, mask and the blockage / blocked-neighbour
/ loop maps on rough unfilled DTMs with ties, pits and NaN regions, in both
normalisation modes; on a plane it equals slm_gradient_uv."""
import numpy as np
import pytest

from oracle import condition_ref as C

pytestmark = pytest.mark.gpu


def _dtms():
    from slm_amd import synth
    rs = np.random.default_rng(4)
    a = synth.make_dtm((150, 131), pixel_size=1.0, relief=20.0, tilt=0.02, seed=3, nan_border=2, fill=False)
    b = np.round(rs.standard_normal((97, 120)) * 2).astype(np.float32)       # ties everywhere
    b[40:45, 60:70] = np.nan
    c = synth.make_dtm(256, pixel_size=2.0, relief=30.0, tilt=0.05, seed=8, nan_border=2, n_valleys=2,
                       cone=True, n_pits=30, fill=False)
    return [("unfilled", a, 1.0), ("ties", b, 0.5), ("cone_pits", c, 2.0)]


@pytest.mark.parametrize("normalize", [True, False])
def test_condition_uv_bit_exact(gpu, normalize):
    import torch
    from slm_amd import hip
    for name, roi, ps in _dtms():
        print(name, roi.shape, flush=True)
        uv_ref, m_ref, b_ref, n_ref, loops_ref = C.conditioned_uv(roi, ps, 2, do_normalize_speed=normalize)
        r = torch.from_numpy(np.ascontiguousarray(roi)).to(gpu)
        uv, m, maps = hip.condition_uv(r, 2, ps, do_normalize_speed=normalize, maps=True, sync=True)
        maps = maps.cpu().numpy()
        assert np.array_equal(maps[0], b_ref), name
        assert np.array_equal(maps[1], n_ref), name
        assert np.array_equal(maps[2][:-1, :-1].astype(bool), loops_ref), name
        assert not maps[2][-1, :].any() and not maps[2][:, -1].any()
        # the entry point returns the merged mask, as slm_gradient_uv does:
        # NaN uv, NaN DTM and the pad ring
        from slm_amd import preprocess
        assert np.array_equal(m.cpu().numpy(), m_ref | preprocess.dtm_mask(roi, 2)), name
        assert np.array_equal(uv.cpu().numpy().view(np.uint32), uv_ref.view(np.uint32)), name
        assert b_ref.any() and loops_ref.any(), name          # the fixes are exercised


def test_condition_plane_equals_gradient_uv(gpu):
    import torch
    from slm_amd import hip
    n = 64
    x, y = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    roi = (-(0.3 * x + 0.1 * y)).astype(np.float32)
    roi[:2] = np.nan
    r = torch.from_numpy(roi).to(gpu)
    uv, m = hip.condition_uv(r, 2, 1.0, sync=True)
    uv0, m0 = hip.gradient_uv(r, 2, 1.0, sync=True)
    assert torch.equal(uv.view(torch.int32), uv0.view(torch.int32)) and torch.equal(m, m0)


def test_condition_bad_arguments(gpu):
    import torch
    from slm_amd import hip
    r = torch.zeros((8, 8), dtype=torch.float32, device=gpu)
    out = torch.empty((12, 12, 2), dtype=torch.float32, device=gpu)
    mask = torch.empty((12, 12), dtype=torch.uint8, device=gpu)
    ws = torch.empty(16, dtype=torch.uint8, device=gpu)        # far below slm_condition_workspace_size
    with pytest.raises(hip.SlmError):
        hip.check(hip.load().slm_condition_uv(hip._ptr(r), 8, 8, 2, 1.0, 1, 0.95, -0.5, 0.0, hip._ptr(out),
                                              hip._ptr(mask), None, hip._ptr(ws), 16, None, 1))
    with pytest.raises(hip.SlmError):                            # np.gradient needs 2 points
        hip.condition_uv(torch.zeros((1, 8), dtype=torch.float32, device=gpu), 2, 1.0)


def test_preprocess_facade(gpu):
    """Preprocess.do() with the JSON defaults (do_condition true): the
    reference's attributes, the uv of the oracle, the 'uv' active mask; with
    do_condition false the plain gradient field."""
    from types import SimpleNamespace
    from slm_amd import preprocess
    from slm_amd.parameters import load_parameters, section
    name, roi, ps = _dtms()[0]
    masks = {}
    params = load_parameters()
    state = section(params, "state")
    state.add_active_mask = masks.update
    geo = SimpleNamespace(roi_array=roi, roi_pixel_size=ps, pad_width=2)
    pre = preprocess.Preprocess(state, section(params, "preprocess"), geo)
    pre.do()
    uv_ref, m_ref, b_ref, n_ref, loops_ref = C.conditioned_uv(roi, ps, 2)
    assert np.array_equal(pre.uv_array.view(np.uint32), uv_ref.view(np.uint32))
    assert pre.n_loops == 4 * int(loops_ref.sum())
    assert np.array_equal(pre.where_blockages_array, np.argwhere(b_ref))
    assert np.array_equal(masks["uv"], m_ref | preprocess.dtm_mask(roi, 2))
    state.do_condition = False
    pre.do()
    uv0, m0 = preprocess.gradient_uv(roi, ps, 2)
    assert np.array_equal(pre.uv_array.view(np.uint32), uv0.view(np.uint32))
