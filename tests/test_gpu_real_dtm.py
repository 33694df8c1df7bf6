"""End to end on the reference's own DTMs (tests/golden/dtm): Geodata reads
the GeoTIFF, Preprocess conditions the uv (loops and blockages, the JSON
default), Trace runs trajectories and fields on the GPU, through the same
classes and JSON keys as the reference (python/streamlines/slm.py order:
geodata, preprocess, trace).  The raw slc/slt of both directions must equal
the CPU oracle's on the same conditioned grids and seeds, bit for bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DTM = os.path.join(HERE, "golden", "dtm")


@pytest.mark.parametrize("name,trace", [
    # the trace settings of the reference's tests/IndianCreek_Test1.json
    ("IndianCreek_res1m_fillsinks0p01_cuttest2.tif", {"subpixel_seed_point_density": 1, "max_length": 30.0,
                                                     "n_trajectory_seed_points": 0}),
    ("Guadalupe_DTM0_res2m_fill_sinks1deg_cuttest1.tif", {"subpixel_seed_point_density": 2, "max_length": 200.0,
                                                         "n_trajectory_seed_points": 500}),
])
def test_real_dtm_end_to_end(gpu, name, trace):
    from oracle import condition_ref, oracle
    from slm_amd.geodata import Geodata, State
    from slm_amd.parameters import load_parameters, section
    from slm_amd.params import kernel_params
    from slm_amd.preprocess import Preprocess, dtm_mask
    from slm_amd.trace import Trace
    params = load_parameters(overrides={"geodata": {"dtm_file": name}, "trace": trace})
    state = State(params["state"])
    geo = Geodata(state, params["geodata"], dtm_path=DTM)
    geo.do()
    pre = Preprocess(state, section(params, "preprocess"), geo)
    pre.do()
    assert pre.n_loops > 0 or len(pre.where_blockages_array) > 0       # conditioning did something
    tr = Trace(state, section(params, "trace"), geo, pre)
    tr.do()
    f = tr.fields
    print(name, geo.roi_array.shape, "steps", f.steps, "loops", pre.n_loops,
          "blockages", len(pre.where_blockages_array), flush=True)
    # oracle: conditioned uv on the CPU, the merged mask, the same seeds and grid crop
    uv_ref, m_ref, _, _, _ = condition_ref.conditioned_uv(geo.roi_array, geo.roi_pixel_size, 2)
    assert np.array_equal(pre.uv_array.view(np.uint32), uv_ref.view(np.uint32))
    mask_full = dtm_mask(geo.roi_array, 2) | m_ref
    assert np.array_equal(state.merge_active_masks(), mask_full)
    d = f.data
    nxp, nyp = d.mask_array.shape
    for k, sign in enumerate((1, -1)):
        f.info.downup_sign = float(sign)
        p = kernel_params(f.info, sign, "INTEGRATE_FIELDS", False)
        c, t, _ = oracle.trace_fields(p, d.seed_point_array, d.mask_array.reshape(-1).astype(np.uint8),
                                      d.uv_array.reshape(-1))
        assert np.array_equal(f.raw_slc[..., k], c.reshape(nxp, nyp)), (name, sign)
        assert np.array_equal(f.raw_slt[..., k], t.reshape(nxp, nyp)), (name, sign)
    assert tr.slc_array.shape == (geo.roi_nx + 4, geo.roi_ny + 4, 2)
