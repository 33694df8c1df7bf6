"""DTM input and map output (slm_amd/geodata.py) on the CPU.

The TIFF reader is checked against Pillow on the reference's own test DTMs
(tests/golden/dtm: Indian Creek 1 m and Guadalupe 2 m cut-outs from the
reference's data folder) and on files it wrote itself; the geotransform
against the files' ModelPixelScale / ModelTiepoint tags; the region of
interest against direct indexing of the file's grid (flipud, then x = column,
y = row from the bottom, as geodata.py:163-166); the masks against their
definitions; and GeoTIFF export by a write / read round trip."""
import os

import numpy as np
import pytest

from slm_amd import geodata as G

HERE = os.path.dirname(os.path.abspath(__file__))
DTM = os.path.join(HERE, "golden", "dtm")
FILES = ["IndianCreek_res1m_fillsinks0p01_cuttest2.tif", "Guadalupe_DTM0_res2m_fill_sinks1deg_cuttest1.tif"]


@pytest.mark.parametrize("name", FILES)
def test_read_tiff_matches_pillow(name):
    from PIL import Image
    a, info = G.read_tiff(os.path.join(DTM, name))
    im = Image.open(os.path.join(DTM, name))
    b = np.array(im)
    assert a.dtype == np.float32 and np.array_equal(a, b, equal_nan=True)
    sx, sy = im.tag_v2[33550][:2]
    i, j, _, X, Y, _ = im.tag_v2[33922][:6]
    assert info.geotransform == (X - i * sx, sx, 0.0, Y + j * sy, 0.0, -sy)
    assert 34735 in info.geo_tags and 34737 in info.geo_tags


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint8, np.int16, np.uint32, np.int32])
def test_tiff_write_read_round_trip(tmp_path, dtype):
    rs = np.random.default_rng(1)
    a = (rs.standard_normal((37, 53)) * 100).astype(dtype)
    gt = (1000.0, 2.0, 0.0, 5000.0, 0.0, -2.0)
    tags = {34737: (2, "WGS 84 / UTM zone 17N|WGS 84|\x00")}
    f = str(tmp_path / "x.tif")
    G.write_tiff(f, a, gt, tags)
    b, info = G.read_tiff(f)
    assert b.dtype == a.dtype and np.array_equal(a, b)
    assert info.geotransform == gt and info.geo_tags[34737][1] == tags[34737][1]
    if dtype in (np.float32, np.uint8, np.int32):
        from PIL import Image
        assert np.array_equal(np.array(Image.open(f)), a)


def _indian_creek(**kw):
    state = G.State({})
    params = {"dtm_file": FILES[0], "no_data_values": [-99999, 0], "pad_width": 2}
    params.update(kw)
    g = G.Geodata(state, params, dtm_path=DTM)
    g.do()
    return state, g


def test_geodata_roi_orientation_and_mask():
    # the reference's tests/IndianCreek_Test1.json region of interest
    state, g = _indian_creek(do_clip_roi=True, roi_x_bounds=[115, 130], roi_y_bounds=[65, 80])
    raw, _ = G.read_tiff(os.path.join(DTM, FILES[0]))
    assert g.roi_array.shape == (15, 15) and (g.roi_nx, g.roi_ny) == (15, 15)
    for i in range(15):
        for j in range(15):
            assert g.roi_array[i, j] == np.flipud(raw)[65 + j, 115 + i]
    assert g.pixel_size == 1.0
    assert np.isclose(g.x_roi_n_pixel_centers[0], 115.5) and np.isclose(g.y_roi_n_pixel_centers[-1], 79.5)
    m = state.merge_active_masks()
    assert m.shape == (19, 19) and m[:2].all() and m[:, -2:].all()
    assert np.array_equal(m[2:-2, 2:-2], np.isnan(g.roi_array))
    # metres
    _, g2 = _indian_creek(do_clip_roi=True, roi_bounds_units="meters", roi_x_bounds_meters=[115.0, 130.0],
                          roi_y_bounds_meters=[65.0, 80.0])
    assert np.array_equal(g2.roi_array, g.roi_array, equal_nan=True)


def test_geodata_no_data_and_h_min():
    state, g = _indian_creek()
    raw, _ = G.read_tiff(os.path.join(DTM, FILES[0]))
    assert g.roi_array.shape == raw.T.shape
    assert np.isnan(g.dtm_array[raw == 0]).all() if (raw == 0).any() else True
    h_min = float(np.nanmedian(g.roi_array))
    state2, g2 = _indian_creek(h_min=h_min)
    m = state2.merge_active_masks()[2:-2, 2:-2]
    with np.errstate(invalid="ignore"):
        assert np.array_equal(m, np.isnan(g.roi_array) | (g.roi_array <= h_min))
    with pytest.raises(ValueError):
        _indian_creek(do_clip_roi=True, roi_x_bounds=[0, 10000], roi_y_bounds=[0, 10])


def test_geotiff_export_round_trip(tmp_path):
    _, g = _indian_creek(do_clip_roi=True, roi_x_bounds=[20, 90], roi_y_bounds=[30, 75])
    pad = 2
    field = np.arange((g.roi_nx + 2 * pad) * (g.roi_ny + 2 * pad), dtype=np.float32).reshape(
        g.roi_nx + 2 * pad, g.roi_ny + 2 * pad)
    sl = np.index_exp[pad:-pad, pad:-pad]
    G.write_geotiff(str(tmp_path), "slc.tif", field, g.roi_nx, g.roi_ny, 1, sl, g)
    back, info = G.read_tiff(str(tmp_path / "slc.tif"))
    assert np.array_equal(back, np.flipud(field[sl].T))
    assert np.allclose(info.geotransform, g.roi_geotransform)
    # the ROI's top-left corner in map coordinates
    gt = g.tiff.GetGeoTransform()
    assert np.isclose(info.geotransform[0], gt[0] + 20 * g.pixel_size)
    assert np.isclose(info.geotransform[3], gt[3] - (g.dtm_array.shape[0] - 75) * g.pixel_size)
    # a boolean map is written as uint8
    G.write_geotiff(str(tmp_path), "m.tif", field > 100, g.roi_nx, g.roi_ny, 1, sl, g)
    assert G.read_tiff(str(tmp_path / "m.tif"))[0].dtype == np.uint8
