"""What the trained surrogate gains over the methods without it, on a DTM it
was not trained on.

The network is trained on traced family DTMs (16 of 1024^2, seeds 1000..,
1024 tiles of 128^2, 30 epochs) and its channel logit calibrated on those
training tiles.  On a held-out 1024^2 family DTM (seed 90000) the truth is
the plain downstream trace (one seed per pixel, streamlines capped at
100 m), and three maps of its channel pixels are scored against it:
  - the network's prediction (256^2 tiles, 16 px halo), logit > 0: its
    threshold was fixed on the training tiles;
  - bounded D8 flow accumulation (paths capped at the same 100 m), at the
    threshold with the best F1 on the held-out truth itself;
  - a 2x coarser DTM (2 m pixels) traced the same way, each coarse pixel's
    density read at its 4 pixels, at its own best threshold.
The baselines get their best thresholds from the truth they are scored on,
an advantage the network does not have.  Its channel F1 must beat both, and
its median relative log-density error must beat the coarse trace's and
D8's (each read as a density in the truth's units)."""
import math

import numpy as np
import pytest

from baselines import best_f1, block_mean, d8_bounded_accumulation, f1

pytestmark = pytest.mark.gpu


def _to_grid(values, grid_origin, grid_shape, roi_shape, f=1):
    """Sample a (coarse) grid's per-pixel values at every pixel of a fine
    grid: fine grid (i, j) -> roi (i + ox, j + oy) -> coarse roi / f -> the
    coarse grid.  values: [nx_c, ny_c] over a grid whose origin (in coarse
    roi pixels) is grid_origin[1]; pixels that fall outside get 0."""
    (ox, oy), (cox, coy) = grid_origin
    nx, ny = grid_shape
    ii, jj = np.mgrid[:nx, :ny]
    ci, cj = (ii + ox) // f - cox, (jj + oy) // f - coy
    ok = (ci >= 0) & (cj >= 0) & (ci < values.shape[0]) & (cj < values.shape[1])
    out = np.zeros((nx, ny), np.float64)
    out[ok] = values[ci[ok], cj[ok]]
    return out


def _median_rel_err(pred_log, truth_log):
    return float(np.median(np.abs(pred_log - truth_log) / np.maximum(truth_log, 1e-6)))


def test_surrogate_beats_d8_and_coarse_trace(gpu):
    import torch
    from slm_amd import gated
    from slm_amd import surrogate as S
    print("training set 16 x 1024^2", flush=True)
    data = gated.training_set(n_dtms=16, n=1024, size=128, seed0=1000, device=gpu)
    const = gated.constant_loss(data)
    net, hist = gated.train(data, epochs=30, batch=32, lr=1e-3, device=gpu)
    cal = gated.calibrate(net, data)
    print("loss: constant %.3f, first epoch %.3f, last %.3f; calibrated threshold %.2f (train F1 %.3f)"
          % (const, hist["loss"][0], hist["loss"][-1], cal[0], cal[1]), flush=True)
    assert hist["loss"][-1] < 0.5 * const
    inf = S.UNet(net.params_numpy(), device=gpu, inference=True)
    theta = data.theta
    del data, net
    torch.cuda.empty_cache()

    roi = gated.family_dtm(1024, 90000)
    grid = gated.DeviceGrid(roi, device=gpu)
    truth, _ = gated.plain_map(grid, theta)
    pred, _ = gated.predict(inf, grid, size=256, halo=16, batch=16)
    torch.cuda.synchronize()
    shape = (grid.nx, grid.ny)
    v = (grid.mask.view(shape) == 0).cpu().numpy()
    t = truth.view(3, *shape).cpu().numpy().astype(np.float64)
    p = pred.view(3, *shape).cpu().numpy().astype(np.float64)
    tc = t[2][v] > 0
    res = {}
    res["network"] = (f1(p[2][v] > 0, tc), best_f1(p[2][v], tc)[0], _median_rel_err(p[1][v], t[1][v]))

    # bounded D8 accumulation on the same DTM, counted per pixel like the traced density
    d8 = d8_bounded_accumulation(roi, grid.max_length_px)
    d8g = _to_grid(d8, (grid.origin, (0, 0)), shape, roi.shape)
    res["d8"] = (best_f1(d8g[v], tc)[0], None, _median_rel_err(np.log1p(d8g[v]), t[1][v]))

    # a 2x coarser DTM traced the same way (2 m pixels, the same 100 m cap)
    coarse = gated.DeviceGrid(block_mean(roi, 2), pixel_size=2.0, device=gpu)
    slc_c, _, _ = coarse.trace_down()
    dens_c = (slc_c.view(coarse.nx, coarse.ny).cpu().numpy().astype(np.float64) / coarse.density ** 2)
    dens_c[coarse.mask.view(coarse.nx, coarse.ny).cpu().numpy() != 0] = 0.0
    # a coarse pixel is 2 m wide and stands for 4 m^2 of seeds: the flow across
    # a 1 m pixel is twice its count in 1 m units
    cg = 2.0 * _to_grid(dens_c, (grid.origin, coarse.origin), shape, roi.shape, f=2)
    res["coarse_trace_2x"] = (best_f1(cg[v], tc)[0], None, _median_rel_err(np.log1p(cg[v]), t[1][v]))

    for k, (f, f_best, err) in res.items():
        print("%-16s channel F1 %.3f%s  median rel. log-density error %.3f"
              % (k, f, "" if f_best is None else " (best threshold %.3f)" % f_best, err), flush=True)
    print("channel pixels in the truth: %d of %d" % (int(tc.sum()), int(v.sum())), flush=True)
    assert math.isfinite(res["network"][2])
    assert res["network"][0] > res["d8"][0] + 0.05
    assert res["network"][0] > res["coarse_trace_2x"][0] + 0.05
    assert res["network"][2] < min(res["d8"][2], res["coarse_trace_2x"][2])
