"""GPU parity of the grids either side of the tracer.

slm_gradient_uv against the host numpy restatement of the reference's
preprocessing (slm_amd.preprocess, built on np.gradient and np.hypot): uv and
mask bit for bit.  slm_fields_finish_raw / _scaled against
oracle/postprocess_ref.py: slt and sla bit for bit; the float64 slc within 2
ulps (device pow against the host libm pow).  And the drop-in Trajectories +
Fields classes end to end against the oracle on the 256^2 workload."""
import numpy as np
import pytest

from slm_amd import preprocess, synth, workloads

pytestmark = pytest.mark.gpu


def _roi(kind):
    if kind == "small_256":
        return synth.make_dtm(256, pixel_size=1.0, relief=20.0, tilt=0.05, seed=0, nan_border=2), 1.0
    rs = np.random.default_rng(11)
    roi = (rs.standard_normal((97, 131)).cumsum(0).cumsum(1) * 0.01).astype(np.float32)
    roi[rs.random(roi.shape) < 0.01] = np.nan          # holes
    roi[40:44, 60:70] = 5.0                             # a flat patch: zero gradient -> speed 1
    return roi, 0.3


@pytest.mark.parametrize("kind,renorm,pad", [("small_256", True, 2), ("holes", True, 2), ("holes", False, 3)])
def test_gradient_uv_bit_exact(gpu, kind, renorm, pad):
    import torch
    from slm_amd import hip
    roi, ps = _roi(kind)
    uv_h, uvmask_h = preprocess.gradient_uv(roi, ps, pad, do_normalize_speed=renorm)
    mask_h = preprocess.dtm_mask(roi, pad) | uvmask_h
    uv_d, mask_d = hip.gradient_uv(torch.from_numpy(roi).to(gpu), pad, ps, do_normalize_speed=renorm, sync=True)
    assert np.array_equal(mask_d.cpu().numpy(), mask_h)
    assert np.array_equal(uv_d.cpu().numpy().view(np.uint32), uv_h.view(np.uint32))


@pytest.mark.parametrize("d", [1, 2, 3])
def test_fields_finish_bit_exact(gpu, d):
    import torch
    from oracle.postprocess_ref import postprocess_raw, postprocess_scaled
    from slm_amd import hip
    rs = np.random.default_rng(d)
    nxp, nyp = 173, 211
    slc = rs.integers(0, 50000, size=(2, nxp * nyp)).astype(np.uint32)
    slc[:, rs.random(nxp * nyp) < 0.2] = 0
    slt = (slc.astype(np.uint64) * rs.integers(1, 900, size=slc.shape)).clip(0, 2**32 - 1).astype(np.uint32)
    slt[:, :5] = [4294967295, 16777217, 3, 0, 1]                  # rounding edges of the float32 cast
    slc[:, :5] = [7, 3, 1, 0, 9]
    c_d = torch.from_numpy(slc.view(np.int32)).to(gpu)
    t_d = torch.from_numpy(slt.view(np.int32)).to(gpu)
    t_f, a_f = hip.fields_finish_raw(c_d, t_d, d)
    to_h = lambda a: a.reshape(2, nxp, nyp).transpose(1, 2, 0)
    c_r, t_r, a_r = postprocess_raw(to_h(slc), to_h(slt).astype(np.float32), d)
    assert np.array_equal(t_f.cpu().numpy().reshape(nxp, nyp, 2).view(np.uint32), t_r.view(np.uint32))
    assert np.array_equal(a_f.cpu().numpy().reshape(nxp, nyp, 2).view(np.uint32), a_r.view(np.uint32))
    dds, uds, ps = np.float32(0.4734), np.float32(0.5129), np.float32(2.0)
    sd, su = np.float32(dds / ps), np.float32(uds / ps)
    c_s = hip.fields_finish_scaled(c_d, t_f, d, sd, su, sync=True)
    c_sr, t_sr = postprocess_scaled(c_r, t_r, d, dds, uds, ps)
    assert np.array_equal(t_f.cpu().numpy().reshape(nxp, nyp, 2).view(np.uint32), t_sr.view(np.uint32))
    got = c_s.cpu().numpy().reshape(nxp, nyp, 2)
    ulps = np.abs(got.view(np.int64) - c_sr.view(np.int64))
    assert ulps.max() <= 2
    assert np.all(got[c_r == 0] == 0.0)


def test_trajectories_and_fields_dropin(gpu):
    """Trajectories.integrate then Fields.integrate (the classes trace.py
    drives) on the 256^2 workload, against the oracle's fields and the
    oracle post-processing with the same trajectory statistics."""
    import copy
    from oracle import oracle
    from oracle.postprocess_ref import postprocess_raw, postprocess_scaled
    from slm_amd.fields import Fields
    from slm_amd.params import kernel_params
    from slm_amd.trajectories import Trajectories
    from slm_amd.useful import Data
    job = workloads.small_256(jitter=1.0)
    nxp, nyp = job.mask.shape
    pad = int(job.info.pad_width)
    bbox = (pad, nxp - pad - 1, pad, nyp - pad - 1)
    info_t = copy.deepcopy(job.info)
    data_t = Data(info=info_t, bbox=bbox, pad=pad, mask_array=job.mask, uv_array=job.uv,
                  mapping_array=np.zeros(job.mask.shape, np.uint32))
    tr = Trajectories(0, 0, info=info_t, data=data_t)
    tr.integrate()
    stats = tr.data.traj_stats_df
    info_f = copy.deepcopy(job.info)
    data_f = Data(info=info_f, bbox=bbox, pad=pad, mask_array=job.mask, uv_array=job.uv,
                  mapping_array=np.zeros(job.mask.shape, np.uint32), traj_stats_df=stats)
    f = Fields(0, 0, info=info_f, data=data_f)
    f.integrate()
    print("steps", f.steps, flush=True)
    # oracle: same seeds, both directions, then the numpy post-processing
    seeds = f.data.seed_point_array
    raw_c, raw_t = [], []
    for sign in (1, -1):
        info_f.downup_sign = float(sign)
        p = kernel_params(info_f, sign, "INTEGRATE_FIELDS", False)
        c, t, _ = oracle.trace_fields(p, seeds, job.mask.reshape(-1).astype(np.uint8), job.uv.reshape(-1))
        raw_c.append(c.reshape(nxp, nyp))
        raw_t.append(t.reshape(nxp, nyp))
    rc = np.stack(raw_c, axis=2)
    rt = np.stack(raw_t, axis=2)
    assert np.array_equal(f.raw_slc, rc) and np.array_equal(f.raw_slt, rt)
    c_r, t_r, a_r = postprocess_raw(rc, rt.astype(np.float32), 1)
    dds = stats["ds"]["downstream", "mean"]
    uds = stats["ds"]["upstream", "mean"]
    c_s, t_s = postprocess_scaled(c_r, t_r, 1, dds, uds, info_f.pixel_size)
    assert np.array_equal(f.data.sla_array.view(np.uint32), a_r.view(np.uint32))
    assert np.array_equal(f.data.slt_array.view(np.uint32), t_s.view(np.uint32))
    assert np.abs(f.data.slc_array.view(np.int64) - c_s.view(np.int64)).max() <= 2
