"""Trajectory packing on the device (slm_paths_offsets, slm_paths_pack), the
step that replaces the per-seed host slices of streamline_arrays_list
(trajectories.py:320-325): offsets against numpy's cumulative sum and the
packed steps against the concatenated row slices, over record counts around
the scan's block size (1024) and past one pass of its block bases (256
blocks), with empty records and counts above max_n_steps (clamped)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,max_steps", [(0, 5), (1, 5), (1023, 7), (1024, 3), (1025, 9), (5000, 40),
                                         (300000, 6)])
def test_paths_pack_matches_host_slices(gpu, n, max_steps):
    import torch
    from slm_amd import hip
    rs = np.random.default_rng(n)
    traj = rs.integers(-128, 128, (n, max_steps, 2)).astype(np.int8)
    nsteps = rs.integers(0, max_steps + 1, n).astype(np.uint16)
    if n > 2:
        nsteps[::7] = 0
        nsteps[1] = max_steps + 5                     # clamped to the row
    print("paths pack n=%d max_steps=%d" % (n, max_steps), flush=True)
    offsets, packed = hip.paths_pack(torch.from_numpy(traj).to(gpu),
                                     torch.from_numpy(nsteps.view(np.int16)).to(gpu), max_steps, sync=True)
    cnt = np.minimum(nsteps.astype(np.int64), max_steps)
    want_off = np.concatenate([[0], np.cumsum(cnt)])
    assert np.array_equal(offsets.cpu().numpy(), want_off)
    want = np.concatenate([traj[k, :cnt[k]] for k in range(n)]) if n else np.zeros((0, 2), np.int8)
    assert np.array_equal(packed.cpu().numpy(), want.reshape(-1, 2))


def test_paths_pack_rejects_mismatched_rows(gpu):
    import torch
    from slm_amd import hip
    with pytest.raises(ValueError):
        hip.paths_pack(torch.zeros((4, 5, 2), dtype=torch.int8, device=gpu),
                       torch.zeros(4, dtype=torch.int16, device=gpu), 6)
