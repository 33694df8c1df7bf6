"""GPU tests of the surrogate's kernels and network (slm_nn_*).

Each kernel is compared with the same operation on the CPU in float32 on the
same bf16 values (products accumulate in fp32 on both sides, in different
orders, so bf16 outputs may differ in the last place); the whole network's
parameter gradients are compared with the CPU oracle
(oracle/surrogate_ref.py) on two 32x32 tiles; and a short training run must
lower the loss well below that of the best constant predictor."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ULP = 2.0 ** -7          # bf16 relative spacing (8 significant bits)


def _bf(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16)


def _close_bf16(got, want, frac=0.99):
    """bf16 results: at least frac exactly equal, all within 2 ulps (relative to
    the larger of |want| and a small floor)."""
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    scale = np.maximum(np.abs(want), 1e-3 * max(1e-6, float(np.abs(want).max())))
    rel = np.abs(got - want) / scale
    assert rel.max() <= 2 * ULP + 1e-7, rel.max()
    assert (got == want).mean() >= frac, (got == want).mean()


# (H, W) = (16, 24) runs the direct kernel; (32, 64), (16, 16) and (32, 32)
# the LDS-tiled kernel (32- and 16-column block tiles)
@pytest.mark.parametrize("taps,relu,cin,cout,ld_extra,H,W", [
    (9, True, 16, 32, 0, 16, 24), (9, False, 64, 64, 32, 16, 24), (1, False, 32, 32, 0, 16, 24),
    (9, True, 96, 128, 0, 16, 24), (9, True, 16, 32, 0, 32, 64), (9, False, 64, 64, 32, 32, 64),
    (9, True, 96, 32, 0, 32, 64), (9, True, 384, 128, 0, 16, 16), (9, False, 32, 64, 16, 16, 16),
    (9, True, 128, 256, 0, 32, 32)])
def test_conv_forward(gpu, taps, relu, cin, cout, ld_extra, H, W):
    import torch
    import torch.nn.functional as F
    from slm_amd import surrogate as S
    rs = np.random.default_rng(taps * 100 + cin)
    N = 2
    x = rs.standard_normal((N, H, W, cin + ld_extra)).astype(np.float32)
    w = (rs.standard_normal((cout, taps, cin)) / np.sqrt(cin * taps)).astype(np.float32)
    b = rs.standard_normal(cout).astype(np.float32) * 0.1
    xb, wb = _bf(x), _bf(w)
    xg = xb.to(gpu).reshape(N * H * W, cin + ld_extra)
    y = torch.empty((N * H * W, cout), dtype=torch.bfloat16, device=gpu)
    S.conv(xg, cin + ld_extra, wb.to(gpu).reshape(-1), torch.from_numpy(b).to(gpu), y, cout, N, H, W, cin, cout,
           taps, relu)
    torch.cuda.synchronize()
    k = 3 if taps == 9 else 1
    xr = xb.float()[..., :cin].permute(0, 3, 1, 2)
    wr = wb.float().reshape(cout, k, k, cin).permute(0, 3, 1, 2)
    ref = F.conv2d(xr, wr, torch.from_numpy(b), padding=1 if taps == 9 else 0)
    if relu:
        ref = F.relu(ref)
    ref = ref.to(torch.bfloat16).float().permute(0, 2, 3, 1).reshape(N * H * W, cout)
    _close_bf16(y.float().cpu().numpy(), ref.numpy())
    # the direct kernel computes the same products (summed in another order)
    from slm_amd import hip
    y2 = torch.empty_like(y)
    wg, bg = wb.to(gpu).reshape(-1), torch.from_numpy(b).to(gpu)       # held while the kernel runs
    hip.check(hip.load().slm_nn_conv(S._p(xg), cin + ld_extra, S._p(wg), S._p(bg), S._p(y2), cout, N, H, W, cin,
                                     cout, taps, 1 if relu else 0, S._s(), 1 | 16))
    _close_bf16(y2.float().cpu().numpy(), y.float().cpu().numpy(), frac=0.98)


# (N, H, W): 16-, 32-, 64- and 128-column images run the LDS-staged
# weight-gradient kernel (chunks of 256 / W rows, the last one partial when
# H is not a multiple of them); 48-column images the im2col-tile kernel
@pytest.mark.parametrize("taps,cin,cout,N,H,W", [
    (9, 16, 32, 2, 16, 16), (9, 64, 128, 2, 16, 16), (1, 32, 32, 2, 16, 16), (9, 96, 32, 2, 16, 16),
    (9, 384, 128, 2, 16, 16), (9, 128, 256, 2, 16, 16), (1, 32, 64, 2, 16, 16), (9, 16, 32, 2, 5, 128),
    (9, 32, 32, 1, 7, 128), (9, 64, 64, 2, 7, 64), (9, 96, 32, 3, 9, 32), (9, 192, 64, 1, 6, 64),
    (9, 16, 64, 2, 3, 64), (9, 32, 32, 2, 4, 48), (1, 32, 32, 1, 5, 128), (1, 64, 32, 2, 4, 48)])
def test_conv_backward(gpu, taps, cin, cout, N, H, W):
    """Backward-weight (fp32) and backward-data (flipped weights through the
    forward kernel) against torch autograd on the CPU."""
    import torch
    import torch.nn.functional as F
    from slm_amd import surrogate as S
    rs = np.random.default_rng(7 + cin + 1000 * W + 10 * H)
    x = _bf(rs.standard_normal((N, H, W, cin)))
    w = _bf(rs.standard_normal((cout, taps, cin)) / np.sqrt(cin * taps))
    dy = _bf(rs.standard_normal((N, H, W, cout)))
    k = 3 if taps == 9 else 1
    xr = x.float().permute(0, 3, 1, 2).requires_grad_(True)
    wr = w.float().reshape(cout, k, k, cin).permute(0, 3, 1, 2).contiguous().requires_grad_(True)
    yr = F.conv2d(xr, wr, padding=1 if taps == 9 else 0)
    yr.backward(dy.float().permute(0, 3, 1, 2))
    dw_ref = wr.grad.permute(0, 2, 3, 1).reshape(cout, taps, cin).numpy()
    dx_ref = xr.grad.to(torch.bfloat16).float().permute(0, 2, 3, 1).reshape(-1, cin).numpy()
    M = N * H * W
    xg, wg, dyg = x.to(gpu).reshape(M, cin), w.to(gpu).reshape(-1), dy.to(gpu).reshape(M, cout)
    dw = torch.zeros(cout * taps * cin, dtype=torch.float32, device=gpu)
    db = torch.zeros(cout, dtype=torch.float32, device=gpu)
    S.wgrad(xg, cin, dyg, cout, dw, db, N, H, W, cin, cout, taps)
    if cin % 32 == 0:          # backward-data has Cin as its output channels (the network skips it for Cin 16)
        wt = torch.empty_like(wg)
        S.wflip(wg, wt, taps, cin, cout)
        dx = torch.empty((M, cin), dtype=torch.bfloat16, device=gpu)
        S.conv(dyg, cout, wt, None, dx, cin, N, H, W, cout, cin, taps, False)
    torch.cuda.synchronize()
    got = dw.cpu().numpy().reshape(cout, taps, cin)
    assert np.abs(got - dw_ref).max() <= 1e-4 * np.abs(dw_ref).max() + 1e-5
    # the direct weight-gradient kernel (test hook) sums the same products
    from slm_amd import hip
    dw2 = torch.zeros_like(dw)
    hip.check(hip.load().slm_nn_conv_wgrad(S._p(xg), cin, S._p(dyg), cout, S._p(dw2), None, N, H, W, cin, cout, taps,
                                           S._s(), 1 | 16))
    assert np.abs(dw2.cpu().numpy().reshape(cout, taps, cin) - got).max() <= 1e-4 * np.abs(dw_ref).max() + 1e-5
    db_ref = dy.float().sum(dim=(0, 1, 2)).numpy()
    assert np.allclose(db.cpu().numpy(), db_ref, rtol=1e-5, atol=1e-4)
    if cin % 32 == 0:
        _close_bf16(dx.float().cpu().numpy(), dx_ref, frac=0.97)


@pytest.mark.parametrize("C,ld,M,staged", [(32, 32, 3 * 1024 + 32, False), (64, 96, 3 * 1024 + 32, False),
                                            (256, 256, 3 * 1024 + 32, False), (32, 32, 3 * 1024 + 32, True),
                                            (64, 96, 40000, True), (256, 256, 9000, True), (32, 96, 300000, True)])
def test_relu_bias_grad(gpu, C, ld, M, staged):
    """ReLU mask and column sums; staged: per-block sums in a partials buffer
    reduced by a second kernel (up to 512 blocks), else atomics."""
    import torch
    from slm_amd import surrogate as S
    rs = np.random.default_rng(C + M)
    y = _bf(rs.standard_normal((M, ld))).to(gpu)
    dy = _bf(rs.standard_normal((M, ld))).to(gpu)
    dy0 = dy.float().cpu().numpy()
    db = torch.full((C,), 0.5, dtype=torch.float32, device=gpu)         # sums are added to what db holds
    partials = torch.full((512 * 256,), float("nan"), dtype=torch.float32, device=gpu) if staged else None
    S.relu_bias_grad(y, ld, dy, ld, db, M, C, partials)
    torch.cuda.synchronize()
    yn = y.float().cpu().numpy()
    want = np.where(yn[:, :C] > 0, dy0[:, :C], 0.0)
    got = dy.float().cpu().numpy()
    assert np.array_equal(got[:, :C], want)
    assert np.array_equal(got[:, C:], dy0[:, C:])          # channels beyond C untouched
    tol = 1e-5 * np.abs(want).sum(0) + 1e-3
    assert np.all(np.abs(db.cpu().numpy() - (0.5 + want.astype(np.float64).sum(0))) <= tol)
    # bias only (no ReLU): dy unchanged
    db.zero_()
    S.relu_bias_grad(None, 0, dy, ld, db, M, C, partials)
    torch.cuda.synchronize()
    assert np.array_equal(dy.float().cpu().numpy(), got)
    assert np.all(np.abs(db.cpu().numpy() - got[:, :C].astype(np.float64).sum(0)) <= tol)


def test_pool_up_add_loss_adam(gpu):
    import torch
    import torch.nn.functional as F
    from slm_amd import hip
    from slm_amd import surrogate as S
    rs = np.random.default_rng(3)
    N, H, W, C = 2, 8, 12, 32
    x = _bf(rs.standard_normal((N, H, W, C)))
    xg = x.to(gpu).reshape(-1, C)
    y = torch.empty((N * H * W // 4, C), dtype=torch.bfloat16, device=gpu)
    S.pool(xg, C, y, C, N, H, W, C)
    ref = F.max_pool2d(x.float().permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1).reshape(-1, C)
    assert torch.equal(y.float().cpu(), ref)
    dy = _bf(rs.standard_normal((N * H * W // 4, C)))
    dx = torch.empty((N * H * W, C), dtype=torch.bfloat16, device=gpu)
    S.pool_bwd(xg, C, y, C, dy.to(gpu), C, dx, C, N, H, W, C)
    xr = x.float().permute(0, 3, 1, 2).requires_grad_(True)
    F.max_pool2d(xr, 2).backward(dy.float().reshape(N, H // 2, W // 2, C).permute(0, 3, 1, 2))
    assert torch.equal(dx.float().cpu(), xr.grad.permute(0, 2, 3, 1).reshape(-1, C))
    # upsample and its backward
    u = torch.empty((N * H * W, C), dtype=torch.bfloat16, device=gpu)
    S.up(y, C, u, C, N, H, W, C)
    uref = F.interpolate(ref.reshape(N, H // 2, W // 2, C).permute(0, 3, 1, 2), scale_factor=2, mode="nearest")
    assert torch.equal(u.float().cpu(), uref.permute(0, 2, 3, 1).reshape(-1, C))
    du = _bf(rs.standard_normal((N * H * W, C)))
    dd = torch.empty((N * H * W // 4, C), dtype=torch.bfloat16, device=gpu)
    S.up_bwd(du.to(gpu), C, dd, C, N, H // 2, W // 2, C)
    dref = du.float().reshape(N, H // 2, 2, W // 2, 2, C).sum(dim=(2, 4)).reshape(-1, C)
    _close_bf16(dd.float().cpu().numpy(), dref.to(torch.bfloat16).float().numpy())
    # loss (sums and gradient) against numpy
    M = N * H * W
    head = _bf(rs.standard_normal((M, 32))).to(gpu)
    tgt = rs.random((M, 3)).astype(np.float32)
    tgt[:, 2] = (tgt[:, 2] > 0.7).astype(np.float32)
    valid = (rs.random(M) > 0.2).astype(np.uint8)
    dh = torch.empty_like(head)
    out = torch.zeros(3, dtype=torch.float32, device=gpu)
    tg, vg = torch.from_numpy(tgt).to(gpu), torch.from_numpy(valid).to(gpu)    # held while the kernel runs
    hip.check(hip.load().slm_nn_loss(S._p(head), 32, S._p(tg), S._p(vg), S._p(dh), 32, M,
                                     S.ctypes.c_float(1.0 / valid.sum()), S.ctypes.c_float(0.5),
                                     S.ctypes.c_float(2.0), S._p(out), S._s(), 1))
    h = head.float().cpu().numpy()
    v = valid.astype(bool)
    z = h[:, 2].astype(np.float64)
    bce = np.maximum(z, 0) - z * tgt[:, 2] + np.log1p(np.exp(-np.abs(z)))
    want = [((h[:, 0] - tgt[:, 0]) ** 2)[v].sum(), ((h[:, 1] - tgt[:, 1]) ** 2)[v].sum(), bce[v].sum()]
    assert np.allclose(out.cpu().numpy(), want, rtol=1e-4)
    g2 = 2.0 * (1 / (1 + np.exp(-z)) - tgt[:, 2]) / v.sum()
    assert np.allclose(dh.float().cpu().numpy()[v, 2], g2[v], rtol=2 * ULP, atol=1e-6)
    g0 = 0.5 * 2.0 * (h[:, 0] - tgt[:, 0]) / v.sum()
    assert np.allclose(dh.float().cpu().numpy()[v, 0], g0[v], rtol=2 * ULP, atol=1e-6)
    assert np.all(dh.float().cpu().numpy()[~v] == 0)
    # Adam step against the closed form
    n = 1000
    p = torch.from_numpy(rs.standard_normal(n).astype(np.float32)).to(gpu)
    g = torch.from_numpy(rs.standard_normal(n).astype(np.float32)).to(gpu)
    m0 = torch.zeros_like(p)
    v0 = torch.zeros_like(p)
    p0 = p.cpu().numpy().copy()
    hip.check(hip.load().slm_nn_adam(S._p(p), S._p(g), S._p(m0), S._p(v0), None, n, S.ctypes.c_float(1e-3),
                                     S.ctypes.c_float(0.9), S.ctypes.c_float(0.999), S.ctypes.c_float(1e-8), 1,
                                     S._s(), 1))
    gg = g.cpu().numpy()
    want_p = p0 - 1e-3 * gg / (np.abs(gg) + 1e-8)      # first step: m/c1 = g, v/c2 = g^2
    assert np.allclose(p.cpu().numpy(), want_p, rtol=1e-5, atol=1e-7)


def test_network_gradients_match_oracle(gpu):
    import torch
    from oracle.surrogate_ref import RefUNet
    from slm_amd import surrogate as S
    rs = np.random.default_rng(11)
    N, H, W = 2, 32, 32
    params = S.init_params(seed=5)
    # non-zero biases so every path is exercised
    params = {k: (w, (rs.standard_normal(b.shape) * 0.05).astype(np.float32)) for k, (w, b) in params.items()}
    x = np.zeros((N, H, W, S.IN_CH), np.float32)
    x[..., :4] = rs.standard_normal((N, H, W, 4))
    tgt = np.stack([rs.random((N, H, W)) * 3, rs.random((N, H, W)) * 3, (rs.random((N, H, W)) > 0.8) * 1.0],
                   -1).astype(np.float32)
    valid = rs.random((N, H, W)) > 0.1
    net = S.UNet(params, device=gpu)
    xb = _bf(x)
    head = net.forward(xb.to(gpu).reshape(-1, S.IN_CH), N, H, W)
    dhead, sums, nv = net.loss(torch.from_numpy(tgt.reshape(-1, 3)).to(gpu),
                               torch.from_numpy(valid.reshape(-1).astype(np.uint8)).to(gpu), w_bce=3.0, w_len=0.25)
    net.backward(dhead)
    torch.cuda.synchronize()
    ref = RefUNet(params, S.LAYERS)
    rh = ref.forward(xb.float().permute(0, 3, 1, 2))
    loss, parts = ref.loss(rh, tgt, valid, w_bce=3.0, w_len=0.25)
    loss.backward()
    gh = head.float().cpu().numpy().reshape(N, H, W, S.OUT_CH)[..., :3]
    rhn = rh.detach().permute(0, 2, 3, 1).numpy()[..., :3]
    assert np.abs(gh - rhn).max() <= 0.05 * np.abs(rhn).max()
    assert np.allclose(sums.cpu().numpy(), [float(t.detach()) for t in parts], rtol=0.02)
    g_ref = ref.grads()
    gflat = net.g.cpu().numpy()
    for name, cin, cout, taps, _ in S.LAYERS:
        o, nw, ob, nb = net.offsets[name]
        gw = gflat[o:o + nw].reshape(cout, taps, cin)
        rw = g_ref[name][0]
        cos = float((gw * rw).sum() / (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-30))
        rel = float(np.linalg.norm(gw - rw) / (np.linalg.norm(rw) + 1e-30))
        print(name, "cos %.5f rel %.4f" % (cos, rel), flush=True)
        assert cos > 0.995 and rel < 0.1, name
        gb, rb = gflat[ob:ob + nb], g_ref[name][1]
        assert np.linalg.norm(gb - rb) <= 0.1 * np.linalg.norm(rb) + 1e-6, name


@pytest.mark.parametrize("N,H,W", [(2, 64, 64), (1, 256, 128), (4, 32, 32)])
def test_inference_forward_fuses_upsample_bitwise(gpu, N, H, W):
    """The inference network (upsample fused into the decoder convs) gives the
    bits of the training network (explicit upsample into the concatenation)."""
    import torch
    from slm_amd import surrogate as S
    rs = np.random.default_rng(N * H)
    params = S.init_params(seed=9)
    params = {k: (w, (rs.standard_normal(b.shape) * 0.05).astype(np.float32)) for k, (w, b) in params.items()}
    x = _bf(rs.standard_normal((N * H * W, S.IN_CH))).to(gpu)
    a = S.UNet(params, device=gpu).forward(x, N, H, W).clone()
    b = S.UNet(params, device=gpu, inference=True).forward(x, N, H, W)
    assert torch.equal(a, b)


def test_training_beats_constant_predictor(gpu):
    """A few hundred Adam steps on tiles whose targets are a known function of
    the input (smoothed height and its threshold): the loss falls and the
    fitted map beats the per-channel constant that minimises the loss."""
    import torch
    from scipy import ndimage
    from slm_amd import surrogate as S
    rs = np.random.default_rng(4)
    n, H, W = 32, 32, 32
    h = ndimage.gaussian_filter(rs.standard_normal((n, H, W)), (0, 3, 3)).astype(np.float32) * 5
    x = np.zeros((n, H, W, S.IN_CH), np.float32)
    x[..., 0] = h
    x[..., 3] = 1.0
    sm = ndimage.gaussian_filter(h, (0, 2, 2))
    tgt = np.stack([np.abs(sm), np.maximum(sm, 0) * 2, (sm > 0.5) * 1.0], -1).astype(np.float32)
    valid = np.ones((n, H, W), bool)
    net = S.UNet(S.init_params(seed=1), device=gpu)
    xs = _bf(x).to(gpu).reshape(-1, S.IN_CH)
    ts = torch.from_numpy(tgt.reshape(-1, 3)).to(gpu)
    vs = torch.from_numpy(valid.reshape(-1).astype(np.uint8)).to(gpu)
    losses = []
    for it in range(300):
        net.forward(xs, n, H, W)
        dh, sums, nv = net.loss(ts, vs)
        net.backward(dh)
        net.adam(lr=2e-3)
        losses.append(float(sums.sum().item()) / nv)
    # the constant predictor: channel means (MSE) and the prevalence logit (BCE)
    mu = tgt.reshape(-1, 3).mean(0)
    pc = min(max(mu[2], 1e-6), 1 - 1e-6)
    const = ((tgt[..., 0] - mu[0]) ** 2).mean() + ((tgt[..., 1] - mu[1]) ** 2).mean() + \
        -(pc * np.log(pc) + (1 - pc) * np.log(1 - pc))
    print("loss first %.4f last %.4f constant %.4f" % (losses[0], losses[-1], const), flush=True)
    assert losses[-1] < 0.5 * losses[0]
    assert losses[-1] < 0.5 * const


def test_pool_bwd_accumulate_and_adam_flip(gpu):
    """The two fused training ops: the max-pool backward added into a
    gradient buffer (the skip connection's) equals the written backward
    added to it; Adam with a flip table updates the parameters as
    slm_nn_adam does and leaves each layer's flipped weights equal to
    slm_nn_wflip of the new bf16 copy."""
    import torch
    from slm_amd import hip
    from slm_amd import surrogate as S
    rs = np.random.default_rng(7)
    N, H, W, C = 2, 8, 12, 32
    xg = _bf(rs.standard_normal((N * H * W, C))).to(gpu)
    y = torch.empty((N * H * W // 4, C), dtype=torch.bfloat16, device=gpu)
    S.pool(xg, C, y, C, N, H, W, C)
    dy = _bf(rs.standard_normal((N * H * W // 4, C))).to(gpu)
    dx = torch.empty((N * H * W, C), dtype=torch.bfloat16, device=gpu)
    S.pool_bwd(xg, C, y, C, dy, C, dx, C, N, H, W, C)
    base = _bf(rs.standard_normal((N * H * W, 2 * C)))
    base[:3] = 0.0
    base[0, C:] = -0.0
    acc = base.to(gpu)
    S.pool_bwd(xg, C, y, C, dy, C, acc[:, C:], 2 * C, N, H, W, C, accumulate=True)
    want = base.clone()
    want[:, C:] = (base[:, C:].float() + dx.float().cpu()).to(torch.bfloat16)
    assert torch.equal(acc.cpu().view(torch.int16), want.view(torch.int16))
    # Adam with flips: two layers (3x3 and 1x1) and a bias gap between them
    shapes = [(9, 16, 32), (1, 32, 8)]
    offs, n = [], 0
    for taps, cin, cout in shapes:
        offs.append(n)
        n += cout * taps * cin + cout
    p0 = rs.standard_normal(n).astype(np.float32)
    g0 = rs.standard_normal(n).astype(np.float32)
    res = []
    for flip in (False, True):
        p = torch.from_numpy(p0.copy()).to(gpu)
        g = torch.from_numpy(g0).to(gpu)
        m, v = torch.zeros_like(p), torch.zeros_like(p)
        pb = torch.empty(n, dtype=torch.bfloat16, device=gpu)
        wts = [torch.empty(cout * taps * cin, dtype=torch.bfloat16, device=gpu) for taps, cin, cout in shapes]
        args = (S._p(p), S._p(g), S._p(m), S._p(v), S._p(pb), n, S.ctypes.c_float(1e-2), S.ctypes.c_float(0.9),
                S.ctypes.c_float(0.999), S.ctypes.c_float(1e-8), 3)
        if flip:
            tab = (hip.SlmNnFlip * 2)()
            for k, ((taps, cin, cout), o) in enumerate(zip(shapes, offs)):
                tab[k].offset, tab[k].taps, tab[k].cin, tab[k].cout = o, taps, cin, cout
                tab[k].wt = wts[k].data_ptr()
            hip.check(hip.load().slm_nn_adam_flip(*args, tab, 2, S._s(), 1))
        else:
            hip.check(hip.load().slm_nn_adam(*args, S._s(), 1))
        res.append((p.cpu(), pb.cpu(), pb, wts))
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1].view(torch.int16), res[1][1].view(torch.int16))
    pb = res[1][2]
    for (taps, cin, cout), o, wt in zip(shapes, offs, res[1][3]):
        ref = torch.empty_like(wt)
        S.wflip(pb[o:o + cout * taps * cin], ref, taps, cin, cout)
        torch.cuda.synchronize()
        assert torch.equal(ref.cpu().view(torch.int16), wt.cpu().view(torch.int16))
