"""BF16 convolutional surrogate of the streamline fields, trained on the GPU.

A 4-level U-Net (32 base channels) maps a DTM tile to log(1+slt),
log(1+slc) and a channel-mask logit, the fields the plain tracer produces
(downstream direction).  Every layer runs through the hand-written HIP
kernels of libslm_hip.so (slm_nn_*: MFMA implicit-GEMM convolutions forward,
backward-data and backward-weight, pooling, upsampling, loss, Adam); torch
only allocates device memory.  The surrogate has no counterpart in the
reference; the plain tracer generates its training data, judges its error and
replaces it on tiles where a check fails (see gated_map).

Layout: activations NHWC bf16; a concatenation is one buffer whose channel
slices the two branches write.  Weights bf16 [Cout][taps][Cin] with fp32
master copies, fp32 gradients and Adam moments in flat buffers.

  level  encoder                 decoder
  H      conv 16->32, 32->32     [up(64) | e0(32)] 96 -> 32, 32->32, head 1x1 32->32 (3 used)
  H/2    conv 32->64, 64->64     [up(128) | e1(64)] 192 -> 64, 64->64
  H/4    conv 64->128, 128->128  [up(256) | e2(128)] 384 -> 128, 128->128
  H/8    bottleneck 128->256, 256->256
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import hip

IN_CH = 16          # input channels (the MFMA K granule): N_FEATURES used, then zeros
N_FEATURES = 7      # h (normalised), u, v, valid, div(uv), Laplacian(h), slope
OUT_CH = 32         # head channels computed (0..2 used)
N_OUT = 3


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _s():
    return hip._stream(None)


# ------------------------------------------------------------------ layer ops

def conv(x, ldx, w, b, y, ldy, N, H, W, cin, cout, taps, relu):
    hip.check(hip.load().slm_nn_conv(_p(x), ldx, _p(w), _p(b) if b is not None else None, _p(y), ldy, N, H, W, cin,
                                     cout, taps, 1 if relu else 0, _s(), 0))


def conv_up(x, ldx, x2, ldx2, c_up, w, b, y, ldy, N, H, W, cin, cout, taps, relu):
    """conv over [up2(x2) channels 0..c_up | x channels c_up..cin]."""
    hip.check(hip.load().slm_nn_conv_up(_p(x), ldx, _p(x2), ldx2, c_up, _p(w), _p(b) if b is not None else None,
                                        _p(y), ldy, N, H, W, cin, cout, taps, 1 if relu else 0, _s(), 0))


def wflip(w, wt, taps, cin, cout):
    hip.check(hip.load().slm_nn_wflip(_p(w), _p(wt), taps, cin, cout, _s(), 0))


def wgrad(x, ldx, dy, lddy, dw, db, N, H, W, cin, cout, taps):
    hip.check(hip.load().slm_nn_conv_wgrad(_p(x), ldx, _p(dy), lddy, _p(dw), _p(db) if db is not None else None,
                                           N, H, W, cin, cout, taps, _s(), 0))


def relu_bwd(y, ldy, dy, lddy, M, C):
    hip.check(hip.load().slm_nn_relu_bwd(_p(y), ldy, _p(dy), lddy, M, C, _s(), 0))


def relu_bias_grad(y, ldy, dy, lddy, db, M, C, partials=None):
    """dy *= (y > 0) when y is given, then db += column sums of dy (per-block
    sums staged in partials, fp32 with at least C elements, when given)."""
    hip.check(hip.load().slm_nn_relu_bias_grad(_p(y) if y is not None else None, ldy, _p(dy), lddy,
                                               _p(db) if db is not None else None, M, C,
                                               _p(partials) if partials is not None else None,
                                               partials.numel() if partials is not None else 0, _s(), 0))


def pool(x, ldx, y, ldy, N, H, W, C):
    hip.check(hip.load().slm_nn_pool(_p(x), ldx, _p(y), ldy, N, H, W, C, _s(), 0))


def pool_bwd(x, ldx, y, ldy, dy, lddy, dx, lddx, N, H, W, C, accumulate=False):
    """accumulate: add the routed gradient to dx instead of writing it."""
    hip.check(hip.load().slm_nn_pool_bwd(_p(x), ldx, _p(y), ldy, _p(dy), lddy, _p(dx), lddx, N, H, W, C, _s(),
                                         hip.SLM_FLAG_NO_ZERO if accumulate else 0))


def up(x, ldx, y, ldy, N, H, W, C):
    hip.check(hip.load().slm_nn_up(_p(x), ldx, _p(y), ldy, N, H, W, C, _s(), 0))


def up_bwd(dy, lddy, dx, lddx, N, H, W, C):
    hip.check(hip.load().slm_nn_up_bwd(_p(dy), lddy, _p(dx), lddx, N, H, W, C, _s(), 0))


def add(a, lda, b, ldb, M, C):
    hip.check(hip.load().slm_nn_add(_p(a), lda, _p(b), ldb, M, C, _s(), 0))


# ------------------------------------------------------------------ the network

# (name, cin, cout, taps, relu)
LAYERS = [
    ("e0a", IN_CH, 32, 9, True), ("e0b", 32, 32, 9, True),
    ("e1a", 32, 64, 9, True), ("e1b", 64, 64, 9, True),
    ("e2a", 64, 128, 9, True), ("e2b", 128, 128, 9, True),
    ("bna", 128, 256, 9, True), ("bnb", 256, 256, 9, True),
    ("d2a", 384, 128, 9, True), ("d2b", 128, 128, 9, True),
    ("d1a", 192, 64, 9, True), ("d1b", 64, 64, 9, True),
    ("d0a", 96, 32, 9, True), ("d0b", 32, 32, 9, True),
    ("head", 32, OUT_CH, 1, False),
]


def init_params(seed=0):
    """He-normal weights (fp32 numpy), zero biases; the head's unused output
    channels are zero.  Returns {name: (w [cout, taps, cin], b [cout])}."""
    rs = np.random.default_rng(seed)
    out = {}
    for name, cin, cout, taps, _ in LAYERS:
        fan_in = cin * taps
        w = (rs.standard_normal((cout, taps, cin)) * math.sqrt(2.0 / fan_in)).astype(np.float32)
        if name == "e0a":
            w[:, :, N_FEATURES:] = 0.0     # the unused input channels carry nothing
        if name == "head":
            w[N_OUT:] = 0.0
            w *= 0.5
        out[name] = (w, np.zeros(cout, np.float32))
    return out


class UNet:
    """Parameters and activation buffers for batches of N tiles of H x W."""

    def __init__(self, params=None, device=None, seed=0, inference=False):
        """inference: allocate forward buffers only (no gradients or Adam state)."""
        import torch
        self.torch = torch
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        params = params if params is not None else init_params(seed)
        # flat fp32 master parameters: per layer w then b
        self.offsets = {}
        sizes = []
        off = 0
        for name, cin, cout, taps, _ in LAYERS:
            nw, nb = cout * taps * cin, cout
            self.offsets[name] = (off, nw, off + nw, nb)
            off += nw + nb
            sizes.append((nw, nb))
        self.n_params = off
        flat = np.zeros(off, np.float32)
        for name, *_ in LAYERS:
            w, b = params[name]
            o, nw, ob, nb = self.offsets[name]
            flat[o:o + nw] = w.reshape(-1)
            flat[ob:ob + nb] = b
        self.inference = inference
        self.p = torch.from_numpy(flat).to(self.dev)
        if not inference:
            self.g = torch.zeros_like(self.p)
            self.m = torch.zeros_like(self.p)
            self.v = torch.zeros_like(self.p)
            # per-block bias-gradient sums: 512 blocks x up to 256 channels
            self.partials = torch.empty(512 * 256, dtype=torch.float32, device=self.dev)
        self.pb = self.p.to(torch.bfloat16)          # bf16 copy (biases unused in bf16)
        self.wt = {name: torch.empty(cout * taps * cin, dtype=torch.bfloat16, device=self.dev)
                   for name, cin, cout, taps, _ in LAYERS}
        if not inference:
            # the backward-data weights: flipped here once, then by every Adam step
            self.flip = (hip.SlmNnFlip * len(LAYERS))()
            for k, (name, cin, cout, taps, _) in enumerate(LAYERS):
                wflip(self.w(name), self.wt[name], taps, cin, cout)
                self.flip[k].offset, self.flip[k].taps = self.offsets[name][0], taps
                self.flip[k].cin, self.flip[k].cout = cin, cout
                self.flip[k].wt = self.wt[name].data_ptr()
        self.step_count = 0
        self.shape = None

    # views
    def w(self, name):
        o, nw, _, _ = self.offsets[name]
        return self.pb[o:o + nw]

    def b(self, name):
        _, _, ob, nb = self.offsets[name]
        return self.p[ob:ob + nb]

    def gw(self, name):
        o, nw, _, _ = self.offsets[name]
        return self.g[o:o + nw]

    def gb(self, name):
        _, _, ob, nb = self.offsets[name]
        return self.g[ob:ob + nb]

    def params_numpy(self):
        flat = self.p.cpu().numpy()
        out = {}
        for name, cin, cout, taps, _ in LAYERS:
            o, nw, ob, nb = self.offsets[name]
            out[name] = (flat[o:o + nw].reshape(cout, taps, cin).copy(), flat[ob:ob + nb].copy())
        return out

    def _alloc(self, N, H, W):
        """Activation (and, for training, gradient) buffers for N tiles of
        H x W.  For inference, buffers made for more tiles of the same size
        serve: a batch is the leading N*H*W rows of each NHWC buffer."""
        if self.shape == (N, H, W):
            return
        if self.inference and self.shape is not None and self.shape[1:] == (H, W) and self.shape[0] >= N:
            return
        t, dev, bf = self.torch, self.dev, self.torch.bfloat16
        e = lambda n, h, w, c: t.empty((n * h * w, c), dtype=bf, device=dev)
        H2, W2, H4, W4, H8, W8 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        A = {}
        A["x"] = e(N, H, W, IN_CH)
        A["e0a"] = e(N, H, W, 32)
        A["cat0"] = e(N, H, W, 96)          # [up(d1) 64 | e0b 32]
        A["p0"] = e(N, H2, W2, 32)
        A["e1a"] = e(N, H2, W2, 64)
        A["cat1"] = e(N, H2, W2, 192)       # [up(d2) 128 | e1b 64]
        A["p1"] = e(N, H4, W4, 64)
        A["e2a"] = e(N, H4, W4, 128)
        A["cat2"] = e(N, H4, W4, 384)       # [up(bnb) 256 | e2b 128]
        A["p2"] = e(N, H8, W8, 128)
        A["bna"] = e(N, H8, W8, 256)
        A["bnb"] = e(N, H8, W8, 256)
        A["d2a"] = e(N, H4, W4, 128)
        A["d2b"] = e(N, H4, W4, 128)
        A["d1a"] = e(N, H2, W2, 64)
        A["d1b"] = e(N, H2, W2, 64)
        A["d0a"] = e(N, H, W, 32)
        A["d0b"] = e(N, H, W, 32)
        A["head"] = e(N, H, W, OUT_CH)
        self.A = A
        self.shape = (N, H, W)
        if self.inference:
            return
        # gradient buffers
        G = {k: t.empty_like(v) for k, v in A.items() if k not in ("x",)}
        self.G = G

    def forward(self, x_nhwc, N, H, W):
        """x_nhwc: bf16 [N*H*W, 16] (may be the input buffer self.A["x"] itself,
        as allocated for (N, H, W)).  Returns the head buffer bf16 [N*H*W, 32]."""
        self._alloc(N, H, W)
        A = self.A
        if x_nhwc.data_ptr() != A["x"].data_ptr():
            A["x"].copy_(x_nhwc)
        H2, W2, H4, W4, H8, W8 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        c = lambda name, xin, ldx, yout, ldy, n, h, w: conv(
            xin, ldx, self.w(name), self.b(name), yout, ldy, n, h, w, *self._io(name))
        cat0_e0 = A["cat0"][:, 64:]
        cat1_e1 = A["cat1"][:, 128:]
        cat2_e2 = A["cat2"][:, 256:]
        c("e0a", A["x"], IN_CH, A["e0a"], 32, N, H, W)
        c("e0b", A["e0a"], 32, cat0_e0, 96, N, H, W)
        pool(cat0_e0, 96, A["p0"], 32, N, H, W, 32)
        c("e1a", A["p0"], 32, A["e1a"], 64, N, H2, W2)
        c("e1b", A["e1a"], 64, cat1_e1, 192, N, H2, W2)
        pool(cat1_e1, 192, A["p1"], 64, N, H2, W2, 64)
        c("e2a", A["p1"], 64, A["e2a"], 128, N, H4, W4)
        c("e2b", A["e2a"], 128, cat2_e2, 384, N, H4, W4)
        pool(cat2_e2, 384, A["p2"], 128, N, H4, W4, 128)
        c("bna", A["p2"], 128, A["bna"], 256, N, H8, W8)
        c("bnb", A["bna"], 256, A["bnb"], 256, N, H8, W8)
        # decoder inputs [up2(coarse) | skip]: in inference the upsample is
        # fused into the conv's input staging (same values, no copy); training
        # keeps the concatenation, which the weight gradients read
        def cat_conv(name, coarse, ldc, cat, ldcat, yout, ldy, n, h, w):
            if self.inference:
                conv_up(cat, ldcat, coarse, ldc, ldc, self.w(name), self.b(name), yout, ldy, n, h, w,
                        *self._io(name))
            else:
                up(coarse, ldc, cat, ldcat, n, h, w, ldc)
                c(name, cat, ldcat, yout, ldy, n, h, w)
        cat_conv("d2a", A["bnb"], 256, A["cat2"], 384, A["d2a"], 128, N, H4, W4)
        c("d2b", A["d2a"], 128, A["d2b"], 128, N, H4, W4)
        cat_conv("d1a", A["d2b"], 128, A["cat1"], 192, A["d1a"], 64, N, H2, W2)
        c("d1b", A["d1a"], 64, A["d1b"], 64, N, H2, W2)
        cat_conv("d0a", A["d1b"], 64, A["cat0"], 96, A["d0a"], 32, N, H, W)
        c("d0b", A["d0a"], 32, A["d0b"], 32, N, H, W)
        c("head", A["d0b"], 32, A["head"], OUT_CH, N, H, W)
        return A["head"]

    def _io(self, name):
        for n, cin, cout, taps, relu in LAYERS:
            if n == name:
                return cin, cout, taps, relu
        raise KeyError(name)

    def _conv_bwd(self, name, xin, ldx, yout, ldy, dy, lddy, dx, lddx, n, h, w, need_dx=True):
        """Backward of one conv layer: dy (the gradient of its output; ReLU
        mask applied here) -> weight/bias grads (+=) and dx (written)."""
        cin, cout, taps, relu = self._io(name)
        M = n * h * w
        relu_bias_grad(yout if relu else None, ldy, dy, lddy, self.gb(name), M, cout, self.partials)
        wgrad(xin, ldx, dy, lddy, self.gw(name), None, n, h, w, cin, cout, taps)
        if need_dx:
            # self.wt[name]: the flipped weights, current since the last Adam step
            conv(dy, lddy, self.wt[name], None, dx, lddx, n, h, w, cout, cin, taps, False)

    def backward(self, dhead):
        """dhead: bf16 [N*H*W, 32] gradient of the loss wrt the head output.
        Accumulates parameter gradients into self.g (zeroed here first)."""
        N, H, W = self.shape
        A, G = self.A, self.G
        H2, W2, H4, W4, H8, W8 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        self.g.zero_()
        G["head"].copy_(dhead)
        cb = self._conv_bwd
        cb("head", A["d0b"], 32, A["head"], OUT_CH, G["head"], OUT_CH, G["d0b"], 32, N, H, W)
        cb("d0b", A["d0a"], 32, A["d0b"], 32, G["d0b"], 32, G["d0a"], 32, N, H, W)
        cb("d0a", A["cat0"], 96, A["d0a"], 32, G["d0a"], 32, G["cat0"], 96, N, H, W)
        up_bwd(G["cat0"], 96, G["d1b"], 64, N, H2, W2, 64)                 # dcat0[:, :64] -> d(d1b)
        cb("d1b", A["d1a"], 64, A["d1b"], 64, G["d1b"], 64, G["d1a"], 64, N, H2, W2)
        cb("d1a", A["cat1"], 192, A["d1a"], 64, G["d1a"], 64, G["cat1"], 192, N, H2, W2)
        up_bwd(G["cat1"], 192, G["d2b"], 128, N, H4, W4, 128)
        cb("d2b", A["d2a"], 128, A["d2b"], 128, G["d2b"], 128, G["d2a"], 128, N, H4, W4)
        cb("d2a", A["cat2"], 384, A["d2a"], 128, G["d2a"], 128, G["cat2"], 384, N, H4, W4)
        up_bwd(G["cat2"], 384, G["bnb"], 256, N, H8, W8, 256)
        cb("bnb", A["bna"], 256, A["bnb"], 256, G["bnb"], 256, G["bna"], 256, N, H8, W8)
        cb("bna", A["p2"], 128, A["bna"], 256, G["bna"], 256, G["p2"], 128, N, H8, W8)
        # e2b's output gradient = skip slice of dcat2 + pool backward
        cat2_e2, gcat2_e2 = A["cat2"][:, 256:], G["cat2"][:, 256:]
        pool_bwd(cat2_e2, 384, A["p2"], 128, G["p2"], 128, gcat2_e2, 384, N, H4, W4, 128, accumulate=True)
        cb("e2b", A["e2a"], 128, cat2_e2, 384, gcat2_e2, 384, G["e2a"], 128, N, H4, W4)
        cb("e2a", A["p1"], 64, A["e2a"], 128, G["e2a"], 128, G["p1"], 64, N, H4, W4)
        cat1_e1, gcat1_e1 = A["cat1"][:, 128:], G["cat1"][:, 128:]
        pool_bwd(cat1_e1, 192, A["p1"], 64, G["p1"], 64, gcat1_e1, 192, N, H2, W2, 64, accumulate=True)
        cb("e1b", A["e1a"], 64, cat1_e1, 192, gcat1_e1, 192, G["e1a"], 64, N, H2, W2)
        cb("e1a", A["p0"], 32, A["e1a"], 64, G["e1a"], 64, G["p0"], 32, N, H2, W2)
        cat0_e0, gcat0_e0 = A["cat0"][:, 64:], G["cat0"][:, 64:]
        pool_bwd(cat0_e0, 96, A["p0"], 32, G["p0"], 32, gcat0_e0, 96, N, H, W, 32, accumulate=True)
        cb("e0b", A["e0a"], 32, cat0_e0, 96, gcat0_e0, 96, G["e0a"], 32, N, H, W)
        cb("e0a", A["x"], IN_CH, A["e0a"], 32, G["e0a"], 32, None, 0, N, H, W, need_dx=False)

    def loss(self, target, valid, w_bce=1.0, n_valid=None, out=None, w_len=1.0):
        """Loss of the last forward against target float32 [N*H*W, 3] and
        valid uint8 [N*H*W]; returns (dhead, [sse0, sse1, bce] sums, n_valid).
        Given n_valid (the number of valid pixels) the call does not wait for
        the device; given out (float32 [3]) the sums are added to it."""
        t = self.torch
        N, H, W = self.shape
        M = N * H * W
        if n_valid is None:
            n_valid = int(valid.sum().item())
        n_valid = max(1, int(n_valid))
        if out is None:
            out = t.zeros(3, dtype=t.float32, device=self.dev)
        dhead = self.G["head"]
        hip.check(hip.load().slm_nn_loss(_p(self.A["head"]), OUT_CH, _p(target), _p(valid), _p(dhead), OUT_CH, M,
                                         ctypes.c_float(1.0 / n_valid), ctypes.c_float(w_len), ctypes.c_float(w_bce),
                                         _p(out), _s(), 0))
        return dhead, out, n_valid

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        hip.check(hip.load().slm_nn_adam_flip(_p(self.p), _p(self.g), _p(self.m), _p(self.v), _p(self.pb),
                                              self.n_params, ctypes.c_float(lr), ctypes.c_float(b1),
                                              ctypes.c_float(b2), ctypes.c_float(eps), self.step_count, self.flip,
                                              len(LAYERS), _s(), 0))

    def flops_per_pixel(self):
        """Forward multiply-adds x 2 per output pixel of the input grid."""
        f = 0.0
        scale = {"e0": 1, "d0": 1, "he": 1, "e1": 0.25, "d1": 0.25, "e2": 1 / 16, "d2": 1 / 16, "bn": 1 / 64}
        for name, cin, cout, taps, _ in LAYERS:
            f += 2.0 * cin * cout * taps * scale[name[:2]]
        return f


# ------------------------------------------------------------------ data

def input_channels(h, uv, mask, h_scale=10.0):
    """Network input for a stack of tiles: h [n, H, W] (metres, NaN allowed),
    uv [n, H, W, 2], mask [n, H, W] bool -> bf16-ready float32 [n, H, W, 16]:
    (h - tile mean)/h_scale, u, v, valid, zeros."""
    n, H, W = h.shape
    x = np.zeros((n, H, W, IN_CH), np.float32)
    valid = ~mask
    hh = np.where(valid, h, np.nan)
    mean = np.nanmean(hh.reshape(n, -1), axis=1)
    mean = np.where(np.isfinite(mean), mean, 0.0)
    x[..., 0] = np.where(valid, (h - mean[:, None, None]) / h_scale, 0.0)
    x[..., 1] = np.where(valid, uv[..., 0], 0.0)
    x[..., 2] = np.where(valid, uv[..., 1], 0.0)
    x[..., 3] = valid.astype(np.float32)
    return x


def targets(slt, slc, mask, channel_slc):
    """Training targets [.., 3]: log1p(slt), log1p(slc), slc >= channel_slc."""
    t = np.zeros(slt.shape + (3,), np.float32)
    t[..., 0] = np.log1p(np.maximum(slt, 0.0))
    t[..., 1] = np.log1p(np.maximum(slc, 0.0))
    t[..., 2] = (slc >= channel_slc).astype(np.float32)
    t[mask] = 0.0
    return t
