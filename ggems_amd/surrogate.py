"""Learned scatter surrogate: a U-Net that maps a low-count Monte Carlo
projection and the ray-traced primary image to the scatter image.

Its 3x3 convolutions run on the matrix cores through the engine library
(ggh_conv3x3_forward / ggh_conv3x3_weight_grad: BF16 operands, FP32
accumulation); PyTorch holds the tensors, streams and optimiser state.
Activations are NHWC BF16; master weights are FP32 [cout][9][cin] and are
cast to BF16 for every call.

Architecture (4 levels, 16-32-64-128 channels): two 3x3 conv + ReLU per
level, 2x2 max-pooling down, nearest-neighbour 2x upsampling, concatenation
with the skip, two 3x3 conv + ReLU per decoder level, and a 1x1 head.

The training pairs come from the engine itself: for a seeded phantom and
gantry angle, a projection of n_low primaries (its counts and its scatter
counts, each over the open-field counts) and the ray-traced transmission are
the inputs; the scatter image of n_high primaries, per primary and per unit
open-field count, is the target.  The scatter image is the reference's
scatter histogram (src/kernels/TrackThroughGGEMSSolidBox.cl:215-225), which
the CT system stores beside the counts.
"""

import ctypes as C
import math

import numpy as np
import torch

from .engine import EngineError, library

_SIGS = {
    "ggh_unet_last_error": ([], C.c_char_p),
    "ggh_conv3x3_forward": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 6 + [C.c_void_p],
                            C.c_int),
    "ggh_conv3x3_weight_grad": ([C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 6 + [C.c_void_p], C.c_int),
    "ggh_relu_bias_backward": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                C.c_int32, C.c_void_p], C.c_int),
    "ggh_conv3x3_weights_bf16": ([C.c_void_p] * 5 + [C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "ggh_pool2x2_forward": ([C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 4 + [C.c_void_p], C.c_int),
    "ggh_pool2x2_backward": ([C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 4 + [C.c_void_p], C.c_int),
    "ggh_upcat_forward": ([C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 5 + [C.c_void_p], C.c_int),
    "ggh_upcat_backward": ([C.c_void_p, C.c_void_p, C.c_void_p] + [C.c_int32] * 5 + [C.c_void_p], C.c_int),
}


def _lib():
    L = library()
    if not getattr(L, "_unet_ready", False):
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        L._unet_ready = True
    return L


def _check(rc):
    if rc < 0:
        raise EngineError(_lib().ggh_unet_last_error().decode(errors="replace") + " (code %d)" % rc)


def _stream(t):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def conv3x3(x, w, bias=None, relu=False):
    """x [N,H,W,Cin] BF16, w [Cout, 9*Cin] BF16, bias [Cout] FP32 -> [N,H,W,Cout] BF16"""
    if x.device.type != "cuda" or x.dtype != torch.bfloat16 or w.dtype != torch.bfloat16:
        raise EngineError("conv3x3 takes BF16 tensors on the GPU")
    x, w = x.contiguous(), w.contiguous()
    n, h, wd, cin = x.shape
    cout = w.shape[0]
    if w.shape[1] != 9 * cin:
        raise EngineError("weights must be [cout, 9*cin]")
    y = torch.empty((n, h, wd, cout), dtype=torch.bfloat16, device=x.device)
    b = None if bias is None else bias.contiguous().float()
    _check(_lib().ggh_conv3x3_forward(x.data_ptr(), w.data_ptr(), None if b is None else b.data_ptr(), y.data_ptr(),
                                      n, h, wd, cin, cout, int(bool(relu)), _stream(x)))
    return y


def conv3x3_weight_grad(x, dy, splits=None, out=None):
    """sum over pixels of dy (x) the shifted x: [Cout, 9*Cin] FP32, added to
    `out` when given (zeroed by the caller), else to a new zero tensor"""
    x, dy = x.contiguous(), dy.contiguous()
    n, h, wd, cin = x.shape
    cout = dy.shape[3]
    dw = torch.zeros((cout, 9 * cin), dtype=torch.float32, device=x.device) if out is None else out
    if splits is None:
        # pixel splits: about 1,024 pixels each, and at least 1,024 blocks in
        # all (cols x splits) so that the small deep layers fill the chip
        cols = (9 * cin + 63) // 64
        p = n * h * wd
        splits = max(1, min(p // 32, max(p // 1024, -(-1024 // cols))))
    _check(_lib().ggh_conv3x3_weight_grad(x.data_ptr(), dy.data_ptr(), dw.data_ptr(), n, h, wd, cin, cout,
                                          int(splits), _stream(x)))
    return dw


def relu_bias_backward(dy, y, relu, db=None):
    """(dy masked where the layer's output y <= 0 when relu, as BF16; the bias
    gradient, the sum of that over the pixels in FP32, added to `db` when
    given (zeroed by the caller)) in one kernel"""
    if dy.dtype not in (torch.bfloat16, torch.float32):
        dy = dy.float()
    dy = dy.contiguous()
    c = dy.shape[-1]
    yc = y.contiguous() if relu else None
    out = torch.empty(dy.shape, dtype=torch.bfloat16, device=dy.device)
    db = torch.zeros(c, dtype=torch.float32, device=dy.device) if db is None else db
    _check(_lib().ggh_relu_bias_backward(dy.data_ptr(), int(dy.dtype == torch.float32),
                                         yc.data_ptr() if relu else None, int(bool(relu)),
                                         out.data_ptr(), db.data_ptr(), dy.numel() // c, c, _stream(dy)))
    return out, db


def flip_transpose(w):
    """weights of the data gradient: [cout][9][cin] -> [cin][9][cout], taps reversed"""
    cout, _, cin = w.shape
    return w.flip(1).permute(2, 1, 0).contiguous()


def weights_bf16(w, zero=None):
    """(w as BF16 [cout][9][cin], flip_transpose of it [cin][9][cout]) from the
    FP32 weights in one kernel, which also sets the FP32 buffers `zero` =
    (dw [cout, 9 cin], db [cout]) to zero when given"""
    cout, _, cin = w.shape
    w = w.detach().contiguous().float()
    wb = torch.empty((cout, 9, cin), dtype=torch.bfloat16, device=w.device)
    wt = torch.empty((cin, 9, cout), dtype=torch.bfloat16, device=w.device)
    dw, db = zero if zero is not None else (None, None)
    _check(_lib().ggh_conv3x3_weights_bf16(w.data_ptr(), wb.data_ptr(), wt.data_ptr(),
                                           None if dw is None else dw.data_ptr(), None if db is None else db.data_ptr(),
                                           cout, cin, _stream(w)))
    return wb, wt


class _Conv3x3(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b, relu):
        cout, _, cin = w.shape
        # the gradient accumulators, zeroed by the weight preparation's pass
        zero = None
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            zero = (torch.empty((cout, 9 * cin), dtype=torch.float32, device=w.device),
                    torch.empty(cout, dtype=torch.float32, device=w.device))
        wb, wt = weights_bf16(w, zero)
        y = conv3x3(x, wb.view(cout, 9 * cin), b, relu)
        ctx.save_for_backward(x, wt, y)
        ctx.relu = relu
        ctx.shape = (cout, cin)
        ctx.zero = zero
        return y

    @staticmethod
    def backward(ctx, dy):
        x, wt, y = ctx.saved_tensors
        cout, cin = ctx.shape
        dw_buf, db_buf = ctx.zero if ctx.zero is not None else (None, None)
        ctx.zero = None  # a second reference would make autograd copy the gradients instead of taking them
        dy, db = relu_bias_backward(dy, y, ctx.relu, db=db_buf)
        dx = None
        if ctx.needs_input_grad[0]:
            dx = conv3x3(dy, wt.view(cin, 9 * cout))
        dw = conv3x3_weight_grad(x, dy, out=dw_buf).view(cout, 9, cin)
        return dx, dw, db, None


class Conv3x3(torch.nn.Module):
    def __init__(self, cin, cout, relu=True, generator=None):
        super().__init__()
        std = math.sqrt(2.0 / (9 * cin))
        self.weight = torch.nn.Parameter(torch.randn(cout, 9, cin, generator=generator) * std)
        self.bias = torch.nn.Parameter(torch.zeros(cout))
        self.relu = relu

    def forward(self, x):
        return _Conv3x3.apply(x, self.weight, self.bias, self.relu)


def _bf16(t):
    return t.to(torch.bfloat16).contiguous()


def pool2x2(x):
    """2 x 2 max pooling of NHWC BF16: (y [N,H/2,W/2,C], the window index of
    each output's first maximum, uint8)"""
    x = _bf16(x)
    n, h, w, c = x.shape
    y = torch.empty((n, h // 2, w // 2, c), dtype=torch.bfloat16, device=x.device)
    arg = torch.empty((n, h // 2, w // 2, c), dtype=torch.uint8, device=x.device)
    _check(_lib().ggh_pool2x2_forward(x.data_ptr(), y.data_ptr(), arg.data_ptr(), n, h, w, c, _stream(x)))
    return y, arg


def pool2x2_backward(dy, arg):
    """the gradient at each window's first maximum, zero elsewhere"""
    dy = _bf16(dy)
    n, ho, wo, c = dy.shape
    dx = torch.empty((n, 2 * ho, 2 * wo, c), dtype=torch.bfloat16, device=dy.device)
    _check(_lib().ggh_pool2x2_backward(dy.data_ptr(), arg.data_ptr(), dx.data_ptr(), n, 2 * ho, 2 * wo, c,
                                       _stream(dy)))
    return dx


def upcat(z, s):
    """(z upsampled 2x by repetition, s) concatenated over channels, NHWC BF16"""
    z, s = _bf16(z), _bf16(s)
    n, h, w, cs = s.shape
    cz = z.shape[3]
    out = torch.empty((n, h, w, cz + cs), dtype=torch.bfloat16, device=s.device)
    _check(_lib().ggh_upcat_forward(z.data_ptr(), s.data_ptr(), out.data_ptr(), n, h, w, cz, cs, _stream(s)))
    return out


def upcat_backward(dout, cz):
    """(dz: the z channels of dout summed over each 2 x 2 block, ds: its s channels)"""
    dout = _bf16(dout)
    n, h, w, ct = dout.shape
    dz = torch.empty((n, h // 2, w // 2, cz), dtype=torch.bfloat16, device=dout.device)
    ds = torch.empty((n, h, w, ct - cz), dtype=torch.bfloat16, device=dout.device)
    _check(_lib().ggh_upcat_backward(dout.data_ptr(), dz.data_ptr(), ds.data_ptr(), n, h, w, cz, ct - cz,
                                     _stream(dout)))
    return dz, ds


class _Pool2x2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        y, arg = pool2x2(x)
        ctx.save_for_backward(arg)
        return y

    @staticmethod
    def backward(ctx, dy):
        (arg,) = ctx.saved_tensors
        return pool2x2_backward(dy, arg)


class _UpCat(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, s):
        ctx.cz = z.shape[3]
        return upcat(z, s)

    @staticmethod
    def backward(ctx, dout):
        return upcat_backward(dout, ctx.cz)


IN_CHANNELS = 8  # the two inputs, zero-padded to the kernels' channel multiple


class UNet(torch.nn.Module):
    """4-level U-Net, 16-32-64-128 channels, on NHWC BF16 tensors"""

    def __init__(self, channels=(16, 32, 64, 128), seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        c = channels
        self.enc = torch.nn.ModuleList()
        cin = IN_CHANNELS
        for ch in c:
            self.enc.append(torch.nn.ModuleList([Conv3x3(cin, ch, generator=g), Conv3x3(ch, ch, generator=g)]))
            cin = ch
        self.dec = torch.nn.ModuleList()
        for skip in reversed(c[:-1]):
            self.dec.append(torch.nn.ModuleList([Conv3x3(cin + skip, skip, generator=g), Conv3x3(skip, skip,
                                                                                                   generator=g)]))
            cin = skip
        self.head = torch.nn.Parameter(torch.randn(c[0], generator=g) * math.sqrt(1.0 / c[0]))
        self.head_bias = torch.nn.Parameter(torch.zeros(1))

    def forward(self, x):
        """x [N, H, W, k <= 8] (any float type) -> [N, H, W] FP32; H, W multiples of 32"""
        n, h, w, k = x.shape
        z = torch.zeros((n, h, w, IN_CHANNELS), dtype=torch.bfloat16, device=x.device)
        z[..., :k] = x.to(torch.bfloat16)
        skips = []
        for i, (a, b) in enumerate(self.enc):
            z = b(a(z))
            if i + 1 < len(self.enc):
                skips.append(z)
                z = _Pool2x2.apply(z)
        for (a, b), s in zip(self.dec, reversed(skips)):
            z = b(a(_UpCat.apply(z, s)))
        # the 16 -> 1 head as a multiply and a channel sum: a matrix-vector
        # product with 16 columns runs far below HBM speed in the BLAS path
        return (z.float() * self.head).sum(dim=3) + self.head_bias


def reference_forward(net, x):
    """The same network with torch's own convolutions (NCHW) in the dtype
    and on the device of its parameters (FP32 or FP64 on the CPU for the
    gradient checks): x [N, H, W, 2] -> [N, H, W]."""
    import torch.nn.functional as Fn

    def conv(mod, z):
        cout, _, cin = mod.weight.shape
        w = mod.weight.view(cout, 3, 3, cin).permute(0, 3, 1, 2)
        y = Fn.conv2d(z, w, mod.bias, padding=1)
        return torch.relu(y) if mod.relu else y

    n, h, w, k = x.shape
    z = torch.zeros((n, IN_CHANNELS, h, w), dtype=net.head.dtype, device=net.head.device)
    z[:, :k] = x.permute(0, 3, 1, 2).to(net.head.dtype)
    skips = []
    for i, (a, b) in enumerate(net.enc):
        z = conv(b, conv(a, z))
        if i + 1 < len(net.enc):
            skips.append(z)
            z = Fn.max_pool2d(z, 2)
    for (a, b), s in zip(net.dec, reversed(skips)):
        z = conv(b, conv(a, torch.cat([Fn.interpolate(z, scale_factor=2, mode="nearest"), s], dim=1)))
    return torch.einsum("nchw,c->nhw", z, net.head) + net.head_bias


# ---------------------------------------------------------------------------
# training data from the engine, training, evaluation
# ---------------------------------------------------------------------------
def _view_engine(keep_detector_solid, seed, detector):
    from .ggems import GGEMS
    g = GGEMS()
    g.initialize(seed)
    eng = g.engine()
    eng.box_elements[keep_detector_solid] = (detector, detector, True)
    return g, eng


def open_field_rate(workdir, n, seed=1, detector=256):
    """Counts per primary of every pixel with no phantom (the normaliser)."""
    from . import configs
    keep = configs.scatter_scene(workdir, None, detector=detector)
    g, eng = _view_engine(0, seed, detector)
    eng.run(n)
    counts, _ = eng.histogram(0)
    g.delete()
    del keep
    return counts.astype(np.float64) / n


def dataset_batched(workdir, phantom_seeds, angles_per_phantom, seed, n_low, n_high, rate, detector=256,
                    shape=(64, 96, 96), log=None):
    """Views of several seeded phantoms at seeded random gantry angles, all
    from three acquisition launches (GGEMS.run_acquisition with one phantom
    per projection): one of n_low primaries per view for the inputs, and the
    two halves of n_high for the target and its variance.  Returns a dict:

    * ``x`` [V, H, W, 3]: the low-count projection's counts and scatter
      counts / open-field counts, and the ray-traced transmission (the
      network's inputs);
    * ``y`` [V, H, W]: high-count scatter / open-field counts (the target);
    * ``low`` [V, H, W]: the low-count run's own scatter image, same scale;
    * ``var`` [V, H, W]: an unbiased estimate of the target's variance from
      its two independent halves: var(y) = E[(y_a - y_b)^2] / 4;
    * ``t_low``, ``t_drr``, ``t_high`` [V]: wall seconds per view of the
      low-count launch, of the ray traces and of the two high-count
      launches (each launch's time divided by the number of views)."""
    import time
    from . import configs
    from .ggems import GGEMS
    seeds = list(phantom_seeds)
    keep = configs.scatter_scenes(workdir, seeds, shape=shape, detector=detector)
    g = GGEMS()
    g.initialize(seed)
    eng = g.engine()
    box = len(seeds)  # the detector follows the phantoms
    eng.box_elements[box] = (detector, detector, True)
    rng = np.random.default_rng(seed)
    angles = rng.uniform(0.0, 360.0, size=(len(seeds), angles_per_phantom))
    rot = np.zeros((angles.size, 3), np.float32)
    rot[:, 2] = angles.ravel()
    only = np.repeat(np.arange(len(seeds), dtype=np.int32), angles_per_phantom)
    n_views = rot.shape[0]
    norm = np.maximum(rate, 1e-12)

    def acquire(n):
        t0 = time.perf_counter()
        g.run_acquisition(rot, np.full(n_views, n, np.uint64), "deg", only)
        images = [eng.projection(box, p) for p in range(n_views)]
        return (time.perf_counter() - t0) / n_views, np.stack([h for h, _ in images]), \
            np.stack([s for _, s in images])

    if log:
        log("%d views of %d phantoms: %d-primary inputs" % (n_views, len(seeds), n_low))
    t_low, low, low_scatter = acquire(n_low)
    t0 = time.perf_counter()
    drr = np.stack([eng.drr(box, p) for p in range(n_views)])
    t_drr = (time.perf_counter() - t0) / n_views
    half = n_high // 2
    if log:
        log("targets: two halves of %d primaries" % n_high)
    t_a, _, scatter_a = acquire(half)
    t_b, _, scatter_b = acquire(n_high - half)
    g.delete()
    del keep
    ya = scatter_a / (norm * half)
    yb = scatter_b / (norm * (n_high - half))
    return {"x": np.stack([low / (norm * n_low), drr, low_scatter / (norm * n_low)], axis=-1).astype(np.float32),
            "y": ((scatter_a + scatter_b) / (norm * n_high)).astype(np.float32),
            "low": (low_scatter / (norm * n_low)).astype(np.float32),
            "var": ((ya - yb) ** 2 / 4.0).astype(np.float32),
            "t_low": np.full(n_views, t_low), "t_drr": np.full(n_views, t_drr), "t_high": np.full(n_views, t_a + t_b)}


def dataset_chunks(workdir, phantom_seeds, angles_per_phantom, seed, n_low, n_high, rate, chunk=5, detector=256,
                   log=None):
    """dataset_batched over groups of `chunk` phantoms (one GGEMS scene and
    three launches per group), concatenated"""
    seeds = list(phantom_seeds)
    parts = []
    for i in range(0, len(seeds), chunk):
        parts.append(dataset_batched(workdir, seeds[i:i + chunk], angles_per_phantom, seed + i, n_low, n_high, rate,
                                     detector=detector, log=log))
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def drr_baseline(x_train, y_train):
    """The method without the network: scatter = k * ray-traced transmission,
    k fitted by least squares on the training views"""
    d = x_train[..., 1].astype(np.float64)
    k = float((d * y_train).sum() / max((d * d).sum(), 1e-300))
    return lambda x: (k * x[..., 1]).astype(np.float32)


def scatter_errors(pred, y, var, low):
    """Errors relative to the mean target, raw and against the truth.

    The raw error compares with the noisy high-count target. Its expected
    square is (error against the truth)^2 + var(y), so subtracting the
    target's estimated variance gives the error against the noise-free
    scatter. The low-count run's scatter is an unbiased estimate whose error
    against the truth is its own noise, which falls as 1 / sqrt(primaries):
    ``mc_primaries_factor`` is how many times more primaries than the
    low-count run plain MC needs to match the prediction's error."""
    mean = float(np.mean(y))
    ms = lambda a: float(np.mean(a.astype(np.float64) ** 2))
    noise = float(np.mean(var.astype(np.float64)))
    pred_ms, low_ms = ms(pred - y), ms(low - y)
    pred_true = max(pred_ms - noise, 0.0)
    low_true = max(low_ms - noise, 0.0)
    out = {"pred": np.sqrt(pred_ms) / mean, "low": np.sqrt(low_ms) / mean, "target_noise": np.sqrt(noise) / mean,
           "pred_vs_truth": np.sqrt(pred_true) / mean, "low_vs_truth": np.sqrt(low_true) / mean}
    out["mc_primaries_factor"] = low_true / pred_true if pred_true > 0 else float("inf")
    return out


def conv_flops(net, h, w):
    """multiply-add FLOPs of the 3x3 convolutions of one forward pass"""
    total, size = 0, (h, w)
    for i, (a, b) in enumerate(net.enc):
        for m in (a, b):
            cout, _, cin = m.weight.shape
            total += 2 * size[0] * size[1] * cout * 9 * cin
        if i + 1 < len(net.enc):
            size = (size[0] // 2, size[1] // 2)
    for a, b in net.dec:
        size = (size[0] * 2, size[1] * 2)
        for m in (a, b):
            cout, _, cin = m.weight.shape
            total += 2 * size[0] * size[1] * cout * 9 * cin
    return total


def train(x, y, steps=1500, batch=8, lr=2e-3, seed=0, device="cuda", log=None):
    """Adam on the mean squared error; returns (net, seconds, losses)"""
    import time
    torch.manual_seed(seed)
    net = UNet(seed=seed).to(device)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    X = torch.from_numpy(x).to(device)
    Y = torch.from_numpy(y).to(device)
    g = torch.Generator(device="cpu").manual_seed(seed)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(steps):
        idx = torch.randint(0, X.shape[0], (batch,), generator=g)
        opt.zero_grad(set_to_none=True)
        loss = ((net(X[idx]) - Y[idx]) ** 2).mean()
        loss.backward()
        opt.step()
        sched.step()
        if step % 50 == 0 or step == steps - 1:
            losses.append(loss.item())
            if log and step % 250 == 0:
                log("step %d loss %.3e" % (step, losses[-1]))
    torch.cuda.synchronize()
    return net, time.perf_counter() - t0, losses


@torch.no_grad()
def predict(net, x, batch=8):
    out = []
    for i in range(0, x.shape[0], batch):
        out.append(net(torch.from_numpy(x[i:i + batch]).to(next(net.parameters()).device)).float().cpu().numpy())
    return np.concatenate(out)
