"""The surrogate's MFMA convolutions against torch's FP32 convolutions on
the CPU (same BF16-rounded operands), the whole U-Net's output and
gradients against the CPU reference, and a few training steps."""

import copy

import pytest
import torch
import torch.nn.functional as Fn

from ggems_amd import surrogate as S

pytestmark = pytest.mark.gpu


def _ref_conv(x, w, b, relu):
    """x [N,H,W,Cin], w [Cout,9,Cin] on the CPU in FP32 -> [N,H,W,Cout]"""
    cout, _, cin = w.shape
    y = Fn.conv2d(x.permute(0, 3, 1, 2), w.view(cout, 3, 3, cin).permute(0, 3, 1, 2), b, padding=1)
    y = torch.relu(y) if relu else y
    return y.permute(0, 2, 3, 1)


@pytest.mark.parametrize("n,h,w,cin,cout,relu", [(2, 32, 32, 8, 16, False), (1, 64, 64, 16, 64, True),
                                                 (2, 32, 64, 48, 128, True), (1, 32, 32, 192, 64, False)])
def test_conv_forward_matches_fp32(n, h, w, cin, cout, relu):
    g = torch.Generator().manual_seed(cin * cout)
    x = torch.randn(n, h, w, cin, generator=g).bfloat16()
    wt = (torch.randn(cout, 9, cin, generator=g) / (3 * cin ** 0.5)).bfloat16()
    b = torch.randn(cout, generator=g)
    y = S.conv3x3(x.cuda(), wt.view(cout, -1).cuda(), b.cuda(), relu).float().cpu()
    ref = _ref_conv(x.float(), wt.float(), b, relu)
    # FP32 accumulation, one BF16 rounding of the output
    assert torch.all((y - ref).abs() <= 2 ** -7 * ref.abs() + 1e-5), float((y - ref).abs().max())


@pytest.mark.parametrize("n,h,w,cin,cout", [(2, 32, 32, 16, 16), (1, 64, 64, 48, 128), (2, 16, 16, 128, 64)])
def test_conv_gradients_match_fp32(n, h, w, cin, cout):
    g = torch.Generator().manual_seed(7 + cin)
    x = torch.randn(n, h, w, cin, generator=g).bfloat16()
    wt = (torch.randn(cout, 9, cin, generator=g) / (3 * cin ** 0.5)).bfloat16()
    dy = torch.randn(n, h, w, cout, generator=g).bfloat16()
    xr = x.float().requires_grad_(True)
    wr = wt.float().requires_grad_(True)
    _ref_conv(xr, wr, None, False).backward(dy.float())
    dw = S.conv3x3_weight_grad(x.cuda(), dy.cuda()).cpu().view(cout, 9, cin)
    # FP32 sums of exact BF16 products, in another order
    scale = wr.grad.abs().max()
    assert torch.all((dw - wr.grad).abs() <= 1e-5 * scale + 1e-4 * wr.grad.abs()), float((dw - wr.grad).abs().max())
    dx = S.conv3x3(dy.cuda(), S.flip_transpose(wt).view(cin, 9 * cout).cuda()).float().cpu()
    assert torch.all((dx - xr.grad).abs() <= 2 ** -7 * xr.grad.abs() + 1e-4), float((dx - xr.grad).abs().max())


@pytest.mark.parametrize("c,relu,f32", [(16, True, False), (64, False, False), (128, True, True), (32, True, True)])
def test_relu_bias_backward_matches_torch(c, relu, f32):
    """The fused ReLU/bias backward against the same steps in torch on the CPU:
    the mask is exact, the bias gradient an FP32 sum in another order."""
    g = torch.Generator().manual_seed(c + 3 * relu)
    dy = torch.randn(3, 40, 48, c, generator=g)
    dy = dy if f32 else dy.bfloat16()
    y = torch.randn(3, 40, 48, c, generator=g).bfloat16()
    got, db = S.relu_bias_backward(dy.cuda(), y.cuda(), relu)
    want = dy.bfloat16() * ((y > 0).bfloat16() if relu else 1)
    assert got.dtype == torch.bfloat16 and torch.equal(got.cpu(), want)
    ref = want.double().sum(dim=(0, 1, 2))
    assert torch.allclose(db.cpu().double(), ref, rtol=1e-5, atol=1e-3), float((db.cpu().double() - ref).abs().max())


def test_unet_forward_and_gradients_match_the_fp32_reference():
    torch.manual_seed(0)
    net = S.UNet(seed=5).cuda()
    x = torch.rand(2, 128, 128, 2)
    out = net(x.cuda())
    target = torch.rand(2, 128, 128)
    loss = ((out - target.cuda()) ** 2).mean()
    loss.backward()
    ref_net = copy.deepcopy(net).cpu().float()
    for p in ref_net.parameters():
        p.grad = None
    ref = S.reference_forward(ref_net, x)
    rel = float((out.detach().cpu() - ref.detach()).norm() / ref.detach().norm())
    assert rel < 0.03, rel
    ((ref - target) ** 2).mean().backward()
    for (name, p), (_, q) in zip(net.named_parameters(), ref_net.named_parameters()):
        a, b = p.grad.detach().cpu().reshape(-1), q.grad.reshape(-1)
        if b.norm() == 0:
            continue
        cos = float(a @ b / (a.norm() * b.norm()))
        assert cos > 0.99, (name, cos)
        assert float((a - b).norm() / b.norm()) < 0.1, name


def test_training_steps_reduce_the_loss():
    torch.manual_seed(1)
    net = S.UNet(seed=2).cuda()
    opt = torch.optim.Adam(net.parameters(), lr=2e-3)
    x = torch.rand(4, 128, 128, 2, device="cuda")
    target = (x[..., 0] * 0.5 + torch.roll(x[..., 1], 3, dims=1) * 0.25).detach()
    first = None
    for step in range(60):
        opt.zero_grad()
        loss = ((net(x) - target) ** 2).mean()
        loss.backward()
        opt.step()
        first = first if first is not None else loss.item()
    assert loss.item() < 0.3 * first, (first, loss.item())


@pytest.mark.parametrize("cout,cin", [(16, 8), (64, 192), (128, 128)])
def test_weights_bf16_equal_cast_and_flip_transpose(cout, cin):
    """the fused weight preparation equals torch's round-to-nearest cast and
    the data gradient's flipped, transposed weights, bit for bit, and sets
    the layer's gradient accumulators to zero"""
    w = torch.randn(cout, 9, cin, device="cuda") * 0.1
    wb, wt = S.weights_bf16(w)
    ref = w.to(torch.bfloat16)
    assert torch.equal(wb.view(torch.int16), ref.view(torch.int16))
    assert torch.equal(wt.view(torch.int16), S.flip_transpose(ref).view(torch.int16))
    dw = torch.full((cout, 9 * cin), 7.0, device="cuda")
    db = torch.full((cout,), 7.0, device="cuda")
    wb2, _ = S.weights_bf16(w, (dw, db))
    assert torch.equal(wb2, wb)
    assert not dw.any() and not db.any()


@pytest.mark.parametrize("n,h,w,c", [(2, 32, 64, 16), (3, 16, 16, 128)])
def test_pool2x2_matches_max_pool2d(n, h, w, c):
    """values and gradients of max_pool2d on the CPU, bit for bit; a window
    of ties (zeros, as after a ReLU) gives the gradient to its first input"""
    g = torch.Generator().manual_seed(h + c)
    x = torch.randn(n, h, w, c, generator=g).bfloat16()
    x[0, :4, :4, :] = 0.0
    xr = x.float().permute(0, 3, 1, 2).contiguous().requires_grad_(True)
    ref = Fn.max_pool2d(xr, 2)
    dy = torch.randn(ref.shape, generator=g).bfloat16()
    ref.backward(dy.float())
    y, arg = S.pool2x2(x.cuda())
    assert torch.equal(y.cpu().float(), ref.detach().permute(0, 2, 3, 1))
    dx = S.pool2x2_backward(dy.permute(0, 2, 3, 1).contiguous().cuda(), arg)
    assert torch.equal(dx.cpu().float(), xr.grad.permute(0, 2, 3, 1))


def test_upcat_matches_cat_of_nearest_upsampling():
    """the decoder input against cat(interpolate(z, nearest), s) on the CPU,
    and its gradients: ds exact, dz the FP32 sum rounded once to BF16"""
    g = torch.Generator().manual_seed(5)
    z = torch.randn(2, 8, 16, 32, generator=g).bfloat16()
    s = torch.randn(2, 16, 32, 16, generator=g).bfloat16()
    zr = z.float().requires_grad_(True)
    sr = s.float().requires_grad_(True)
    up = Fn.interpolate(zr.permute(0, 3, 1, 2), scale_factor=2, mode="nearest").permute(0, 2, 3, 1)
    ref = torch.cat([up, sr], dim=3)
    dout = torch.randn(ref.shape, generator=g).bfloat16()
    ref.backward(dout.float())
    out = S.upcat(z.cuda(), s.cuda())
    assert torch.equal(out.cpu().float(), ref.detach())
    dz, ds = S.upcat_backward(dout.cuda(), 32)
    assert torch.equal(ds.cpu().float(), sr.grad)
    assert torch.equal(dz.cpu(), zr.grad.bfloat16())
