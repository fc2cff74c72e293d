"""The CPU reference of the surrogate U-Net, pinned by finite differences
in FP64 (torch.autograd.gradcheck), and the NHWC weight conventions against
direct sums."""

import copy

import numpy as np
import torch

from ggems_amd import surrogate as S


def test_reference_gradients_match_finite_differences():
    torch.manual_seed(0)
    net = copy.deepcopy(S.UNet(channels=(4, 8), seed=3)).double()
    x = torch.rand(1, 8, 8, 2, dtype=torch.float64)
    names = [n for n, _ in net.named_parameters()]

    class Wrap(torch.nn.Module):
        def __init__(self, inner):
            super().__init__()
            self.net = inner

        def forward(self, z):
            return S.reference_forward(self.net, z)

    wrap = Wrap(net)

    def f(*ps):
        return torch.func.functional_call(wrap, {"net." + n: p for n, p in zip(names, ps)}, (x,))

    inputs = [p.detach().clone().requires_grad_(True) for _, p in net.named_parameters()]
    assert torch.autograd.gradcheck(f, inputs, eps=1e-6, atol=1e-6, rtol=1e-4)


def test_weight_layout_is_tap_major_cross_correlation():
    """w[co][3 (dy + 1) + (dx + 1)][ci] multiplies x[y + dy][x + dx][ci]"""
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 6, 5, 3))
    w = rng.normal(size=(2, 9, 3))
    conv = S.Conv3x3(3, 2, relu=False).double()
    conv.weight.data = torch.tensor(w)
    net = S.UNet(channels=(4,), seed=0).double()
    got = torch.nn.functional.conv2d(torch.tensor(x).permute(0, 3, 1, 2),
                                     conv.weight.view(2, 3, 3, 3).permute(0, 3, 1, 2), padding=1)
    for co in range(2):
        for yy in range(6):
            for xx in range(5):
                s = 0.0
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if 0 <= yy + dy < 6 and 0 <= xx + dx < 5:
                            s += float(w[co, 3 * (dy + 1) + (dx + 1)] @ x[0, yy + dy, xx + dx])
                assert abs(float(got[0, co, yy, xx]) - s) < 1e-12
    # the data gradient's weights: taps reversed, channels swapped
    wt = S.flip_transpose(torch.tensor(w))
    assert wt.shape == (3, 9, 2)
    assert torch.equal(wt[1, 0, 0], torch.tensor(w)[0, 8, 1])
    del net
