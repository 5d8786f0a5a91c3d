"""FP64 CPU oracle of the warm-start U-Net (test infrastructure only).

The network has no counterpart in the reference.  This restatement follows
the library's definition (qes-public_amd/csrc/qesmc_unet.hip): parameter
layout, channels-last volumes, 2x2x2 average pooling, nearest 2x upsampling,
skip concatenation [skip, up], ReLU after every 3x3x3 convolution, 1x1x1
head.  It is built on torch.nn.functional on the CPU in float64 (an
independent library implementation of the convolutions, with autograd for
the gradients), and is itself pinned by central finite differences
(tests/test_unet.py).

With bf16=True it rounds to BF16 where the device stores BF16: the input
features, the conv weights, every conv / pool output.  The head's weights
and the biases stay FP32, as on the device.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

IN_CHANNELS = 8


def layout(levels: int, c0: int):
    """[(name, cin, cout, woff, boff)] in the library's parameter order, and
    the head offset and total count."""
    layers, off = [], 0

    def mk(name, cin, cout):
        nonlocal off
        woff = off
        off += 27 * cin * cout
        boff = off
        off += cout
        layers.append((name, cin, cout, woff, boff))

    for lv in range(levels):
        c = c0 << lv
        cin = IN_CHANNELS if lv == 0 else c0 << (lv - 1)
        mk(f"enc1_{lv}", cin, c)
        mk(f"enc2_{lv}", c, c)
    for lv in range(levels - 2, -1, -1):
        c = c0 << lv
        mk(f"dec1_{lv}", 3 * c, c)
        mk(f"dec2_{lv}", c, c)
    head = off
    off += c0 + 1
    return layers, head, off


def _bf(t, on):
    return t.to(torch.bfloat16).to(torch.float64) if on else t


class _RoundGrad(torch.autograd.Function):
    """Identity whose backward rounds the incoming gradient to BF16: placed
    where the device stores an activation gradient in BF16."""

    @staticmethod
    def forward(ctx, t):
        return t.view_as(t)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(torch.float64)


def forward(params, x, levels, c0, bf16=False):
    """params: torch float64 vector (may require grad); x: [dz, dy, dx, 8]
    tensor.  Returns the [dz, dy, dx] output."""
    layers, head, n = layout(levels, c0)
    assert params.numel() == n
    rg = _RoundGrad.apply if bf16 else (lambda t: t)
    conv = {}
    for name, cin, cout, woff, boff in layers:
        W = params[woff:woff + 27 * cin * cout].reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3)
        conv[name] = (_bf(W, bf16), params[boff:boff + cout])

    def c3(name, t):
        W, b = conv[name]
        # the data gradient of a conv is stored in BF16
        return _bf(F.relu(F.conv3d(rg(t), W, b, padding=1)), bf16)

    h = _bf(x.permute(3, 0, 1, 2).unsqueeze(0), bf16)  # [1, C, dz, dy, dx]
    skips = []
    for lv in range(levels):
        # the encoder output's gradient (skip + pooled part) is stored in BF16
        h = rg(c3(f"enc2_{lv}", c3(f"enc1_{lv}", h)))
        skips.append(h)
        if lv + 1 < levels:
            h = _bf(F.avg_pool3d(h, 2), bf16)
    for lv in range(levels - 2, -1, -1):
        up = F.interpolate(rg(h), scale_factor=2, mode="nearest")
        h = torch.cat([skips[lv], up], dim=1)
        h = c3(f"dec2_{lv}", c3(f"dec1_{lv}", h))
    wh = params[head:head + c0].reshape(1, c0, 1, 1, 1)
    bh = params[head + c0:head + c0 + 1]
    return F.conv3d(rg(h), wh, bh)[0, 0]


def forward_np(params, x, levels, c0, bf16=False):
    with torch.no_grad():
        p = torch.as_tensor(np.asarray(params, np.float64))
        return forward(p, torch.as_tensor(np.asarray(x, np.float64)), levels, c0, bf16).numpy()


def param_grad_np(params, x, grad_out, levels, c0, bf16=False):
    """dL/dparams for L = sum(output * grad_out)."""
    p = torch.tensor(np.asarray(params, np.float64), requires_grad=True)
    y = forward(p, torch.as_tensor(np.asarray(x, np.float64)), levels, c0, bf16)
    (y * torch.as_tensor(np.asarray(grad_out, np.float64))).sum().backward()
    return p.grad.numpy()
