"""Plain PyTorch (CPU, FP32) restatement of the warm-start U-Net, the reference of its HIP kernels.

Same architecture and flat parameter layout as fft_method_test_amd/warmstart.py: per layer the
weights [cout][taps][cin] (tap t = (dz+1)*9 + (dy+1)*3 + (dx+1)) then the biases; periodic 3x3x3
convolutions (circular padding) with ReLU, 2x2x2 average pooling, nearest upsampling, channel
concatenation [upsampled | skip], a 1x1x1 head whose channels 0..2 are the prediction.
bf16_acts=True rounds every stored activation to BF16 as the device kernels do.
"""
import numpy as np
import torch
import torch.nn.functional as F

from fft_method_test_amd.warmstart import LAYERS


def unflatten(params):
    p = torch.as_tensor(np.asarray(params, dtype=np.float32))
    out, off = [], 0
    for taps, c0, c1, co in LAYERS:
        cin = c0 + c1
        w = p[off:off + co * taps * cin].reshape(co, taps, cin)
        off += co * taps * cin
        b = p[off:off + co]
        off += co
        k = 3 if taps == 27 else 1
        out.append([w.permute(0, 2, 1).reshape(co, cin, k, k, k).clone(), b.clone()])
    assert off == p.numel()
    return out


def flatten_grads(layers):
    parts = []
    for w, b in layers:
        co, cin, k = w.shape[0], w.shape[1], w.shape[2]
        parts.append(w.grad.reshape(co, cin, k ** 3).permute(0, 2, 1).reshape(-1))
        parts.append(b.grad.reshape(-1))
    return torch.cat(parts).numpy()


def features(mask, n):
    """[1, 8, n, n, n]: solid indicator, interface support (solid with a fluid 26-neighbour)."""
    fluid = mask.reshape(n, n, n).astype(bool)
    solid = ~fluid
    nb = np.zeros_like(fluid)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dz == dy == dx == 0:
                    continue
                nb |= np.roll(fluid, (-dz, -dy, -dx), axis=(0, 1, 2))
    x = np.zeros((1, 8, n, n, n), dtype=np.float32)
    x[0, 0] = solid
    x[0, 1] = solid & nb
    return torch.from_numpy(x)


def forward(layers, x, bf16_acts=False):
    def q(t):
        return t.to(torch.bfloat16).to(torch.float32) if bf16_acts else t

    def conv(i, t, relu=True):
        w, b = layers[i]
        if w.shape[2] == 3:
            t = F.pad(t, (1, 1, 1, 1, 1, 1), mode="circular")
        y = F.conv3d(t, w, b)
        return q(F.relu(y)) if relu else y

    pool = lambda t: q(F.avg_pool3d(t, 2))
    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    e0 = conv(1, conv(0, x))
    e1 = conv(3, conv(2, pool(e0)))
    e2 = conv(5, conv(4, pool(e1)))
    b3 = conv(7, conv(6, pool(e2)))
    d2 = conv(9, conv(8, torch.cat([up(b3), e2], 1)))
    d1 = conv(11, conv(10, torch.cat([up(d2), e1], 1)))
    d0 = conv(13, conv(12, torch.cat([up(d1), e0], 1)))
    return conv(14, d0, relu=False)[:, :9]


def loss(pred, target, support):
    """Mean squared error over the support voxels of the 9 predicted components (three load
    directions x three components); target [1, 9, n, n, n]."""
    m = support.to(pred.dtype)
    return ((pred - target) ** 2 * m).sum() / (9.0 * m.sum())


def field_to_ncdhw(field, n):
    """(n^3, 9) voxel-major -> [1, 9, n, n, n]."""
    return torch.as_tensor(np.asarray(field, dtype=np.float32)).reshape(n, n, n, 9).permute(3, 0, 1, 2)[None]


def ncdhw_to_field(t):
    return t[0].permute(1, 2, 3, 0).reshape(-1, 9).detach().numpy()
