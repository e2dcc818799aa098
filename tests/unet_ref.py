"""FP32 torch reference of the predictor network (test helper, CPU).

The same flat parameter vector as the HIP engine (per conv layer W[27][co][ci], b[co];
head W[16], b), the same BF16 rounding points (inputs, weights, every activation and
pooled tensor), periodic padding, nearest upsampling, 2x2x2 average pooling.  Tap t
of W is the torch kernel index (kz, ky, kx) = (t // 9, (t // 3) % 3, t % 3).
"""
import numpy as np
import torch
import torch.nn.functional as F

SPEC = [(16, 16, 0, 0, 0), (16, 16, 0, 0, 0), (32, 16, 0, 0, 1), (32, 32, 0, 0, 1), (64, 32, 0, 0, 2),
        (64, 64, 0, 0, 2), (32, 64, 32, 1, 1), (32, 32, 0, 0, 1), (16, 32, 16, 1, 0), (16, 16, 0, 0, 0)]


def bf(t, on=True):
    return t.to(torch.bfloat16).to(torch.float32) if on else t


def unpack(w):
    layers, off = [], 0
    for co, c1, c2, up, lv in SPEC:
        ci = c1 + c2
        W = torch.from_numpy(np.array(w[off : off + 27 * co * ci], dtype=np.float32)).reshape(3, 3, 3, co, ci)
        off += 27 * co * ci
        b = torch.from_numpy(np.array(w[off : off + co], dtype=np.float32))
        off += co
        layers.append((W.permute(3, 4, 0, 1, 2).contiguous(), b))
    head = torch.from_numpy(np.array(w[off : off + 17], dtype=np.float32))
    return layers, head


def forward(w, x, shape, rounding=True, params=None):
    """x: (npts, 4) channels-last; shape (nz, ny, nx).  Returns out (npts,)."""
    layers, head = params if params is not None else unpack(w)
    nz, ny, nx = shape
    r = lambda t: bf(t, rounding)
    h = torch.zeros((1, 16, nz, ny, nx))
    h[0, :4] = r(x.T.reshape(4, nz, ny, nx))

    def conv(t, k, relu=True):
        W, b = layers[k]
        y = F.conv3d(F.pad(t, (1, 1, 1, 1, 1, 1), mode="circular"), r(W), b)
        return r(F.relu(y) if relu else y)

    e1a = conv(h, 0)
    e1b = conv(e1a, 1)
    p1 = r(F.avg_pool3d(e1b, 2))
    e2a = conv(p1, 2)
    e2b = conv(e2a, 3)
    p2 = r(F.avg_pool3d(e2b, 2))
    e3a = conv(p2, 4)
    e3b = conv(e3a, 5)
    d2a = conv(torch.cat([F.interpolate(e3b, scale_factor=2, mode="nearest"), e2b], 1), 6)
    d2b = conv(d2a, 7)
    d1a = conv(torch.cat([F.interpolate(d2b, scale_factor=2, mode="nearest"), e1b], 1), 8)
    d1b = conv(d1a, 9)
    out = torch.einsum("bczyx,c->bzyx", d1b, head[:16]) + head[16]
    return out.reshape(-1)


def loss_and_grads(w, x, target, shape):
    """MSE loss and its gradient with respect to the flat parameters (no rounding in the graph)."""
    layers, head = unpack(w)
    ps = []
    for W, b in layers:
        W.requires_grad_(True)
        b.requires_grad_(True)
        ps += [W, b]
    head.requires_grad_(True)
    out = forward(None, x, shape, rounding=False, params=(layers, head))
    loss = torch.mean((out - target) ** 2)
    loss.backward()
    flat = []
    for W, b in layers:
        flat += [W.grad.permute(2, 3, 4, 0, 1).reshape(-1), b.grad]
    flat.append(head.grad)
    return float(loss), torch.cat(flat).numpy()


def layer_outputs(w, acts, x0, shape):
    """Each layer recomputed from the engine's own input activations (dict index -> (npts, C)
    channels-last arrays, index -1 = packed input): returns dict index -> reference output."""
    layers, _ = unpack(w)
    nz, ny, nx = shape

    def cf(a, lv):  # channels-last -> (1, C, z, y, x)
        d = (nz >> lv, ny >> lv, nx >> lv)
        return torch.from_numpy(np.ascontiguousarray(a)).T.reshape(1, a.shape[1], *d)

    def cl(t):
        return t[0].permute(1, 2, 3, 0).reshape(-1, t.shape[1]).numpy()

    def conv(t, k):
        W, b = layers[k]
        return bf(F.relu(F.conv3d(F.pad(t, (1, 1, 1, 1, 1, 1), mode="circular"), bf(W), b)))

    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    a = {i: acts[i] for i in acts}
    out = {}
    out[0] = cl(conv(cf(x0, 0), 0))
    out[1] = cl(conv(cf(a[0], 0), 1))
    out[2] = cl(bf(F.avg_pool3d(cf(a[1], 0), 2)))
    out[3] = cl(conv(cf(a[2], 1), 2))
    out[4] = cl(conv(cf(a[3], 1), 3))
    out[5] = cl(bf(F.avg_pool3d(cf(a[4], 1), 2)))
    out[6] = cl(conv(cf(a[5], 2), 4))
    out[7] = cl(conv(cf(a[6], 2), 5))
    out[8] = cl(conv(torch.cat([up(cf(a[7], 2)), cf(a[4], 1)], 1), 6))
    out[9] = cl(conv(cf(a[8], 1), 7))
    out[10] = cl(conv(torch.cat([up(cf(a[9], 1)), cf(a[1], 0)], 1), 8))
    out[11] = cl(conv(cf(a[10], 0), 9))
    return out
