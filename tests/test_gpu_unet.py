"""Predictor network (BF16 MFMA engine) against an FP32 torch reference of the same network.

Per layer (the parity test proper): every activation the engine produces equals the
reference operation (BF16-rounded weights, FP32 sums, BF16-rounded output) applied to
the engine's own input activations, except where the two FP32 sums (different
summation order) straddle a BF16 rounding boundary: at most 0.1% of the elements, by
at most 2 BF16 ulps.  End to end the rare flips propagate through the later
BF16-rounded layers, so the outputs agree to 3e-2 relative L2 only (stated bound).
Gradients: the engine's backward carries BF16 gradient tensors, the reference none;
the flat gradient agrees to 6e-2 relative L2 and every layer's gradient points the
same way (cosine > 0.995).  Training reduces the loss of a fixed smooth target.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


def _inputs(shape, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    feats = [np.sin(2 * np.pi * (x / nx + 2 * y / ny)), np.cos(2 * np.pi * z / nz), rng.standard_normal(shape) * 0.3,
             np.sin(2 * np.pi * (x / nx - z / nz))]
    X = np.stack([f.ravel() for f in feats], axis=1).astype(np.float32)
    T = (0.5 * feats[0] * feats[1] + 0.2 * feats[3]).ravel().astype(np.float32)
    return X, T


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 32, 32), (8, 12, 16)])
def test_forward_matches_reference(torch_dev, shape):
    torch = torch_dev
    from salmon2_amd.unet import UNet
    from tests import unet_ref

    print(f"[gpu] unet forward shape={shape}", flush=True)
    nz, ny, nx = shape
    net = UNet(nx, ny, nz, seed=11, stream=torch.cuda.current_stream())
    X, _ = _inputs(shape, 1)
    xd = torch.from_numpy(X).cuda()
    out = torch.zeros(X.shape[0], dtype=torch.float32, device="cuda")
    net.forward(xd, out)
    torch.cuda.synchronize()
    w = net.get_params()
    acts = {i: net.activation(i) for i in range(-1, 12)}
    x0 = np.zeros((X.shape[0], 16), dtype=np.float32)
    x0[:, :4] = unet_ref.bf(torch.from_numpy(X)).numpy()
    assert np.array_equal(acts[-1], x0)
    ref = unet_ref.layer_outputs(w, acts, x0, shape)
    for i in range(12):
        g, r = acts[i], ref[i]
        diff = g != r
        # a flip moves a value by one BF16 ulp (2^-8 relative to the value), at most two; near
        # zero the ReLU can turn a tiny value into 0, allowed up to 1e-5 of the layer's scale
        ulp = np.maximum(np.abs(r[diff]), np.abs(g[diff])) * 2.0 ** -7 + 1e-5 * np.abs(r).max()
        worst = float((np.abs(g[diff] - r[diff]) / ulp).max()) if diff.any() else 0.0
        print(f"[parity] unet {shape} layer {i}: differing {diff.mean():.2e}, worst {worst:.2f} x 2 ulp", flush=True)
        assert diff.mean() <= 1e-3 and worst <= 1.0
    got = out.cpu().numpy()
    head = w[-17:]
    assert np.abs(got - (acts[11] @ head[:16] + head[16])).max() <= 1e-6 * np.abs(got).max()
    e2e = unet_ref.forward(w, torch.from_numpy(X), shape).numpy()
    el2 = float(np.linalg.norm(got - e2e) / np.linalg.norm(e2e))
    print(f"[parity] unet forward {shape} end to end: l2 {el2:.3e}", flush=True)
    assert el2 < 3e-2
    net.close()


def test_gradients_match_reference(torch_dev):
    torch = torch_dev
    from salmon2_amd.unet import UNet
    from tests import unet_ref

    shape = (16, 16, 16)
    print("[gpu] unet gradients", flush=True)
    net = UNet(16, 16, 16, seed=5, stream=torch.cuda.current_stream())
    X, T = _inputs(shape, 2)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    net.train_sample(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), loss)
    torch.cuda.synchronize()
    g = net.get_grads()
    lref, gref = unet_ref.loss_and_grads(net.get_params(), torch.from_numpy(X), torch.from_numpy(T), shape)
    print(f"[parity] unet loss {float(loss.item()):.6e} ref {lref:.6e}", flush=True)
    assert abs(float(loss.item()) - lref) < 2e-3 * lref
    el2 = float(np.linalg.norm(g - gref) / np.linalg.norm(gref))
    print(f"[parity] unet grad l2 {el2:.3e}", flush=True)
    assert el2 < 6e-2
    off = 0
    for co, c1, c2, up, lv in unet_ref.SPEC:
        n = 27 * co * (c1 + c2) + co
        a, b = g[off : off + n], gref[off : off + n]
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        assert cos > 0.995, (off, cos)
        off += n
    net.close()


def test_training_reduces_loss(torch_dev):
    torch = torch_dev
    from salmon2_amd.unet import UNet

    print("[gpu] unet training", flush=True)
    shape = (16, 16, 16)
    net = UNet(16, 16, 16, seed=3, stream=torch.cuda.current_stream())
    data = [_inputs(shape, s) for s in range(4)]
    xs = [torch.from_numpy(x).cuda() for x, _ in data]
    ts = [torch.from_numpy(t).cuda() for _, t in data]
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    hist = []
    for it in range(150):
        tot = 0.0
        for x, t in zip(xs, ts):
            net.train_sample(x, t, loss)
            tot += float(loss.item())
        net.adam_step(2e-3, len(xs))
        hist.append(tot / len(xs))
    print(f"[parity] unet training loss {hist[0]:.4e} -> {hist[-1]:.4e}", flush=True)
    assert hist[-1] < 0.1 * hist[0]
    net.close()
