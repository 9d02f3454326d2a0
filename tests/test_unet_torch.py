"""The U-Net oracle (oracle/unet_oracle.c, exact FP64 mode) against an
independent torch-CPU float64 build of the same network: conv3d with circular
padding in x and y and zero padding in z, bias + ReLU after every 3³ layer,
2×2×2 mean pooling, nearest upsampling, concatenation [upsampled, skip], the
1×1 head, the output times the source's RMS, and the mean-square loss against
the target scaled by 1/RMS. The forward output and the gradient of every
parameter (torch autograd) must agree. This pins the oracle's convolution
index convention, its skip/concatenation order and its transposed-sum input
gradient by a route that shares no code with it."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
F = torch.nn.functional

SPEC = [(0, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]


def _unpack(params, inc):
    """(weights as (cout, cin, 3, 3, 3), biases, head weights, head bias) from the
    oracle's vector: per layer W[tap][c][n], tap = (dx+1) + 3(dy+1) + 9(dz+1)."""
    off, Ws, bs = 0, [], []
    for l, (cin, cout) in enumerate(SPEC):
        cin = inc if l == 0 else cin
        w = params[off:off + 27 * cin * cout].reshape(3, 3, 3, cin, cout)  # [dz][dy][dx][c][n]
        off += 27 * cin * cout
        Ws.append(torch.tensor(np.ascontiguousarray(w.transpose(4, 3, 0, 1, 2)), dtype=torch.float64,
                               requires_grad=True))
        bs.append(torch.tensor(params[off:off + cout], dtype=torch.float64, requires_grad=True))
        off += cout
    wh = torch.tensor(params[off:off + 16], dtype=torch.float64, requires_grad=True)
    bh = torch.tensor(params[off + 16:off + 17], dtype=torch.float64, requires_grad=True)
    return Ws, bs, wh, bh


def _conv(x, W, b):
    x = F.pad(x, (1, 1, 1, 1, 0, 0), mode="circular")   # x (W) and y (H) periodic
    x = F.pad(x, (0, 0, 0, 0, 1, 1))                     # z (D) zero-padded
    return F.relu(F.conv3d(x, W, b))


def _net(src, mask, Ws, bs, wh, bh):
    """src, mask: (nx, ny, nz) arrays -> (y in the scaled space, rms)."""
    s_np = np.asarray(src)
    rms = float(np.sqrt(np.mean(s_np ** 2)))
    chans = [torch.tensor(s_np.transpose(2, 1, 0) / rms, dtype=torch.float64)]
    if mask is not None:
        chans.append(torch.tensor(np.asarray(mask).transpose(2, 1, 0), dtype=torch.float64))
    x = torch.stack(chans)[None]                         # (1, C, nz, ny, nx)
    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    a0 = _conv(x, Ws[0], bs[0])
    a1 = _conv(a0, Ws[1], bs[1])
    a2 = _conv(F.avg_pool3d(a1, 2), Ws[2], bs[2])
    a3 = _conv(a2, Ws[3], bs[3])
    a4 = _conv(F.avg_pool3d(a3, 2), Ws[4], bs[4])
    a5 = _conv(a4, Ws[5], bs[5])
    a6 = _conv(torch.cat([up(a5), a3], 1), Ws[6], bs[6])
    a7 = _conv(a6, Ws[7], bs[7])
    a8 = _conv(torch.cat([up(a7), a1], 1), Ws[8], bs[8])
    a9 = _conv(a8, Ws[9], bs[9])
    y = torch.einsum("c,bczyx->bzyx", wh, a9)[0] + bh
    return y, rms


def _pack_grad(Ws, bs, wh, bh):
    g = []
    for W, b in zip(Ws, bs):
        g.append(W.grad.numpy().transpose(2, 3, 4, 1, 0).ravel())  # back to [dz][dy][dx][c][n]
        g.append(b.grad.numpy())
    g += [wh.grad.numpy(), bh.grad.numpy()]
    return np.concatenate(g)


@pytest.mark.parametrize("inc", [1, 2])
def test_unet_oracle_matches_torch(orc, inc):
    rng = np.random.default_rng(40 + inc)
    n = orc.unet_num_params(inc)
    params = (0.15 * rng.standard_normal(n)).astype(np.float32)
    params[-17:-1] *= 2.0
    shape = (8, 12, 8)
    src = rng.standard_normal(shape)
    mask = (rng.uniform(size=shape) > 0.3).astype(np.float64) if inc == 2 else None
    target = rng.standard_normal(shape)
    Ws, bs, wh, bh = _unpack(params.astype(np.float64), inc)
    y, rms = _net(src, mask, Ws, bs, wh, bh)
    out_t = (y * rms).detach().numpy().transpose(2, 1, 0)
    out_o = orc.unet_forward(params, src, mask, bf16=False)
    assert np.abs(out_o - out_t).max() <= 1e-12 * np.abs(out_t).max()
    tgt = torch.tensor(target.transpose(2, 1, 0) / rms, dtype=torch.float64)
    loss = torch.mean((y - tgt) ** 2)
    loss.backward()
    g_t = _pack_grad(Ws, bs, wh, bh)
    loss_o, g_o = orc.unet_gradients(params, src, target, mask, bf16=False)
    assert abs(loss_o - loss.item()) <= 1e-12 * loss.item()
    assert g_o.shape == g_t.shape
    assert np.abs(g_o - g_t).max() <= 1e-10 * np.abs(g_t).max()
