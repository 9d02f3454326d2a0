"""The network oracle (CPU): its gradient against finite differences of its
own loss (exact FP64 mode), its periodic-x/y translation equivariance, and
the size of its BF16 rounding. These pin the restatement before it is used as
the checker of the GPU kernels."""
import numpy as np
import pytest

from helpers import rel_err


def _params(orc, inc, seed):
    n = orc.unet_num_params(inc)
    rng = np.random.default_rng(seed)
    p = (0.2 * rng.standard_normal(n)).astype(np.float32)
    return p


def _sample(shape, seed):
    rng = np.random.default_rng(seed)
    src = rng.standard_normal(shape)
    mask = (rng.random(shape) > 0.3).astype(np.float64)
    tgt = rng.standard_normal(shape)
    return src, mask, tgt


def test_unet_gradient_matches_finite_differences(orc):
    shape = (8, 8, 8)
    p = _params(orc, 2, 1)
    src, mask, tgt = _sample(shape, 2)
    loss, g = orc.unet_gradients(p, src, tgt, mask, bf16=False)
    rng = np.random.default_rng(3)
    # indices in every layer's weights and biases, and the head
    n = p.size
    idx = sorted(set(rng.integers(0, n, 60).tolist() + [n - 1, n - 2, n - 17]))
    worst = 0.0
    checked = 0
    for i in idx:
        h = np.float32(1e-3)
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = orc.unet_gradients(pp, src, tgt, mask, bf16=False)
        lm, _ = orc.unet_gradients(pm, src, tgt, mask, bf16=False)
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        err = abs(fd - g[i]) / max(abs(g[i]), 1e-3 * np.abs(g).max())
        worst = max(worst, err)
        checked += 1
        assert err < 5e-3, (i, fd, g[i])
    print(f"{checked} parameters: worst |fd - grad| / |grad| = {worst:.2e}")


def test_unet_translation_equivariance(orc):
    # periodic in x and y: a shift by 4 cells (two poolings) shifts the output
    shape = (8, 8, 8)
    p = _params(orc, 2, 5)
    src, mask, _ = _sample(shape, 6)
    out = orc.unet_forward(p, src, mask, bf16=False)
    out_s = orc.unet_forward(p, np.roll(src, 4, axis=0), np.roll(mask, 4, axis=0), bf16=False)
    assert np.abs(out_s - np.roll(out, 4, axis=0)).max() < 1e-12 * np.abs(out).max()
    out_s = orc.unet_forward(p, np.roll(src, -4, axis=1), np.roll(mask, -4, axis=1), bf16=False)
    assert np.abs(out_s - np.roll(out, -4, axis=1)).max() < 1e-12 * np.abs(out).max()


def test_unet_output_scales_with_source(orc):
    # the source is normalised by its RMS and the output scaled back: N(a·b) = a·N(b), a > 0
    shape = (8, 8, 4)
    p = _params(orc, 1, 7)
    src, _, _ = _sample(shape, 8)
    o1 = orc.unet_forward(p, src, bf16=False)
    o2 = orc.unet_forward(p, 1000.0 * src, bf16=False)
    assert rel_err(o2, 1000.0 * o1) < 1e-12


def test_unet_bf16_rounding_is_small(orc):
    shape = (8, 8, 8)
    p = _params(orc, 2, 9)
    src, mask, tgt = _sample(shape, 10)
    o_ex = orc.unet_forward(p, src, mask, bf16=False)
    o_bf = orc.unet_forward(p, src, mask, bf16=True)
    e = np.linalg.norm(o_bf - o_ex) / np.linalg.norm(o_ex)
    print(f"BF16 storage vs exact: relative L2 {e:.2e}")
    assert e < 5e-2
