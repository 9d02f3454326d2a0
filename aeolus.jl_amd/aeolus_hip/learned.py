"""The learned preconditioner: a 3-D U-Net trained on the GPU by the library's
own forward and backward kernels, and the flexible CG that uses it.

No reference counterpart: the reference has no network (its nearest hook is
src/ML.jl:40-56). The solver it serves is MGPreconditionedCG
(src/Multigrid.jl:879-972); that solver stays the fallback and the judge of
the residual (the flexible CG runs on its operator and stops on the same
FP64 residual ‖r‖ ≤ tol·‖b‖).
"""
import ctypes

import numpy as np

from . import _lib as L
from .fields import DeviceArray
from .pressure import PressureSolver

MODES = {"net": 0, "net+vcycle": 1, "vcycle": 2}


class UNet:
    """A 3-level 3-D U-Net on an (nx, ny, nz) grid (each divisible by 4): BF16
    weights and activations on MFMA, FP32 accumulation and FP32 master
    weights with Adam. in_channels 1 (the source) or 2 (the source and the
    fluid mask). Periodic in x and y, zero-padded in z."""

    def __init__(self, shape, in_channels=2, seed=0):
        self.shape = tuple(int(s) for s in shape)
        self.in_channels = in_channels
        h = ctypes.c_void_p()
        L.check(L.init().aeo_unet_create(*self.shape, in_channels, ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self._h = h.value
        n = ctypes.c_longlong()
        L.check(L.init().aeo_unet_num_params(self.handle, ctypes.byref(n)))
        self.num_params = n.value

    @property
    def handle(self):
        return ctypes.c_void_p(self._h)

    def params(self):
        p = np.empty(self.num_params, dtype=np.float32)
        L.check(L.init().aeo_unet_get_params(self.handle, p.ctypes.data_as(ctypes.c_void_p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError(f"{p.size} parameters given, the network has {self.num_params}")
        L.check(L.init().aeo_unet_set_params(self.handle, p.ctypes.data_as(ctypes.c_void_p)))

    def _mask(self, mask):
        if self.in_channels == 2 and mask is None:
            raise ValueError("this network takes the fluid mask as its second input")
        return None if mask is None else L.dptr(mask)

    def forward(self, src, mask, out):
        """out = N(src, mask) on FP64 interior-sized DeviceArrays."""
        L.check(L.init().aeo_unet_forward(self.handle, L.dptr(src), self._mask(mask), L.dptr(out)))
        return out

    def __call__(self, src, mask=None):
        """Convenience: numpy in, numpy out."""
        s = DeviceArray.from_numpy(np.asarray(src, dtype=np.float64))
        m = None if mask is None else DeviceArray.from_numpy(np.asarray(mask, dtype=np.float64))
        out = DeviceArray(self.shape)
        self.forward(s, m, out)
        return out.download()

    def gradients(self, src, target, mask=None):
        """(loss, gradient) of one sample (DeviceArrays), without an update."""
        g = np.empty(self.num_params, dtype=np.float32)
        loss = ctypes.c_float()
        L.check(L.init().aeo_unet_gradients(self.handle, L.dptr(src), self._mask(mask), L.dptr(target),
                                            g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss)))
        return loss.value, g

    def train_step(self, src, target, mask=None, lr=1e-3, want_loss=True):
        loss = ctypes.c_float()
        L.check(L.init().aeo_unet_train_step(self.handle, L.dptr(src), self._mask(mask), L.dptr(target),
                                             ctypes.c_float(lr), ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def __del__(self):
        try:
            if self._h and L._lib is not None and L._initialised:
                L._lib.aeo_unet_destroy(ctypes.c_void_p(self._h))
        except Exception:
            pass


class FlexibleCG(PressureSolver):
    """Flexible CG on the operator of an MGPreconditionedCG `base`, with the
    network as initial guess (guess=True) and preconditioner: mode "net",
    "net+vcycle" (the network, then one V-cycle of `base` on the residual it
    leaves) or "vcycle" (the V-cycle alone, the method without the network)."""

    kind = "fcg"

    def __init__(self, base, net=None, mask=None, mode="net+vcycle", guess=True, tol=1e-8, maxit=200):
        h = ctypes.c_void_p()
        L.check(L.init().aeo_poisson_fcg_create(base.handle, None if net is None else net.handle,
                                                None if mask is None else L.dptr(mask), MODES[mode], int(guess),
                                                float(tol), int(maxit), ctypes.byref(h)))
        self._keep = (base, net, mask)
        self.axis_bcs = base.axis_bcs
        super().__init__(h.value, (base.Nx, base.Ny, base.Nz))


def train(net, samples, epochs=1, lr=1e-3, mask=None, shuffle_seed=0, log=None):
    """Adam over (source, target[, mask]) DeviceArray samples; returns the mean
    loss of each epoch."""
    rng = np.random.default_rng(shuffle_seed)
    out = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for i in order:
            s = samples[i]
            m = s[2] if len(s) > 2 else mask
            tot += net.train_step(s[0], s[1], m, lr=lr)
        out.append(tot / len(samples))
        if log:
            log(f"epoch {ep}: mean loss {out[-1]:.4e}")
    return out
