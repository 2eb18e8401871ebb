"""The far-field surrogate network (lpm_mlp_*): FP32 training on the GPU,
FP16-MFMA inference.  Parameters use the layout W[out][in] then b[out] per layer."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from ._lib import LpmError

_pf = ctypes.POINTER(ctypes.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


ACT_TANH, ACT_ISRU = 1, 2


class Mlp:
    def __init__(self, ctx, nin=13, width=128, nhidden=3, nout=3, seed=1, act=ACT_TANH):
        self.ctx = ctx
        self.nin, self.width, self.nhidden, self.nout, self.act = nin, width, nhidden, nout, act
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.lpm_mlp_new(ctx.handle, ctypes.byref(h), nin, width, nhidden, nout, act, seed))
        self.handle = h
        ctx.adopt(self)
        n = ctypes.c_int64()
        ctx.check(ctx._lib.lpm_mlp_param_count(h, ctypes.byref(n)))
        self.n_params = n.value

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx._lib.lpm_mlp_get_params(self.handle, p.ctypes.data_as(_pf), self.n_params))
        return p

    def set_params(self, p):
        p = _f32(p)
        if p.shape[0] != self.n_params:
            raise LpmError(L.LPM_ERR_ARG, "wrong parameter count")
        self.ctx.check(self.ctx._lib.lpm_mlp_set_params(self.handle, p.ctypes.data_as(_pf), self.n_params))

    def forward(self, x, mfma=False):
        x = _f32(x).reshape(-1, self.nin)
        out = np.zeros((x.shape[0], self.nout), np.float32)
        self.ctx.check(self.ctx._lib.lpm_mlp_forward(self.handle, x.shape[0], x.ctypes.data_as(_pf),
                                                     out.ctypes.data_as(_pf), int(bool(mfma))))
        return out

    def loss_grad(self, x, t):
        x = _f32(x).reshape(-1, self.nin)
        t = _f32(t).reshape(-1, self.nout)
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_double()
        self.ctx.check(self.ctx._lib.lpm_mlp_loss_grad(self.handle, x.shape[0], x.ctypes.data_as(_pf),
                                                       t.ctypes.data_as(_pf), g.ctypes.data_as(_pf),
                                                       ctypes.byref(loss)))
        return loss.value, g

    def train(self, x, t, epochs=1, batch=4096, lr=1e-3, seed=0):
        x = _f32(x).reshape(-1, self.nin)
        t = _f32(t).reshape(-1, self.nout)
        loss = ctypes.c_double()
        self.ctx.check(self.ctx._lib.lpm_mlp_train(self.handle, x.shape[0], x.ctypes.data_as(_pf),
                                                   t.ctypes.data_as(_pf), epochs, batch, lr, seed,
                                                   ctypes.byref(loss)))
        return loss.value

    def delete(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.lpm_mlp_delete(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.delete()
        except Exception:
            pass
