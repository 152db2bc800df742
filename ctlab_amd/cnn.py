"""Learned proximal CNN on BF16 MFMA (libctlab_mi355x.so): 5 layers of 3x3x3
convolution 1 -> 32 -> 32 -> 32 -> 32 -> 1, ReLU after layers 1-4, output
x + net(x).  Training (forward, backward, Adam) and inference run in the
library's own HIP kernels; tensors are float32 torch tensors in HBM with the
volume's (Z,X,Y) layout.  No counterpart in CTLAB."""
import ctypes as C

import numpy as np

from . import _lib
from .dd import _stream_handle

CH = [1, 32, 32, 32, 32, 1]


def num_params():
    return _lib.lib().ctlab_cnn_num_params()


def init_params(seed):
    """He-normal weights (last layer scaled by 0.1 so the net starts near the
    identity map), zero biases; flat FP32 vector W1 b1 ... W5 b5, W[27][cout][cin]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for l in range(5):
        cin, cout = CH[l], CH[l + 1]
        std = np.sqrt(2.0 / (27 * cin)) * (0.1 if l == 4 else 1.0)
        parts.append(rng.standard_normal(27 * cout * cin) * std)
        parts.append(np.zeros(cout))
    return np.concatenate(parts).astype(np.float32)


def _p(t):
    return C.c_void_p(t.data_ptr())


class CNN:
    def __init__(self, shape, training=False, params=None, gpunum=0):
        """shape = (YN, XN, ZN) of the volumes this object will process."""
        YN, XN, ZN = shape
        self.shape = tuple(shape)
        h = _lib.lib().ctlab_cnn_create(XN, YN, ZN, int(training), gpunum)
        if not h:
            _lib.check()
            raise _lib.CtlabError(-1, "cnn creation failed")
        self.h = C.c_void_p(h)
        self.training = training
        self.set_params(init_params(0) if params is None else params)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == num_params()
        _lib.check(_lib.lib().ctlab_cnn_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def params(self):
        p = np.zeros(num_params(), dtype=np.float32)
        _lib.check(_lib.lib().ctlab_cnn_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def __call__(self, x, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(x)
        _lib.check(_lib.lib().ctlab_cnn_forward(self.h, _p(x), _p(out), _stream_handle(stream)))
        return out

    def loss_grad(self, x, target, want_grad=False, out=None, stream=None):
        g = np.zeros(num_params(), dtype=np.float32) if want_grad else None
        loss = _lib.lib().ctlab_cnn_loss_grad(self.h, _p(x), _p(target), _p(out) if out is not None else None,
                                              g.ctypes.data_as(C.POINTER(C.c_float)) if want_grad else None,
                                              _stream_handle(stream))
        _lib.check()
        return (loss, g) if want_grad else loss

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        _lib.check(_lib.lib().ctlab_cnn_adam_step(self.h, lr, beta1, beta2, eps, _stream_handle(stream)))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ctlab_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
