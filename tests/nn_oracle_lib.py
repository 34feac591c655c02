"""ctypes loader of oracle/libnn_oracle.so: the CPU FP64 restatement of the
learned correction's network (forward, loss, gradient).  Test infrastructure
only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "libnn_oracle.so")
_lib = C.CDLL(PATH)
_P = C.c_void_p
_lib.nno_num_params.restype = C.c_int
_lib.nno_forward.argtypes = [_P, C.c_int, _P, _P, C.c_int, _P]
_lib.nno_forward.restype = None
_lib.nno_loss_grad.argtypes = [_P, C.c_int, _P, _P, _P, C.c_int, C.POINTER(C.c_double), _P]
_lib.nno_loss_grad.restype = None
NPARAM = _lib.nno_num_params()


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def forward(w, x, mask, round_bf16=True):
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    mask = np.ascontiguousarray(mask, np.uint8)
    out = np.zeros((x.shape[0], 4096))
    _lib.nno_forward(_p(w), x.shape[0], _p(x), _p(mask), 1 if round_bf16 else 0, _p(out))
    return out


def loss_grad(w, x, target, mask, round_bf16=True):
    w = np.ascontiguousarray(w, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    mask = np.ascontiguousarray(mask, np.uint8)
    loss = C.c_double()
    g = np.zeros(NPARAM)
    _lib.nno_loss_grad(_p(w), x.shape[0], _p(x), _p(target), _p(mask), 1 if round_bf16 else 0, C.byref(loss), _p(g))
    return loss.value, g
