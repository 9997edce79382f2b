"""Host side of the path-guiding network (mi_guide_*): parameters, BF16 MFMA
inference and Adam training on device arrays owned by torch."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check

NUM_PARAMS = 24800


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _dev(a, cols=None) -> torch.Tensor:
    t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a, np.float32))
    t = t.to(device="cuda", dtype=torch.float32).contiguous()
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"expected (n, {cols}) array")
    return t


def _column(w):
    return w.reshape(-1, 1) if isinstance(w, torch.Tensor) else np.asarray(w).reshape(-1, 1)


class Guide:
    def __init__(self, seed: int = 1):
        h = ctypes.c_uint32(_lib.MI_NULL)
        check(_lib.load().mi_guide_create(seed, ctypes.byref(h)), "mi_guide_create")
        self.handle = h.value

    def release(self) -> None:
        if self.handle != _lib.MI_NULL:
            check(_lib.load().mi_guide_release(self.handle), "mi_guide_release")
            self.handle = _lib.MI_NULL

    def set_params(self, flat: np.ndarray) -> None:
        p = np.ascontiguousarray(flat, np.float32)
        assert p.size == NUM_PARAMS
        check(_lib.load().mi_guide_set_params(self.handle, p.ctypes.data), "mi_guide_set_params")

    def set_frequencies(self, freq) -> None:
        """The encoding multipliers, (9, 12): feature pair (i, j) is sin / cos
        of pi freq[i][j] x_i."""
        f = np.ascontiguousarray(freq, np.float32)
        assert f.size == 9 * 12
        check(_lib.load().mi_guide_set_frequencies(self.handle, f.ctypes.data),
              "mi_guide_set_frequencies")

    def params(self) -> np.ndarray:
        p = np.zeros(NUM_PARAMS, np.float32)
        check(_lib.load().mi_guide_get_params(self.handle, p.ctypes.data), "mi_guide_get_params")
        return p

    def forward(self, x) -> torch.Tensor:
        xd = _dev(x, 9)
        out = torch.empty((xd.shape[0], 32), dtype=torch.float32, device="cuda")
        check(_lib.load().mi_guide_forward(self.handle, _ptr(xd), xd.shape[0], _ptr(out),
                                           _stream()), "mi_guide_forward")
        return out

    def gradients(self, x, dirs, weights) -> torch.Tensor:
        xd, dd, wd = _dev(x, 9), _dev(dirs, 3), _dev(_column(weights), 1)
        g = torch.empty(NUM_PARAMS, dtype=torch.float32, device="cuda")
        check(_lib.load().mi_guide_gradients(self.handle, _ptr(xd), _ptr(dd), _ptr(wd),
                                             xd.shape[0], _ptr(g), _stream()),
              "mi_guide_gradients")
        return g

    def train_step(self, x, dirs, weights, lr: float = 1e-3) -> None:
        xd, dd, wd = _dev(x, 9), _dev(dirs, 3), _dev(_column(weights), 1)
        check(_lib.load().mi_guide_train_step(self.handle, _ptr(xd), _ptr(dd), _ptr(wd),
                                              xd.shape[0], lr, _stream()), "mi_guide_train_step")


def flatten(params) -> np.ndarray:
    """oracle.guide parameter list -> the flat device layout."""
    out = []
    for w, b in params:
        out.append(np.asarray(w, np.float64).reshape(-1))
        out.append(np.asarray(b, np.float64).reshape(-1))
    return np.concatenate(out).astype(np.float32)


def unflatten(flat, shapes=((224, 64), (64, 64), (64, 64), (64, 32))) -> list:
    params, o = [], 0
    for i, n in shapes:
        w = np.asarray(flat[o:o + i * n], np.float64).reshape(i, n)
        o += i * n
        b = np.asarray(flat[o:o + n], np.float64)
        o += n
        params.append([w, b])
    return params
