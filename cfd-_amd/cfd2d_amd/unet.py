"""The U-Net of the learned pressure preconditioner, on the GPU.

The network has no counterpart in the reference: it is the learned part of
the flexible PCG (``pressure_solver="cg_learned"``).  Four levels of 3x3
convolutions with channels 16-32-64-64, ReLU, 2x2 average pooling and nearest
upsampling with skip concatenation; BF16 weights and activations with FP32
accumulation on the matrix cores; FP32 master parameters trained with Adam by
the library's own forward and backward kernels (``cfd_unet_*`` in
include/cfd2d_hip.h).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

# type, cin1, cin2, cout: 0 = 1->C input layer, 1 = matrix-core convolution,
# 2 = C->1 linear output layer (the table of csrc/unet_net.hip)
LAYERS = [
    (0, 1, 0, 16), (1, 16, 0, 16), (1, 16, 0, 32), (1, 32, 0, 32), (1, 32, 0, 64),
    (1, 64, 0, 64), (1, 64, 0, 64), (1, 64, 0, 64), (1, 64, 64, 64), (1, 64, 0, 64),
    (1, 64, 32, 32), (1, 32, 0, 32), (1, 32, 16, 16), (1, 16, 0, 16), (2, 16, 0, 1),
]


def param_count() -> int:
    return sum(9 * (c1 + c2) * co + co for _t, c1, c2, co in LAYERS)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (std sqrt(2 / fan_in); sqrt(1 / fan_in) for the
    linear output layer), zero biases, as one flat float32 vector."""
    rng = np.random.default_rng(seed)
    out = []
    for t, c1, c2, co in LAYERS:
        fan_in = 9 * (c1 + c2)
        std = np.sqrt((1.0 if t == 2 else 2.0) / fan_in)
        out.append(rng.standard_normal(9 * (c1 + c2) * co) * std)
        out.append(np.zeros(co))
    return np.concatenate(out).astype(np.float32)


class UNet:
    """A network instance for H x W images (multiples of 8) on the device."""

    def __init__(self, H: int, W: int, params: np.ndarray | None = None, seed: int = 0):
        L = _lib.lib()
        if L.cfd_unet_param_count() != param_count():
            raise RuntimeError("library and host disagree on the U-Net layout")
        self.H, self.W = int(H), int(W)
        h = ctypes.c_void_p()
        _lib.check(L.cfd_unet_create(self.H, self.W, ctypes.byref(h)), "U-Net create")
        self._h = h
        _lib.track(self)
        self.set_params(init_params(seed) if params is None else params)

    # ---- parameters ----
    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != param_count():
            raise ValueError(f"expected {param_count()} parameters, got {p.size}")
        _lib.check(_lib.lib().cfd_unet_set_params(self._h, _lib.ptr(p)), "U-Net set_params")

    def get_params(self) -> np.ndarray:
        out = np.empty(param_count(), dtype=np.float32)
        _lib.check(_lib.lib().cfd_unet_get_params(self._h, _lib.ptr(out)), "U-Net get_params")
        return out

    def get_grad(self) -> np.ndarray:
        out = np.empty(param_count(), dtype=np.float32)
        _lib.check(_lib.lib().cfd_unet_get_grad(self._h, _lib.ptr(out)), "U-Net get_grad")
        return out

    def zero_grad(self) -> None:
        _lib.check(_lib.lib().cfd_unet_zero_grad(self._h), "U-Net zero_grad")

    # ---- compute (device arrays of float32, shape (H, W)) ----
    def _check_img(self, a: _lib.DeviceArray) -> None:
        if a.dtype != np.float32 or tuple(a.shape) != (self.H, self.W):
            raise ValueError(f"expected a float32 ({self.H}, {self.W}) device array")

    def forward(self, x: _lib.DeviceArray, y: _lib.DeviceArray) -> None:
        self._check_img(x)
        self._check_img(y)
        _lib.check(_lib.lib().cfd_unet_forward(self._h, x.ptr, y.ptr), "U-Net forward")

    def loss_grad(self, x: _lib.DeviceArray, t: _lib.DeviceArray,
                  mask: _lib.DeviceArray | None = None, tnorm2: float = 1.0) -> float:
        """sum mask (net(x) - t)^2 / tnorm2; adds its gradient to the network's."""
        self._check_img(x)
        self._check_img(t)
        if mask is not None:
            self._check_img(mask)
        loss = ctypes.c_double()
        _lib.check(_lib.lib().cfd_unet_loss_grad(self._h, x.ptr, t.ptr,
                                                 None if mask is None else mask.ptr,
                                                 float(tnorm2), ctypes.byref(loss)),
                   "U-Net loss_grad")
        return float(loss.value)

    def set_loss(self, kind: str = "mse", region: tuple[int, int] | None = None,
                 eps: float = 0.0) -> None:
        """'mse': masked squared error; 'energy': (y - t).A(y - t) on the
        region (rows, cols) with A the unit 5-point Laplacian (zero-flux
        sides) + eps I."""
        k = {"mse": 0, "energy": 1}[kind]
        nx, ny = region if region is not None else (self.H, self.W)
        _lib.check(_lib.lib().cfd_unet_set_loss(self._h, k, int(nx), int(ny), float(eps)),
                   "U-Net set_loss")

    def adam(self, lr: float, scale: float = 1.0) -> None:
        _lib.check(_lib.lib().cfd_unet_adam(self._h, float(lr), float(scale)), "U-Net adam")

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            _lib.lib().cfd_unet_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        if _lib.closing_allowed():
            try:
                self.close()
            except Exception:
                pass


def mfma_layout_check(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """A @ B through one bf16 matrix-core instruction: 16 x 16 operands on
    v_mfma_f32_16x16x16_bf16, or A 16 x 32 and B 32 x 16 on
    v_mfma_f32_16x16x32_bf16 fed with two 16-deep K steps as the
    convolutions feed it."""
    a = np.ascontiguousarray(A, dtype=np.float32)
    b = np.ascontiguousarray(B, dtype=np.float32)
    d = np.empty((16, 16), dtype=np.float32)
    if a.shape == (16, 16) and b.shape == (16, 16):
        fn = _lib.lib().cfd_mfma_layout_check
    elif a.shape == (16, 32) and b.shape == (32, 16):
        fn = _lib.lib().cfd_mfma_layout_check_k32
    else:
        raise ValueError("operands 16 x 16 and 16 x 16, or 16 x 32 and 32 x 16 expected")
    _lib.check(fn(_lib.ptr(a), _lib.ptr(b), _lib.ptr(d)), "MFMA layout check")
    return d
