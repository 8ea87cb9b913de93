"""PNG files for the stereo pipeline's host images, over the C ABI
(rts_png_*): the replacement of the reference driver's OpenCV imread /
imwrite (seq/Main.cpp:72-73, 87-88, 275-282) and its display scaling
(seq/Main.cpp:335-360)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, lib


def _take(ptr: C.c_void_p, W: int, H: int, channels: int) -> np.ndarray:
    shape = (H, W) if channels == 1 else (H, W, 3)
    try:
        buf = (C.c_uint8 * (W * H * channels)).from_address(ptr.value)
        return np.frombuffer(buf, np.uint8).reshape(shape).copy()
    finally:
        lib.rts_png_free(ptr)


def read_png(path: str, channels: int = 3) -> np.ndarray:
    """uint8 [H][W][3] in B,G,R order (channels=3) or gray [H][W] (channels=1)."""
    ptr, W, H = C.c_void_p(), C.c_int(), C.c_int()
    call("rts_png_read", path.encode(), channels, C.byref(ptr), C.byref(W), C.byref(H))
    return _take(ptr, W.value, H.value, channels)


def decode_png(data: bytes, channels: int = 3) -> np.ndarray:
    ptr, W, H = C.c_void_p(), C.c_int(), C.c_int()
    buf = np.frombuffer(data, np.uint8)
    call("rts_png_decode", buf.ctypes.data, len(data), channels, C.byref(ptr), C.byref(W), C.byref(H))
    return _take(ptr, W.value, H.value, channels)


def write_png(path: str, img: np.ndarray) -> None:
    """Writes a gray [H][W] or B,G,R [H][W][3] uint8 image."""
    a = np.ascontiguousarray(img, np.uint8)
    if a.ndim == 2:
        ch = 1
    elif a.ndim == 3 and a.shape[2] == 3:
        ch = 3
    else:
        raise ValueError(f"image of shape {a.shape} is neither gray nor 3-channel")
    call("rts_png_write", path.encode(), a.ctypes.data, a.shape[1], a.shape[0], ch)


def disp_scale(disp: np.ndarray, coeff: float = 0.0) -> np.ndarray:
    """The reference's display scaling of a disparity map (a new array):
    min(round(v * coeff), 255), coeff <= 0 meaning 255 / max."""
    a = np.ascontiguousarray(disp, np.uint8).copy()
    call("rts_disp_scale", a.ctypes.data, a.shape[1], a.shape[0], float(coeff))
    return a
