"""numpy restatement of the reference's grid <-> image transforms
(TEST INFRASTRUCTURE ONLY).

MS4/kernel/cpu/fft/fft_dyn_padded.cpp:26-115 wraps FFTW (a system library,
unversioned in the reference, absent here): fftshift_even, then an
unnormalised guru transform in place, then fftshift_even, with rows `pitch`
elements apart for both the real and the complex view of one buffer.  FFTW
itself cannot run here, so these functions state the mathematics with
numpy.fft (pocketfft): parity of the FFT arithmetic is pinned by numpy's
own transforms, not by reference vectors (none exist for this path).

The in-place buffer is a (size, pitch) complex128 array; its real view is
the first size*pitch doubles read as (size, pitch).
"""
from __future__ import annotations

import numpy as np


def fftshift_even(a: np.ndarray) -> np.ndarray:
    """Quadrant swap of an even square array (fft_dyn_padded.cpp:26-48),
    equal to numpy.fft.fftshift over both axes for even sizes."""
    n = a.shape[0]
    h = n // 2
    out = np.empty_like(a)
    out[:h, :h], out[h:, h:] = a[h:, h:], a[:h, :h]
    out[:h, h:], out[h:, :h] = a[h:, :h], a[:h, h:]
    return out


def real_view(buf: np.ndarray, size: int, pitch: int) -> np.ndarray:
    return buf.reshape(-1).view(np.float64)[: size * pitch].reshape(size, pitch)


def fft_inplace_even(buf: np.ndarray, sign: int, size: int) -> np.ndarray:
    """Returns the buffer after the transform (a copy); buf is (size, pitch)
    complex128.  sign -1/+1: complex forward/backward over the size x size
    region; 0: real (size x size of the real view) -> half spectrum in
    columns 0..size/2 (the other columns keep their bytes); 2: half spectrum
    -> real image in the real view."""
    out = np.array(buf, dtype=np.complex128, copy=True)
    pitch = out.shape[1]
    n = size
    if sign in (-1, 1):
        a = fftshift_even(out[:n, :n])
        a = np.fft.fft2(a) if sign == -1 else np.fft.ifft2(a) * (n * n)
        out[:n, :n] = fftshift_even(a)
        return out
    if sign == 0:
        rv = real_view(out, n, pitch)
        r = fftshift_even(rv[:n, :n].copy())
        rv[:n, :n] = r
        h = np.fft.rfft2(r)
        out[:n, : n // 2 + 1] = h
        out[:n, :n] = fftshift_even(out[:n, :n].copy())
        return out
    if sign == 2:
        s = fftshift_even(out[:n, :n].copy())
        out[:n, :n] = s
        img = np.fft.irfft2(s[:, : n // 2 + 1], s=(n, n)) * (n * n)
        rv = real_view(out, n, pitch)
        rv[:n, :n] = fftshift_even(img)
        return out
    raise ValueError(sign)
