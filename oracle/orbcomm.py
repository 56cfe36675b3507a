"""Restatement of the ORBCOMM delay-search helpers of
src/utils/orbcomm_utils_gpu.py in numpy (complex128).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
import numpy as np


def coarse_xcorr(f1, f2, dN):
    """orbcomm_utils_gpu.py:39-52: each channel's spectra zero-padded to
    2 Nsmall points, ``ifft(fft(a) * conj(fft(b)))``, fftshift, lags
    ``[-dN, dN)``, divided by ``Nsmall - |lag|``.  Returns ``(nchans, 2 dN)``."""
    a = np.asarray(f1, dtype=np.complex128)
    b = np.asarray(f2, dtype=np.complex128)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
        b = b.reshape(-1, 1)
    N, nch = a.shape
    big1 = np.zeros((nch, 2 * N), dtype=np.complex128)
    big2 = np.zeros((nch, 2 * N), dtype=np.complex128)
    big1[:, :N] = a.T
    big2[:, :N] = b.T
    xc = np.fft.ifft(np.fft.fft(big1, axis=1) * np.conj(np.fft.fft(big2, axis=1)), axis=1)
    norm = N - np.abs(np.arange(-dN, dN))
    return np.fft.fftshift(xc, axes=1)[:, N - dN:N + dN] / norm[None, :]


def lag_sums_brute(f1, f2, dN):
    """The same lags as direct sums ``sum_t f1[t + lag] conj(f2[t]) / (N - |lag|)``
    (small inputs only)."""
    a = np.asarray(f1, dtype=np.complex128)
    b = np.asarray(f2, dtype=np.complex128)
    N, nch = a.shape
    out = np.zeros((nch, 2 * dN), dtype=np.complex128)
    for i, lag in enumerate(range(-dN, dN)):
        t0, t1 = max(0, -lag), min(N, N - lag)
        out[:, i] = np.sum(a[t0 + lag:t1 + lag] * np.conj(b[t0:t1]), axis=0) / (N - abs(lag))
    return out


def apply_delay(arr, delay, freqs):
    """orbcomm_utils_gpu.py:13-37: ``arr * exp(-2j pi freqs[None, :] delay[:, None])``
    in complex128."""
    return np.asarray(arr) * np.exp(-2j * np.pi * np.asarray(freqs)[None, :] * np.asarray(delay)[:, None])
