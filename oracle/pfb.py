"""Float64 restatement of the reference PFB / inverse-PFB path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference computes in float32 (CuPy + cuFFT).  The oracle computes the
same steps in float64 from the same float32 inputs (the window is rounded
to float32 exactly as the reference callers do, helper_gpu_stream.py:74-75
and pfb_gpu_utils.py:126).
"""
import numpy as np


def sinc_hamming(ntap, lblock):
    """pfb_utils.py:4-7: ``hamming(N) * sinc((arange(N) - N/2) / lblock)``."""
    N = ntap * lblock
    w = np.arange(0, N) - N / 2
    return np.hamming(N) * np.sinc(w / lblock)


def sinc_hanning(ntap, lblock):
    """pfb_utils.py:9-12."""
    N = ntap * lblock
    w = np.arange(0, N) - N / 2
    return np.hanning(N) * np.sinc(w / lblock)


def window_f32(ntap, lblock):
    """The float32 window the reference hands to the GPU
    (pfb_gpu_utils.py:126, helper_gpu_stream.py:74-75)."""
    return sinc_hamming(ntap, lblock).astype(np.float32)


def matft_fp64(nslice, nchan=2049, ntap=4, win=None):
    """pfb_gpu_utils.py:123-137 ``get_matft`` in float64.

    ``mat[nslice, lblock]`` zeros with rows ``0..ntap-1`` = the window
    reshaped ``(ntap, lblock)``; transpose; ``rfft`` along nslice.
    Returns complex128 ``[lblock, nslice//2 + 1]``.
    """
    nn = 2 * (nchan - 1)
    if win is None:
        win = window_f32(ntap, nn)
    w = np.asarray(win, dtype=np.float64).reshape(ntap, nn)
    mat = np.zeros((nslice, nn), dtype=np.float64)
    mat[:ntap] = w
    return np.fft.rfft(mat.T, axis=1)


def matft_closed_form(nslice, nchan=2049, ntap=4, win=None):
    """``W[k, f] = sum_{j<ntap} w[j*lblock + k] exp(-2 pi i f j / nslice)``."""
    nn = 2 * (nchan - 1)
    if win is None:
        win = window_f32(ntap, nn)
    w = np.asarray(win, dtype=np.float64).reshape(ntap, nn)
    f = np.arange(nslice // 2 + 1)
    ph = np.exp(-2j * np.pi * np.outer(np.arange(ntap), f) / nslice)  # [ntap, F]
    return w.T @ ph


def wiener_filter(matft, thresh):
    """pfb_gpu_utils.py:29-36 ``calculate_filter`` (= pfb_cpu_utils.py:58-75):
    ``1/conj(W)``, times ``|W|^2/(t^2+|W|^2)*(1+t^2)`` when ``thresh > 0``."""
    W = np.asarray(matft, dtype=np.complex128)
    filt = 1 / np.conj(W)
    if thresh > 0:
        a = np.abs(W) ** 2
        filt = filt * (a / (thresh ** 2 + a) * (1.0 + thresh ** 2))
    return filt


def ipfb(dat, filt):
    """pfb_gpu_utils.py:56-91 ``cupy_ipfb`` / pfb_cpu_utils.py:90-121.

    ``dd = irfft(dat, axis=1)``; ``ddft = rfft(dd.T, axis=1)``;
    ``ddft *= filt``; ``res = irfft(ddft, axis=1).T`` -> ``[nslice, lblock]``.
    The first ``irfft`` ignores the imaginary parts of the DC and Nyquist
    bins (numpy / C2R convention).
    """
    dd = np.fft.irfft(np.asarray(dat, dtype=np.complex128), axis=1)
    ddft = np.fft.rfft(dd.T, axis=1)
    ddft = ddft * filt
    return np.fft.irfft(ddft, n=dd.shape[0], axis=1).T


def pfb(timestream, win, nchan=2049, ntap=4):
    """pfb_gpu_utils.py:105-121 ``cupy_pfb`` / pfb_cpu_utils.py:123-172.

    ``lblock = 2(nchan-1)``, ``nblock = len/lblock - (ntap-1)``,
    ``y[i,k] = sum_j ts[(i+j)*lblock + k] * win[j*lblock + k]``,
    ``rfft(y, axis=1)`` -> complex ``[nblock, nchan]``.
    """
    lblock = 2 * (nchan - 1)
    ts = np.asarray(timestream, dtype=np.float64)
    nblock = ts.size // lblock - (ntap - 1)
    ts = ts[: (nblock + ntap - 1) * lblock].reshape(-1, lblock)
    w = np.asarray(win, dtype=np.float64).reshape(ntap, lblock)
    y = np.zeros((nblock, lblock))
    for j in range(ntap):
        y += ts[j : nblock + j] * w[j]
    return np.fft.rfft(y, axis=1)


def pfb_channels(timestream, win, nchan, ntap, c0, c1):
    """``pfb(...)[:, c0:c1]`` (the slice the streaming driver keeps,
    helper_gpu_stream.py:222)."""
    return pfb(timestream, win, nchan, ntap)[:, c0:c1]


def dft_brute(x, sign=-1):
    """Brute-force O(n^2) DFT for tiny sizes (pins the FFT convention)."""
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[-1]
    k = np.arange(n)
    M = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return x @ M.T


def irfft_brute(X, n):
    """Brute-force inverse real DFT with the 1/n normalisation of
    pycufft.py:153 and numpy (imag of DC / Nyquist ignored)."""
    X = np.asarray(X, dtype=np.complex128).copy()
    X[..., 0] = X[..., 0].real
    if n % 2 == 0:
        X[..., n // 2] = X[..., n // 2].real
    full = np.zeros(X.shape[:-1] + (n,), dtype=np.complex128)
    full[..., : X.shape[-1]] = X
    for c in range(1, (n + 1) // 2):
        full[..., n - c] = np.conj(X[..., c])
    return dft_brute(full, sign=+1).real / n
