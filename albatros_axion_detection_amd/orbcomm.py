"""ORBCOMM delay search on the GPU -- the call surface of
src/utils/orbcomm_utils_gpu.py: the per-channel FFT lag correlation of two
antennas' coarse spectra (``coarse_xcorr``) and the time-dependent phase of a
trial delay (``apply_delay``), both on libalbatros_hip.
"""
import numpy as np

from . import _lib


def coarse_xcorr(f1, f2, dN, Npfb=4096, window=None):
    """orbcomm_utils_gpu.py:39-52: lag cross-correlation per channel.

    ``f1``, ``f2``: complex64 ``(Nsmall, nchans)`` spectra (1-D: one
    channel).  Returns complex64 ``(nchans, 2 dN)``: column i holds lag
    ``i - dN``, ``sum_t f1[t + lag] conj(f2[t]) / (Nsmall - |lag|)``, which
    the reference computes as the fftshifted inverse FFT of
    ``fft(f1) conj(fft(f2))`` over ``2 Nsmall`` zero-padded points.  Here the
    padding is the power of two >= ``Nsmall + dN`` (no kept lag wraps either
    way, so the values are the same).  ``Npfb`` and ``window`` are accepted
    and unused, as in the reference."""
    t = _lib.torch()
    isnp = _lib.is_numpy(f1)
    a = _lib.to_device(f1, t.complex64)
    b = _lib.to_device(f2, t.complex64)
    if a.dim() == 1:
        a = a.reshape(-1, 1)
    if b.dim() == 1:
        b = b.reshape(-1, 1)
    if tuple(a.shape) != tuple(b.shape):
        raise ValueError("f1 and f2 must have the same shape")
    N, nchan = int(a.shape[0]), int(a.shape[1])
    dN = int(dN)
    if not 1 <= dN <= N:
        raise ValueError("dN must lie in [1, Nsmall]")
    P = 64
    while P < N + dN:
        P *= 2
    if P > (1 << 24):
        raise NotImplementedError("lag correlation longer than 2^24 points")
    work = t.empty(3 * nchan * P, dtype=t.complex64, device=a.device)
    out = t.empty((nchan, 2 * dN), dtype=t.complex64, device=a.device)
    _lib.call("alb_lag_xcorr", _lib.ptr(a), _lib.ptr(b), N, nchan, dN, P, _lib.ptr(work), _lib.ptr(out),
              _lib.stream())
    return _lib.host(out) if isnp else out


def apply_delay(arr, delay, freqs, out=None):
    """orbcomm_utils_gpu.py:13-37: ``arr * exp(-2j pi freqs[None, :] delay[:, None])``
    for ``arr`` complex64 ``(nspec, nchan)`` (the phase in float64, as the
    reference's complex128 expression; the product rounded to complex64).
    ``out``: a complex64 device tensor of the same shape (may be ``arr``)."""
    t = _lib.torch()
    isnp = _lib.is_numpy(arr)
    a = _lib.to_device(arr, t.complex64)
    if a.dim() != 2:
        raise ValueError("arr must be (nspec, nchan)")
    nspec, nchan = int(a.shape[0]), int(a.shape[1])
    d = _lib.to_device(np.asarray(delay, dtype=np.float64) if not isinstance(delay, t.Tensor) else delay, t.float64)
    f = _lib.to_device(np.asarray(freqs, dtype=np.float64) if not isinstance(freqs, t.Tensor) else freqs, t.float64)
    if d.numel() != nspec or f.numel() != nchan:
        raise ValueError("delay must have nspec entries and freqs nchan")
    if out is None:
        o = t.empty_like(a)
    else:
        if tuple(out.shape) != (nspec, nchan) or out.dtype != t.complex64 or not out.is_contiguous():
            raise ValueError("invalid out buffer")
        o = out
    _lib.call("alb_apply_delay", _lib.ptr(a), _lib.ptr(d), _lib.ptr(f), nspec, nchan, _lib.ptr(o), _lib.stream())
    if out is not None:
        return out
    return _lib.host(o) if isnp else o
