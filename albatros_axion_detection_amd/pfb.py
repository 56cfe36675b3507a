"""PFB, inverse PFB and FFTs on the GPU -- the call surface of
src/utils/pfb_utils.py, src/utils/pfb_gpu_utils.py, src/utils/pfb_cpu_utils.py
and src/utils/pycufft.py.

Tensors are torch tensors on the GPU (numpy inputs are copied in and the
results returned as numpy).  The window spectrum and the Wiener filter are
computed in float64 on the GPU and stored as complex64.
"""
import numpy as np

from . import _lib


# ------------------------------------------------------------------ windows
def sinc_hamming(ntap, lblock):
    """pfb_utils.py:4-7 (float64 host window; callers cast to float32)."""
    N = ntap * lblock
    w = np.arange(0, N) - N / 2
    return np.hamming(ntap * lblock) * np.sinc(w / lblock)


def sinc_hanning(ntap, lblock):
    """pfb_utils.py:9-12."""
    N = ntap * lblock
    w = np.arange(0, N) - N / 2
    return np.hanning(ntap * lblock) * np.sinc(w / lblock)


def _sinc_hamming_f32(ntap, lblock, chunk=1 << 24):
    """``sinc_hamming(ntap, lblock).astype(float32)`` evaluated a chunk at a
    time: numpy's hamming is ``0.54 + 0.46 cos(pi n / (N - 1))`` with
    ``n = 1 - N, 3 - N, ...`` and its sinc ``sin(y) / y``, ``y = pi x`` (x = 0
    as 1e-20), so the same float64 operations per element give the same
    values without the full-length float64 temporaries (the osamp 65536 window
    has 2^30 taps)."""
    N = ntap * lblock
    out = np.empty(N, dtype=np.float32)
    if N == 1:
        out[0] = np.float32(np.sinc(-0.5 / lblock))
        return out
    for a in range(0, N, chunk):
        i = np.arange(a, min(N, a + chunk))
        n = 1 - N + 2 * i
        ham = 0.54 + 0.46 * np.cos(np.pi * n / (N - 1))
        x = (i - N / 2) / lblock
        y = np.pi * np.where(x == 0, 1.0e-20, x)
        out[a:a + len(i)] = ham * (np.sin(y) / y)
    return out


def window(ntap, lblock, device_tensor=True):
    """float32 sinc-Hamming window as the reference hands it to the GPU
    (helper_gpu_stream.py:74-75)."""
    w = _sinc_hamming_f32(ntap, lblock)
    return _lib.to_device(w) if device_tensor else w


# -------------------------------------------------------- window spectrum
def get_matft(nslice, nchan=2049, ntap=4):
    """pfb_gpu_utils.py:123-137: complex64 ``[2(nchan-1), nslice//2+1]``,
    ``W[k,f] = sum_j w[j*lblock+k] exp(-2 pi i f j / nslice)`` (float64 math)."""
    t = _lib.torch()
    nn = 2 * (nchan - 1)
    w = window(ntap, nn)
    out = t.empty((nn, nslice // 2 + 1), dtype=t.complex64, device=w.device)
    _lib.call("alb_matft_filter", _lib.ptr(w), ntap, nn, int(nslice), 0.0, _lib.ptr(out), _lib.ptr(None),
              _lib.stream())
    return out


def get_matft_cpu(nslice, nchan=2049, ntap=4, n_workers=None):
    """pfb_cpu_utils.py:175-203 call surface (computed on the GPU, numpy out)."""
    return _lib.host(get_matft(nslice, nchan, ntap))


def get_filter(nslice, thresh, nchan=2049, ntap=4, win=None):
    """Wiener filter straight from the window in float64:
    ``W (1+t^2) / (|W|^2 + t^2)`` (= ``calculate_filter(get_matft(nslice), thresh)``
    without the float32 rounding of W)."""
    t = _lib.torch()
    nn = 2 * (nchan - 1)
    w = window(ntap, nn) if win is None else _lib.to_device(win, t.float32)
    out = t.empty((nn, nslice // 2 + 1), dtype=t.complex64, device=w.device)
    _lib.call("alb_matft_filter", _lib.ptr(w), ntap, nn, int(nslice), float(thresh), _lib.ptr(None), _lib.ptr(out),
              _lib.stream())
    return out


def calculate_filter(matft, thresh):
    """pfb_gpu_utils.py:29-36: ``1/conj(W)``, times ``|W|^2/(t^2+|W|^2)(1+t^2)``
    for ``thresh > 0`` (float64 arithmetic, complex64 result)."""
    t = _lib.torch()
    isnp = _lib.is_numpy(matft)
    W = _lib.to_device(matft, t.complex64)
    out = t.empty_like(W)
    _lib.call("alb_filter_from_matft", _lib.ptr(W), W.numel(), float(thresh), _lib.ptr(out), _lib.stream())
    return _lib.host(out) if isnp else out


compute_filter = calculate_filter  # pfb_cpu_utils.py:58-75 (same formula)


# ---------------------------------------------------------------- pycufft
def _lastaxis(x, axis):
    if axis != x.ndim - 1 and axis != -1:
        x = x.swapaxes(axis, -1)
    return x


def rfft(inp, axis=1, use_cache=True, copy=False):
    """pycufft.py:120-137: unnormalised real FFT along ``axis``."""
    t = _lib.torch()
    isnp = _lib.is_numpy(inp)
    x = _lib.to_device(_lastaxis(inp, axis), t.float32) if isnp else _lastaxis(inp, axis).to(t.float32).contiguous()
    n = x.shape[-1]
    rows = x.numel() // n
    out = t.empty(x.shape[:-1] + (n // 2 + 1,), dtype=t.complex64, device=x.device)
    _lib.call("alb_rfft_rows", _lib.ptr(x), n, _lib.ptr(out), n // 2 + 1, rows, n, 0, n // 2 + 1, _lib.stream())
    out = _lastaxis(out, axis)
    return _lib.host(out) if isnp else out


def irfft(inp, axis=1, use_cache=True, copy=False):
    """pycufft.py:139-161: inverse real FFT along ``axis`` with the 1/n of
    pycufft.py:153; output length ``2*(m-1)``."""
    t = _lib.torch()
    isnp = _lib.is_numpy(inp)
    x = _lib.to_device(_lastaxis(inp, axis), t.complex64) if isnp else _lastaxis(inp, axis).to(t.complex64).contiguous()
    m = x.shape[-1]
    n = 2 * (m - 1)
    rows = x.numel() // m
    out = t.empty(x.shape[:-1] + (n,), dtype=t.float32, device=x.device)
    _lib.call("alb_irfft_rows", _lib.ptr(x), m, _lib.ptr(out), n, rows, n, _lib.stream())
    out = _lastaxis(out, axis)
    return _lib.host(out) if isnp else out


def _fft(inp, axis=1, direction=-1):
    t = _lib.torch()
    isnp = _lib.is_numpy(inp)
    x = _lib.to_device(_lastaxis(inp, axis), t.complex64) if isnp else _lastaxis(inp, axis).to(t.complex64).contiguous()
    n = x.shape[-1]
    out = t.empty_like(x)
    if n > 16384:
        # four-step form above one workgroup's 16384 points (2^15..2^24)
        work = t.empty_like(x)
        _lib.call("alb_fft_large", _lib.ptr(x), _lib.ptr(work), _lib.ptr(out), n, x.numel() // n, int(direction),
                  _lib.stream())
    else:
        _lib.call("alb_fft_rows", _lib.ptr(x), n, _lib.ptr(out), n, x.numel() // n, n, int(direction), _lib.stream())
    out = _lastaxis(out, axis)
    return _lib.host(out) if isnp else out


def fft(inp, axis=1, use_cache=True, copy=False):
    """pycufft.py:114-115 (complex forward, unnormalised)."""
    return _fft(inp, axis, -1)


def ifft(inp, axis=1, use_cache=True, copy=False):
    """pycufft.py:117-118 (complex inverse with 1/n)."""
    return _fft(inp, axis, 1)


# ------------------------------------------------------------ inverse PFB
_BLUESTEIN = {}


def bluestein_tables(n):
    """Chirp-z tables of an n-point DFT for the column pass at non power-of-two
    nslice: chirp ``exp(-i pi m^2 / n)`` and the P-point FFT of its conjugate
    filter (float64 on the host, stored complex64 on the device)."""
    dev = _lib.device()
    key = (n, str(dev))
    if key not in _BLUESTEIN:
        P = 1
        while P < 2 * n - 1:
            P *= 2
        m = np.arange(n, dtype=np.int64)
        ang = -np.pi * ((m * m) % (2 * n)).astype(np.float64) / n
        chirp = np.exp(1j * ang)
        b = np.zeros(P, dtype=np.complex128)
        b[:n] = np.conj(chirp)
        b[P - n + 1:] = np.conj(chirp[1:])[::-1]
        bhat = np.fft.fft(b)
        _BLUESTEIN[key] = (_lib.to_device(chirp.astype(np.complex64)), _lib.to_device(bhat.astype(np.complex64)), P)
    return _BLUESTEIN[key]


def ipfb_columns(y, ldy, tile_stride, ntiles, nslice, ncols, filt, cut, out, ldo, out_tile_stride):
    """Column pass of the inverse PFB (pfb_gpu_utils.py:84-91) over ``ntiles``
    tiles: power-of-two nslice on the radix-16 path, any even nslice on the
    chirp-z path."""
    n = int(nslice)
    if n & (n - 1) == 0:
        _lib.call("alb_ipfb_columns", _lib.ptr(y), ldy, tile_stride, ntiles, n, ncols, _lib.ptr(filt), filt.shape[1],
                  int(cut), _lib.ptr(out), ldo, out_tile_stride, _lib.stream())
    else:
        chirp, bhat, P = bluestein_tables(n)
        _lib.call("alb_ipfb_columns_any", _lib.ptr(y), ldy, tile_stride, ntiles, n, ncols, _lib.ptr(filt),
                  filt.shape[1], int(cut), _lib.ptr(out), ldo, out_tile_stride, _lib.ptr(chirp), _lib.ptr(bhat), P,
                  _lib.stream())


def cupy_ipfb(dat, filt, thresh=None):
    """pfb_gpu_utils.py:56-91: ``irfft`` rows -> ``rfft`` along time ->
    ``* filt`` -> ``irfft`` along time; returns float32 ``(nslice, 2(nchan-1))``.

    Also accepts the older ``cupy_ipfb(dat, matft, thresh=...)`` form
    (scripts/ipfb/upchannelize.py:94): then ``filt`` is a window spectrum and
    the filter is derived from it.
    """
    t = _lib.torch()
    isnp = _lib.is_numpy(dat)
    d = _lib.to_device(dat, t.complex64)
    f = _lib.to_device(filt, t.complex64)
    if thresh is not None:
        f = calculate_filter(f, thresh)
    nslice, nchan = d.shape
    n = 2 * (nchan - 1)
    if tuple(f.shape) != (n, nslice // 2 + 1):
        raise ValueError("filter shape %s does not match (%d, %d)" % (tuple(f.shape), n, nslice // 2 + 1))
    if nslice % 2:
        raise ValueError("nslice must be even (the inverse real FFT along time returns 2*(nslice//2) rows, "
                         "pycufft.py:146)")
    y = t.empty((nslice, n), dtype=t.float32, device=d.device)
    _lib.call("alb_irfft_rows", _lib.ptr(d), nchan, _lib.ptr(y), n, nslice, n, _lib.stream())
    ipfb_columns(y, n, 0, 1, nslice, n, f, 0, y, n, 0)
    return _lib.host(y) if isnp else y


ipfb = cupy_ipfb


def cpu_ipfb(dat, out, filt, cut=0, n_workers=None):
    """pfb_cpu_utils.py:90-121 call surface: writes rows ``[cut:-cut]`` of the
    inverse PFB, transposed back and ravelled, into ``out``."""
    res = cupy_ipfb(dat, filt)
    nslice = res.shape[0]
    kept = res[cut:nslice - cut]
    if _lib.is_numpy(out):
        out[:] = (kept if _lib.is_numpy(kept) else _lib.host(kept)).ravel()
    else:
        out.copy_(kept.reshape(-1) if not _lib.is_numpy(kept) else _lib.to_device(kept).reshape(-1))
    return out


# ------------------------------------------------------------ forward PFB
def pfb_channels(timestream, win, nchan=2049, ntap=4, c0=0, c1=None, out=None, method="auto"):
    """``cupy_pfb(timestream, win, nchan, ntap)[:, c0:c1]`` computed directly
    (pfb_gpu_utils.py:105-121; the slice of helper_gpu_stream.py:222).
    ``method``: "small" (one workgroup per spectrum, lblock <= 16384),
    "repfb" (four-step, lblock = 4096*osamp) or "auto"."""
    t = _lib.torch()
    lblock = 2 * (nchan - 1)
    if c1 is None:
        c1 = nchan
    ts = _lib.to_device(timestream, t.float32).reshape(-1)
    w = _lib.to_device(win, t.float32).reshape(-1)
    if w.numel() != ntap * lblock:
        raise ValueError("window length %d != ntap*lblock %d" % (w.numel(), ntap * lblock))
    nblock = ts.numel() // lblock - (ntap - 1)
    if nblock <= 0:
        return t.empty((0, c1 - c0), dtype=t.complex64, device=ts.device)
    if out is None:
        out = t.empty((nblock, c1 - c0), dtype=t.complex64, device=ts.device)
    if method == "small" or (method == "auto" and lblock <= 16384):
        _lib.call("alb_pfb_small", _lib.ptr(ts), nblock, lblock, _lib.ptr(w), ntap, int(c0), int(c1), _lib.ptr(out),
                  out.stride(0), _lib.stream())
    else:
        osamp = lblock // 4096
        if lblock != 4096 * osamp:
            raise NotImplementedError("re-PFB block length must be 4096*osamp")
        ws = t.empty(int(_lib.load().alb_repfb_workspace_bytes(nblock, osamp)) // 8, dtype=t.complex64,
                     device=ts.device)
        _lib.call("alb_repfb", _lib.ptr(ts), _lib.ptr(w), ntap, nblock, osamp, _lib.ptr(ws), int(c0), int(c1),
                  _lib.ptr(out), out.stride(0), out.stride(1), nblock, 0, _lib.stream())
    return out


def cupy_pfb(timestream, win, out=None, nchan=2049, ntap=4):
    """pfb_gpu_utils.py:105-121: complex64 ``(nblock, nchan)``,
    ``nblock = len/lblock - (ntap-1)``."""
    isnp = _lib.is_numpy(timestream)
    r = pfb_channels(timestream, win, nchan, ntap, 0, nchan)
    if out is not None:
        assert tuple(out.shape) == tuple(r.shape)
        out[...] = _lib.host(r) if _lib.is_numpy(out) else r
        return out
    return _lib.host(r) if isnp else r


pfb = cupy_pfb


def cpu_pfb(timestream, win, scratch=None, nchan=2049, ntap=4, n_workers=None):
    """pfb_cpu_utils.py:143-172 call surface (numpy out)."""
    r = cupy_pfb(timestream, win, nchan=nchan, ntap=ntap)
    return r if _lib.is_numpy(r) else _lib.host(r)
