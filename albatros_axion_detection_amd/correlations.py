"""Correlators on the GPU -- the call surface of src/correlations/correlations.py
and src/correlations/correlations_gpu.py.

Integer correlators return exact int64 sums from the device; the averages are
formed exactly as the reference forms them on the host (float32 sums, then the
numpy ``complex64 / int`` division, which is ``a * fl32(1/n)``).
"""
import numpy as np

from . import _lib


def _finish(x, isnp):
    return _lib.host(x) if isnp else x


def _sums_to_c64(sums):
    """int64 [n][2] device sums -> numpy complex64 of the float32 sums
    (what correlations_cpu.c accumulates into its float output)."""
    s = _lib.host(sums).astype(np.float32)
    return (s[..., 0] + 1j * s[..., 1]).astype(np.complex64)


def avg_autocorr_4bit(data, specnums):
    """correlations.py:233-262 (C: correlations_cpu.c:39-93): int64 sum of
    ``r^2 + i^2`` per channel over ``len(specnums)`` rows, divided by the row
    count (float64); NaN for an empty block."""
    t = _lib.torch()
    rowcount = len(specnums)
    if rowcount == 0:
        return np.nan
    d = _lib.to_device(data, t.uint8)
    ncol = d.shape[1]
    out = t.empty((ncol, 2), dtype=t.int64, device=d.device)
    _lib.call("alb_xcorr_4bit_pair", _lib.ptr(d), _lib.ptr(d), rowcount, ncol, _lib.ptr(None), 1, _lib.ptr(out),
              _lib.stream())
    return _lib.host(out)[:, 0] / rowcount


def avg_xcorr_4bit(data0, data1, specnums):
    """correlations.py:296-327 (C: correlations_cpu.c:128-186): time average of
    ``x0 * conj(x1)`` per channel over ``len(specnums)`` rows -> complex64."""
    t = _lib.torch()
    assert data0.shape[1] == data1.shape[1]
    assert data0.shape[0] == data1.shape[0]
    rowcount = len(specnums)
    if rowcount == 0:
        return np.nan
    d0 = _lib.to_device(data0, t.uint8)
    d1 = _lib.to_device(data1, t.uint8)
    ncol = d0.shape[1]
    out = t.empty((ncol, 2), dtype=t.int64, device=d0.device)
    _lib.call("alb_xcorr_4bit_pair", _lib.ptr(d0), _lib.ptr(d1), rowcount, ncol, _lib.ptr(None), 0, _lib.ptr(out),
              _lib.stream())
    return _sums_to_c64(out) / rowcount


def _match(specnum0, specnum1, idx0, idx1):
    t = _lib.torch()
    s0 = _lib.to_device(np.asarray(specnum0, dtype=np.int64), t.int64)
    s1 = _lib.to_device(np.asarray(specnum1, dtype=np.int64), t.int64)
    match = t.empty(max(len(specnum0), 1), dtype=t.int32, device=s0.device)
    count = t.empty(1, dtype=t.int32, device=s0.device)
    _lib.call("alb_match_rows", _lib.ptr(s0), _lib.ptr(s1), len(specnum0), len(specnum1), int(idx0), int(idx1),
              _lib.ptr(match), _lib.ptr(count), _lib.stream())
    return match, int(count.item())


def avg_xcorr_4bit_2ant(data0, data1, specnum0, specnum1, start_idx0, start_idx1):
    """correlations.py:330-383 (C: correlations_cpu.c:188-271): cross
    correlation of two antennas over the spectra present in both
    (``specnum - start_idx`` equal); NaN if none."""
    t = _lib.torch()
    assert data0.shape[1] == data1.shape[1]
    if len(specnum0) == 0 or len(specnum1) == 0:
        return np.nan
    match, rowcount = _match(specnum0, specnum1, start_idx0, start_idx1)
    if rowcount == 0:
        return np.nan
    d0 = _lib.to_device(data0, t.uint8)
    d1 = _lib.to_device(data1, t.uint8)
    ncol = d0.shape[1]
    out = t.empty((ncol, 2), dtype=t.int64, device=d0.device)
    _lib.call("alb_xcorr_4bit_pair", _lib.ptr(d0), _lib.ptr(d1), len(specnum0), ncol, _lib.ptr(match), 0,
              _lib.ptr(out), _lib.stream())
    return _sums_to_c64(out) / rowcount


def avg_xcorr_1bit(data0, data1, specnums, nchannels):
    """correlations.py:386-408 (C: correlations_cpu.c:273-433) on sortpols
    1-bit rows -> complex64 per channel."""
    t = _lib.torch()
    assert data0.shape[0] == data1.shape[0]
    assert data0.shape[1] == data1.shape[1]
    rowcount = len(specnums)
    if rowcount == 0:
        return np.nan
    d0 = _lib.to_device(data0, t.uint8)
    d1 = _lib.to_device(data1, t.uint8)
    out = t.empty((nchannels, 2), dtype=t.int64, device=d0.device)
    _lib.call("alb_xcorr_1bit_pair", _lib.ptr(d0), _lib.ptr(d1), rowcount, d0.shape[1], int(nchannels),
              _lib.ptr(None), 0, _lib.ptr(out), _lib.stream())
    return _sums_to_c64(out) / rowcount


def _vv(d0, d1, nrows, nchannels, match):
    t = _lib.torch()
    out = t.empty((4, nchannels), dtype=t.int64, device=d0.device)
    _lib.call("alb_xcorr_1bit_pair", _lib.ptr(d0), _lib.ptr(d1), nrows, d0.shape[1], int(nchannels), _lib.ptr(match),
              1, _lib.ptr(out), _lib.stream())
    return [x.astype(np.float32) for x in _lib.host(out)]


def avg_xcorr_1bit_vanvleck(data0, data1, specnums, nchannels):
    """correlations.py:411-443 (C: correlations_cpu.c:435-531):
    ``[R0, R1, IM0, IM1] / rowcount`` (float32)."""
    t = _lib.torch()
    assert data0.shape[0] == data1.shape[0]
    assert data0.shape[1] == data1.shape[1]
    rowcount = len(specnums)
    if rowcount == 0:
        return [np.full(nchannels, np.nan, dtype=np.float32) for _ in range(4)]
    d0 = _lib.to_device(data0, t.uint8)
    d1 = _lib.to_device(data1, t.uint8)
    return [x / rowcount for x in _vv(d0, d1, rowcount, nchannels, None)]


def avg_xcorr_1bit_vanvleck_2ant(data0, data1, nchannels, specnum0, specnum1, idxstart0, idxstart1):
    """correlations.py:446-489 (C: correlations_cpu.c:533-639): un-normalised
    ``[R0, R1, IM0, IM1]`` over matched spectra and the matched row count."""
    t = _lib.torch()
    assert data0.shape[0] == data1.shape[0]
    assert data0.shape[1] == data1.shape[1]
    nan = [np.full(nchannels, np.nan, dtype=np.float32) for _ in range(4)]
    if len(specnum0) == 0 or len(specnum1) == 0:
        return nan, 0
    match, rowcount = _match(specnum0, specnum1, idxstart0, idxstart1)
    if rowcount == 0:
        return nan, 0
    d0 = _lib.to_device(data0, t.uint8)
    d1 = _lib.to_device(data1, t.uint8)
    return _vv(d0, d1, len(specnum0), nchannels, match), rowcount


def van_vleck_correction(R0, R1, I0, I1, rowcount):
    """correlations.py:491-497."""
    h = np.pi / 2 / rowcount
    return np.sin(R0 * h) + np.sin(R1 * h) + 1j * (np.sin(I1 * h) - np.sin(I0 * h))


PRECISION_MODES = {"fp32": 0, "bf16x3": 1, "fp16": 2, "fp16x3": 3}


def _same_array(a, b):
    if a is b:
        return True
    try:
        return (np.asarray(a).__array_interface__["data"][0] == np.asarray(b).__array_interface__["data"][0]
                and isinstance(a, np.ndarray) and isinstance(b, np.ndarray))
    except (TypeError, KeyError):
        return False


def avg_xcorr_4bit_float_gpu(data0, data1, specnum0, specnum1):
    """correlations.py:180-204: one baseline of unpacked (float) spectra,
    ``mean_t data0[t] * conj(data1[t])`` per channel over the
    ``N = len(specnum0)`` present rows (packed at the top, as the float
    BasebandFileIterator yields them), when ``specnum0`` and ``specnum1`` are
    the same array (one antenna).  Returns ``(complex64 [nchan], N)``; the
    reference leaves the two-antenna case unimplemented and returns None, as
    does this.  The sum runs on the matrix-core correlator (two inputs); no
    present rows gives NaN, as the reference's 0/0."""
    if not _same_array(specnum0, specnum1):
        return None
    t = _lib.torch()
    N = len(specnum0)
    d0 = _lib.to_device(data0, t.complex64)
    d1 = _lib.to_device(data1, t.complex64)
    if d0.dim() != 2 or tuple(d0.shape[1:]) != tuple(d1.shape[1:]) or d0.shape[0] < N or d1.shape[0] < N:
        raise ValueError("data0 and data1 must be (n_spectra >= len(specnum), n_channels)")
    nchan = int(d0.shape[1])
    if N == 0:
        res = t.full((nchan,), complex("nan"), dtype=t.complex64, device=d0.device)
    else:
        # correlator input [channel][time][input] (F-order (2, N, nchan))
        x = t.stack([d0[:N], d1[:N]], dim=-1).permute(1, 0, 2).contiguous()
        o = t.empty((nchan, 2, 2), dtype=t.complex64, device=x.device)
        scale = np.float32(1.0) / np.float32(N)
        _lib.call("alb_xcorr_all_ant", _lib.ptr(x), 2, N, nchan, _lib.ptr(o), float(scale), _lib.stream())
        res = o[:, 1, 0].contiguous()   # out[i + 2 (j + 2 f)] = C_ij: baseline (0, 1)
    return (_lib.host(res) if _lib.is_numpy(data0) else res), N


def avg_xcorr_all_ant_gpu(x, nant, npol, ntime, nfreq, split=1, scratch=None, out=None, precision="fp32"):
    """correlations_gpu.py:95-121: ``out[i,j,f] = sum_k x[i,k,f] conj(x[j,k,f]) / K``.

    ``x``: complex64 F-order ``(nant*npol, ntime, nfreq)`` (a torch tensor whose
    memory is column-major, e.g. ``t.permute(2,1,0)`` of a C-order
    ``(nfreq, ntime, M)`` tensor, or a numpy F-order array).  Returns the full
    Hermitian ``(M, M, nfreq)`` complex64 in F-order.

    ``precision`` (not in the reference, which runs a cuBLAS FP32 CGEMM):
    ``"fp32"`` (default), ``"bf16x3"`` (BF16 hi/lo split, three products),
    ``"fp16"`` or ``"fp16x3"`` (FP16 hi/lo split), all accumulating in FP32 on
    the matrix cores; the reduced forms take ``nant*npol <= 8``.
    """
    t = _lib.torch()
    M = nant * npol
    K = ntime
    if K % split != 0:
        raise ValueError("split should be a divisor of ntime")
    if split > 1:
        raise NotImplementedError()
    isnp = _lib.is_numpy(x)
    if isnp:
        xd = _lib.to_device(np.asarray(x, dtype=np.complex64).reshape(-1, order="F"))
    else:
        if tuple(x.shape) != (M, K, nfreq) or x.dtype != t.complex64:
            raise ValueError("x must be complex64 of shape (M, ntime, nfreq)")
        # F-order memory == C-order memory of the reversed shape
        xd = x.permute(2, 1, 0).contiguous()
    if out is None:
        o = t.empty((nfreq, M, M), dtype=t.complex64, device=xd.device)
    else:
        if tuple(out.shape) != (M, M, nfreq) or out.dtype != t.complex64:
            raise ValueError("invalid out buffer")
        o = out.permute(2, 1, 0)
        if not o.is_contiguous():
            raise ValueError("invalid out buffer")
    if precision not in PRECISION_MODES:
        raise ValueError("precision must be one of %s" % sorted(PRECISION_MODES))
    scale = np.float32(1.0) / np.float32(K)
    _lib.call("alb_xcorr_all_ant_prec", _lib.ptr(xd), M, K, nfreq, _lib.ptr(o), float(scale),
              PRECISION_MODES[precision], _lib.stream())
    res = o.permute(2, 1, 0)
    if isnp:
        return np.asfortranarray(_lib.host(o).transpose(2, 1, 0))
    return res


def avg_xcorr_all_ant_cpu(x, nant, npol, ntime, nfreq, split=1, out=None):
    """correlations.py:115-151: ``out[i,j,f] = sum_k x[i,k,f] conj(x[j,k,f]) / K``
    with the reference's arithmetic -- every product rounded to complex64, the
    sum and the division by K in float64 (its Numba loop accumulates into
    ``0.0+0.0j``), the lower triangle the conjugate -- computed on the GPU by
    ``alb_xcorr_all_ant_f64acc``.  ``x``: complex64 ``(M, ntime, nfreq)`` in any
    memory order (numpy or a device tensor).  Returns a C-order
    ``(M, M, nfreq)`` complex64 like the reference (numpy for numpy input);
    ``out``, when given, receives it."""
    t = _lib.torch()
    M = nant * npol
    K = ntime
    if K % split != 0:
        raise ValueError("split should be a divisor of ntime")
    if split > 1:
        raise NotImplementedError("Splitting not yet implemented for CPU version")
    isnp = _lib.is_numpy(x)
    if isnp:
        xd = _lib.to_device(np.asarray(x, dtype=np.complex64))
    else:
        xd = x if x.dtype == t.complex64 else x.to(t.complex64)
        if xd.device != _lib.device():
            xd = xd.to(_lib.device())
    if tuple(xd.shape) != (M, K, nfreq):
        raise ValueError("x must have shape (nant*npol, ntime, nfreq) = (%d, %d, %d)" % (M, K, nfreq))
    sm, sk, sf = (int(v) for v in xd.stride())
    o = t.empty((M, M, nfreq), dtype=t.complex64, device=xd.device)
    _lib.call("alb_xcorr_all_ant_f64acc", _lib.ptr(xd), M, K, nfreq, sm, sk, sf, _lib.ptr(o), M * nfreq, nfreq, 1,
              _lib.stream())
    r = _lib.host(o) if isnp else o
    if out is not None:
        if tuple(out.shape) != (M, M, nfreq):
            raise ValueError("invalid out buffer")
        out[...] = r if (_lib.is_numpy(out) or not isnp) else t.from_numpy(r)
        return out
    return r


def _fortran_device(a, shape, name):
    """numpy array (any order; its F-order element sequence is used) or a
    device tensor of ``shape`` whose memory is column-major -> contiguous
    device tensor holding the column-major element sequence."""
    t = _lib.torch()
    if _lib.is_numpy(a):
        a = np.asarray(a, dtype=np.complex64)
        if a.shape != tuple(shape):
            raise ValueError("%s must have shape %s" % (name, tuple(shape)))
        return _lib.to_device(a.reshape(-1, order="F"))
    if tuple(a.shape) != tuple(shape) or a.dtype != t.complex64:
        raise ValueError("%s must be complex64 of shape %s" % (name, tuple(shape)))
    return a.permute(*reversed(range(a.dim()))).contiguous().to(_lib.device())


def cgemm_strided_batched(A, B, M, N, K, batch_count, out=None):
    """cgemm_batch.cu:8-54 ``C = A B^H`` per batch (column-major, alpha 1):
    ``A`` (M, K, batch) and ``B`` (N, K, batch) complex64, numpy arrays or
    device tensors with column-major memory.  Returns ``C`` (M, N, batch) in
    F-order (numpy for numpy ``A``); ``out`` (a column-major device tensor),
    when given, receives it."""
    t = _lib.torch()
    Ad = _fortran_device(A, (M, K, batch_count), "A")
    Bd = _fortran_device(B, (N, K, batch_count), "B")
    if out is not None:
        if tuple(out.shape) != (M, N, batch_count) or out.dtype != t.complex64:
            raise ValueError("invalid out buffer")
        o = out.permute(2, 1, 0)
        if not o.is_contiguous():
            raise ValueError("invalid out buffer")
    else:
        o = t.empty((batch_count, N, M), dtype=t.complex64, device=Ad.device)
    _lib.call("alb_cgemm_strided_batched", _lib.ptr(Ad), _lib.ptr(Bd), _lib.ptr(o), M, N, K, batch_count,
              _lib.stream())
    if out is not None:
        return out
    if _lib.is_numpy(A):
        return np.asfortranarray(_lib.host(o).transpose(2, 1, 0))
    return o.permute(2, 1, 0)


def pair_index(ninput):
    """Baselines (i, j), i <= j, in the order of the all-pairs outputs."""
    return [(i, j) for i in range(ninput) for j in range(i, ninput)]


_INT_METHODS = {"auto": 0, "valu": 1, "mfma": 2}


class _int_method:
    """Select the all-pairs integer form for one call ("auto": int8 matrix
    cores whenever the shape allows; "valu": the VALU / XOR form; "mfma")."""

    def __init__(self, method):
        if method not in _INT_METHODS:
            raise ValueError("method must be one of %s" % sorted(_INT_METHODS))
        self.v = _INT_METHODS[method]

    def __enter__(self):
        lib = _lib.load()
        self.old = lib.alb_tune_get(b"xcorr_int_method")
        _lib.call("alb_tune_set", b"xcorr_int_method", self.v)

    def __exit__(self, *exc):
        _lib.call("alb_tune_set", b"xcorr_int_method", self.old)


def xcorr_4bit_allpairs(inputs, blocklen, method="auto"):
    """Exact block sums of every baseline of 4-bit inputs.

    ``inputs``: uint8 ``[ninput, T, nchan]`` per-pol bytes (sortpols layout),
    device tensor or numpy.  Returns int64 ``[T//blocklen, npair, nchan, 2]``
    (re, im of ``x_i conj(x_j)``, pairs in :func:`pair_index` order).
    ``method``: "auto" (int8 matrix cores when ``blocklen % 64 == 0``),
    "valu" or "mfma".
    """
    t = _lib.torch()
    isnp = _lib.is_numpy(inputs)
    d = _lib.to_device(inputs, t.uint8)
    n, T, C = d.shape
    nb = T // blocklen
    npair = n * (n + 1) // 2
    out = t.empty((nb, npair, C, 2), dtype=t.int64, device=d.device)
    ptrs = (_lib._vp * n)(*[d[i].data_ptr() for i in range(n)])
    with _int_method(method):
        _lib.call("alb_xcorr_4bit_allpairs", ptrs, n, T, C, int(blocklen), _lib.ptr(out), _lib.stream())
    return _finish(out, isnp)


def xcorr_1bit_allpairs(inputs, nchan, blocklen, method="auto"):
    """Exact block sums of every baseline of 1-bit inputs (sortpols 1-bit
    layout ``[ninput, T, ncol >= ceil(nchan/4)]``) -> int64
    ``[T//blocklen, npair, nchan, 2]``.  ``method``: "auto" (int8 matrix cores
    when ``blocklen % 64 == 0`` and ``ncol % 4 == 0``), "valu" (XOR form) or "mfma"."""
    t = _lib.torch()
    isnp = _lib.is_numpy(inputs)
    d = _lib.to_device(inputs, t.uint8)
    n, T, ncol = d.shape
    nb = T // blocklen
    npair = n * (n + 1) // 2
    out = t.empty((nb, npair, nchan, 2), dtype=t.int64, device=d.device)
    ptrs = (_lib._vp * n)(*[d[i].data_ptr() for i in range(n)])
    with _int_method(method):
        _lib.call("alb_xcorr_1bit_allpairs", ptrs, n, T, ncol, int(nchan), int(blocklen), _lib.ptr(out),
                  _lib.stream())
    return _finish(out, isnp)


def block_averages(sums, rowcount):
    """int64 block sums [..., 2] -> complex64 averages ``fl32(sum) * fl32(1/n)``
    (numpy's complex64 / int, correlations.py:327)."""
    s = np.asarray(sums).astype(np.float32)
    c = (s[..., 0] + 1j * s[..., 1]).astype(np.complex64)
    return c / rowcount
