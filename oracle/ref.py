"""ctypes bindings of the reference C sources compiled into ``oracle/_ref``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  ``oracle/Makefile``
builds the two libraries from unpacking.c and correlations_cpu.c where the
reference tree is mounted; the built files travel with the repository.
``available()`` is False when they are absent and the tests that need them
skip.
"""
import ctypes
import os

import numpy as np

_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref")
_U = os.path.join(_DIR, "libref_unpacking.so")
_C = os.path.join(_DIR, "libref_correlations.so")
_libs = {}


def available():
    return os.path.exists(_U) and os.path.exists(_C)


def _lib(path):
    if path not in _libs:
        _libs[path] = ctypes.cdll.LoadLibrary(path)
    return _libs[path]


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


_i = ctypes.c_int


def unpack_4bit_float(data, rowstart, rowend, channels, nchan):
    """unpacking.c:155."""
    lib = _lib(_U)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    ch = np.ascontiguousarray(channels, dtype=np.int64)
    n = rowend - rowstart
    p0 = np.empty((n, len(ch)), dtype=np.complex64)
    p1 = np.empty((n, len(ch)), dtype=np.complex64)
    lib.unpack_4bit_float(_p(data), _p(p0), _p(p1), _i(rowstart), _i(rowend), _p(ch),
                          _i(len(ch)), _i(nchan))
    return p0, p1


def unpack_1bit_float(data, rowstart, rowend, chanstart, chanend, nchan):
    """unpacking.c:207."""
    lib = _lib(_U)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    n = rowend - rowstart
    p0 = np.empty((n, chanend - chanstart), dtype=np.complex64)
    p1 = np.empty((n, chanend - chanstart), dtype=np.complex64)
    lib.unpack_1bit_float(_p(data), _p(p0), _p(p1), _i(rowstart), _i(rowend), _i(chanstart),
                          _i(chanend), _i(nchan))
    return p0, p1


def sortpols(data, length_channels, bit_mode, rowstart, rowend, chanstart, chanend):
    """unpacking.c:248 with the ncols rule of unpacking.py:347-356."""
    lib = _lib(_U)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    n = rowend - rowstart
    ncols = chanend - chanstart if bit_mode == 4 else int(np.ceil((chanend - chanstart) / 4))
    p0 = np.empty((n, ncols), dtype=np.uint8)
    p1 = np.empty((n, ncols), dtype=np.uint8)
    lib.sortpols(_p(data), _p(p0), _p(p1), _i(rowstart), _i(rowend), _i(ncols),
                 _i(length_channels), ctypes.c_short(bit_mode), _i(chanstart), _i(chanend))
    return p0, p1


def hist(data, rowstart, rowend, nchan, bit_depth, mode):
    """unpacking.c:7 / :76.  The reference adds one element past its output
    (``k <= hist_len``), so one spare element is allocated and dropped."""
    lib = _lib(_U)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    nbins = 2 ** bit_depth - 1
    h = np.zeros((nbins + 1) * nchan + 1, dtype=np.uint64)
    fn = lib.hist_4bit if bit_depth == 4 else lib.hist_1bit
    fn(_p(data), _p(h), _i(rowstart), _i(rowend), _i(nchan), _i(nbins), _i(mode))
    return h[:-1].reshape(nbins + 1, nchan)


def avg_autocorr_4bit(data, nrows):
    """correlations_cpu.c:39 (int64 sums)."""
    lib = _lib(_C)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.empty(data.shape[1], dtype=np.int64)
    lib.avg_autocorr_4bit(_p(data), _p(out), _i(nrows), _i(data.shape[1]))
    return out


def avg_xcorr_4bit(data0, data1, nrows):
    """correlations_cpu.c:128 (float output of int32 partial sums)."""
    lib = _lib(_C)
    d0 = np.ascontiguousarray(data0, dtype=np.uint8)
    d1 = np.ascontiguousarray(data1, dtype=np.uint8)
    out = np.empty(d0.shape[1], dtype=np.complex64)
    lib.avg_xcorr_4bit(_p(d0), _p(d1), _p(out), _i(nrows), _i(d0.shape[1]))
    return out


def avg_xcorr_4bit_2ant(data0, data1, specnum0, specnum1, idx0, idx1):
    """correlations_cpu.c:188 -> (sums as complex64, row_count)."""
    lib = _lib(_C)
    lib.avg_xcorr_4bit_2ant.restype = ctypes.c_int
    d0 = np.ascontiguousarray(data0, dtype=np.uint8)
    d1 = np.ascontiguousarray(data1, dtype=np.uint8)
    s0 = np.ascontiguousarray(specnum0, dtype=np.int64)
    s1 = np.ascontiguousarray(specnum1, dtype=np.int64)
    out = np.empty(d0.shape[1], dtype=np.complex64)
    n = lib.avg_xcorr_4bit_2ant(_p(d0), _p(d1), _p(out), _p(s0), _p(s1), ctypes.c_int64(idx0),
                                ctypes.c_int64(idx1), _i(len(s0)), _i(len(s1)), _i(d0.shape[1]))
    return out, n


def avg_xcorr_1bit(data0, data1, nchan, nspec):
    """correlations_cpu.c:273."""
    lib = _lib(_C)
    d0 = np.ascontiguousarray(data0, dtype=np.uint8)
    d1 = np.ascontiguousarray(data1, dtype=np.uint8)
    out = np.empty(nchan, dtype=np.complex64)
    lib.avg_xcorr_1bit(_p(d0), _p(d1), _p(out), _i(nchan), ctypes.c_uint32(nspec),
                       ctypes.c_uint32(d0.shape[1]))
    return out


def avg_xcorr_1bit_vanvleck(data0, data1, nchan, nrows):
    """correlations_cpu.c:435 (its printf output goes to stdout)."""
    lib = _lib(_C)
    d0 = np.ascontiguousarray(data0, dtype=np.uint8)
    d1 = np.ascontiguousarray(data1, dtype=np.uint8)
    full = 4 * d0.shape[1]
    outs = [np.zeros(max(nchan, full), dtype=np.float32) for _ in range(4)]
    lib.avg_xcorr_1bit_vanvleck(_p(d0), _p(d1), *[_p(o) for o in outs], _i(nchan),
                                ctypes.c_uint32(nrows), ctypes.c_uint32(d0.shape[1]))
    return [o[:nchan] for o in outs]


def avg_xcorr_1bit_vanvleck_2ant(data0, data1, specnum0, specnum1, idx0, idx1, nchan):
    """correlations_cpu.c:533."""
    lib = _lib(_C)
    lib.avg_xcorr_1bit_vanvleck_2ant.restype = ctypes.c_int
    d0 = np.ascontiguousarray(data0, dtype=np.uint8)
    d1 = np.ascontiguousarray(data1, dtype=np.uint8)
    s0 = np.ascontiguousarray(specnum0, dtype=np.int64)
    s1 = np.ascontiguousarray(specnum1, dtype=np.int64)
    outs = [np.zeros(max(nchan, 4 * d0.shape[1]), dtype=np.float32) for _ in range(4)]
    n = lib.avg_xcorr_1bit_vanvleck_2ant(_p(d0), _p(d1), *[_p(o) for o in outs], _p(s0), _p(s1),
                                         ctypes.c_int64(idx0), ctypes.c_int64(idx1),
                                         ctypes.c_uint32(len(s0)), ctypes.c_uint32(len(s1)),
                                         ctypes.c_uint32(d0.shape[1]), _i(nchan))
    return [o[:nchan] for o in outs], n
