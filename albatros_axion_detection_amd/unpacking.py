"""Baseband unpacking on the GPU -- the call surface of src/correlations/unpacking.py.

Inputs may be numpy arrays (copied to the device) or torch tensors on the GPU.
Outputs follow the input kind: numpy in -> numpy out, torch in -> torch out
(device tensors, like the CuPy arrays of the reference's GPU path).
All arithmetic runs in libalbatros_hip.
"""
import numpy as np

from . import _lib

_backend = "hip"


def set_backend(backend):
    """unpacking.py:8-16.  Accepted for source compatibility; the computation
    always runs on the GPU ("numpy" only selects numpy outputs)."""
    global _backend
    if backend not in ("numpy", "cupy", "torch", "hip"):
        raise ValueError("Invalid backend. Choose either 'numpy' or 'cupy'.")
    _backend = backend


def _out(x, like_numpy):
    return _lib.host(x) if (like_numpy or _backend == "numpy") else x


def _check_rows(d, rowstart, rowend, bytes_per_spectrum, name):
    """Rows [rowstart, rowend) of spectra of ``bytes_per_spectrum`` bytes must
    lie inside ``d`` (checked before any launch: the kernels index the raw
    buffer by row and would otherwise read past it)."""
    if rowstart < 0 or rowend < rowstart:
        raise ValueError("%s: need 0 <= rowstart <= rowend (got %d, %d)" % (name, rowstart, rowend))
    if rowend > rowstart and d.numel() < rowend * bytes_per_spectrum:
        raise ValueError("%s: rowend %d needs %d bytes of data, the array holds %d" %
                         (name, rowend, rowend * bytes_per_spectrum, d.numel()))


def _outputs(out, nrows, ncols, dtype, device, name):
    """(pol0, pol1): new device tensors, or the caller's ``out`` pair checked
    for shape, dtype and row-contiguous memory."""
    t = _lib.torch()
    if out is None:
        return (t.empty((nrows, ncols), dtype=dtype, device=device),
                t.empty((nrows, ncols), dtype=dtype, device=device))
    p0, p1 = out
    for p in (p0, p1):
        if tuple(p.shape) != (nrows, ncols) or p.dtype != dtype or not p.is_contiguous() or p.device != device:
            raise ValueError("%s: out must be two contiguous %s device tensors of shape (%d, %d)"
                             % (name, dtype, nrows, ncols))
    return p0, p1


def _check_channel_indices(ch, length_channels, name):
    """Column indices must lie in [0, length_channels)."""
    if ch.numel() == 0:
        return
    lo, hi = int(ch.min().item()), int(ch.max().item())
    if lo < 0 or hi >= length_channels:
        raise ValueError("%s: channel indices must lie in [0, %d) (got %d..%d)" % (name, length_channels, lo, hi))


def unpack_4bit(data, rowstart, rowend, channels, length_channels, out=None):
    """unpacking.py:193-245 (GPU form :140-155; C form unpacking.c:155-205).

    ``data``: raw spectra, ``2*length_channels`` bytes each; ``channels``:
    column indices.  Returns complex64 (pol0, pol1) of shape
    ``(rowend-rowstart, len(channels))``; nibble 8 decodes to +8.  ``out``
    (not in the reference): a pair of device tensors to write instead.
    """
    t = _lib.torch()
    isnp = _lib.is_numpy(data)
    d = _lib.to_device(data, t.uint8)
    ch = _lib.to_device(np.asarray(channels, dtype=np.int64) if not isinstance(channels, t.Tensor) else channels,
                        t.int64)
    _check_rows(d, int(rowstart), int(rowend), 2 * int(length_channels), "unpack_4bit")
    _check_channel_indices(ch, int(length_channels), "unpack_4bit")
    nrows = int(rowend - rowstart)
    ncols = int(ch.numel())
    pol0, pol1 = _outputs(out, nrows, ncols, t.complex64, d.device, "unpack_4bit")
    _lib.call("alb_unpack_4bit_float", _lib.ptr(d), _lib.ptr(pol0), _lib.ptr(pol1), int(rowstart), int(rowend),
              _lib.ptr(ch), ncols, int(length_channels), _lib.stream())
    if out is not None:
        return pol0, pol1
    return _out(pol0, isnp), _out(pol1, isnp)


def unpack_1bit(data, rowstart, rowend, channels, length_channels, out=None):
    """unpacking.py:248-302 (C form unpacking.c:207-246).

    ``channels``: a continuous range of channel indices starting at an even
    index with an even count (baseband_data_classes.py:334-341).
    """
    t = _lib.torch()
    ch = np.asarray(channels, dtype=np.int64)
    chanstart = int(ch[0])
    chanend = int(ch[-1]) + 1
    if np.any(np.diff(ch) != 1):
        raise ValueError("A continuous range of channels must be provided in 1-bit mode.")
    if chanstart < 0 or chanend > length_channels:
        raise ValueError("unpack_1bit: channels must lie in [0, %d)" % length_channels)
    isnp = _lib.is_numpy(data)
    d = _lib.to_device(data, t.uint8)
    _check_rows(d, int(rowstart), int(rowend), int(length_channels) // 2, "unpack_1bit")
    nrows = int(rowend - rowstart)
    ncols = chanend - chanstart
    pol0, pol1 = _outputs(out, nrows, ncols, t.complex64, d.device, "unpack_1bit")
    _lib.call("alb_unpack_1bit_float", _lib.ptr(d), _lib.ptr(pol0), _lib.ptr(pol1), int(rowstart), int(rowend),
              chanstart, chanend, int(length_channels), _lib.stream())
    if out is not None:
        return pol0, pol1
    return _out(pol0, isnp), _out(pol1, isnp)


def sortpols(data, length_channels, bit_mode, rowstart, rowend, chanstart, chanend, out=None):
    """unpacking.py:305-377 / unpacking.c:248-368.

    4-bit: per-pol bytes ``(nrows, chanend-chanstart)``.  1-bit: per pol, four
    channels per byte, ``ceil((chanend-chanstart)/4)`` bytes per row.
    """
    t = _lib.torch()
    if chanend is None:
        chanstart, chanend = 0, length_channels
    elif chanend > length_channels:
        raise ValueError("ERROR: Trying to unpack more channels than present. Check chanend.")
    nrows = int(rowend - rowstart)
    if bit_mode == 4:
        ncols = chanend - chanstart
    elif bit_mode == 1:
        if chanstart % 2 > 0:
            raise ValueError("ERROR: Start channel index must be even.")
        ncols = int(np.ceil((chanend - chanstart) / 4))
    else:
        raise ValueError("bit_mode must be 1 or 4")
    if chanstart < 0 or chanend <= chanstart:
        raise ValueError("sortpols: need 0 <= chanstart < chanend (got %d, %d)" % (chanstart, chanend))
    isnp = _lib.is_numpy(data)
    d = _lib.to_device(data, t.uint8)
    _check_rows(d, int(rowstart), int(rowend), 2 * int(length_channels) if bit_mode == 4 else int(length_channels) // 2,
                "sortpols")
    pol0, pol1 = _outputs(out, nrows, ncols, t.uint8, d.device, "sortpols")
    _lib.call("alb_sortpols", _lib.ptr(d), _lib.ptr(pol0), _lib.ptr(pol1), int(rowstart), int(rowend), int(ncols),
              int(length_channels), int(bit_mode), int(chanstart), int(chanend), _lib.stream())
    if out is not None:
        return pol0, pol1
    return _out(pol0, isnp), _out(pol1, isnp)


def hist(data, rowstart, rowend, length_channels, bit_depth, mode):
    """unpacking.py:82-138: uint64 ``(2**bit_depth, length_channels)`` counts of
    raw quantisation levels; mode 0 = pol0, 1 = pol1, -1 = both."""
    t = _lib.torch()
    assert bit_depth in (1, 4), f"bit_depth set to {bit_depth}, only takes values 1 or 4"
    assert mode in (0, 1, -1), f"mode set to {mode}, only takes values 0,1,-1"
    isnp = _lib.is_numpy(data)
    d = _lib.to_device(data, t.uint8)
    _check_rows(d, int(rowstart), int(rowend), 2 * int(length_channels) if bit_depth == 4 else int(length_channels) // 2,
                "hist")
    nbins = 2 ** bit_depth
    h = t.empty((nbins, int(length_channels)), dtype=t.int64, device=d.device)
    _lib.call("alb_hist", _lib.ptr(d), _lib.ptr(h), int(rowstart), int(rowend), int(length_channels), int(bit_depth),
              int(mode), _lib.stream())
    if isnp or _backend == "numpy":
        return _lib.host(h).astype(np.uint64)
    return h


def dropped_packets(data, spec_num, spectra_per_packet, nchan, bit_depth):
    """unpacking_gpu.cu:24-72 ``dropped_packets``: zero every byte of packet i
    when ``spec_num[i+1] - spec_num[i] != spectra_per_packet`` (i < num_packets
    - 1, num_packets = len(spec_num)) and return how many packets that is.

    ``data`` holds ``num_packets`` packets of ``2*nchan*spectra_per_packet``
    (4-bit), ``nchan*spectra_per_packet`` (2-bit) or
    ``nchan*spectra_per_packet/2`` (1-bit) bytes and is changed in place: a
    numpy array as the reference's host buffer (copied to the device and
    back), a device tensor on the device.  Like the reference, fewer than two
    packets return 0 after a message, and another bit depth returns 0."""
    t = _lib.torch()
    spec = np.asarray(spec_num, dtype=np.int64) if not isinstance(spec_num, t.Tensor) else spec_num
    npk = int(spec.shape[0])
    if npk < 2:
        print("There are fewer than 2 packets in the dropped packets function")
        return 0
    if bit_depth not in (1, 2, 4):
        return 0
    spp, nchan = int(spectra_per_packet), int(nchan)
    bpp = {4: 2 * nchan * spp, 2: nchan * spp, 1: nchan * spp // 2}[bit_depth]
    isnp = _lib.is_numpy(data)
    if isnp and not (data.flags.c_contiguous and data.dtype == np.uint8):
        raise ValueError("dropped_packets: data must be a C-contiguous uint8 array")
    d = _lib.to_device(data, t.uint8)
    if not isnp and d.data_ptr() != data.data_ptr():
        raise ValueError("dropped_packets: data must be a contiguous uint8 device tensor (it is changed in place)")
    if d.numel() < npk * bpp:
        raise ValueError("dropped_packets: %d packets of %d bytes need %d bytes of data, the array holds %d"
                         % (npk, bpp, npk * bpp, d.numel()))
    s = _lib.to_device(spec, t.int64)
    nd = t.empty(1, dtype=t.int64, device=d.device)
    _lib.call("alb_dropped_packets", _lib.ptr(d), _lib.ptr(s), npk, spp, nchan, int(bit_depth), _lib.ptr(nd),
              _lib.stream())
    if isnp:
        data.reshape(-1)[:] = _lib.host(d).reshape(-1)
    return int(nd.item())
