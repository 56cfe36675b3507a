"""Integer restatement of the reference unpack / repack / histogram code.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

All functions work on numpy arrays and are exact (integer arithmetic, the
complex outputs hold small integers).
"""
import numpy as np


def nibble_decode(v):
    """Signed 4-bit rule of the reference: ``v > 8 ? v - 16 : v``.

    ``8`` decodes to ``+8`` (not -8).  unpacking.c:189-202,
    unpacking.py:145-152, tests/mars2019/albatrostools.py:246-249.
    """
    v = np.asarray(v).astype(np.int64)
    return np.where(v > 8, v - 16, v)


def decode_4bit_bytes(b):
    """Byte -> (real, imag) ints: real = high nibble, imag = low nibble.

    unpacking.c:187-188 (``im = d & 15``, ``r = (d & 240) >> 4``).
    """
    b = np.asarray(b, dtype=np.uint8)
    return nibble_decode((b >> 4) & 0xF), nibble_decode(b & 0xF)


def unpack_4bit(data, rowstart, rowend, channels, length_channels):
    """unpacking.c:155-205 ``unpack_4bit_float`` / unpacking.py:193-245.

    ``data`` holds raw spectra of ``2*length_channels`` bytes
    ``[ch0 pol0][ch0 pol1][ch1 pol0]...``; ``channels`` are column indices.
    Returns complex128 (pol0, pol1) of shape (rowend-rowstart, len(channels)).
    """
    d = np.asarray(data, dtype=np.uint8).reshape(-1, 2 * length_channels)
    ch = np.asarray(channels, dtype=np.int64)
    rows = d[rowstart:rowend]
    b0 = rows[:, 2 * ch]
    b1 = rows[:, 2 * ch + 1]
    r0, i0 = decode_4bit_bytes(b0)
    r1, i1 = decode_4bit_bytes(b1)
    return r0 + 1j * i0, r1 + 1j * i1


def unpack_1bit(data, rowstart, rowend, chanstart, chanend, nchan):
    """unpacking.c:207-246 ``unpack_1bit_float``.

    Raw 1-bit byte = two channels x two pols, bits 7..0 =
    ``r0c0 i0c0 r1c0 i1c0 r0c1 i0c1 r1c1 i1c1`` (p = pol, c = channel);
    bit 1 -> +1, bit 0 -> -1.  ``nchan`` channels -> ``nchan/2`` bytes per
    spectrum; columns ``chanstart/2 ...`` for ``(chanend-chanstart)/2`` bytes.
    """
    d = np.asarray(data, dtype=np.uint8).reshape(-1)
    bps = nchan // 2
    nrows = rowend - rowstart
    nb = (chanend - chanstart) // 2
    idx = (np.arange(nrows)[:, None] + rowstart) * bps + chanstart // 2 + np.arange(nb)[None, :]
    b = d[idx].astype(np.int64)
    bit = lambda s: 2 * ((b >> s) & 1) - 1
    pol0 = np.empty((nrows, 2 * nb), dtype=np.complex128)
    pol1 = np.empty((nrows, 2 * nb), dtype=np.complex128)
    pol0[:, 0::2] = bit(7) + 1j * bit(6)
    pol1[:, 0::2] = bit(5) + 1j * bit(4)
    pol0[:, 1::2] = bit(3) + 1j * bit(2)
    pol1[:, 1::2] = bit(1) + 1j * bit(0)
    return pol0, pol1


def unpack_1bit_gpu_semantics(data, rowstart, rowend, channels, length_channels):
    """unpacking.py:165-191 ``_unpack_1bit_float_gpu`` (channel index list,
    columns ``channels[::2]//2``)."""
    ch = np.asarray(channels, dtype=np.int64)
    d = np.asarray(data, dtype=np.uint8).reshape(-1, length_channels // 2)
    cols = ch[::2] // 2
    b = d[rowstart:rowend][:, cols].astype(np.int64)
    bit = lambda s: 2 * ((b >> s) & 1) - 1
    n = len(ch)
    pol0 = np.empty((rowend - rowstart, n), dtype=np.complex128)
    pol1 = np.empty((rowend - rowstart, n), dtype=np.complex128)
    pol0[:, 0::2] = bit(7) + 1j * bit(6)
    pol0[:, 1::2] = bit(3) + 1j * bit(2)
    pol1[:, 0::2] = bit(5) + 1j * bit(4)
    pol1[:, 1::2] = bit(1) + 1j * bit(0)
    return pol0, pol1


def sortpols(data, length_channels, bit_mode, rowstart, rowend, chanstart, chanend):
    """unpacking.py:305-377 wrapper + unpacking.c:248-368 ``sortpols``.

    4-bit: per-pol bytes ``pol[i, k-chanstart] = data[row, 2k + pol]``
    (unpacking.c:263-275).
    1-bit: per pol, 4 channels per byte, channel-major, (re, im) MSB first:
    ``pol0 = p0c0 + (p0c1<<2) + (p0c2>>4) + (p0c3>>2)``,
    ``pol1 = (p1c0<<2) + (p1c1<<4) + (p1c2>>2) + p1c3`` with the ragged last
    byte of unpacking.c:306-365.
    """
    if chanend is None:
        chanstart, chanend = 0, length_channels
    elif chanend > length_channels:
        raise ValueError("ERROR: Trying to unpack more channels than present. Check chanend.")
    nrows = rowend - rowstart
    d = np.asarray(data, dtype=np.uint8).reshape(-1)
    if bit_mode == 4:
        ncols = chanend - chanstart
        dd = d.reshape(-1, 2 * length_channels)[rowstart:rowend]
        k = np.arange(chanstart, chanend)
        return dd[:, 2 * k].copy(), dd[:, 2 * k + 1].copy()
    if bit_mode != 1:
        raise ValueError("bit_mode must be 1 or 4")
    if chanstart % 2 > 0:
        raise ValueError("ERROR: Start channel index must be even.")
    ncols = int(np.ceil((chanend - chanstart) / 4))
    pol0 = np.zeros((nrows, ncols), dtype=np.uint8)
    pol1 = np.zeros((nrows, ncols), dtype=np.uint8)
    bps = length_channels // 2          # ceil(nchan/2) of integer nchan/2
    cndn = (chanend - chanstart) % 4
    rows = np.arange(nrows)[:, None] + rowstart
    j = np.arange(ncols)[None, :]
    base = bps * rows + 2 * j + chanstart // 2
    # number of raw bytes consumed for the last column (unpacking.c:310-365)
    last_two = cndn in (0, 3)
    b1 = d[base].astype(np.int64)
    has_b2 = np.ones((1, ncols), dtype=bool)
    if not last_two:
        has_b2[0, -1] = False
    b2 = np.where(has_b2, d[np.where(has_b2, base + 1, base)], 0).astype(np.int64)
    p0c0, p1c0, p0c1, p1c1 = b1 & 192, b1 & 48, b1 & 12, b1 & 3
    p0c2, p1c2, p0c3, p1c3 = b2 & 192, b2 & 48, b2 & 12, b2 & 3
    full0 = p0c0 + (p0c1 << 2) + (p0c2 >> 4) + (p0c3 >> 2)
    full1 = (p1c0 << 2) + (p1c1 << 4) + (p1c2 >> 2) + p1c3
    # tail cases for the last column
    if cndn == 3:
        full0[:, -1] = (p0c0 + (p0c1 << 2) + (p0c2 >> 4))[:, -1]
        full1[:, -1] = ((p1c0 << 2) + (p1c1 << 4) + (p1c2 >> 2))[:, -1]
    elif cndn == 2:
        full0[:, -1] = (p0c0 + (p0c1 << 2))[:, -1]
        full1[:, -1] = ((p1c0 << 2) + (p1c1 << 4))[:, -1]
    elif cndn == 1:
        full0[:, -1] = p0c0[:, -1]
        full1[:, -1] = (p1c0 << 2)[:, -1]
    pol0[:] = full0 & 0xFF
    pol1[:] = full1 & 0xFF
    return pol0, pol1


def hist_4bit(data, rowstart, rowend, nchan, mode):
    """unpacking.c:7-74 ``hist_4bit`` (16 bins of raw nibble value x nchan).

    Counts raw nibble values 0..15 of the real and imaginary parts per
    channel; ``mode`` 0 = pol0, 1 = pol1, -1 = both.  Returns uint64
    ``[16, nchan]`` (the reference writes one element past the end of its
    output, unpacking.c:71; that element is not part of the result).
    """
    d = np.asarray(data, dtype=np.uint8).reshape(-1, 2 * nchan)[rowstart:rowend]
    hist = np.zeros((16, nchan), dtype=np.uint64)
    pols = {0: [0], 1: [1], -1: [0, 1]}[mode]
    for p in pols:
        b = d[:, p::2]
        for v in (b >> 4, b & 15):
            for bin_ in range(16):
                hist[bin_] += (v == bin_).sum(axis=0).astype(np.uint64)
    return hist


def hist_1bit(data, rowstart, rowend, nchan, mode):
    """unpacking.c:76-153 ``hist_1bit`` (2 bins x nchan)."""
    d = np.asarray(data, dtype=np.uint8).reshape(-1, nchan // 2)[rowstart:rowend].astype(np.int64)
    hist = np.zeros((2, nchan), dtype=np.uint64)
    shifts = {0: [(7, 6, 3, 2)], 1: [(5, 4, 1, 0)], -1: [(7, 6, 3, 2), (5, 4, 1, 0)]}[mode]
    for (sr0, si0, sr1, si1) in shifts:
        for s, off in ((sr0, 0), (si0, 0), (sr1, 1), (si1, 1)):
            v = (d >> s) & 1
            for bin_ in range(2):
                hist[bin_, off::2] += (v == bin_).sum(axis=0).astype(np.uint64)
    return hist
