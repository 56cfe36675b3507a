"""Integer and complex restatement of the reference correlators.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
import numpy as np

from .unpack import nibble_decode


def _decode(pol):
    b = np.asarray(pol, dtype=np.uint8).astype(np.int64)
    return nibble_decode(b >> 4), nibble_decode(b & 15)


def autocorr_4bit_sums(data, nrows):
    """correlations_cpu.c:39-93 ``avg_autocorr_4bit``: int64 ``sum r^2 + i^2``
    per channel over the first ``nrows`` rows of a per-pol byte array."""
    r, i = _decode(np.asarray(data)[:nrows])
    return (r * r + i * i).sum(axis=0).astype(np.int64)


def xcorr_4bit_sums(data0, data1, nrows):
    """correlations_cpu.c:128-186 ``avg_xcorr_4bit``.

    ``Re += r0*r1 + i0*i1``, ``Im += r1*i0 - r0*i1`` (correlations_cpu.c:172-173),
    i.e. ``x0 * conj(x1)``.  Returns int64 (re, im) per channel.
    """
    r0, i0 = _decode(np.asarray(data0)[:nrows])
    r1, i1 = _decode(np.asarray(data1)[:nrows])
    re = (r0 * r1 + i0 * i1).sum(axis=0)
    im = (r1 * i0 - r0 * i1).sum(axis=0)
    return re.astype(np.int64), im.astype(np.int64)


def intersect_rows(specnum0, specnum1, idxstart0, idxstart1):
    """correlations_cpu.c:201-214: merge of two sorted spec_num lists on
    ``specnum - idxstart``; returns matched row index pairs."""
    a = np.asarray(specnum0, dtype=np.int64) - idxstart0
    b = np.asarray(specnum1, dtype=np.int64) - idxstart1
    i = j = 0
    r0, r1 = [], []
    while i < len(a) and j < len(b):
        if a[i] == b[j]:
            r0.append(i)
            r1.append(j)
            i += 1
            j += 1
        elif a[i] > b[j]:
            j += 1
        else:
            i += 1
    return np.asarray(r0, dtype=np.int64), np.asarray(r1, dtype=np.int64)


def xcorr_4bit_2ant_sums(data0, data1, specnum0, specnum1, idxstart0, idxstart1):
    """correlations_cpu.c:188-271 ``avg_xcorr_4bit_2ant``: a4 on the rows
    whose ``specnum - idxstart`` agree.  Returns (re, im, row_count)."""
    r0s, r1s = intersect_rows(specnum0, specnum1, idxstart0, idxstart1)
    d0 = np.asarray(data0)[r0s]
    d1 = np.asarray(data1)[r1s]
    re, im = xcorr_4bit_sums(d0, d1, len(r0s))
    return re, im, len(r0s)


def _bits_1bit(packed, nchan):
    """Per-channel (r, i) bits of the sortpols 1-bit format: byte j holds
    channels 4j..4j+3 as ``r c0, i c0, r c1, i c1, ...`` MSB first
    (correlations_cpu.c:309-316)."""
    p = np.asarray(packed, dtype=np.uint8).astype(np.int64)
    nrows, ncol = p.shape
    r = np.zeros((nrows, 4 * ncol), dtype=np.int64)
    i = np.zeros((nrows, 4 * ncol), dtype=np.int64)
    for c in range(4):
        r[:, c::4] = (p >> (7 - 2 * c)) & 1
        i[:, c::4] = (p >> (6 - 2 * c)) & 1
    return r[:, :nchan], i[:, :nchan]


def xcorr_1bit_sums(data0, data1, nchan, nspec):
    """correlations_cpu.c:273-433 ``avg_xcorr_1bit``.

    ``Re += 2 - 2[(r0^r1) + (i0^i1)]``, ``Im += -2[(r1^i0) - (r0^i1)]``
    (correlations_cpu.c:327-334) which is the exact +-1 product
    ``x0 * conj(x1)``.  The last byte covers ``nchan % 4`` channels
    (correlations_cpu.c:340-420).  Returns int64 (re, im).
    """
    r0, i0 = _bits_1bit(np.asarray(data0)[:nspec], nchan)
    r1, i1 = _bits_1bit(np.asarray(data1)[:nspec], nchan)
    re = (2 - 2 * ((r0 ^ r1) + (i0 ^ i1))).sum(axis=0)
    im = (-2 * ((r1 ^ i0) - (r0 ^ i1))).sum(axis=0)
    return re.astype(np.int64), im.astype(np.int64)


def vanvleck_sums(data0, data1, nchan, nrows, rows0=None, rows1=None):
    """correlations_cpu.c:435-531 ``avg_xcorr_1bit_vanvleck`` (and the
    row-matched ``_2ant`` form, :533-639).

    Per channel: ``R0 += 1 - 2(r0^r1)``, ``R1 += 1 - 2(i0^i1)``,
    ``IM0 += 1 - 2(r0^i1)``, ``IM1 += 1 - 2(r1^i0)``.  All ``4*ncol``
    channels of the packed rows are summed (no tail handling in the
    reference), then the first ``nchan`` are returned.
    """
    d0 = np.asarray(data0)
    d1 = np.asarray(data1)
    if rows0 is None:
        d0 = d0[:nrows]
        d1 = d1[:nrows]
    else:
        d0 = d0[rows0]
        d1 = d1[rows1]
    full = 4 * d0.shape[1]
    r0, i0 = _bits_1bit(d0, full)
    r1, i1 = _bits_1bit(d1, full)
    R0 = (1 - 2 * (r0 ^ r1)).sum(axis=0)
    R1 = (1 - 2 * (i0 ^ i1)).sum(axis=0)
    IM0 = (1 - 2 * (r0 ^ i1)).sum(axis=0)
    IM1 = (1 - 2 * (r1 ^ i0)).sum(axis=0)
    return [x[:nchan].astype(np.int64) for x in (R0, R1, IM0, IM1)]


def van_vleck_correction(R0, R1, I0, I1, rowcount):
    """correlations.py:491-497."""
    h = np.pi / 2 / rowcount
    return np.sin(R0 * h) + np.sin(R1 * h) + 1j * (np.sin(I1 * h) - np.sin(I0 * h))


def avg_xcorr_all_ant(x, ntime):
    """correlations.py:115-151 ``avg_xcorr_all_ant_cpu`` (complex128
    accumulation) with the convention of correlations_gpu.py:95-121:
    ``out[i, j, f] = sum_k x[i, k, f] * conj(x[j, k, f]) / K``, full
    Hermitian ``(M, M, F)``."""
    xx = np.asarray(x, dtype=np.complex128)
    return np.einsum("ikf,jkf->ijf", xx, np.conj(xx)) / ntime


def avg_xcorr_all_ant_c64_products(x, ntime):
    """correlations.py:134-149 ``avg_xcorr_all_ant_cpu`` arithmetic step by
    step: each product ``x[i,k,f] * conj(x[j,k,f])`` of complex64 values is
    formed in float32 (``ar*br + ai*bi`` and ``ai*br - ar*bi``, each multiply
    and add rounded, as the complex64 product of the Numba loop), the products
    are summed in float64 (``corr_sum = 0.0 + 0.0j``), divided by K in
    float64 and stored as complex64; the lower triangle is the conjugate.
    Returns complex64 ``(M, M, F)``."""
    xx = np.asarray(x, dtype=np.complex64)
    M, K, F = xx.shape
    re = np.ascontiguousarray(xx.real, dtype=np.float32)
    im = np.ascontiguousarray(xx.imag, dtype=np.float32)
    out = np.empty((M, M, F), dtype=np.complex64)
    for i in range(M):
        for j in range(i, M):
            pr = np.add(np.multiply(re[i], re[j]), np.multiply(im[i], im[j]), dtype=np.float32)
            pi = np.subtract(np.multiply(im[i], re[j]), np.multiply(re[i], im[j]), dtype=np.float32)
            sr = pr.astype(np.float64).sum(axis=0) / K
            si = pi.astype(np.float64).sum(axis=0) / K
            out[i, j] = (sr + 1j * si).astype(np.complex64)
            out[j, i] = np.conj(out[i, j])
    return out


def pair_list(ninput):
    """Pairs (i, j) with i <= j in row-major order (autos included)."""
    return [(i, j) for i in range(ninput) for j in range(i, ninput)]


def allpairs_4bit_block_sums(inputs, blocklen):
    """All-pair 4-bit block sums for the direct correlator: inputs
    ``[ninput, T, nchan]`` per-pol bytes -> int64 ``[nblock, npair, nchan, 2]``
    of ``x_i * conj(x_j)`` (correlations_cpu.c:172-173 per pair)."""
    inp = np.asarray(inputs, dtype=np.uint8)
    n, T, C = inp.shape
    nb = T // blocklen
    r, im = _decode(inp[:, : nb * blocklen])
    r = r.reshape(n, nb, blocklen, C)
    im = im.reshape(n, nb, blocklen, C)
    pairs = pair_list(n)
    out = np.zeros((nb, len(pairs), C, 2), dtype=np.int64)
    for p, (i, j) in enumerate(pairs):
        out[:, p, :, 0] = (r[i] * r[j] + im[i] * im[j]).sum(axis=1)
        out[:, p, :, 1] = (r[j] * im[i] - r[i] * im[j]).sum(axis=1)
    return out


def allpairs_1bit_block_sums(inputs, nchan, blocklen):
    """All-pair 1-bit block sums: inputs ``[ninput, T, ncol]`` sortpols 1-bit
    bytes -> int64 ``[nblock, npair, nchan, 2]`` (correlations_cpu.c:327-334
    per pair)."""
    inp = np.asarray(inputs, dtype=np.uint8)
    n, T, ncol = inp.shape
    nb = T // blocklen
    bits = [_bits_1bit(inp[k, : nb * blocklen], nchan) for k in range(n)]
    pairs = pair_list(n)
    out = np.zeros((nb, len(pairs), nchan, 2), dtype=np.int64)
    for p, (i, j) in enumerate(pairs):
        r0, i0 = bits[i]
        r1, i1 = bits[j]
        re = (2 - 2 * ((r0 ^ r1) + (i0 ^ i1))).reshape(nb, blocklen, nchan).sum(axis=1)
        im = (-2 * ((r1 ^ i0) - (r0 ^ i1))).reshape(nb, blocklen, nchan).sum(axis=1)
        out[:, p, :, 0] = re
        out[:, p, :, 1] = im
    return out
