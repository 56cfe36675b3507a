"""Known-answer inputs of the reference's own tests, rebuilt as data.

Each builder reproduces the fixture values of the cited reference test
(src/correlations/tests/...), so the oracle and the GPU path can be checked
against the same expected values the reference asserts.
"""
import numpy as np


def str_to_bits(s):
    """test_1bit_unpack.py:6-8."""
    return sum(int(s[i]) * 2 ** (7 - i) for i in range(8))


def real_im_4bit_pol0():
    """test_4bit_float_unpack.py:28-35: 2 spectra x 10 channels."""
    r = np.ones(10, dtype="int64").reshape(-1, 10)
    r = np.vstack([r, 0 * r])
    r[1, 3] = 4
    im = -1 * np.ones(10, dtype="int64").reshape(-1, 10)
    im = np.vstack([im, 0 * im])
    return r, im


def real_im_4bit_pol1():
    """test_4bit_float_unpack.py:38-45."""
    r = 7 * np.ones(10, dtype="int64").reshape(-1, 10)
    r = np.vstack([r, 0 * r])
    r[1, 3] = -4
    im = -7 * np.ones(10, dtype="int64").reshape(-1, 10)
    im = np.vstack([im, 0 * im])
    return r, im


def packed_4bit():
    """test_4bit_float_unpack.py:57-79: raw interleaved pol0/pol1 bytes."""
    out = []
    for r, im in (real_im_4bit_pol0(), real_im_4bit_pol1()):
        r = r.copy()
        im = im.copy()
        r[r < 0] += 16
        im[im < 0] += 16
        out.append(((r << 4) + im).astype("uint8"))
    p0, p1 = out
    packed = np.zeros((1, p0.size + p1.size), dtype="uint8")
    packed[0, 0::2] = np.ravel(p0)
    packed[0, 1::2] = np.ravel(p1)
    return packed


def expected_complex_4bit():
    """test_4bit_float_unpack.py:48-54."""
    a, b = real_im_4bit_pol0(), real_im_4bit_pol1()
    return a[0] + 1j * a[1], b[0] + 1j * b[1]


def real_im_1bit_pol0():
    """test_1bit_unpack.py:11-19: 2 spectra x 8 channels of +-1."""
    r = np.ones(8, dtype="int64").reshape(-1, 8)
    r = np.vstack([r, -1 * r])
    r[1, 3] = 1
    im = -1 * np.ones(8, dtype="int64").reshape(-1, 8)
    im = np.vstack([im, -1 * im])
    return r, im


def real_im_1bit_pol1():
    """test_1bit_unpack.py:22-29."""
    r = -1 * np.ones(8, dtype="int64").reshape(-1, 8)
    r = np.vstack([r, -1 * r])
    r[1, 3] = -1
    im = np.ones(8, dtype="int64").reshape(-1, 8)
    im = np.vstack([im, -1 * im])
    return r, im


def expected_complex_1bit():
    a, b = real_im_1bit_pol0(), real_im_1bit_pol1()
    return a[0] + 1j * a[1], b[0] + 1j * b[1]


def packed_1bit():
    """test_1bit_unpack.py:41-65: raw 1-bit bytes (2 channels x 2 pols per byte)."""
    p0r, p0i = [x.copy() for x in real_im_1bit_pol0()]
    p1r, p1i = [x.copy() for x in real_im_1bit_pol1()]
    for a in (p0r, p0i, p1r, p1i):
        a[a < 0] += 1
    packed = np.zeros((1, 8), dtype="uint8")
    for i in range(2):
        for j in range(4):
            packed[0, i * 4 + j] = ((p0r[i, 2 * j] << 7) + (p0i[i, 2 * j] << 6) + (p1r[i, 2 * j] << 5)
                                    + (p1i[i, 2 * j] << 4) + (p0r[i, 2 * j + 1] << 3) + (p0i[i, 2 * j + 1] << 2)
                                    + (p1r[i, 2 * j + 1] << 1) + p1i[i, 2 * j + 1])
    return packed


def ant_1bit(nspec=10, im_sign=1):
    """test_1bit_xcorr_2ant.py:13-30: all-ones real, imag = +-1."""
    r = np.ones(8 * nspec, dtype="int64").reshape(-1, 8)
    return r, im_sign * r.copy()


def packed_1bit_ant(ant):
    """test_1bit_xcorr_2ant.py:51-76: both pols equal."""
    p0r, p0im = ant[0].copy(), ant[1].copy()
    p1r, p1im = p0r.copy(), p0im.copy()
    nchan = ant[0].shape[1]
    nspec = ant[0].shape[0]
    packed = np.zeros((1, nchan // 2 * nspec), dtype="uint8")
    for a in (p0r, p1r, p0im, p1im):
        a[a < 0] += 1
    for i in range(nspec):
        for j in range(nchan // 2):
            packed[0, i * 4 + j] = ((p0r[i, 2 * j] << 7) + (p0im[i, 2 * j] << 6) + (p1r[i, 2 * j] << 5)
                                    + (p1im[i, 2 * j] << 4) + (p0r[i, 2 * j + 1] << 3) + (p0im[i, 2 * j + 1] << 2)
                                    + (p1r[i, 2 * j + 1] << 1) + p1im[i, 2 * j + 1])
    return packed


def xcorr_constants():
    """test_gpu_all_ant_complex_xcorr.py:35-68: 3 antennas x 1 pol, two
    channels of constants 3,5,7 and 4,6,8 over 10000 times; expected
    products 9, 15, 21, 35 / 16, 24, 32, 48."""
    nant, npol, ntime, nfreq = 3, 1, 10000, 2
    x = np.empty((nant * npol, ntime, nfreq), dtype="complex64", order="F")
    x[0, :, 0] = 3
    x[1, :, 0] = 5
    x[2, :, 0] = 7
    x[0, :, 1] = 4
    x[1, :, 1] = 6
    x[2, :, 1] = 8
    expected = {(0, 0, 0): 9, (0, 1, 0): 15, (0, 2, 0): 21, (1, 2, 0): 35,
                (0, 0, 1): 16, (0, 1, 1): 24, (0, 2, 1): 32, (1, 2, 1): 48}
    return x, nant, npol, ntime, nfreq, expected


def make_continuous_case():
    """test_bdc_helpers.py:5-25 (current 7-argument signature)."""
    nspec, nchans = 100, 2049
    channels = [100, 200]
    specnum = np.asarray([10, 20, 30, 40, 50], dtype="int64")
    return nspec, nchans, channels, specnum


def write_raw(path, chans, spp, specnums, bits=4, missing=(), payload=None):
    """A .raw file in the format of generate_test_baseband.py:52-152; the
    spectra are that script's pattern, or ``payload[p]`` (uint8, one row of
    bytes_per_packet - 4 per packet) when given."""
    bps = len(chans)
    bpp = bps * spp + 4
    hdr = np.asarray([8 * 10 + 8 * len(chans), bpp, len(chans), spp, bits, 0], dtype=">Q")
    with open(path, "wb") as f:
        hdr.tofile(f)
        np.asarray(chans, dtype=">Q").tofile(f)
        np.asarray([0, 1234567], dtype=">Q").tofile(f)
        np.asarray([1.0, 2.0, 3.0], dtype=">d").tofile(f)
        for p, s in enumerate(specnums):
            if p in missing:
                continue
            if payload is None:
                spectrum = np.ones(bps, dtype="u1")
                spectrum[2:] = spectrum[2:] << 4
                body = np.tile(spectrum, spp)
            else:
                body = payload[p]
            np.asarray((int(s) % 2 ** 32, body),
                       dtype=[("spec_num", ">I"), ("spectra", "%dB" % (bpp - 4))]).tofile(f)


# missing fraction: hand-traced cases of helper_gpu_stream.py:157-233.
# Units are abstract sample counts (lblock, lchunk); acclen = 100, so a chunk
# with n present spectra is (1 - n/100) * 100 % missing.  One antenna, so the
# reference's shared list (helper_gpu_stream.py:125) and the per-antenna lists
# (helper_gpu_stream_clean.py:204) agree.  Each entry: (lchunk, nblock, lblock,
# present counts, planner output, expected per-job percentages).
MISSING_CASES = [
    # szblock 60, stride 30: chunk 2 spills 15 samples into job 1; the list
    # entries of chunks 0-2 reach jobs 0-1 (counters 3, 3, 2), chunk 3 jobs 1-2
    (25, 3, 10, [100, 90, 80, 100, 50, 70], [(0, 3), (3, 4), (4, 5), (5, 6), (6, None)],
     [10.0, 7.5, 15.0, 30.0, np.nan]),
    # chunks of 95 samples: the carried remainder alone fills jobs 1, 3, 4, 6, 7
    # (first branch, nothing appended); a job whose list is empty is 0/0 = nan
    (95, 3, 10, [90, 80, 60], [(0, 1), (1, 1), (1, 2), (2, 2), (2, 2), (2, 3), (3, 3), (3, 3), (3, None)],
     [10.0, np.nan, 20.0, 20.0, np.nan, 40.0, 40.0, np.nan, np.nan]),
    # every chunk exactly two blocks: counters all 1, each job reports its own chunks
    (20, 2, 10, [100, 90, 80, 100, 50, 100, 70], [(0, 3), (3, 4), (4, 5), (5, 6), (6, 7), (7, None)],
     [10.0, 0.0, 50.0, 0.0, 30.0, np.nan]),
]
