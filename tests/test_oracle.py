"""Pins the CPU oracle: against the reference's own fixtures and asserts,
against the reference C sources compiled into oracle/_ref, and against
closed forms / brute force for what those leave open."""
import numpy as np
import pytest

from oracle import corr as ocorr
from oracle import pfb as opfb
from oracle import ref as oref
from oracle import stream as ostream
from oracle import unpack as ounpack

from tests import fixtures as fx
from tests.conftest import rel_err

needs_ref = pytest.mark.skipif(not oref.available(), reason="oracle/_ref not built (reference tree absent)")


# ------------------------------------------------------------------- unpack
def test_unpack_4bit_reference_fixture():
    # test_4bit_float_unpack.py:81-125
    packed = fx.packed_4bit()
    t0, t1 = fx.expected_complex_4bit()
    p0, p1 = ounpack.unpack_4bit(packed, 0, 2, np.arange(0, 10), 10)
    assert np.array_equal(p0, t0) and np.array_equal(p1, t1)
    p0, p1 = ounpack.unpack_4bit(packed, 0, 2, np.arange(3, 4), 10)
    assert np.array_equal(p0, t0[0:2, 3:4]) and np.array_equal(p1, t1[0:2, 3:4])
    p0, p1 = ounpack.unpack_4bit(packed, 0, 2, np.asarray([2, 5]), 10)
    assert np.array_equal(p0, t0[0:2, [2, 5]]) and np.array_equal(p1, t1[0:2, [2, 5]])


def test_nibble_rule_exhaustive():
    # v > 8 -> v - 16: 8 stays +8 (unpacking.c:189-202); the mars2019 legacy
    # decoder (albatrostools.py:246-249) is the independent statement
    v = np.arange(16)
    assert list(ounpack.nibble_decode(v)) == [0, 1, 2, 3, 4, 5, 6, 7, 8, -7, -6, -5, -4, -3, -2, -1]
    b = np.arange(256, dtype=np.uint8)
    r, i = ounpack.decode_4bit_bytes(b)
    assert np.array_equal(r, ounpack.nibble_decode(b >> 4)) and np.array_equal(i, ounpack.nibble_decode(b & 15))
    # packing then decoding is the identity on [-7, 8]
    vals = np.arange(-7, 9)
    assert np.array_equal(ounpack.nibble_decode(np.where(vals < 0, vals + 16, vals)), vals)


def test_unpack_1bit_reference_fixture():
    # test_1bit_unpack_gpu.py:186-218 (channels 2..5; and 2..3 of row 1)
    packed = fx.packed_1bit()
    t0, t1 = fx.expected_complex_1bit()
    p0, p1 = ounpack.unpack_1bit(packed, 0, 2, 2, 6, 8)
    assert np.array_equal(p0, t0[:, 2:6]) and np.array_equal(p1, t1[:, 2:6])
    p0, p1 = ounpack.unpack_1bit_gpu_semantics(packed, 1, 2, np.arange(2, 4), 8)
    assert p0[0, 1] == 1 + 1j and p1[0, 1] == -1 - 1j


def test_sortpols_1bit_reference_asserts():
    # test_1bit_unpack.py:81-107
    packed = fx.packed_1bit()
    p0, p1 = ounpack.sortpols(packed, 8, 1, 0, 2, 2, 7)
    assert p0[1, 0] == fx.str_to_bits("01110101")
    assert p0[0, 0] == fx.str_to_bits("10101010")
    assert p0[1, 1] == fx.str_to_bits("01000000")
    assert p1[1, 0] == fx.str_to_bits("10001010")
    assert p1[0, 0] == fx.str_to_bits("01010101")
    assert p1[1, 1] == fx.str_to_bits("10000000")
    p0, p1 = ounpack.sortpols(packed, 8, 1, 1, 2, 2, 7)
    assert p0[0, 0] == fx.str_to_bits("01110101") and p0[0, 1] == fx.str_to_bits("01000000")
    assert p1[0, 0] == fx.str_to_bits("10001010") and p1[0, 1] == fx.str_to_bits("10000000")
    # test_1bit_xcorr_2ant.py:76-94
    a1 = fx.packed_1bit_ant(fx.ant_1bit(im_sign=1))
    a2 = fx.packed_1bit_ant(fx.ant_1bit(im_sign=-1))
    p0, p1 = ounpack.sortpols(a1, 8, 1, 5, 6, 2, 3)
    assert p0[0, 0] == 192 and p1[0, 0] == 192
    p0, p1 = ounpack.sortpols(a2, 8, 1, 9, 10, 4, 7)
    assert p0[0, 0] == 168 and p1[0, 0] == 168


def test_hist_reference_asserts():
    # test_4bit_float_unpack.py:128-170
    packed = fx.packed_4bit()
    r0, im0 = fx.real_im_4bit_pol0()
    r1, im1 = fx.real_im_4bit_pol1()
    h = ounpack.hist_4bit(packed, 0, 2, 10, 0)
    assert h[0, 3] == 1 and h[4, 3] == 1
    assert h[0].sum() == (r0 == 0).sum() + (im0 == 0).sum()
    h = ounpack.hist_4bit(packed, 0, 2, 10, 1)
    assert h[0, 3] == 1 and h[12, 3] == 1
    assert h[7].sum() == (r1 == 7).sum() + (im1 == 7).sum()
    h = ounpack.hist_4bit(packed, 0, 2, 10, -1)
    assert h[12, 3] == 1 and h[4, 3] == 1
    # test_1bit_unpack.py:110-134
    p1 = fx.packed_1bit()
    h = ounpack.hist_1bit(p1, 0, 2, 8, 0)
    assert h[0, 3] == 1 and h[1, 3] == 3 and h[0, 4] == 2
    h = ounpack.hist_1bit(p1, 0, 2, 8, 1)
    assert h[0, 3] == 3 and h[1, 3] == 1 and h[0, 4] == 2
    h = ounpack.hist_1bit(p1, 0, 2, 8, -1)
    assert h[0, 3] == 4 and h[1, 3] == 4 and h[0, 4] == 4


@needs_ref
def test_unpack_family_vs_reference_c(rng):
    nchan, T = 37, 64
    raw = rng.integers(0, 256, size=(T, 2 * nchan), dtype=np.uint8)
    ch = np.array([0, 3, 5, 36, 7], dtype=np.int64)
    a = ounpack.unpack_4bit(raw, 3, 40, ch, nchan)
    b = oref.unpack_4bit_float(raw, 3, 40, ch, nchan)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    raw1 = rng.integers(0, 256, size=(T, 8), dtype=np.uint8)
    a = ounpack.unpack_1bit(raw1, 2, 30, 4, 12, 16)
    b = oref.unpack_1bit_float(raw1, 2, 30, 4, 12, 16)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    for cs, ce in [(0, 16), (2, 7), (4, 13), (0, 14), (2, 16), (6, 11), (0, 1), (2, 3)]:
        a = ounpack.sortpols(raw1, 16, 1, 1, 40, cs, ce)
        b = oref.sortpols(raw1, 16, 1, 1, 40, cs, ce)
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), (cs, ce)
    a = ounpack.sortpols(raw, nchan, 4, 1, 40, 3, 30)
    b = oref.sortpols(raw, nchan, 4, 1, 40, 3, 30)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    for mode in (0, 1, -1):
        assert np.array_equal(ounpack.hist_4bit(raw, 2, 45, nchan, mode), oref.hist(raw, 2, 45, nchan, 4, mode))
        assert np.array_equal(ounpack.hist_1bit(raw1, 2, 45, 16, mode), oref.hist(raw1, 2, 45, 16, 1, mode))


# -------------------------------------------------------------- correlators
def test_xcorr_4bit_brute_force(rng):
    d0 = rng.integers(0, 256, size=(50, 9), dtype=np.uint8)
    d1 = rng.integers(0, 256, size=(50, 9), dtype=np.uint8)
    r0, i0 = ounpack.decode_4bit_bytes(d0)
    r1, i1 = ounpack.decode_4bit_bytes(d1)
    x0 = r0 + 1j * i0
    x1 = r1 + 1j * i1
    re, im = ocorr.xcorr_4bit_sums(d0, d1, 50)
    ref = (x0 * np.conj(x1)).sum(axis=0)
    assert np.array_equal(re, ref.real.astype(np.int64)) and np.array_equal(im, ref.imag.astype(np.int64))
    assert np.array_equal(ocorr.autocorr_4bit_sums(d0, 50), (np.abs(x0) ** 2).sum(axis=0).astype(np.int64))
    # a plausible mistake -- conjugating the wrong operand -- flips Im
    assert not np.array_equal(im, (np.conj(x0) * x1).sum(axis=0).imag.astype(np.int64))


def test_xcorr_1bit_identity(rng):
    # 2 - 2[(r0^r1) + (i0^i1)] and -2[(r1^i0) - (r0^i1)] equal the +-1 product
    for nch in (16, 15, 14, 13):
        p = rng.integers(0, 256, size=(40, 4), dtype=np.uint8)
        q = rng.integers(0, 256, size=(40, 4), dtype=np.uint8)
        re, im = ocorr.xcorr_1bit_sums(p, q, nch, 40)
        r0, i0 = ocorr._bits_1bit(p, nch)
        r1, i1 = ocorr._bits_1bit(q, nch)
        x0 = (2 * r0 - 1) + 1j * (2 * i0 - 1)
        x1 = (2 * r1 - 1) + 1j * (2 * i1 - 1)
        ref = (x0 * np.conj(x1)).sum(axis=0)
        assert np.array_equal(re, ref.real.astype(np.int64)) and np.array_equal(im, ref.imag.astype(np.int64))


def test_allpairs_match_pairwise(rng):
    inp = rng.integers(0, 256, size=(3, 64, 8), dtype=np.uint8)
    out = ocorr.allpairs_4bit_block_sums(inp, 16)
    for p, (i, j) in enumerate(ocorr.pair_list(3)):
        for b in range(4):
            re, im = ocorr.xcorr_4bit_sums(inp[i, 16 * b:16 * b + 16], inp[j, 16 * b:16 * b + 16], 16)
            assert np.array_equal(out[b, p, :, 0], re) and np.array_equal(out[b, p, :, 1], im)
    inp1 = rng.integers(0, 256, size=(3, 64, 4), dtype=np.uint8)
    out = ocorr.allpairs_1bit_block_sums(inp1, 15, 32)
    for p, (i, j) in enumerate(ocorr.pair_list(3)):
        re, im = ocorr.xcorr_1bit_sums(inp1[i, 32:], inp1[j, 32:], 15, 32)
        assert np.array_equal(out[1, p, :, 0], re) and np.array_equal(out[1, p, :, 1], im)


def test_vanvleck_reference_property(rng):
    # test_1bit_xcorr_2ant.py:253-304 with a seeded generator
    R0 = 2 * (rng.standard_normal((10, 8)) > 0) - 1
    I0 = 2 * (rng.standard_normal((10, 8)) > 0) - 1
    R1 = 2 * (rng.standard_normal((10, 8)) > 0) - 1
    I1 = 2 * (rng.standard_normal((10, 8)) > 0) - 1
    pa = ounpack.sortpols(fx.packed_1bit_ant((R0, I0)), 8, 1, 0, 10, 0, 8)[0]
    pb = ounpack.sortpols(fx.packed_1bit_ant((R1, I1)), 8, 1, 0, 10, 0, 8)[0]
    m0 = np.ones(10, bool)
    m1 = np.ones(10, bool)
    m0[0:4] = False
    m1[6:8] = False
    t0 = np.zeros_like(pa)
    t1 = np.zeros_like(pa)
    t0[: m0.sum()] = pa[m0]
    t1[: m1.sum()] = pb[m1]
    s0 = np.arange(10)[m0]
    s1 = np.arange(10)[m1]
    rows0, rows1 = ocorr.intersect_rows(s0, s1, 0, 0)
    m = m0 & m1
    assert len(rows0) == m.sum()
    got = ocorr.vanvleck_sums(t0, t1, 8, None, rows0, rows1)
    true = (np.sum(R0[m] * R1[m], axis=0), np.sum(I0[m] * I1[m], axis=0), np.sum(R0[m] * I1[m], axis=0),
            np.sum(R1[m] * I0[m], axis=0))
    for a, b in zip(got, true):
        assert np.array_equal(a, b)


def test_intersect_rows_brute_force(rng):
    s0 = np.sort(rng.choice(200, 80, replace=False))
    s1 = np.sort(rng.choice(200, 90, replace=False))
    r0, r1 = ocorr.intersect_rows(s0, s1, 5, 2)
    common = sorted(set((s0 - 5).tolist()) & set((s1 - 2).tolist()))
    assert np.array_equal(s0[r0] - 5, common) and np.array_equal(s1[r1] - 2, common)


@needs_ref
def test_correlators_vs_reference_c(rng, capfd):
    T, nch = 60, 37
    raw = rng.integers(0, 256, size=(T, 2 * nch), dtype=np.uint8)
    p0, p1 = ounpack.sortpols(raw, nch, 4, 0, T, 0, nch)
    assert np.array_equal(ocorr.autocorr_4bit_sums(p0, 44), oref.avg_autocorr_4bit(p0, 44))
    re, im = ocorr.xcorr_4bit_sums(p0, p1, 44)
    o = oref.avg_xcorr_4bit(p0, p1, 44)
    assert np.array_equal(re, o.real.astype(np.int64)) and np.array_equal(im, o.imag.astype(np.int64))
    s0 = np.sort(rng.choice(100, 50, replace=False))
    s1 = np.sort(rng.choice(100, 50, replace=False))
    re, im, n = ocorr.xcorr_4bit_2ant_sums(p0, p1, s0, s1, 0, 0)
    o, n2 = oref.avg_xcorr_4bit_2ant(p0, p1, s0, s1, 0, 0)
    assert n == n2 and np.array_equal(re, o.real.astype(np.int64)) and np.array_equal(im, o.imag.astype(np.int64))
    raw1 = rng.integers(0, 256, size=(T, 8), dtype=np.uint8)
    for n1 in (16, 15, 14, 13):
        q0, _ = ounpack.sortpols(raw1, 16, 1, 0, T, 0, n1)
        q1, _ = ounpack.sortpols(raw1[::-1].copy(), 16, 1, 0, T, 0, n1)
        re, im = ocorr.xcorr_1bit_sums(q0, q1, n1, T)
        o = oref.avg_xcorr_1bit(q0, q1, n1, T)
        assert np.array_equal(re, o.real.astype(np.int64)) and np.array_equal(im, o.imag.astype(np.int64))
    q0, _ = ounpack.sortpols(raw1, 16, 1, 0, T, 0, 16)
    q1, _ = ounpack.sortpols(raw1[::-1].copy(), 16, 1, 0, T, 0, 16)
    a = ocorr.vanvleck_sums(q0, q1, 16, T)
    b = oref.avg_xcorr_1bit_vanvleck(q0, q1, 16, T)
    assert all(np.array_equal(x, y.astype(np.int64)) for x, y in zip(a, b))
    s0, s1 = s0[:T], s1[:T]
    r0, r1 = ocorr.intersect_rows(s0, s1, 0, 0)
    a = ocorr.vanvleck_sums(q0, q1, 16, None, r0, r1)
    b, n = oref.avg_xcorr_1bit_vanvleck_2ant(q0, q1, s0, s1, 0, 0, 16)
    assert n == len(r0) and all(np.array_equal(x, y.astype(np.int64)) for x, y in zip(a, b))
    capfd.readouterr()


def test_all_ant_reference_constants():
    # test_gpu_all_ant_complex_xcorr.py:35-68
    x, nant, npol, ntime, nfreq, exp = fx.xcorr_constants()
    out = ocorr.avg_xcorr_all_ant(x, ntime)
    for (i, j, f), v in exp.items():
        assert out[i, j, f] == v
        assert out[j, i, f] == np.conj(v)


def test_numpy_complex64_division_is_reciprocal_multiply():
    # the averaging convention: complex64 / int == a * fl32(1/n) (correlations.py:327)
    rng = np.random.default_rng(3)
    a = (rng.integers(-2 ** 20, 2 ** 20, 1000) + 1j * rng.integers(-2 ** 20, 2 ** 20, 1000)).astype(np.complex64)
    for n in (3, 7, 4093, 4095, 4096):
        r = np.float32(1) / np.float32(n)
        assert np.array_equal(a / n, (a.real * r + 1j * (a.imag * r)).astype(np.complex64))


# ---------------------------------------------------------------------- PFB
def test_window_closed_form():
    w = opfb.sinc_hamming(4, 64)
    N = 256
    n = np.arange(N)
    ham = 0.54 - 0.46 * np.cos(2 * np.pi * n / (N - 1))
    x = (n - N / 2) / 64
    sinc = np.where(x == 0, 1.0, np.sin(np.pi * x) / np.where(x == 0, 1, np.pi * x))
    assert np.allclose(w, ham * sinc, rtol=0, atol=1e-15)
    assert w[N // 2] == pytest.approx(np.hamming(N)[N // 2])


def test_matft_closed_form_vs_rfft_construction():
    for nslice in (16, 64, 4096):
        a = opfb.matft_fp64(nslice, nchan=33)
        b = opfb.matft_closed_form(nslice, nchan=33)
        assert np.max(np.abs(a - b)) < 1e-12
    # W[k, 0] is the sum of the four taps of phase k
    w = opfb.window_f32(4, 64).astype(np.float64).reshape(4, 64)
    assert np.allclose(opfb.matft_closed_form(16, nchan=33)[:, 0], w.sum(axis=0))


def test_wiener_filter_identities():
    W = opfb.matft_closed_form(64, nchan=33)
    f0 = opfb.wiener_filter(W, 0.0)
    assert np.allclose(f0 * np.conj(W), 1.0)
    t = 0.45
    ft = opfb.wiener_filter(W, t)
    assert np.allclose(ft, W * (1 + t * t) / (np.abs(W) ** 2 + t * t))
    assert np.max(np.abs(ft)) <= (1 + t * t) / (2 * t) + 1e-12


def test_ipfb_inverts_pfb_at_zero_threshold():
    # a circulant forward PFB followed by the unregularised iPFB is the identity
    nchan, ntap, nslice = 33, 4, 64
    lblock = 2 * (nchan - 1)
    rng = np.random.default_rng(7)
    ts = rng.standard_normal(nslice * lblock)
    ext = np.concatenate([ts, ts[: (ntap - 1) * lblock]])  # periodic extension
    win = opfb.window_f32(ntap, lblock)
    spec = opfb.pfb(ext, win, nchan, ntap)
    assert spec.shape == (nslice, nchan)
    filt = opfb.wiener_filter(opfb.matft_fp64(nslice, nchan, ntap, win), 0.0)
    rec = opfb.ipfb(spec, filt)
    assert rel_err(rec.ravel(), ts) < 1e-9


def test_fft_conventions_brute_force():
    rng = np.random.default_rng(9)
    x = rng.standard_normal((3, 32))
    assert np.allclose(opfb.dft_brute(x)[:, :17], np.fft.rfft(x, axis=1))
    X = np.fft.rfft(x, axis=1)
    X[:, 0] += 1j * 5  # imaginary DC part is ignored by the inverse real FFT
    assert np.allclose(opfb.irfft_brute(X, 32), np.fft.irfft(X, axis=1))
    assert np.allclose(opfb.irfft_brute(X, 32), x)


def test_pfb_brute_force_and_tone():
    nchan, ntap = 17, 4
    L = 32
    rng = np.random.default_rng(11)
    ts = rng.standard_normal(10 * L)
    win = opfb.window_f32(ntap, L)
    out = opfb.pfb(ts, win, nchan, ntap)
    y = np.array([sum(ts[(i + j) * L:(i + j + 1) * L] * win.astype(np.float64)[j * L:(j + 1) * L]
                      for j in range(ntap)) for i in range(7)])
    assert np.allclose(out, opfb.dft_brute(y)[:, :nchan])
    # a tone at channel centre c lands in channel c with the summed window gain
    c = 5
    n = np.arange(10 * L)
    out = opfb.pfb(np.cos(2 * np.pi * c * n / L), win, nchan, ntap)
    assert np.argmax(np.abs(out[3])) == c
    # y[k] = cos(2 pi c k / L) * Wsum[k]  ->  X[c] = (sum Wsum + sum Wsum e^{-4 pi i c k / L}) / 2
    wsum = win.astype(np.float64).reshape(ntap, L).sum(axis=0)
    k = np.arange(L)
    expect = 0.5 * wsum.sum() + 0.5 * (wsum * np.exp(-4j * np.pi * c * k / L)).sum()
    assert abs(out[3, c] - expect) < 1e-12


# ------------------------------------------------------------------ stream
def test_chunk_planner_worked_cases():
    # helper_gpu_stream.py:256-284 traced by hand
    # stride 10, need 25: 30 samples -> job, 20 left; +10 -> job; 20 left, no chunks -> tail job
    assert ostream.get_chunkier(4, 10, 2, 5, 4) == [(0, 3), (3, 4), (4, None)]
    # stride 20, need 50, chunks of 100
    assert ostream.get_chunkier(3, 100, 2, 10, 4) == [(0, 1), (1, 1), (1, 1), (1, 2), (2, 2), (2, 2), (2, 2),
                                                      (2, 2), (2, 3), (3, 3), (3, 3), (3, 3), (3, 3), (3, None)]
    assert ostream.get_chunkier(0, 10, 2, 5, 4) == []
    # property: every complete job s has consumed >= s*stride + need samples
    lchunk, nblock, lblock, ntap = 7, 3, 2, 4
    jobs = ostream.get_chunkier(20, lchunk, nblock, lblock, ntap)
    for s, (c0, c1) in enumerate(jobs):
        if c1 is not None:
            assert c1 * lchunk >= s * nblock * lblock + (nblock + ntap - 1) * lblock


def test_make_continuous_reference_case():
    # test_bdc_helpers.py:5-25
    nspec, nchans, channels, specnum = fx.make_continuous_case()
    rng = np.random.default_rng(0)
    spec = rng.standard_normal((len(specnum), len(channels))).astype(np.complex64)
    out = ostream.make_continuous(spec, specnum, 0, channels, nspec, nchans)
    assert out[10, 100] == spec[0, 0] and out[50, 200] == spec[4, 1]
    assert out[49, 200] == 0 and out[50, 199] == 0 and out[9, 100] == 0 and out[10, 101] == 0
    ones = np.ones((len(specnum), len(channels)), np.complex64) * (1 + 1j)
    out = ostream.make_continuous(ones, specnum, 0, channels, nspec, nchans)
    assert out.sum() == 10 + 10j and out.sum(axis=0)[100] == 5 + 5j


def test_stream_is_continuous_away_from_cut_edges():
    # concatenated kept rows of successive chunks = iPFB of the whole stream
    # (away from the chunk edges, where the finite cut leaves an edge error)
    nch, acclen, cut = 2049, 60, 2
    rng = np.random.default_rng(5)
    nchunks = 3
    specs = rng.standard_normal((nchunks * acclen, nch)) + 1j * rng.standard_normal((nchunks * acclen, nch))
    chunks = [{"pol0": specs[c * acclen:(c + 1) * acclen], "specnums": np.arange(c * acclen, (c + 1) * acclen)}
              for c in range(nchunks)]
    win = opfb.window_f32(4, 4096)
    filt = opfb.wiener_filter(opfb.matft_fp64(acclen + 2 * cut, nch, 4, win), 0.45)
    st = ostream.ipfb_stream(chunks, 0, np.arange(nch), acclen, cut, filt, 0)
    assert st.shape == (nchunks * acclen * 4096,)
    # the first kept rows of chunk 1 equal rows [cut:] of an iPFB whose carried rows are chunk 0's tail
    tile = np.concatenate([specs[acclen - 2 * cut: acclen], specs[acclen: 2 * acclen]])
    ref = opfb.ipfb(tile, filt)[cut:-cut].ravel()
    assert np.allclose(st[acclen * 4096: 2 * acclen * 4096], ref)


def test_make_continuous_repeated_channel_keeps_last_column():
    # numpy's fancy assignment (baseband_data_classes.py:72) keeps the last
    # column of a repeated channel number; the device paths follow it
    from oracle import stream as ostream
    spec = np.array([[1 + 1j, 2 + 2j, 3 + 3j], [4 + 4j, 5 + 5j, 6 + 6j]])
    out = ostream.make_continuous(spec, [10, 11], 10, [5, 7, 5], 4, nchans=9)
    assert out[0, 5] == 3 + 3j and out[1, 5] == 6 + 6j and out[0, 7] == 2 + 2j
    assert np.count_nonzero(out) == 4


# ---------------------------------------------------------------------------
# missing fraction: hand-traced cases (tests/fixtures.py MISSING_CASES)
MISSING_CASES = fx.MISSING_CASES


@pytest.mark.parametrize("case", range(len(MISSING_CASES)))
def test_missing_fraction_hand_traced(case):
    lchunk, nblock, lblock, present, jobs_expected, expected = MISSING_CASES[case]
    nchunks = len(present)
    jobs = ostream.get_chunkier(nchunks, lchunk, nblock, lblock, 4)
    assert jobs == jobs_expected
    chunks = [[{"specnums": np.arange(n)} for n in present]]
    got = ostream.missing_fraction(chunks, 100, nchunks, nblock, lblock, lchunk, 4, jobs)
    assert got.shape == (1, len(jobs))
    assert np.allclose(got[0], expected, equal_nan=True, rtol=0, atol=1e-12)


def test_missing_fraction_antennas_keep_their_own_lists():
    # two antennas: each row equals the one-antenna result of that antenna
    # (the per-antenna lists of helper_gpu_stream_clean.py:204)
    lchunk, nblock, lblock, present, _, exp0 = MISSING_CASES[0]
    other = [50, 100, 100, 80, 100, 100]
    jobs = ostream.get_chunkier(len(present), lchunk, nblock, lblock, 4)
    chunks = [[{"specnums": np.arange(n)} for n in present], [{"specnums": np.arange(n)} for n in other]]
    got = ostream.missing_fraction(chunks, 100, len(present), nblock, lblock, lchunk, 4, jobs)
    alone = ostream.missing_fraction(chunks[1:], 100, len(present), nblock, lblock, lchunk, 4, jobs)
    assert np.allclose(got[0], exp0, equal_nan=True)
    assert np.allclose(got[1], alone[0], equal_nan=True)
    # antenna 1 by hand: job 0 (0, 50, 0)/3, job 1 (0, 50, 0, 20)/4, job 2 (0, 50, 20, 0)/4
    assert np.allclose(got[1][:3], [50 / 3, 17.5, 17.5])


def test_van_vleck_correction_closed_forms():
    # sin(pi/2 * R/n): R = n -> 1, R = n/3 -> 1/2, R = 0 -> 0; imag = I1 - I0 terms
    n = 3000
    v = ocorr.van_vleck_correction(np.float32(n), np.float32(n / 3), np.float32(0), np.float32(n), n)
    assert abs(v - (1.5 + 1j * (1.0 - 0.0))) < 1e-12
    v = ocorr.van_vleck_correction(0.0, 0.0, np.float32(-n), np.float32(n / 3), n)
    assert abs(v - (0.0 + 1j * (0.5 + 1.0))) < 1e-12


def test_van_vleck_correction_recovers_gaussian_correlation():
    # arcsine law: sign-bit products of unit-variance Gaussians with
    # correlation rho average to (2/pi) asin(rho); the correction returns the
    # complex correlation E[x0 conj(x1)] = 2 a e^{i phi} of the generator
    rng = np.random.default_rng(7)
    n, a, phi = 400000, 0.3, 0.7
    z = (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    w0 = (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    w1 = (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    x0 = np.sqrt(a) * z + np.sqrt(1 - a) * w0
    x1 = np.sqrt(a) * np.exp(-1j * phi) * z + np.sqrt(1 - a) * w1
    s = lambda v: np.sign(v)  # noqa: E731
    R0 = np.sum(s(x0.real) * s(x1.real))
    R1 = np.sum(s(x0.imag) * s(x1.imag))
    I0 = np.sum(s(x0.real) * s(x1.imag))
    I1 = np.sum(s(x1.real) * s(x0.imag))
    v = ocorr.van_vleck_correction(R0, R1, I0, I1, n)
    assert abs(v - 2 * a * np.exp(1j * phi)) < 0.01
    # the uncorrected sign correlation is biased low (2/pi asin < identity)
    raw = (R0 + R1 + 1j * (I1 - I0)) / n
    assert abs(raw) < abs(v) - 0.02


def test_complex64_product_oracle():
    # the step-by-step arithmetic of avg_xcorr_all_ant_cpu: the reference
    # constants exactly, the complex128 form to product rounding, Hermitian
    x, nant, npol, ntime, nfreq, exp = fx.xcorr_constants()
    got = ocorr.avg_xcorr_all_ant_c64_products(x, ntime)
    for (i, j, f), v in exp.items():
        assert got[i, j, f] == v
    rng = np.random.default_rng(3)
    y = (rng.standard_normal((5, 300, 7)) + 1j * rng.standard_normal((5, 300, 7))).astype(np.complex64)
    a = ocorr.avg_xcorr_all_ant_c64_products(y, 300)
    b = ocorr.avg_xcorr_all_ant(y, 300)
    assert a.dtype == np.complex64
    assert np.max(np.abs(a - b)) < 1e-6 * np.max(np.abs(b))
    assert np.array_equal(a, np.conj(a.transpose(1, 0, 2)))
    # the products are rounded: not bit-identical to complex128 products
    assert not np.array_equal(a, b.astype(np.complex64))
    # the timed CPU port (channel chunks on threads) is the same arithmetic,
    # for any chunking; F-order input as the correlator takes it
    from oracle import cpu_port
    yf = np.asfortranarray(y)
    for workers in (1, 3):
        assert np.array_equal(cpu_port.avg_xcorr_all_ant(yf, workers), a)


def test_orbcomm_coarse_xcorr_equals_direct_lag_sums():
    # the padded-FFT form of orbcomm_utils_gpu.py:39-52 against direct lag
    # sums; a copy of f1 delayed by 7 spectra peaks at lag +7
    from oracle import orbcomm as oorb
    rng = np.random.default_rng(4)
    N, nch, dN = 300, 3, 40
    f1 = rng.standard_normal((N, nch)) + 1j * rng.standard_normal((N, nch))
    f2 = rng.standard_normal((N, nch)) + 1j * rng.standard_normal((N, nch))
    a = oorb.coarse_xcorr(f1, f2, dN)
    assert a.shape == (nch, 2 * dN)
    assert np.max(np.abs(a - oorb.lag_sums_brute(f1, f2, dN))) < 1e-12
    g = np.zeros_like(f1)
    g[7:] = f1[:-7]                                         # f2[t] = f1[t - 7] -> peak at lag -7
    c = oorb.coarse_xcorr(f1, g, dN)
    assert np.all(np.argmax(np.abs(c), axis=1) == dN - 7)
    d = oorb.apply_delay(f1, np.zeros(N), np.linspace(1e6, 2e6, nch))
    assert np.array_equal(d, f1)
    e = oorb.apply_delay(np.ones((2, 4)), np.array([1e-6, 2e-6]), np.array([0.0, 1e5, 2e5, 3e5]))
    assert np.allclose(np.abs(e), 1.0) and np.allclose(np.angle(e[1, 1]), -2 * np.pi * 0.2)
