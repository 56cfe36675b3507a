"""GPU parity of the unpack family against the oracle (bit exact)."""
import os
import numpy as np
import pytest

from oracle import ref as oref
from oracle import stream as ostream
from oracle import unpack as ounpack
from tests import fixtures as fx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def unpk():
    from albatros_axion_detection_amd import unpacking
    return unpacking


def test_unpack_4bit_reference_fixture(unpk):
    # test_4bit_float_unpack.py:28-125 and test_4bit_unpack_gpu.py:6-89: the
    # reference's packed fixture, all channels, one channel, and channels [2, 5]
    packed = fx.packed_4bit()
    t0, t1 = fx.expected_complex_4bit()
    p0, p1 = unpk.unpack_4bit(packed, 0, 2, np.arange(0, 10), 10)
    assert p0.dtype == np.complex64 and p0.shape == t0.shape
    assert np.array_equal(p0, t0) and np.array_equal(p1, t1)
    p0, p1 = unpk.unpack_4bit(packed, 0, 2, np.arange(3, 4), 10)
    assert np.array_equal(p0, t0[0:2, 3:4]) and np.array_equal(p1, t1[0:2, 3:4])
    p0, p1 = unpk.unpack_4bit(packed, 0, 2, np.asarray([2, 5], dtype="int64"), 10)
    assert np.array_equal(p0, t0[0:2, [2, 5]]) and np.array_equal(p1, t1[0:2, [2, 5]])


def test_unpack_4bit_exhaustive_and_random(unpk, rng):
    # every byte value in both pols, then a random block with a channel gather
    allb = np.arange(256, dtype=np.uint8)
    raw = np.stack([allb, allb[::-1]], axis=1).reshape(1, -1)  # 256 channels, one spectrum
    p0, p1 = unpk.unpack_4bit(raw, 0, 1, np.arange(256), 256)
    e0, e1 = ounpack.unpack_4bit(raw, 0, 1, np.arange(256), 256)
    assert np.array_equal(p0, e0) and np.array_equal(p1, e1)
    nchan = 2049
    raw = rng.integers(0, 256, size=(333, 2 * nchan), dtype=np.uint8)
    ch = np.sort(rng.choice(nchan, 300, replace=False)).astype(np.int64)
    p0, p1 = unpk.unpack_4bit(raw, 17, 300, ch, nchan)
    e0, e1 = ounpack.unpack_4bit(raw, 17, 300, ch, nchan)
    assert np.array_equal(p0, e0) and np.array_equal(p1, e1)
    # empty row range
    p0, _ = unpk.unpack_4bit(raw, 5, 5, ch, nchan)
    assert p0.shape == (0, len(ch))


def test_unpack_1bit(unpk, rng):
    packed = fx.packed_1bit()
    t0, t1 = fx.expected_complex_1bit()
    p0, p1 = unpk.unpack_1bit(packed, 0, 2, np.arange(2, 6), 8)
    assert np.array_equal(p0, t0[:, 2:6]) and np.array_equal(p1, t1[:, 2:6])
    p0, p1 = unpk.unpack_1bit(packed, 1, 2, np.arange(2, 4), 8)
    assert p0[0, 1] == 1 + 1j and p1[0, 1] == -1 - 1j
    raw = rng.integers(0, 256, size=(100, 512), dtype=np.uint8)
    p0, p1 = unpk.unpack_1bit(raw, 3, 99, np.arange(64, 900), 1024)
    e0, e1 = ounpack.unpack_1bit(raw, 3, 99, 64, 900, 1024)
    assert np.array_equal(p0, e0) and np.array_equal(p1, e1)
    with pytest.raises(ValueError):
        unpk.unpack_1bit(raw, 0, 2, np.array([2, 5]), 1024)


def test_sortpols_all_tails(unpk, rng):
    packed = fx.packed_1bit()
    p0, p1 = unpk.sortpols(packed, 8, 1, 0, 2, 2, 7)
    assert p0[1, 0] == fx.str_to_bits("01110101") and p0[0, 0] == fx.str_to_bits("10101010")
    assert p0[1, 1] == fx.str_to_bits("01000000") and p1[1, 0] == fx.str_to_bits("10001010")
    assert p1[0, 0] == fx.str_to_bits("01010101") and p1[1, 1] == fx.str_to_bits("10000000")
    a1 = fx.packed_1bit_ant(fx.ant_1bit(im_sign=1))
    a2 = fx.packed_1bit_ant(fx.ant_1bit(im_sign=-1))
    p0, p1 = unpk.sortpols(a1, 8, 1, 5, 6, 2, 3)
    assert p0[0, 0] == 192 and p1[0, 0] == 192
    p0, p1 = unpk.sortpols(a2, 8, 1, 9, 10, 4, 7)
    assert p0[0, 0] == 168 and p1[0, 0] == 168
    raw = rng.integers(0, 256, size=(200, 512), dtype=np.uint8)
    # (0, 1024), (8, 1000), (16, 528): four output bytes per lane; the rest byte-wise
    for cs, ce in [(0, 1024), (8, 1000), (16, 528), (2, 1023), (4, 1022), (6, 1021), (0, 1), (10, 13), (4, 1020)]:
        g = unpk.sortpols(raw, 1024, 1, 7, 190, cs, ce)
        e = ounpack.sortpols(raw, 1024, 1, 7, 190, cs, ce)
        assert np.array_equal(g[0], e[0]) and np.array_equal(g[1], e[1]), (cs, ce)
        if oref.available():
            r = oref.sortpols(raw, 1024, 1, 7, 190, cs, ce)
            assert np.array_equal(g[0], r[0]) and np.array_equal(g[1], r[1])
    raw4 = rng.integers(0, 256, size=(200, 2 * 300), dtype=np.uint8)
    g = unpk.sortpols(raw4, 300, 4, 3, 199, 10, 250)
    e = ounpack.sortpols(raw4, 300, 4, 3, 199, 10, 250)
    assert np.array_equal(g[0], e[0]) and np.array_equal(g[1], e[1])
    with pytest.raises(ValueError):
        unpk.sortpols(raw4, 300, 4, 0, 10, 0, 301)
    with pytest.raises(ValueError):
        unpk.sortpols(raw, 1024, 1, 0, 10, 1, 9)


def test_hist(unpk, rng):
    packed = fx.packed_4bit()
    h = unpk.hist(packed, 0, 2, 10, 4, 0)
    assert h[0, 3] == 1 and h[4, 3] == 1
    h = unpk.hist(packed, 0, 2, 10, 4, 1)
    assert h[0, 3] == 1 and h[12, 3] == 1
    p1 = fx.packed_1bit()
    h = unpk.hist(p1, 0, 2, 8, 1, -1)
    assert h[0, 3] == 4 and h[1, 3] == 4 and h[0, 4] == 4
    raw = rng.integers(0, 256, size=(700, 2 * 61), dtype=np.uint8)
    for mode in (0, 1, -1):
        assert np.array_equal(unpk.hist(raw, 3, 650, 61, 4, mode), ounpack.hist_4bit(raw, 3, 650, 61, mode))
    raw1 = rng.integers(0, 256, size=(700, 32), dtype=np.uint8)
    for mode in (0, 1, -1):
        assert np.array_equal(unpk.hist(raw1, 3, 650, 64, 1, mode), ounpack.hist_1bit(raw1, 3, 650, 64, mode))


def test_make_continuous(rng):
    from albatros_axion_detection_amd import baseband_data_classes as bdc
    nspec, nchans, channels, specnum = fx.make_continuous_case()
    spec = rng.standard_normal((len(specnum), len(channels))).astype(np.complex64)
    out = bdc.make_continuous_gpu(spec, specnum, 0, channels, nspec, nchans)
    assert out[10, 100] == spec[0, 0] and out[50, 200] == spec[4, 1]
    assert out[49, 200] == 0 and out[50, 199] == 0 and out[9, 100] == 0 and out[10, 101] == 0
    ones = np.ones((len(specnum), len(channels)), np.complex64) * (1 + 1j)
    out = bdc.make_continuous_gpu(ones, specnum, 0, channels, nspec, nchans)
    assert out.sum() == 10 + 10j and out.sum(axis=0)[100] == 5 + 5j
    # modular rows and packed-at-top semantics vs the oracle
    acclen = 64
    sn = np.sort(rng.choice(acclen, 40, replace=False)) + 1000
    sp = np.zeros((acclen, 30), np.complex64)
    sp[:40] = rng.standard_normal((40, 30)) + 1j * rng.standard_normal((40, 30))
    ch = np.arange(500, 530)
    g = bdc.make_continuous_gpu(sp, sn, 1000 - 3 * acclen, ch, acclen, 2049)
    e = ostream.make_continuous(sp, sn, 1000 - 3 * acclen, ch, acclen, 2049)
    assert np.array_equal(g, e.astype(np.complex64))
    # a repeated channel number keeps its last column
    ch2 = ch.copy()
    ch2[3] = ch2[20]
    g = bdc.make_continuous_gpu(sp, sn, 1000 - 3 * acclen, ch2, acclen, 2049)
    e = ostream.make_continuous(sp, sn, 1000 - 3 * acclen, ch2, acclen, 2049)
    assert np.array_equal(g, e.astype(np.complex64))
    # spectra on one row (a repeated spec_num, spec_nums acclen apart) keep
    # the last one, as numpy's fancy assignment does
    sn2 = sn.copy()
    sn2[5] = sn2[2]
    sn2[9] = sn2[30] + acclen
    g = bdc.make_continuous_gpu(sp, sn2, 1000 - 3 * acclen, ch, acclen, 2049)
    e = ostream.make_continuous(sp, sn2, 1000 - 3 * acclen, ch, acclen, 2049)
    assert np.array_equal(g, e.astype(np.complex64))
    assert np.array_equal(g[(sn2[30] - 1000 + 3 * acclen) % acclen, ch], sp[30])
    # channel numbers outside the output's columns are skipped (no write past a row)
    ch3 = ch.copy()
    ch3[4] = 2049
    ch3[9] = 5000
    g = bdc.make_continuous_gpu(sp, sn, 1000 - 3 * acclen, ch3, acclen, 2049)
    ok = np.ones(len(ch), dtype=bool)
    ok[[4, 9]] = False
    e = ostream.make_continuous(sp[:, ok], sn, 1000 - 3 * acclen, ch[ok], acclen, 2049)
    assert np.array_equal(g, e.astype(np.complex64))


def test_dropped_packets(rng):
    import torch
    from albatros_axion_detection_amd import _lib
    spp, nchan, npk = 5, 4, 7
    data = rng.integers(1, 256, size=(npk, 2 * nchan * spp), dtype=np.uint8)
    spec = np.array([0, 5, 10, 20, 25, 30, 35], dtype=np.int64)  # jump after packet 2
    d = _lib.to_device(data)
    s = _lib.to_device(spec)
    nd = torch.empty(1, dtype=torch.int64, device=d.device)
    _lib.call("alb_dropped_packets", _lib.ptr(d), _lib.ptr(s), npk, spp, nchan, 4, _lib.ptr(nd), _lib.stream())
    out = _lib.host(d)
    exp = data.copy()
    exp[2] = 0
    assert np.array_equal(out, exp) and int(nd.item()) == 1
    # the Python mirror: host array changed in place, device tensor in place
    from albatros_axion_detection_amd import unpacking
    h = data.copy()
    assert unpacking.dropped_packets(h, spec, spp, nchan, 4) == 1 and np.array_equal(h, exp)
    dd = _lib.to_device(data)
    assert unpacking.dropped_packets(dd, s, spp, nchan, 4) == 1 and np.array_equal(_lib.host(dd), exp)
    h1 = rng.integers(1, 256, size=(npk, nchan * spp // 2 + 1), dtype=np.uint8)[:, :nchan * spp // 2].copy()
    exp1 = h1.copy()
    exp1[2] = 0
    assert unpacking.dropped_packets(h1, spec, spp, 2 * (nchan // 2), 1) == 1 and np.array_equal(h1, exp1)
    assert unpacking.dropped_packets(h, spec[:1], spp, nchan, 4) == 0
    with pytest.raises(ValueError):
        unpacking.dropped_packets(data[:3].copy(), spec, spp, nchan, 4)


@pytest.mark.gpu
def test_library_loaded_before_torch_in_a_fresh_process():
    # torch and the library each bring a HIP runtime: loading the library
    # first and touching the GPU through torch afterwards must still leave the
    # library's launches working
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys\n"
            "sys.path.insert(0, %r)\n"
            "from albatros_axion_detection_amd import _lib\n"
            "_lib.load()\n"
            "import torch\n"
            "x = torch.ones(4096, device='cuda')\n"
            "_lib.call('alb_fill_zero', _lib.ptr(x), x.numel() * 4, _lib.stream())\n"
            "torch.cuda.synchronize()\n"
            "assert float(x.abs().sum()) == 0.0\n"
            "print('launch ok')\n") % root
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "launch ok" in r.stdout, r.stdout + r.stderr


def test_file_iterator_spec_num_wraps(tmp_path):
    # test_specnum_wrap.py:30-69 on files built as generate_test_baseband.py:
    # file 1 starts 10 spectra before the 2^32 wrap (5 spectra per packet,
    # packet 2 lost), file 2 continues it, file 3 wraps again mid-file.  The
    # iterator yields acclen = 10 spectra per step across files, gaps and
    # wraps, present rows first in zero-filled (acclen, ncols) arrays.
    from albatros_axion_detection_amd import baseband_data_classes as bdc
    chans = np.ravel(np.column_stack((np.arange(452, 454), np.arange(452, 454))))
    files = [str(tmp_path / ("specnum_wrap%d.raw" % k)) for k in (1, 2, 3)]
    spec_start = 2 ** 32 - 10
    fx.write_raw(files[0], chans, 5, [spec_start + 5 * i for i in range(7)], missing=(2,))
    fx.write_raw(files[1], chans, 5, [25 + 5 * i for i in range(3)])
    fx.write_raw(files[2], chans, 5, [40, 5, 10])
    it = bdc.BasebandFileIterator(files, 0, 0, 10)
    steps, overflowed = [], []
    for _ in range(6):
        steps.append(next(it))
        overflowed.append(it.obj._overflowed)
    lens = [len(d["specnums"]) for d in steps]
    assert overflowed[0] and not overflowed[3]               # file 1 wraps, file 2 does not
    print("[gpu-case] file iterator: present spectra per step", lens, flush=True)
    # the reference's asserts
    assert lens[0] == 10 and steps[0]["specnums"][-1] == spec_start + 9
    assert lens[1] == 5 and steps[1]["specnums"][-1] == spec_start + 19
    assert steps[3]["specnums"][-1] == spec_start + 39      # spans files 1 and 2
    assert it._OVERFLOW_CTR == 2 and lens[5] == 5            # second wrap in file 3
    # every step: present rows carry the written bytes (channel 452 -> 0x01,
    # 453 -> 0x10 in both pols), the rest of the chunk is zero
    for d, n in zip(steps, lens):
        for pol in ("pol0", "pol1"):
            a = _lib_host(d[pol])
            assert a.shape == (10, 2)
            assert np.all(a[:n] == np.array([1, 16], dtype=np.uint8)) and np.all(a[n:] == 0)
        assert np.all(np.diff(np.asarray(d["specnums"])) > 0)


def _lib_host(x):
    from albatros_axion_detection_amd import _lib
    return _lib.host(x) if not isinstance(x, np.ndarray) else x


def test_float_file_iterator_and_single_baseline_average(tmp_path):
    # test_specnum_wrap_gpu.py:7-97: the float iterator over the same three
    # files (channel 452 = 0+1j, 453 = 1+0j in both pols), and
    # avg_xcorr_4bit_float_gpu over the present rows of one antenna
    import torch
    from albatros_axion_detection_amd import baseband_data_classes as bdc
    from albatros_axion_detection_amd import correlations as cr
    chans = np.ravel(np.column_stack((np.arange(452, 454), np.arange(452, 454))))
    files = [str(tmp_path / ("specnum_wrap%d.raw" % k)) for k in (1, 2, 3)]
    spec_start = 2 ** 32 - 10
    fx.write_raw(files[0], chans, 5, [spec_start + 5 * i for i in range(7)], missing=(2,))
    fx.write_raw(files[1], chans, 5, [25 + 5 * i for i in range(3)])
    fx.write_raw(files[2], chans, 5, [40, 5, 10])
    it = bdc.BasebandFileIterator(files, 0, 0, 10, type="float")
    for step in range(6):
        d = next(it)
        n = len(d["specnums"])
        p0, p1, sn = d["pol0"], d["pol1"], d["specnums"]
        assert isinstance(p0, torch.Tensor) and p0.is_cuda and isinstance(sn, np.ndarray)
        assert torch.allclose(p0[:n, 0] + p0[:n, 1], torch.tensor(1 + 1j, dtype=torch.complex64, device=p0.device))
        assert torch.allclose(p0[:n, 0] + p1[:n, 1], torch.tensor(1 + 1j, dtype=torch.complex64, device=p0.device))
        assert torch.count_nonzero(p0[n:]) == 0 and torch.count_nonzero(p1[n:]) == 0
        for a, b, want in ((p0, p1, 1), (p0, p0, 1), (p0[:, 0:1], p0[:, 1:], 1j)):
            xc, nr = cr.avg_xcorr_4bit_float_gpu(a, b, sn, sn)
            assert nr == n and torch.allclose(xc, torch.full_like(xc, want), atol=1e-7), (step, want)
        if step == 0:
            assert it.obj._overflowed and n == 10 and sn[-1] == spec_start + 9
        if step == 1:
            assert n == 5 and sn[-1] == spec_start + 19
        if step == 3:
            assert not it.obj._overflowed and sn[-1] == spec_start + 39
        if step == 5:
            assert it._OVERFLOW_CTR == 2 and n == 5
    # another antenna's spec_num array: the reference returns None (unimplemented)
    assert cr.avg_xcorr_4bit_float_gpu(p0, p1, sn, sn.copy()) is None


@pytest.mark.parametrize("bits", [4, 1])
def test_partial_read_matches_full_read(tmp_path, bits):
    # test_partial_read.py:47-56: reading the first packets of a file
    # (readlen) gives the leading rows of a full read, for 4-bit and 1-bit
    # files (random payload, a gap in the spec_nums)
    import torch
    from albatros_axion_detection_amd import baseband_data_classes as bdc
    rng = np.random.default_rng(bits)
    spp, npk = 5, 40
    if bits == 4:
        chans = np.ravel(np.column_stack((np.arange(452, 462), np.arange(452, 462))))   # 2 bytes per channel
    else:
        chans = np.arange(452, 472, 2)                                                 # 2 channels per byte
    payload = rng.integers(0, 256, size=(npk, len(chans) * spp), dtype=np.uint8)
    specnums = [1000 + spp * i + (spp * 3 if i >= 20 else 0) for i in range(npk)]
    f = str(tmp_path / "part.raw")
    fx.write_raw(f, chans, spp, specnums, bits=bits, payload=payload)
    full = bdc.BasebandPacked(f)
    part = bdc.BasebandPacked(f, readlen=25)
    en = spp * 25
    assert np.array_equal(np.asarray(part.spec_idx), np.asarray(full.spec_idx)[:en])
    for pol in ("pol0", "pol1"):
        a, b = getattr(full, pol), getattr(part, pol)
        a = a if isinstance(a, torch.Tensor) else torch.as_tensor(a)
        b = b if isinstance(b, torch.Tensor) else torch.as_tensor(b)
        assert b.shape[0] == en and torch.equal(a[:en].cpu(), b.cpu()), pol


@pytest.mark.parametrize("bits", [4, 1])
def test_file_readers_float_and_packed_agree(tmp_path, bits):
    # test_4bit_unpack_with_file.py:22-33, test_4bit_xcorr_with_file.py:31-96,
    # test_1bit_xcorr_with_file.py:36-96 on a generated file (the reference's
    # .raw fixtures are not in its tree): the float reader matches an
    # independent decode of the packet bytes, the packed reader matches the
    # oracle's sortpols, and the packed correlators equal the float reader's
    # data correlated in float64, for the whole file and for the reference's
    # row / channel windows
    import torch
    from albatros_axion_detection_amd import baseband_data_classes as bdc
    from albatros_axion_detection_amd import correlations as cr
    rng = np.random.default_rng(40 + bits)
    spp, npk = 7, 30
    if bits == 4:
        chans = np.ravel(np.column_stack((np.arange(452, 468), np.arange(452, 468))))   # 16 channels
    else:
        chans = np.arange(452, 468, 2)                                                 # 16 channels
    payload = rng.integers(0, 256, size=(npk, len(chans) * spp), dtype=np.uint8)
    f = str(tmp_path / "file.raw")
    fx.write_raw(f, chans, spp, [5000 + spp * i for i in range(npk)], bits=bits, payload=payload)
    fl = bdc.BasebandFloat(f)
    pk = bdc.BasebandPacked(f)
    nrows, nchan = npk * spp, fl.length_channels
    assert nchan == 16 and fl.pol0.shape == (nrows, nchan) and pk.pol0.shape[0] == nrows
    raw = payload.reshape(nrows, -1)
    if bits == 4:
        want = ounpack.unpack_4bit(raw, 0, nrows, np.arange(nchan), nchan)
    else:
        want = ounpack.unpack_1bit(raw, 0, nrows, 0, nchan, nchan)
    f0, f1 = _lib_host(fl.pol0).astype(np.complex128), _lib_host(fl.pol1).astype(np.complex128)
    assert np.array_equal(f0, want[0]) and np.array_equal(f1, want[1])
    sp = ounpack.sortpols(raw, nchan, bits, 0, nrows, 0, nchan)
    assert np.array_equal(_lib_host(pk.pol0), sp[0]) and np.array_equal(_lib_host(pk.pol1), sp[1])

    def close(got, ref):
        got = np.asarray(got)
        assert np.all(np.abs(got - ref) <= 2e-7 * np.maximum(1.0, np.abs(ref))), (got, ref)

    windows = [(0, nrows, 0, nchan)] + ([(10, 20, 4, 12)] if bits == 4 else [(10, 20, 6, 10), (10, 20, 6, 12)])
    for r0, r1, c0, c1 in windows:
        if (r0, c0) == (0, 0):
            p0, p1 = pk.pol0, pk.pol1
        else:
            w = bdc.BasebandPacked(f, rowstart=r0, rowend=r1, chanstart=c0, chanend=c1)
            p0, p1 = w.pol0, w.pol1
            assert p0.shape[0] == r1 - r0
        x0, x1 = f0[r0:r1, c0:c1], f1[r0:r1, c0:c1]
        n = r1 - r0
        specnums = np.arange(n)
        ref = np.sum(x0 * np.conj(x1), axis=0) / n
        if bits == 4:
            close(cr.avg_xcorr_4bit(p0, p1, specnums), ref)
            for p, x in ((p0, x0), (p1, x1)):
                close(cr.avg_autocorr_4bit(p, specnums), np.sum(np.real(x * np.conj(x)).astype(int), axis=0) / n)
        else:
            close(cr.avg_xcorr_1bit(p0, p1, specnums, c1 - c0), ref)


def test_hist_1bit_matches_packed_bits(tmp_path):
    # test_histogram_1bit.py:8-66: counting the bits of the sortpols-packed
    # 1-bit pols (4 channels a byte, (re, im) of channel 4j+k at bits
    # 7-2k, 6-2k) gives the file's per-pol histogram, and bits are conserved
    from albatros_axion_detection_amd import baseband_data_classes as bdc
    rng = np.random.default_rng(17)
    spp, npk = 9, 40
    chans = np.arange(452, 468, 2)                                                     # 16 channels
    payload = rng.integers(0, 256, size=(npk, len(chans) * spp), dtype=np.uint8)
    f = str(tmp_path / "h1.raw")
    fx.write_raw(f, chans, spp, [300 + spp * i for i in range(npk)], bits=1, payload=payload)
    pk = bdc.BasebandPacked(f)
    assert pk.pol0.shape[0] == len(pk.spec_idx)
    for mode, pol in ((0, pk.pol0), (1, pk.pol1)):
        b = _lib_host(pol).astype(np.int64)
        nrow, ncol = b.shape
        want = np.zeros((2, 4 * ncol), dtype=np.int64)
        for k in range(4):
            for s in (7 - 2 * k, 6 - 2 * k):
                ones = ((b >> s) & 1).sum(axis=0)
                want[1, k::4] += ones
                want[0, k::4] += nrow - ones
        got = np.asarray(_lib_host(pk.get_hist(mode=mode))).astype(np.int64)
        assert want.sum() == nrow * ncol * 8 and got.sum() == nrow * ncol * 8
        assert np.array_equal(got, want), mode


@pytest.mark.parametrize("bits", [4, 1])
@pytest.mark.parametrize("split", [120, 130])
@pytest.mark.parametrize("acclen,nchunks", [(1000, 3), (3000, 1)])
def test_file_iterator_across_files_matches_full_file(tmp_path, bits, split, acclen, nchunks):
    # test_missing_block_bw_files.py:42-66 on generated files: a recording
    # with a 2500-spectrum gap, read whole and as two files split before or
    # at the gap; every iterator step holds the full file's rows whose
    # spec_nums fall in the step, in order, and zeros after them (a step
    # inside the gap is empty).  The start (idxstart spectra after the first
    # file's first spectrum) lies inside the first file, as the reference's
    # iterator requires (baseband_data_classes.py:634)
    import torch
    from albatros_axion_detection_amd import baseband_data_classes as bdc
    rng = np.random.default_rng(bits * 1000 + split)
    spp, npk, gap_at, gap = 5, 400, 130, 2500
    if bits == 4:
        chans = np.ravel(np.column_stack((np.arange(452, 468), np.arange(452, 468))))
    else:
        chans = np.arange(452, 468, 2)
    payload = rng.integers(0, 256, size=(npk, len(chans) * spp), dtype=np.uint8)
    specs = [1000 + spp * p + (gap if p >= gap_at else 0) for p in range(npk)]
    full_path = str(tmp_path / "full.raw")
    fx.write_raw(full_path, chans, spp, specs, bits=bits, payload=payload)
    parts = [str(tmp_path / "part1.raw"), str(tmp_path / "part2.raw")]
    fx.write_raw(parts[0], chans, spp, specs[:split], bits=bits, payload=payload[:split])
    fx.write_raw(parts[1], chans, spp, specs[split:], bits=bits, payload=payload[split:])
    full = bdc.BasebandPacked(full_path)
    sidx = np.asarray(full.spec_idx)
    f0, f1 = _lib_host(full.pol0), _lib_host(full.pol1)
    idxstart = 550
    it = bdc.BasebandFileIterator(parts, 0, idxstart, acclen, nchunks=nchunks)
    counts = []
    for k in range(nchunks):
        d = next(it)
        lo = sidx[0] + idxstart + k * acclen
        sel = np.where((sidx >= lo) & (sidx < lo + acclen))[0]
        n = len(sel)
        counts.append(n)
        assert np.array_equal(d["specnums"], sidx[sel]), k
        p0, p1 = _lib_host(d["pol0"]), _lib_host(d["pol1"])
        assert p0.shape == (acclen, f0.shape[1])
        assert np.array_equal(p0[:n], f0[sel]) and np.array_equal(p1[:n], f1[sel]), k
        assert not p0[n:].any() and not p1[n:].any(), k
    assert counts == ([100, 0, 400] if acclen == 1000 else [500])
    with pytest.raises(StopIteration):
        next(it)


def test_unpack_row_tiled_forms(unpk, rng):
    # shapes that take the row-tiled kernels (even column count / 4-aligned
    # channel span, 16-byte aligned pols) with more rows than the 4096-row
    # grid, beside shapes that keep the per-element kernels
    nrows, nchan = 5000, 256
    raw4 = rng.integers(0, 256, size=(nrows + 3, 2 * nchan), dtype=np.uint8)
    for ch in (np.arange(0, 256), np.sort(rng.choice(nchan, 90, replace=False)), np.arange(7, 100)):
        p0, p1 = unpk.unpack_4bit(raw4, 3, nrows + 3, ch, nchan)
        e0, e1 = ounpack.unpack_4bit(raw4, 3, nrows + 3, ch, nchan)
        assert np.array_equal(p0, e0) and np.array_equal(p1, e1), len(ch)
    for cs, ce in ((0, 256), (12, 252), (10, 250), (4, 8)):
        g = unpk.sortpols(raw4, nchan, 4, 3, nrows + 3, cs, ce)
        e = ounpack.sortpols(raw4, nchan, 4, 3, nrows + 3, cs, ce)
        assert np.array_equal(g[0], e[0]) and np.array_equal(g[1], e[1]), (cs, ce)
    raw1 = rng.integers(0, 256, size=(nrows + 2, nchan // 2), dtype=np.uint8)
    for cs, ce in ((0, 256), (6, 200), (2, 4)):
        p0, p1 = unpk.unpack_1bit(raw1, 2, nrows + 2, np.arange(cs, ce), nchan)
        e0, e1 = ounpack.unpack_1bit(raw1, 2, nrows + 2, cs, ce, nchan)
        assert np.array_equal(p0, e0) and np.array_equal(p1, e1), (cs, ce)
    for cs, ce in ((0, 256), (2, 255), (6, 201), (8, 10)):
        g = unpk.sortpols(raw1, nchan, 1, 2, nrows + 2, cs, ce)
        e = ounpack.sortpols(raw1, nchan, 1, 2, nrows + 2, cs, ce)
        assert np.array_equal(g[0], e[0]) and np.array_equal(g[1], e[1]), (cs, ce)


def test_unpack_row_blocks_longer_than_a_batch(unpk, rng):
    # 9 * 4096 + 5 rows: each of the 4096 row blocks holds 10 rows, so the
    # 4-bit LDS form (4 spectra per batch) runs two full batches and a partial
    # one, and the 1-bit form walks 10 rows per lane; the last blocks are
    # short or empty.  Full, gathered (repeats, descending) and edge channel sets.
    nrows, nchan = 9 * 4096 + 5, 256
    raw4 = rng.integers(0, 256, size=(nrows + 1, 2 * nchan), dtype=np.uint8)
    gathered = rng.choice(nchan, 90, replace=True)
    for ch in (np.arange(0, 256), gathered, np.arange(255, -1, -2), np.array([0, 255])):
        print("[gpu-case] unpack_4bit rows %d cols %d" % (nrows, len(ch)), flush=True)
        p0, p1 = unpk.unpack_4bit(raw4, 1, nrows + 1, ch, nchan)
        e0, e1 = ounpack.unpack_4bit(raw4, 1, nrows + 1, ch, nchan)
        assert np.array_equal(p0, e0) and np.array_equal(p1, e1), len(ch)
    raw1 = rng.integers(0, 256, size=(nrows + 1, nchan // 2), dtype=np.uint8)
    for cs, ce in ((0, 256), (6, 200)):
        print("[gpu-case] unpack_1bit rows %d channels %d..%d" % (nrows, cs, ce), flush=True)
        p0, p1 = unpk.unpack_1bit(raw1, 1, nrows + 1, np.arange(cs, ce), nchan)
        e0, e1 = ounpack.unpack_1bit(raw1, 1, nrows + 1, cs, ce, nchan)
        assert np.array_equal(p0, e0) and np.array_equal(p1, e1), (cs, ce)


def test_dropped_packets_more_than_65536_packets(rng):
    # unpacking_gpu.cu:24 takes any num_packets: 70 000 packets, jumps after
    # packets 3, 65 540 and 69 998 (the last pair)
    import torch
    from albatros_axion_detection_amd import _lib
    spp, nchan, npk = 2, 3, 70000
    bpp = 2 * nchan * spp
    data = rng.integers(1, 256, size=(npk, bpp), dtype=np.uint8)
    spec = spp * np.arange(npk, dtype=np.int64)
    jumps = [3, 65540, 69998]
    for j in jumps:
        spec[j + 1:] += 7
    d = _lib.to_device(data)
    s = _lib.to_device(spec)
    nd = torch.empty(1, dtype=torch.int64, device=d.device)
    _lib.call("alb_dropped_packets", _lib.ptr(d), _lib.ptr(s), npk, spp, nchan, 4, _lib.ptr(nd), _lib.stream())
    exp = data.copy()
    exp[jumps] = 0
    assert int(nd.item()) == len(jumps)
    assert np.array_equal(_lib.host(d), exp)


def test_unpack_rejects_out_of_range_rows_and_channels(unpk, rng):
    # host checks before any launch: a channel index outside the spectrum, or
    # rows past the end of the raw data, raise instead of reading beyond it
    nchan = 8
    data = rng.integers(0, 256, size=(10, 2 * nchan), dtype=np.uint8)
    with pytest.raises(ValueError):
        unpk.unpack_4bit(data, 0, 10, [0, 8], nchan)
    with pytest.raises(ValueError):
        unpk.unpack_4bit(data, 0, 10, [-1, 2], nchan)
    with pytest.raises(ValueError):
        unpk.unpack_4bit(data, 0, 11, [0, 1], nchan)
    with pytest.raises(ValueError):
        unpk.unpack_4bit(data, 5, 4, [0, 1], nchan)
    with pytest.raises(ValueError):
        unpk.sortpols(data, nchan, 4, 0, 11, 0, nchan)
    with pytest.raises(ValueError):
        unpk.hist(data, 0, 11, nchan, 4, -1)
    d1 = rng.integers(0, 256, size=(10, nchan // 2), dtype=np.uint8)
    with pytest.raises(ValueError):
        unpk.unpack_1bit(d1, 0, 11, np.arange(0, nchan), nchan)
    with pytest.raises(ValueError):
        unpk.unpack_1bit(d1, 0, 10, np.arange(2, nchan + 2), nchan)
    with pytest.raises(ValueError):
        unpk.sortpols(d1, nchan, 1, 0, 11, 0, nchan)
    # the full valid ranges still work
    p0, p1 = unpk.unpack_4bit(data, 0, 10, [0, nchan - 1], nchan)
    assert p0.shape == (10, 2)
    q0, _ = unpk.unpack_1bit(d1, 0, 10, np.arange(0, nchan), nchan)
    assert q0.shape == (10, nchan)
