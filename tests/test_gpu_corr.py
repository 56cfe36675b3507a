"""GPU parity of the correlators: integer sums bit exact against the oracle
and the reference C (oracle/_ref); float correlator within a stated tolerance
against complex128."""
import numpy as np
import pytest

from oracle import corr as ocorr
from oracle import ref as oref
from oracle import unpack as ounpack
from tests import fixtures as fx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cr():
    from albatros_axion_detection_amd import correlations
    return correlations


def test_autocorr_xcorr_4bit(cr, rng):
    raw = rng.integers(0, 256, size=(3000, 2 * 130), dtype=np.uint8)
    p0, p1 = ounpack.sortpols(raw, 130, 4, 0, 3000, 0, 130)
    sn = np.arange(2900)
    got = cr.avg_autocorr_4bit(p0, sn)
    assert np.array_equal(got, ocorr.autocorr_4bit_sums(p0, 2900) / 2900)  # int64 / rowcount
    re, im = ocorr.xcorr_4bit_sums(p0, p1, 2900)
    exp = (re.astype(np.float32) + 1j * im.astype(np.float32)).astype(np.complex64) / 2900
    got = cr.avg_xcorr_4bit(p0, p1, sn)
    assert got.dtype == np.complex64 and np.array_equal(got, exp)
    if oref.available():
        assert np.array_equal(got, oref.avg_xcorr_4bit(p0, p1, 2900) / 2900)
    assert np.isnan(cr.avg_xcorr_4bit(p0, p1, []))
    assert np.isnan(cr.avg_autocorr_4bit(p0, []))


def test_xcorr_4bit_2ant(cr, rng):
    raw = rng.integers(0, 256, size=(500, 2 * 64), dtype=np.uint8)
    p0, p1 = ounpack.sortpols(raw, 64, 4, 0, 500, 0, 64)
    s0 = np.sort(rng.choice(900, 480, replace=False)) + 50
    s1 = np.sort(rng.choice(900, 470, replace=False)) + 20
    re, im, n = ocorr.xcorr_4bit_2ant_sums(p0, p1, s0, s1, 50, 20)
    got = cr.avg_xcorr_4bit_2ant(p0, p1, s0, s1, 50, 20)
    exp = (re.astype(np.float32) + 1j * im.astype(np.float32)).astype(np.complex64) / n
    assert np.array_equal(got, exp)
    if oref.available():
        o, n2 = oref.avg_xcorr_4bit_2ant(p0, p1, s0, s1, 50, 20)
        assert n2 == n and np.array_equal(got, o / n2)
    assert np.isnan(cr.avg_xcorr_4bit_2ant(p0, p1, s0, s1 + 10 ** 6, 50, 20))


def test_xcorr_1bit_and_vanvleck(cr, rng):
    raw = rng.integers(0, 256, size=(5000, 256), dtype=np.uint8)
    for nch in (512, 511, 510, 509):
        q0, _ = ounpack.sortpols(raw, 512, 1, 0, 5000, 0, nch)
        q1, _ = ounpack.sortpols(raw[::-1].copy(), 512, 1, 0, 5000, 0, nch)
        got = cr.avg_xcorr_1bit(q0, q1, np.arange(4999), nch)
        re, im = ocorr.xcorr_1bit_sums(q0, q1, nch, 4999)
        exp = (re.astype(np.float32) + 1j * im.astype(np.float32)).astype(np.complex64) / 4999
        assert np.array_equal(got, exp), nch
    q0, _ = ounpack.sortpols(raw, 512, 1, 0, 5000, 0, 512)
    q1, _ = ounpack.sortpols(raw[::-1].copy(), 512, 1, 0, 5000, 0, 512)
    got = cr.avg_xcorr_1bit_vanvleck(q0, q1, np.arange(4000), 512)
    exp = ocorr.vanvleck_sums(q0, q1, 512, 4000)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e.astype(np.float32) / 4000)
    s0 = np.sort(rng.choice(7000, 5000, replace=False))
    s1 = np.sort(rng.choice(7000, 5000, replace=False))
    got, n = cr.avg_xcorr_1bit_vanvleck_2ant(q0, q1, 512, s0, s1, 0, 0)
    r0, r1 = ocorr.intersect_rows(s0, s1, 0, 0)
    exp = ocorr.vanvleck_sums(q0, q1, 512, None, r0, r1)
    assert n == len(r0)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e.astype(np.float32))
    if oref.available():
        ref, n2 = oref.avg_xcorr_1bit_vanvleck_2ant(q0, q1, s0, s1, 0, 0, 512)
        assert n2 == n and all(np.array_equal(g, r) for g, r in zip(got, ref))


@pytest.mark.parametrize("method", ["valu", "mfma"])
def test_allpairs_4bit(cr, rng, method):
    from albatros_axion_detection_amd import synth
    inp = synth.direct4(T=8192, nchan=128, seed=4)
    got = cr.xcorr_4bit_allpairs(inp, 4096, method=method)
    exp = ocorr.allpairs_4bit_block_sums(inp, 4096)
    assert got.shape == (2, 10, 128, 2) and np.array_equal(got, exp)
    # edge: nibble 8 (+8) everywhere, the largest products (16- and 64-byte column groups)
    for nc in (16, 64):
        eights = np.full((3, 4096, nc), 0x88, dtype=np.uint8)
        got = cr.xcorr_4bit_allpairs(eights, 1024, method=method)
        assert np.all(got[..., 0] == 128 * 1024) and np.all(got[..., 1] == 0)
    # every byte value, 8 inputs (a full 16-row product), 32 channels
    raw = rng.integers(0, 256, size=(8, 640, 32), dtype=np.uint8)
    got = cr.xcorr_4bit_allpairs(raw, 320, method=method)
    assert np.array_equal(got, ocorr.allpairs_4bit_block_sums(raw, 320))


def test_allpairs_4bit_full_size(cr):
    # the direct 4-bit workload at full size: 2 antennas x 2 pols x 2^17
    # spectra x 128 channels, 4096-spectrum blocks, 10 pairs; both forms
    # bit-exact against the oracle
    from albatros_axion_detection_amd import synth
    inp = synth.direct4(T=2 ** 17, nchan=128, seed=0)
    exp = ocorr.allpairs_4bit_block_sums(inp, 4096)
    for method in ("mfma", "valu"):
        got = cr.xcorr_4bit_allpairs(inp, 4096, method=method)
        print("[gpu-case] direct 4-bit full size", method, flush=True)
        assert got.shape == (32, 10, 128, 2) and np.array_equal(got, exp), method


def test_allpairs_4bit_ragged_columns(cr, rng):
    # 12 channels: the VALU form serves "auto", the matrix-core form refuses
    raw = rng.integers(0, 256, size=(5, 640, 12), dtype=np.uint8)
    got = cr.xcorr_4bit_allpairs(raw, 320)
    assert np.array_equal(got, ocorr.allpairs_4bit_block_sums(raw, 320))
    with pytest.raises(ValueError):
        cr.xcorr_4bit_allpairs(raw, 320, method="mfma")


@pytest.mark.parametrize("method", ["valu", "mfma"])
def test_allpairs_1bit(cr, method):
    from albatros_axion_detection_amd import synth
    inp = synth.direct1(ninput=5, T=4096, nchan=64, seed=6)
    got = cr.xcorr_1bit_allpairs(inp, 64, 2048, method=method)
    exp = ocorr.allpairs_1bit_block_sums(inp, 64, 2048)
    assert got.shape == exp.shape and np.array_equal(got, exp)
    # autocorrelation of +-1 data is exactly 2 per spectrum
    assert np.all(got[:, 0, :, 0] == 2 * 2048) and np.all(got[:, 0, :, 1] == 0)
    # 16 inputs (every row of the product), channel count short of the last byte
    inp = synth.direct1(ninput=16, T=1024, nchan=64, seed=7)
    got = cr.xcorr_1bit_allpairs(inp, 62, 512, method=method)
    assert np.array_equal(got, ocorr.allpairs_1bit_block_sums(inp, 62, 512))


def test_allpairs_1bit_full_size(cr):
    # the direct 1-bit workload at full size: 16 inputs x 2^20 spectra x 1024
    # channels (4 GiB), 8192-spectrum blocks, 136 pairs: the matrix-core and
    # XOR forms agree bit for bit everywhere, autocorrelations are exactly
    # 2 per spectrum, and the first block's first 64 channels of every pair
    # match the oracle
    import torch
    from albatros_axion_detection_amd import _lib, synth
    inp = synth.direct1(ninput=16, T=2 ** 20, nchan=1024, seed=3, device=_lib.device())
    a = cr.xcorr_1bit_allpairs(inp, 1024, 8192, method="mfma")
    b = cr.xcorr_1bit_allpairs(inp, 1024, 8192, method="valu")
    print("[gpu-case] direct 1-bit full size", flush=True)
    assert a.shape == (128, 136, 1024, 2) and torch.equal(a, b)
    autos = [p for p, (i, j) in enumerate(ocorr.pair_list(16)) if i == j]
    assert torch.all(a[:, autos, :, 0] == 2 * 8192) and torch.all(a[:, autos, :, 1] == 0)
    part = inp[:, :8192, :16].cpu().numpy()
    exp = ocorr.allpairs_1bit_block_sums(part, 64, 8192)
    assert np.array_equal(a[:1, :, :64].cpu().numpy(), exp)


def _with_tune(key, value, fn):
    from albatros_axion_detection_amd import _lib
    lib = _lib.load()
    old = lib.alb_tune_get(key)
    lib.alb_tune_set(key, value)
    try:
        return fn()
    finally:
        lib.alb_tune_set(key, old)


def test_allpairs_1bit_matrix_core_forms(cr, rng):
    # Two matrix-core forms: one 16 x 16 tile per channel (x = +-1 as e2m1
    # nibbles on the FP4 matrix cores, T = Re + Im; 64-byte column groups) and
    # three int8 products per channel on the bits (any other shape; the bits
    # weighed by their place value up to 2^16 spectra, shifted to place 0
    # above).  At 2^16 and 2^17 spectra per block both agree with the XOR form
    # everywhere and with the oracle on the first 16 channels; constant bytes
    # give closed forms (x = +-1 +- i: 0xFF -> 1 + i, 0x00 -> -1 - i, 0xAA -> 1 - i).
    T = 2 ** 17
    inp = rng.integers(0, 256, size=(3, T, 64), dtype=np.uint8)
    for blocklen in (2 ** 16, 2 ** 17):
        ref = cr.xcorr_1bit_allpairs(inp, 256, blocklen, method="valu")
        exp = ocorr.allpairs_1bit_block_sums(inp[:, :, :4], 16, blocklen)
        assert np.array_equal(ref[:, :, :16], exp)
        for form in (0, 1, 2):
            got = _with_tune(b"xcorr1_form", form, lambda: cr.xcorr_1bit_allpairs(inp, 256, blocklen, method="mfma"))
            print("[gpu-case] 1-bit allpairs blocklen", blocklen, "form", form, flush=True)
            assert np.array_equal(got, ref), (blocklen, form)
    with pytest.raises(ValueError):   # the one-tile form takes 64-byte column groups only
        _with_tune(b"xcorr1_form", 1, lambda: cr.xcorr_1bit_allpairs(inp[:, :, :48], 192, 2 ** 16, method="mfma"))
    const = np.empty((3, T, 64), dtype=np.uint8)
    const[0], const[1], const[2] = 0xFF, 0x00, 0xAA
    x = np.array([1 + 1j, -1 - 1j, 1 - 1j])
    for blocklen, form in ((2 ** 16, 1), (2 ** 16, 2), (2 ** 17, 1), (2 ** 17, 2)):
        got = _with_tune(b"xcorr1_form", form, lambda: cr.xcorr_1bit_allpairs(const, 256, blocklen, method="mfma"))
        for p, (i, j) in enumerate(ocorr.pair_list(3)):
            v = x[i] * np.conj(x[j]) * blocklen
            assert np.all(got[:, p, :, 0] == v.real) and np.all(got[:, p, :, 1] == v.imag), (blocklen, form, i, j)


def test_allpairs_1bit_one_tile_inputs_and_ragged_channels(cr):
    # the one-tile form for 1, 5 and 16 inputs, 250 of 256 channels, two
    # 64-byte column groups, against the oracle
    from albatros_axion_detection_amd import synth
    for n in (1, 5, 16):
        inp = synth.direct1(ninput=n, T=4096, nchan=512, seed=10 + n)
        exp = ocorr.allpairs_1bit_block_sums(inp, 500, 2048)
        got = _with_tune(b"xcorr1_form", 1, lambda: cr.xcorr_1bit_allpairs(inp, 500, 2048, method="mfma"))
        assert got.shape == exp.shape and np.array_equal(got, exp), n


def test_allpairs_4bit_every_input_count(cr, rng):
    # 1..8 inputs (4 loaded slots and two channels per product up to 4
    # inputs, 8 slots above), 128 channels (64-byte column groups) and 32
    # (16-byte groups), with the workgroup target swept so the block splits
    # change (1 .. 16 workgroups per block, down to one 64-spectrum step
    # each); all exact against the oracle
    from albatros_axion_detection_amd import _lib
    lib = _lib.load()
    for ncol in (128, 32):
        for n in range(1, 9):
            raw = rng.integers(0, 256, size=(n, 2048, ncol), dtype=np.uint8)
            exp = ocorr.allpairs_4bit_block_sums(raw, 1024)
            for target in (0, 1, 8, 4096):
                lib.alb_tune_set(b"xcorr4_workgroups", target)
                try:
                    got = cr.xcorr_4bit_allpairs(raw, 1024, method="mfma")
                finally:
                    lib.alb_tune_set(b"xcorr4_workgroups", 0)
                assert np.array_equal(got, exp), (ncol, n, target)


def test_allpairs_1bit_ragged_columns(cr):
    # ncol not a multiple of 4: the VALU form serves "auto", the matrix-core form refuses
    from albatros_axion_detection_amd import synth
    inp = synth.direct1(ninput=5, T=4096, nchan=60, seed=6)
    got = cr.xcorr_1bit_allpairs(inp, 60, 2048)
    assert np.array_equal(got, ocorr.allpairs_1bit_block_sums(inp, 60, 2048))
    with pytest.raises(ValueError):
        cr.xcorr_1bit_allpairs(inp, 60, 2048, method="mfma")


def test_xcorr_all_ant_reference_constants(cr):
    x, nant, npol, ntime, nfreq, exp = fx.xcorr_constants()
    out = cr.avg_xcorr_all_ant_gpu(x, nant, npol, ntime, nfreq)
    for (i, j, f), v in exp.items():
        assert out[i, j, f] == v
    assert np.all(out[:, :, 0].real.astype(int) % 2 == 1) and np.all(out[:, :, 0].imag == 0)
    assert np.all(out[:, :, 1].real.astype(int) % 2 == 0) and np.all(out[:, :, 1].imag == 0)
    with pytest.raises(NotImplementedError):
        cr.avg_xcorr_all_ant_gpu(x, nant, npol, ntime, nfreq, split=16)
    with pytest.raises(ValueError):
        cr.avg_xcorr_all_ant_gpu(x, nant, npol, ntime, nfreq, split=3)


@pytest.mark.parametrize("M,K,F", [(8, 1024, 64), (4, 1000, 33), (1, 10000, 2), (10, 256, 17), (2, 4099, 5)])
def test_xcorr_all_ant_random(cr, M, K, F):
    # tolerance: |err| <= 1e-5 * sqrt(C_ii C_jj) (FP32 products, FP32 accumulation)
    from albatros_axion_detection_amd import synth
    x = synth.precision_xcorr(M=M, T=K, F=F, seed=M + K)
    got = cr.avg_xcorr_all_ant_gpu(x, M, 1, K, F)
    ref = ocorr.avg_xcorr_all_ant(x, K)
    d = np.sqrt(np.abs(np.einsum("iif->if", ref)))
    norm = d[:, None, :] * d[None, :, :]
    assert np.max(np.abs(got - ref) / norm) < 1e-5
    assert np.max(np.abs(got - np.conj(got.transpose(1, 0, 2))) / norm) < 1e-6


def test_xcorr_all_ant_torch_fortran_view(cr):
    import torch
    from albatros_axion_detection_amd import synth
    x = synth.precision_xcorr(M=8, T=128, F=40, seed=1, device="cuda")
    assert x.permute(2, 1, 0).is_contiguous()
    out = cr.avg_xcorr_all_ant_gpu(x, 4, 2, 128, 40)
    ref = ocorr.avg_xcorr_all_ant(x.cpu().numpy(), 128)
    assert isinstance(out, torch.Tensor) and tuple(out.shape) == (8, 8, 40)
    d = np.sqrt(np.abs(np.einsum("iif->if", ref)))
    assert np.max(np.abs(out.cpu().numpy() - ref) / (d[:, None] * d[None])) < 1e-5


@pytest.mark.parametrize("M,K,F", [(8, 1024, 64), (5, 1000, 33), (2, 77, 9), (4, 300, 7), (3, 64, 5), (1, 40, 3),
                                   (8, 1024, 6400)])
def test_xcorr_all_ant_precision_forms(cr, M, K, F):
    # diagonal-normalised error |dC_ij| / sqrt(C_ii C_jj) against the float64
    # oracle for each arithmetic; the FP16 form must show its 11-bit rounding
    # (so it is not silently an FP32 product) and the split forms must recover
    # close to FP32 accuracy; (8, 1024, 6400) is the precision-correlator
    # workload at full size
    from albatros_axion_detection_amd import synth
    x = synth.precision_xcorr(M=M, T=K, F=F, seed=3 * M + K)
    ref = ocorr.avg_xcorr_all_ant(x, K)
    d = np.sqrt(np.abs(np.einsum("iif->if", ref)))
    norm = d[:, None, :] * d[None, :, :]
    err = {}
    for prec in ("fp32", "bf16x3", "fp16", "fp16x3"):
        got = cr.avg_xcorr_all_ant_gpu(x, M, 1, K, F, precision=prec)
        err[prec] = float(np.max(np.abs(got - ref) / norm))
        assert np.max(np.abs(got - np.conj(got.transpose(1, 0, 2))) / norm) < 1e-6, prec
    print("[gpu-case] precision errors", M, K, F, err, flush=True)
    assert err["fp32"] < 1e-5 and err["fp16x3"] < 1e-5
    assert err["bf16x3"] < 1e-5
    assert 1e-5 < err["fp16"] < 5e-3
    assert err["bf16x3"] < err["fp16"] / 10 and err["fp16x3"] < err["fp16"] / 10


def test_xcorr_precision_rejects_large_m(cr):
    from albatros_axion_detection_amd import synth
    x = synth.precision_xcorr(M=10, T=16, F=2, seed=0)
    with pytest.raises(NotImplementedError):
        cr.avg_xcorr_all_ant_gpu(x, 10, 1, 16, 2, precision="fp16")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1, 3])
def test_strided_correlator_matches_f_order(rng, mode):
    # [S][K][F][M] input through the strided entry == the same data in the
    # reference's F-order (M, K, F) per job through alb_xcorr_all_ant(_prec)
    import torch
    from albatros_axion_detection_amd import _lib
    M, K, F, S = 4, 100, 37, 3
    x = (rng.standard_normal((S, K, F, M)) + 1j * rng.standard_normal((S, K, F, M))).astype(np.complex64)
    xs = _lib.to_device(x)
    xf = _lib.to_device(np.ascontiguousarray(x.transpose(0, 2, 1, 3)))       # [S][F][K][M]
    a = torch.empty((S, F, M, M), dtype=torch.complex64, device=xs.device)
    b = torch.empty_like(a)
    scale = float(np.float32(1.0) / np.float32(K))
    print("[gpu-case] strided correlator mode", mode, flush=True)
    _lib.call("alb_xcorr_all_ant_strided", _lib.ptr(xs), M, K, F, S, F * M, M, K * F * M, _lib.ptr(a), scale, mode,
              _lib.stream())
    _lib.call("alb_xcorr_all_ant_prec", _lib.ptr(xf), M, K, F * S, _lib.ptr(b), scale, mode, _lib.stream())
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    ref = np.einsum("skfi,skfj->sfji", x.astype(np.complex128), np.conj(x.astype(np.complex128))) / K
    got = a.cpu().numpy()
    tol = {0: 2e-6, 1: 1e-5, 3: 2e-6}[mode]      # FP32, BF16 hi/lo, FP16 hi/lo
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < tol


@pytest.mark.parametrize("M", [1, 2, 3, 4, 8])
@pytest.mark.parametrize("mode", [0, 3])
def test_correlator_channel_independent_of_wave_partners(M, mode):
    # for M <= 4 two or four channels share one matrix product; every channel
    # gives the same bits alone (no partner) as next to its partners
    import torch
    from albatros_axion_detection_amd import _lib
    rng = np.random.default_rng(M)
    K, F = 96, 7
    x = (rng.standard_normal((F, K, M)) + 1j * rng.standard_normal((F, K, M))).astype(np.complex64)
    xd = _lib.to_device(x)
    allc = torch.empty((F, M, M), dtype=torch.complex64, device=xd.device)
    s = _lib.stream()
    _lib.call("alb_xcorr_all_ant_strided", _lib.ptr(xd), M, K, F, 1, M, M * K, 0, _lib.ptr(allc), 1.0 / K, mode, s)
    for f in range(F):
        one = torch.empty((1, M, M), dtype=torch.complex64, device=xd.device)
        _lib.call("alb_xcorr_all_ant_strided", _lib.ptr(xd[f:]), M, K, 1, 1, M, M * K, 0, _lib.ptr(one), 1.0 / K,
                  mode, s)
        assert torch.equal(one[0], allc[f]), (M, mode, f)
    ref = np.einsum("fki,fkj->fji", x.astype(np.complex128), np.conj(x.astype(np.complex128))) / K
    assert np.max(np.abs(allc.cpu().numpy() - ref)) < 1e-5 * np.max(np.abs(ref))


def test_avg_xcorr_all_ant_cpu_reference_arithmetic(cr, rng):
    # correlations.py:115-151: complex64 products summed in float64, divided by
    # K in float64, C-order (M, M, F).  Against the step-by-step oracle the only
    # difference is the float64 summation order, so every element is within
    # one float32 ulp of it (most are bit-identical); against complex128
    # products the rounding of each product shows (tolerance 1e-6 of the
    # diagonal-normalised value).
    import torch
    M, K, F = 6, 700, 37
    x = (rng.standard_normal((M, K, F)) + 1j * rng.standard_normal((M, K, F))).astype(np.complex64)
    got = cr.avg_xcorr_all_ant_cpu(x, 3, 2, K, F)
    assert isinstance(got, np.ndarray) and got.dtype == np.complex64 and got.flags.c_contiguous
    emu = ocorr.avg_xcorr_all_ant_c64_products(x, K)
    for part in ("real", "imag"):
        g, e = getattr(got, part), getattr(emu, part)
        assert np.all(np.abs(g - e) <= np.spacing(np.abs(e).astype(np.float32))), part
    assert np.mean(got == emu) > 0.9
    ref = ocorr.avg_xcorr_all_ant(x, K)
    d = np.sqrt(np.abs(np.einsum("iif->if", ref)))
    assert np.max(np.abs(got - ref) / (d[:, None] * d[None])) < 1e-6
    assert np.array_equal(got, np.conj(got.transpose(1, 0, 2)))
    # the reference's constants, exactly
    xc, nant, npol, ntime, nfreq, exp = fx.xcorr_constants()
    oc = cr.avg_xcorr_all_ant_cpu(xc, nant, npol, ntime, nfreq)
    for (i, j, f), v in exp.items():
        assert oc[i, j, f] == v
    # any memory order of the input, device tensors in, out= filled
    xt = torch.from_numpy(np.asfortranarray(x)).cuda()
    xt = xt.permute(2, 1, 0).contiguous().permute(2, 1, 0)      # F-order memory
    o = torch.empty((M, M, F), dtype=torch.complex64, device="cuda")
    r = cr.avg_xcorr_all_ant_cpu(xt, 3, 2, K, F, out=o)
    assert r is o and np.array_equal(o.cpu().numpy(), got)
    with pytest.raises(NotImplementedError):
        cr.avg_xcorr_all_ant_cpu(x, 3, 2, K, F, split=7 * 5 * 2 * 10)
    with pytest.raises(ValueError):
        cr.avg_xcorr_all_ant_cpu(x, 3, 2, K, F, split=3)


def test_van_vleck_on_gpu_sums_recovers_correlation(cr):
    # 1-bit samples of correlated complex Gaussians through the GPU van Vleck
    # sums (avg_xcorr_1bit_vanvleck_2ant: un-normalised R0, R1, IM0, IM1 and
    # the matched row count), then van_vleck_correction (correlations.py:491-497):
    # per channel the arcsine law returns the generator's correlation
    # 2 a e^{i phi} to the sampling error (~3 / sqrt(n))
    from albatros_axion_detection_amd import synth
    rng = np.random.default_rng(21)
    n, nchan, a = 1 << 16, 16, 0.35
    phi = np.linspace(-2.5, 2.5, nchan)
    cn = lambda: rng.standard_normal((n, nchan)) + 1j * rng.standard_normal((n, nchan))  # noqa: E731
    z = cn()
    x0 = np.sqrt(a) * z + np.sqrt(1 - a) * cn()
    x1 = np.sqrt(a) * np.exp(-1j * phi) * z + np.sqrt(1 - a) * cn()
    p0 = synth.pack_1bit((x0.real > 0).astype(np.uint8), (x0.imag > 0).astype(np.uint8))
    p1 = synth.pack_1bit((x1.real > 0).astype(np.uint8), (x1.imag > 0).astype(np.uint8))
    sn = np.arange(n, dtype=np.int64)
    (R0, R1, I0, I1), rows = cr.avg_xcorr_1bit_vanvleck_2ant(p0, p1, nchan, sn, sn, 0, 0)
    assert rows == n
    v = cr.van_vleck_correction(R0.astype(np.float64), R1.astype(np.float64), I0.astype(np.float64),
                                I1.astype(np.float64), rows)
    assert np.max(np.abs(v - 2 * a * np.exp(1j * phi))) < 3 * 2 / np.sqrt(n) * 2
    # the GPU sums equal the oracle's exact integer sums
    ref = ocorr.vanvleck_sums(p0, p1, nchan, n)
    for g, r in zip((R0, R1, I0, I1), ref):
        assert np.array_equal(np.asarray(g, dtype=np.int64), np.asarray(r, dtype=np.int64))


@pytest.mark.parametrize("M,N,K,B", [(5, 3, 77, 9), (1, 16, 33, 4), (20, 7, 1, 2), (17, 17, 100, 3)])
def test_cgemm_strided_batched(cr, rng, M, N, K, B):
    # cgemm_batch.cu:8-54: C = A B^H per column-major batch; FP32 products
    # and sums (tolerance 1e-5 of the largest |C|, float64 reference)
    import torch
    A = (rng.standard_normal((M, K, B)) + 1j * rng.standard_normal((M, K, B))).astype(np.complex64)
    Bm = (rng.standard_normal((N, K, B)) + 1j * rng.standard_normal((N, K, B))).astype(np.complex64)
    C = cr.cgemm_strided_batched(A, Bm, M, N, K, B)
    ref = np.einsum("mkb,nkb->mnb", A.astype(np.complex128), np.conj(Bm.astype(np.complex128)))
    assert C.shape == (M, N, B) and C.flags.f_contiguous
    assert np.max(np.abs(C - ref)) < 1e-5 * np.max(np.abs(ref))
    # device tensors with column-major memory, out= filled
    Ad = torch.from_numpy(np.ascontiguousarray(A.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    Bd = torch.from_numpy(np.ascontiguousarray(Bm.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    o = torch.empty((B, N, M), dtype=torch.complex64, device="cuda").permute(2, 1, 0)
    r = cr.cgemm_strided_batched(Ad, Bd, M, N, K, B, out=o)
    assert r is o and np.array_equal(o.cpu().numpy(), C)
