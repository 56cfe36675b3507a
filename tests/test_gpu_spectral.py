"""GPU parity of the spectral-order inverse PFB: the transform along time
taken first on the channel columns (alb_ipfb_chan_fft_packed4), the spectra
along the samples (alb_ipfb_rows_spectral) and the column pass that is left
(alb_ipfb_columns_repfb_spectral), against float64 numpy restatements of
cupy_ipfb's two forward transforms (pfb_gpu_utils.py:80-86: irfft along the
channels, then rfft along time) and against the reference-order kernels.

Tolerance: ``max|gpu - ref| / rms(ref)`` (tests.conftest.rel_err); float32
transforms of 4096 points stay below 1e-6 of rms, the path's bar is 1e-5.
"""
import numpy as np
import pytest

from oracle import pfb as opfb
from oracle import stream as ostream
from oracle import unpack as ounpack
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _tile_spectra(packed, rowmap, chans, acclen, cut, ntiles):
    """Zero-filled complex128 spectra [ntiles][nslice][2049] of the tiles (rows
    as alb_ipfb_rows_packed4 takes them; the last copy of a repeated channel
    wins, as numpy's fancy assignment keeps it)."""
    nslice = acclen + 2 * cut
    r, i = ounpack.decode_4bit_bytes(packed)
    vals = (r + 1j * i).astype(np.complex128)
    keep = [c for c in range(len(chans)) if 0 <= chans[c] <= 2048]
    spec = np.zeros((ntiles, nslice, 2049), dtype=np.complex128)
    for t in range(ntiles):
        for row in range(nslice):
            chunk, crow = (t - 1, acclen - 2 * cut + row) if row < 2 * cut else (t, row - 2 * cut)
            if chunk < 0 or rowmap[chunk, crow] < 0:
                continue
            spec[t, row, chans[keep]] = vals[rowmap[chunk, crow], keep]
    return spec


def test_spectral_steps_against_fp64(rng):
    # channels 0 and 2048 (imaginary parts dropped, no mirror bin), a channel
    # listed twice (the earlier copy marked -1), one above 2048 (skipped),
    # lost spectra and a first tile whose carried rows are zeros
    import torch
    from albatros_axion_detection_amd import _lib
    acclen, cut, ntiles = 4088, 4, 2
    nslice = acclen + 2 * cut
    chans = np.concatenate([[0, 2048, 2047, 1, 3000], rng.choice(np.arange(2, 2047), 40, replace=False)]).astype(np.int32)
    chans = np.concatenate([chans, chans[7:8]])        # repeated: the last column wins
    marked = chans.copy()
    marked[7] = -1
    ncols = len(chans)
    packed = rng.integers(0, 256, size=(ntiles * acclen, ncols), dtype=np.uint8)
    rowmap = np.arange(ntiles * acclen, dtype=np.int32).reshape(ntiles, acclen)
    rowmap[rng.random(rowmap.shape) < 0.05] = -1
    dev = _lib.device()
    pk, rm, ch = _lib.to_device(packed), _lib.to_device(rowmap), _lib.to_device(marked)
    E = torch.empty((ntiles, nslice, ncols), dtype=torch.complex64, device=dev)
    D = torch.empty((ntiles, 1024, nslice // 2 + 1, 4), dtype=torch.complex64, device=dev)
    s = _lib.stream()
    _lib.call("alb_ipfb_chan_fft_packed4", _lib.ptr(pk), _lib.ptr(rm), _lib.ptr(ch), ncols, acclen, cut, 0, ntiles,
              _lib.ptr(E), s)
    _lib.call("alb_ipfb_rows_spectral", _lib.ptr(E), _lib.ptr(ch), ncols, nslice, ntiles, _lib.ptr(D), s)
    spec = _tile_spectra(packed, rowmap, marked, acclen, cut, ntiles)
    # step 1: transform along time of each kept column (real part for 0 / 2048)
    cols = spec[:, :, np.clip(marked, 0, 2048)]
    cols[:, :, (marked == 0) | (marked == 2048)] = cols[:, :, (marked == 0) | (marked == 2048)].real
    cols[:, :, (marked < 0) | (marked > 2048)] = 0
    ref_e = np.fft.fft(cols, axis=1)
    got_e = E.cpu().numpy()
    # step 2: 4096 * rfft(irfft(spec, axis=2), axis=1), in the strip layout
    got_d = D.cpu().numpy().transpose(0, 2, 1, 3).reshape(ntiles, nslice // 2 + 1, 4096)
    ref_d = 4096.0 * np.fft.rfft(np.fft.irfft(spec, n=4096, axis=2), axis=1)
    # The time-DC row (f = 0) holds the column sums of 4-bit values whose mean
    # is 1/2, ~7x the other rows, so errors are measured per row f: max over f
    # of max |gpu - ref| / rms(ref) of that row (E: the rows of one column).
    e1 = max(rel_err(got_e[t, :, c], ref_e[t, :, c]) for t in range(ntiles) for c in range(ncols)
             if 0 <= marked[c] <= 2048)
    e2 = max(rel_err(got_d[t, f], ref_d[t, f]) for t in range(ntiles) for f in range(nslice // 2 + 1))
    e_all = rel_err(got_d, ref_d)
    print("[gpu-case] spectral steps vs float64: E %.3g, D %.3g (per row), D %.3g of the whole tile's rms"
          % (e1, e2, e_all), flush=True)
    assert np.all(got_e[:, :, (marked < 0) | (marked > 2048)] == 0)
    assert e1 < 1e-6 and e2 < 2e-6


@pytest.mark.parametrize("ncols", [700, 2049])
def test_spectral_steps_many_channels(rng, ncols):
    # more channel items per thread than the small cases (the 4- and 9-item
    # forms of the spectral row kernel): 700 distinct channels, and every
    # channel 0..2048 in a shuffled order, one tile, against float64
    import torch
    from albatros_axion_detection_amd import _lib
    acclen, cut, ntiles = 4088, 4, 1
    nslice = acclen + 2 * cut
    chans = rng.permutation(2049)[:ncols].astype(np.int32)
    packed = rng.integers(0, 256, size=(ntiles * acclen, ncols), dtype=np.uint8)
    rowmap = np.arange(ntiles * acclen, dtype=np.int32).reshape(ntiles, acclen)
    rowmap[rng.random(rowmap.shape) < 0.05] = -1
    dev = _lib.device()
    pk, rm, ch = _lib.to_device(packed), _lib.to_device(rowmap), _lib.to_device(chans)
    E = torch.empty((ntiles, nslice, ncols), dtype=torch.complex64, device=dev)
    D = torch.empty((ntiles, 1024, nslice // 2 + 1, 4), dtype=torch.complex64, device=dev)
    s = _lib.stream()
    _lib.call("alb_ipfb_chan_fft_packed4", _lib.ptr(pk), _lib.ptr(rm), _lib.ptr(ch), ncols, acclen, cut, 0, ntiles,
              _lib.ptr(E), s)
    _lib.call("alb_ipfb_rows_spectral", _lib.ptr(E), _lib.ptr(ch), ncols, nslice, ntiles, _lib.ptr(D), s)
    got_d = D.cpu().numpy().transpose(0, 2, 1, 3).reshape(ntiles, nslice // 2 + 1, 4096)
    del D, E
    spec = _tile_spectra(packed, rowmap, chans, acclen, cut, ntiles)
    ref_d = 4096.0 * np.fft.rfft(np.fft.irfft(spec, n=4096, axis=2), axis=1)
    del spec
    e2 = max(rel_err(got_d[0, f], ref_d[0, f]) for f in range(nslice // 2 + 1))
    print("[gpu-case] spectral steps, %d channels: D %.3g (per row) vs float64" % (ncols, e2), flush=True)
    # dense spectra (every bin filled) go through two float32 4096-point
    # transforms with no zero bins: 5e-6 per row, the bound of the dense
    # row-pass test (test_gpu_pfb.py::test_row_pass_every_channel)
    assert e2 < (5e-6 if ncols > 1024 else 2e-6)


@pytest.mark.parametrize("osamp,nblock", [(64, 20), (32, 30)])
def test_spectral_form_matches_direct_form_and_oracle(rng, osamp, nblock):
    # the fused streaming spectra of both iPFB orders: same stream rows, same
    # stage-1 blocks to float32 rounding, and the spectral form against the
    # float64 oracle at the path's bar
    from albatros_axion_detection_amd import _lib, repfb
    acclen, cut, nchunks = 4088, 4, 3
    chan = np.arange(900, 1200)
    chunks = []
    for c in range(nchunks):
        keep = np.sort(rng.choice(acclen, int(acclen * 0.97), replace=False))
        raw = rng.integers(0, 256, size=(acclen, len(chan)), dtype=np.uint8)
        raw[len(keep):] = 0
        chunks.append({"pol0": raw, "specnums": 1000 + c * acclen + keep})
    rm = repfb.rowmaps_from_specnums([c["specnums"] for c in chunks], 1000, acclen)
    packed = _lib.to_device(repfb.stack_packed(chunks, "pol0", acclen, len(chan)))
    outs = {}
    for form in ("spectral", "direct"):
        pipe = repfb.RePFBPipeline(1, chan, acclen, cut, osamp, nblock, 900 * osamp, 1200 * osamp, ipfb_form=form,
                                   tile_batch=2)
        assert pipe.ipfb_form == form
        outs[form] = pipe.spectra(packed, rm, nchunks).cpu().numpy()
    d = rel_err(outs["spectral"], outs["direct"])
    jobs, nbt, _ = pipe.plan(nchunks)
    dec = []
    for chk in chunks:
        r, i = ounpack.decode_4bit_bytes(chk["pol0"])
        dec.append({"pol0": r + 1j * i, "specnums": chk["specnums"]})
    filt = opfb.wiener_filter(opfb.matft_fp64(acclen + 2 * cut), 0.45)
    st = ostream.ipfb_stream(dec, 1000, chan, acclen, cut, filt, 0)
    need = (nbt + 3) * 4096 * osamp
    buf = np.zeros(max(need, st.size))
    buf[: st.size] = st
    ref = opfb.pfb(buf[:need], opfb.window_f32(4, 4096 * osamp), 2048 * osamp + 1, 4)[:, 900 * osamp:1200 * osamp]
    e = rel_err(outs["spectral"], ref)
    print("[gpu-case] spectral vs direct %.3g, spectral vs float64 %.3g" % (d, e), flush=True)
    assert outs["spectral"].shape == ref.shape and d < 2e-6 and e < TOL


def test_spectral_steps_reject_bad_shapes():
    # host checks, before any launch
    from albatros_axion_detection_amd import _lib
    lib = _lib.load()
    assert lib.alb_ipfb_chan_fft_packed4(None, None, None, 10, 4088, 4, 0, 1, None, None) != 0
    assert lib.alb_ipfb_rows_spectral(None, None, 10, 4096, 1, None, None) != 0
    assert lib.alb_ipfb_columns_repfb_spectral(None, 1, 2048, None, 1025, 4, None, 0, 2040, None, 4, 64, 10, None,
                                               None) != 0
