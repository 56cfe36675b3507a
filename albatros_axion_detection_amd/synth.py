"""Seeded synthetic inputs for the benchmark workloads (no files, no network).

* direct 4-bit correlator: 2 antennas x 2 pols, complex Gaussian sigma 2.5 per
  component, antenna B = 0.3 A + 0.95 independent noise, rounded half-even and
  clipped to [-7, 7], high nibble real (sortpols 4-bit layout);
* rechannelizer: float64 noise + unit tone at 0.2371 fs, forward coarse PFB
  (ntap 4, lblock 4096, 2049 channels), coarse channels kept, rescaled to
  sigma 2.5 and quantised to 4 bits;
* precision correlator: 8 inputs of complex64 noise plus a shared 0.4 sky term;
* direct 1-bit correlator: sign bits of pairwise-correlated Gaussians packed
  4 channels per byte, (re, im) MSB first.
"""
import numpy as np


def quantize_4bit(re, im):
    """Round half-even, clip [-7, 7], pack (re << 4) | (im & 15)."""
    r = np.clip(np.rint(re), -7, 7).astype(np.int64)
    i = np.clip(np.rint(im), -7, 7).astype(np.int64)
    return (((r & 15) << 4) | (i & 15)).astype(np.uint8)


def direct4(T=2 ** 17, nchan=128, seed=0, sigma=2.5):
    """[4, T, nchan] uint8: inputs ant0-pol0, ant0-pol1, ant1-pol0, ant1-pol1."""
    rng = np.random.default_rng(seed)
    out = np.empty((4, T, nchan), dtype=np.uint8)
    for pol in range(2):
        a = sigma * (rng.standard_normal((T, nchan)) + 1j * rng.standard_normal((T, nchan)))
        n = sigma * (rng.standard_normal((T, nchan)) + 1j * rng.standard_normal((T, nchan)))
        b = 0.3 * a + 0.95 * n
        out[pol] = quantize_4bit(a.real, a.imag)
        out[2 + pol] = quantize_4bit(b.real, b.imag)
    return out


def coarse_pfb_fp64(ts, win, nchan=2049, ntap=4):
    """Forward coarse PFB in float64 (the data generator's, pfb_gpu_utils.py:105-121 form)."""
    lblock = 2 * (nchan - 1)
    nblock = ts.size // lblock - (ntap - 1)
    x = ts[: (nblock + ntap - 1) * lblock].reshape(-1, lblock)
    w = np.asarray(win, dtype=np.float64).reshape(ntap, lblock)
    y = np.zeros((nblock, lblock))
    for j in range(ntap):
        y += x[j: nblock + j] * w[j]
    return np.fft.rfft(y, axis=1)


def rechannelizer_input(n_coarse=2 ** 14, chan0=64, nkeep=100, seed=1, sigma=2.5, tone=0.2371):
    """Coarse 4-bit spectra [n_coarse, nkeep] (uint8) of one input, plus the
    coarse channel numbers.  Timestream length (n_coarse + 3) * 4096."""
    rng = np.random.default_rng(seed)
    L = (n_coarse + 3) * 4096
    n = np.arange(L)
    ts = rng.standard_normal(L) + np.cos(2 * np.pi * tone * n)
    N = 4 * 4096
    k = np.arange(N) - N / 2
    win = (np.hamming(N) * np.sinc(k / 4096)).astype(np.float32)
    spec = coarse_pfb_fp64(ts, win)[:, chan0: chan0 + nkeep]
    s = sigma / np.std(spec.real)
    return quantize_4bit(spec.real * s, spec.imag * s), np.arange(chan0, chan0 + nkeep)


def precision_xcorr(M=8, T=1024, F=6400, seed=2, sky=0.4, device=None):
    """complex64 F-order (M, T, F): x_i = n_i + sky * s (torch tensor on
    ``device`` if given, else numpy)."""
    if device is not None:
        import torch
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        shape = (F, T, M)
        n = torch.randn(shape + (2,), generator=g, device=device, dtype=torch.float32)
        s = torch.randn((F, T, 1, 2), generator=g, device=device, dtype=torch.float32)
        x = torch.view_as_complex((n + sky * s).contiguous())
        return x.permute(2, 1, 0)  # (M, T, F) view with F-order memory
    rng = np.random.default_rng(seed)
    n = rng.standard_normal((M, T, F)) + 1j * rng.standard_normal((M, T, F))
    s = rng.standard_normal((1, T, F)) + 1j * rng.standard_normal((1, T, F))
    return np.asfortranarray((n + sky * s).astype(np.complex64))


def _quantize_4bit_torch(re, im):
    import torch
    r = torch.clamp(torch.round(re), -7, 7).to(torch.int32)
    i = torch.clamp(torch.round(im), -7, 7).to(torch.int32)
    return (((r & 15) << 4) | (i & 15)).to(torch.uint8)


def batched_stream(nant=2, nchan=300, nchunks=64, lmin=2 ** 10, lmax=2 ** 14, loss=0.005, run=5, seed=4,
                   sigma=2.5, device=None):
    """Uneven arrival chunks of 4-bit data with packet loss.

    ``nchunks`` arrival chunks with seeded lengths uniform in [lmin, lmax]
    spectra; missing spectra come in runs of ``run`` consecutive spec_num
    values (about ``loss`` of the total).  Antenna b = 0.3 a + 0.95 noise per
    pol (as the direct 4-bit workload).  Returns
    ``(packed [nant*2][P][nchan] uint8, specnums [P] int64, lengths)`` where
    the P present spectra are stored consecutively (every input has the same
    gaps) and specnums start at 0.
    """
    rng = np.random.default_rng(seed)
    lengths = rng.integers(lmin, lmax + 1, size=nchunks)
    total = int(lengths.sum())
    nruns = rng.poisson(loss * total / run)
    starts = np.sort(rng.choice(max(1, total - run), size=nruns, replace=False))
    present = np.ones(total, dtype=bool)
    for s in starts:
        present[s:s + run] = False
    specnums = np.nonzero(present)[0].astype(np.int64)
    P = len(specnums)
    if device is not None:
        # same construction drawn with torch on the device (torch.round is
        # half-to-even like np.rint); a different random stream than numpy
        import torch
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        packed = torch.empty((2 * nant, P, nchan), dtype=torch.uint8, device=device)
        step = 1 << 16
        for p0 in range(0, P, step):
            n = min(step, P - p0)
            for pol in range(2):
                a = sigma * torch.randn((n, nchan, 2), generator=g, device=device)
                for ant in range(nant):
                    x = a if ant == 0 else 0.3 * a + 0.95 * sigma * torch.randn((n, nchan, 2), generator=g,
                                                                                device=device)
                    packed[ant * 2 + pol, p0:p0 + n] = _quantize_4bit_torch(x[..., 0], x[..., 1])
        return packed, specnums, lengths
    packed = np.empty((2 * nant, P, nchan), dtype=np.uint8)
    step = 1 << 16
    for p0 in range(0, P, step):
        n = min(step, P - p0)
        for pol in range(2):
            a = sigma * (rng.standard_normal((n, nchan)) + 1j * rng.standard_normal((n, nchan)))
            for ant in range(nant):
                if ant == 0:
                    x = a
                else:
                    x = 0.3 * a + 0.95 * sigma * (rng.standard_normal((n, nchan))
                                                  + 1j * rng.standard_normal((n, nchan)))
                packed[ant * 2 + pol, p0:p0 + n] = quantize_4bit(x.real, x.imag)
    return packed, specnums, lengths


def pack_1bit(re_bits, im_bits):
    """Sign bits [..., nchan] (1 = positive) -> sortpols 1-bit bytes
    [..., ceil(nchan/4)]: channel 4j+c at bits (7-2c, 6-2c) of byte j."""
    nchan = re_bits.shape[-1]
    ncol = (nchan + 3) // 4
    pad = 4 * ncol - nchan
    r = np.pad(re_bits.astype(np.uint8), [(0, 0)] * (re_bits.ndim - 1) + [(0, pad)])
    i = np.pad(im_bits.astype(np.uint8), [(0, 0)] * (im_bits.ndim - 1) + [(0, pad)])
    r = r.reshape(r.shape[:-1] + (ncol, 4))
    i = i.reshape(i.shape[:-1] + (ncol, 4))
    out = np.zeros(r.shape[:-1], dtype=np.uint8)
    for c in range(4):
        out |= (r[..., c] << (7 - 2 * c)) | (i[..., c] << (6 - 2 * c))
    return out


def direct1(ninput=16, T=2 ** 20, nchan=1024, rho=0.2, seed=3, device=None):
    """[ninput, T, nchan/4] uint8 1-bit inputs; torch on ``device`` if given."""
    if device is not None:
        import torch
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        ncol = nchan // 4
        out = torch.empty((ninput, T, ncol), dtype=torch.uint8, device=device)
        step = max(1, (1 << 24) // nchan)
        a, b = np.sqrt(rho), np.sqrt(1 - rho)
        shifts = torch.tensor([7, 6, 5, 4, 3, 2, 1, 0], dtype=torch.uint8, device=device)
        for t0 in range(0, T, step):
            tt = min(step, T - t0)
            s = torch.randn((tt, nchan, 2), generator=g, device=device)
            for u in range(ninput):
                x = a * s + b * torch.randn((tt, nchan, 2), generator=g, device=device)
                bits = (x > 0).to(torch.uint8).reshape(tt, ncol, 8)  # (c0 r, c0 i, c1 r, ...)
                out[u, t0:t0 + tt] = (bits << shifts).sum(dim=-1, dtype=torch.uint8)
        return out
    rng = np.random.default_rng(seed)
    s = rng.standard_normal((T, nchan, 2))
    out = []
    for _ in range(ninput):
        x = np.sqrt(rho) * s + np.sqrt(1 - rho) * rng.standard_normal((T, nchan, 2))
        out.append(pack_1bit(x[..., 0] > 0, x[..., 1] > 0))
    return np.stack(out)
