"""CPU restatement of the reference's CPU rechannelizer, for the timed CPU
baseline (``cpu_baseline`` kind "port" in bench.py).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows src/utils/pfb_cpu_utils.py:90-172: ``cpu_ipfb`` (scipy.fft irfft ->
transpose -> rfft -> filter -> irfft -> transpose with cut) and ``cpu_pfb``
(polyphase then rfft), in float32 with ``scipy.fft`` workers = cores.  The
Numba loops of the reference (numba is not installed) are vectorised numpy.
"""
import os

import numpy as np

try:
    import scipy.fft as sfft
except ImportError:  # pragma: no cover - scipy is in the image
    sfft = None


def ncores():
    return len(os.sched_getaffinity(0))


def cpu_ipfb(dat, filt, cut=0, workers=None):
    """pfb_cpu_utils.py:90-121 -> ravelled rows [cut, nslice-cut)."""
    w = workers or ncores()
    dd = sfft.irfft(dat, axis=1, workers=w)
    ddft = sfft.rfft(dd.T, axis=1, workers=w)
    ddft *= filt
    res = sfft.irfft(ddft, axis=1, workers=w)
    n = res.shape[1]
    return np.ascontiguousarray(res.T[cut:n - cut]).ravel()


def cpu_pfb(timestream, win, nchan=2049, ntap=4, c0=0, c1=None, workers=None):
    """pfb_cpu_utils.py:123-172 (polyphase + rfft), optional channel slice."""
    w = workers or ncores()
    lblock = 2 * (nchan - 1)
    nblock = timestream.size // lblock - (ntap - 1)
    x = timestream[: (nblock + ntap - 1) * lblock].reshape(-1, lblock)
    wv = win.reshape(ntap, lblock)
    y = x[0:nblock] * wv[0]
    for j in range(1, ntap):
        y += x[j:nblock + j] * wv[j]
    out = sfft.rfft(y, axis=1, workers=w)
    return out[:, c0:c1] if c1 is not None else out


def rechannelize(chunks_c64, filt, win, acclen, cut, osamp, c0, c1, workers=None):
    """Unpacked (zero-filled, 2049-channel complex64) chunks -> iPFB stream ->
    re-PFB fine channels of every complete block, as helper_cpu.py:150-176."""
    carried = np.zeros((2 * cut, 2049), dtype=np.complex64)
    parts = []
    for ch in chunks_c64:
        tile = np.concatenate([carried, ch])
        carried = tile[-2 * cut:]
        parts.append(cpu_ipfb(tile, filt, cut, workers))
    ts = np.concatenate(parts)
    return cpu_pfb(ts, win, 2048 * osamp + 1, 4, c0, c1, workers)


def avg_xcorr_all_ant(x, workers=None):
    """correlations.py:115-151 ``avg_xcorr_all_ant_cpu``: for every channel f
    and pair i <= j, the complex64 products ``x[i,k,f] * conj(x[j,k,f])``
    summed over k in complex128, divided by K and stored as complex64, the
    lower triangle the conjugate.  ``x`` is ``(M, K, F)`` complex64; returns
    ``(M, M, F)`` complex64.  The Numba ``prange`` over channels becomes
    channel chunks on a thread pool (numpy releases the GIL in its loops);
    each chunk runs the step-by-step arithmetic of
    ``oracle.corr.avg_xcorr_all_ant_c64_products``."""
    from concurrent.futures import ThreadPoolExecutor
    from . import corr
    x = np.asarray(x, dtype=np.complex64)
    M, K, F = x.shape
    workers = int(workers or ncores())
    out = np.empty((M, M, F), dtype=np.complex64)
    nchunk = max(1, min(F, 4 * workers))
    edges = np.linspace(0, F, nchunk + 1).astype(np.int64)

    def one(c):
        f0, f1 = int(edges[c]), int(edges[c + 1])
        if f1 > f0:
            out[:, :, f0:f1] = corr.avg_xcorr_all_ant_c64_products(x[:, :, f0:f1], K)

    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(one, range(nchunk)))
    return out
