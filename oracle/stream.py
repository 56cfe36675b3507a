"""Restatement of the streaming re-PFB + correlate driver on in-memory chunks.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows scripts/xcorr/helper_gpu_stream.py:62-284 with two documented
deviations taken from the reference's own refactor:
* correlator input row = ``ant*npol + pol`` (helper_gpu_stream_clean.py:375;
  helper_gpu_stream.py:222 writes ``j*nant + k``, wrong for nant > 2);
* one missing-count list per antenna (helper_gpu_stream_clean.py:204;
  helper_gpu_stream.py:125 aliases one list across antennas).
The zero-fill channel map is the selected channel numbers
(helper_gpu_stream.py:196), not the clean version's ``channels[:]``.
"""
import numpy as np

from . import pfb as opfb
from .corr import avg_xcorr_all_ant


def get_chunkier(nchunks, lchunk, nblock, lblock, ntap):
    """helper_gpu_stream.py:256-284 job planner: (first_chunk, end_chunk)
    ranges of iPFB chunks per PFB job; the last may end with ``None``."""
    ranges = []
    stride = nblock * lblock
    total = stride + (ntap - 1) * lblock
    remainder = 0
    cur = 0
    while True:
        start = cur
        while remainder < total and cur < nchunks:
            remainder += lchunk
            cur += 1
        if remainder >= total:
            ranges.append((start, cur))
            remainder -= stride
        else:
            if remainder > 0:
                ranges.append((start, None))
            break
    return ranges


def make_continuous(spec, chan_specnum, start_specnum, channels, nspec, nchans=2049):
    """baseband_data_classes.py:48-73 ``make_continuous_gpu``: scatter the
    rows present (packed at the top of ``spec``) to
    ``(specnum - start) % nspec`` and the columns to channel numbers
    ``channels``; everything else is zero."""
    rows = (np.asarray(chan_specnum, dtype=np.int64) - start_specnum) % nspec
    spec = np.asarray(spec)
    if rows.shape[0] == nspec and len(channels) == nchans:
        return spec.astype(np.complex128)
    out = np.zeros((nspec, nchans), dtype=np.complex128)
    out[np.ix_(rows, np.asarray(channels, dtype=np.int64))] = spec[: len(rows)]
    return out


def ipfb_stream(chunks, start_specnum, chan_numbers, acclen, cut, filt, pol):
    """Per (antenna, pol) timestream: for every chunk, prepend the carried
    ``2*cut`` rows (zeros for the first chunk), iPFB, keep rows
    ``[cut:-cut]`` and ravel (helper_gpu_stream.py:195-199)."""
    cut_pol = np.zeros((2 * cut, 2049), dtype=np.complex128)
    pieces = []
    for ch in chunks:
        cont = make_continuous(ch["pol%d" % pol], ch["specnums"], start_specnum,
                               chan_numbers, acclen, 2049)
        tile = np.concatenate([cut_pol, cont], axis=0)
        cut_pol = tile[-2 * cut:]
        res = opfb.ipfb(tile, filt)
        pieces.append(res[cut: tile.shape[0] - cut].ravel())
    return np.concatenate(pieces) if pieces else np.zeros(0)


def repfb_xcorr_avg(chunks, start_specnums, chan_numbers, repfb_chanstart, repfb_chanend,
                    acclen, nchunks, nblock, osamp, cut, filt, win, ntap=4, npol=2,
                    return_xin=False):
    """helper_gpu_stream.py:62-254 restated on arrays.

    ``chunks[j]`` is the list of BasebandFileIterator outputs of antenna j
    (dicts with ``pol0``, ``pol1``, ``specnums``).  Returns
    ``(vis, missing_fraction, freqs)`` with ``vis`` complex128
    ``[M, M, nchan, njobs]``.
    """
    nant = len(chunks)
    lblock = 4096 * osamp
    szblock = (nblock + ntap - 1) * lblock
    lchunk = 4096 * acclen
    nchan = repfb_chanend - repfb_chanstart
    jobs = get_chunkier(nchunks, lchunk, nblock, lblock, ntap)
    M = nant * npol
    streams = [[ipfb_stream(chunks[j][:nchunks], start_specnums[j], chan_numbers, acclen, cut,
                            filt, k) for k in range(npol)] for j in range(nant)]
    vis = np.zeros((M, M, nchan, len(jobs)), dtype=np.complex128)
    xins = []
    for s in range(len(jobs)):
        xin = np.zeros((M, nblock, nchan), dtype=np.complex128)
        for j in range(nant):
            for k in range(npol):
                buf = np.zeros(szblock)
                seg = streams[j][k][s * nblock * lblock: s * nblock * lblock + szblock]
                buf[: len(seg)] = seg
                xin[j * npol + k] = opfb.pfb(buf, win, 2048 * osamp + 1, ntap)[:, repfb_chanstart:repfb_chanend]
        vis[:, :, :, s] = avg_xcorr_all_ant(xin, nblock)
        if return_xin:
            xins.append(xin)
    missing = missing_fraction(chunks, acclen, nchunks, nblock, lblock, lchunk, ntap, jobs)
    freqs = np.arange(repfb_chanstart, repfb_chanend)
    if return_xin:
        return vis, missing, freqs, xins
    return vis, missing, freqs


def missing_fraction(chunks, acclen, nchunks, nblock, lblock, lchunk, ntap, jobs):
    """Missing-data bookkeeping of helper_gpu_stream.py:157-233 with one list
    per antenna: every chunk contributes its missing percentage to the jobs
    its samples reach."""
    nant = len(chunks)
    szblock = (nblock + ntap - 1) * lblock
    out = np.zeros((nant, len(jobs)), dtype=np.float64)
    for j in range(nant):
        lst = []
        pfb_idx = 0
        rem_idx = 0
        for s, (c0, c1) in enumerate(jobs):
            if s > 0:
                pfb_idx = (ntap - 1) * lblock
            sub = chunks[j][c0:c1] if c1 is not None else chunks[j][c0:nchunks]
            if pfb_idx + rem_idx > szblock:
                rem_idx -= szblock - pfb_idx
                pfb_idx = szblock
            else:
                pfb_idx += rem_idx
                rem_idx = 0
                for ch in sub:
                    final_idx = pfb_idx + lchunk
                    lst.append([(1 - len(ch["specnums"]) / acclen) * 100,
                                (final_idx // lblock) % nblock + 1])
                    if final_idx > szblock:
                        rem_idx = final_idx - szblock
                        pfb_idx = szblock
                    else:
                        rem_idx = 0
                        pfb_idx = final_idx
            with np.errstate(invalid="ignore", divide="ignore"):
                tot = sum(m[0] for m in lst)
                out[j, s] = tot / len(lst) if lst else np.nan
            for m in lst:
                m[1] -= 1
            lst = [m for m in lst if m[1] > 0]
    return out
