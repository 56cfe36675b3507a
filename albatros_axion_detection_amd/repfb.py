"""Streaming re-PFB + cross-correlation -- the call surface of
scripts/xcorr/helper_gpu_stream.py (``repfb_xcorr_avg``, ``get_chunkier``).

The reference loops job -> antenna -> pol -> chunk with a handful of CuPy /
cuFFT / cuBLAS launches per chunk and a host sync per job.  Here one pass over
the whole stream (or over a group of jobs, when memory is bounded) is:

1. ``alb_build_rowmap``: spec_num -> row map of the fixed iPFB tiles, so
   uneven arrival chunks with lost spectra are re-blocked without copies
   (helper_gpu_stream.py:183-199; the ``2*cut`` rows carried from the
   previous chunk are read in place);
2. per input, per batch of tiles: ``alb_ipfb_rows_packed4_blocked`` (4-bit
   unpack + zero fill + channel scatter + inverse FFT of each spectrum) and
   ``alb_ipfb_columns_repfb`` (transform along time, Wiener filter, inverse
   transform, and re-PFB stage 1 of every block inside the tile); or, in the
   spectral order, ``alb_ipfb_chan_fft_packed4`` + ``alb_ipfb_rows_spectral``
   + ``alb_ipfb_columns_repfb_spectral``;
3. per input: ``alb_repfb_stage1_list`` for the blocks that straddle tiles
   (job s covers stream samples ``[s*nblock*lblock, s*nblock*lblock +
   (nblock+ntap-1)*lblock)``, helper_gpu_stream.py:148-150, 217-222); blocks
   that read only the zero-filled stream tail get their channels from one
   fill instead;
4. ``alb_repfb_stage2_inputs``: the kept fine channels of every input and
   block, written into the correlator layout;
5. ``alb_xcorr_all_ant_strided``: every job's ``M x M`` visibilities.

``RePFBPipeline.correlate_per_chunk`` keeps the reference's per-chunk,
per-job launch structure in plain FP32 (the form the batched pass is
measured against).  Row ``ant*npol + pol`` of the correlator input follows
helper_gpu_stream_clean.py:375 (helper_gpu_stream.py:222 indexes ``j*nant+k``).
"""
import ctypes

import numpy as np

from . import _lib
from . import pfb as _pfb

# largest re-PFB workspace (all blocks of one input at once) the fused path allocates
FUSED_WORKSPACE_LIMIT = 16 << 30
# stage-1 workspaces of every input kept at once for one multi-input stage 2
# (a larger pass goes through in groups of jobs)
INPUTS_WORKSPACE_LIMIT = 64 << 30
# default time-stream window of the file-driven entry point: jobs go through in
# groups whose stream rows fit in it
STREAM_WINDOW_BYTES = 8 << 30
# re-PFB workspace of one unfused alb_repfb call (blocks are batched up to it)
BATCH_WORKSPACE_LIMIT = 8 << 30


def get_chunkier(nchunks, lchunk, nblock, lblock, ntap):
    """helper_gpu_stream.py:256-284: which iPFB chunks each re-PFB job consumes."""
    ranges = []
    stride_sz = nblock * lblock
    total_needed = stride_sz + (ntap - 1) * lblock
    remainder = 0
    cur_chunk = 0
    while True:
        start_chunk = cur_chunk
        while remainder < total_needed and cur_chunk < nchunks:
            remainder += lchunk
            cur_chunk += 1
        if remainder >= total_needed:
            ranges.append((start_chunk, cur_chunk))
            remainder -= stride_sz
        else:
            if remainder > 0:
                ranges.append((start_chunk, None))
            break
    return ranges


def missing_fraction(nspec_present, acclen, nblock, lblock, lchunk, ntap, jobs):
    """Missing-data percentage per (antenna, job): helper_gpu_stream.py:157-233
    bookkeeping with one list per antenna (helper_gpu_stream_clean.py:204).
    ``nspec_present[j][c]`` = spectra present in chunk c of antenna j."""
    nant = len(nspec_present)
    nchunks = len(nspec_present[0]) if nant else 0
    szblock = (nblock + ntap - 1) * lblock
    out = np.zeros((nant, len(jobs)), dtype=np.float64)
    for j in range(nant):
        lst = []
        pfb_idx = 0
        rem_idx = 0
        for s, (c0, c1) in enumerate(jobs):
            if s > 0:
                pfb_idx = (ntap - 1) * lblock
            sub = nspec_present[j][c0:c1] if c1 is not None else nspec_present[j][c0:nchunks]
            if pfb_idx + rem_idx > szblock:
                rem_idx -= szblock - pfb_idx
                pfb_idx = szblock
            else:
                pfb_idx += rem_idx
                rem_idx = 0
                for npres in sub:
                    final_idx = pfb_idx + lchunk
                    lst.append([(1 - npres / acclen) * 100, (final_idx // lblock) % nblock + 1])
                    if final_idx > szblock:
                        rem_idx = final_idx - szblock
                        pfb_idx = szblock
                    else:
                        rem_idx = 0
                        pfb_idx = final_idx
            out[j, s] = (sum(m[0] for m in lst) / len(lst)) if lst else np.nan
            for m in lst:
                m[1] -= 1
            lst = [m for m in lst if m[1] > 0]
    return out


class RePFBPipeline:
    """Device-resident unpack -> iPFB -> re-PFB -> correlate engine.

    Parameters mirror helper_gpu_stream.py:62: ``acclen`` spectra per chunk,
    ``cut`` carried/dropped rows, ``osamp`` upchannelization, ``nblock`` re-PFB
    spectra per job, fine channels ``[repfb_chanstart, repfb_chanend)``,
    ``chan_numbers`` = coarse channel number of each packed column
    (helper_gpu_stream.py:196).  ``tile_batch`` iPFB tiles go through one row
    and column launch; ``lanes`` inputs run their stage 1 concurrently in
    :meth:`correlate_into` (one HIP stream and stream buffer each).  With
    ``stage2_overlap`` each input's re-PFB stage 2 runs on a stream of its
    own as soon as that input's stage 1 is done, beside the next inputs'
    stage 1 (off by default: on MI355X the overlapped kernels evict each
    other's working sets, 67.4 against 63.4 ms on the batched workload;
    :meth:`autotune` tries both).  ``ipfb_form`` picks the iPFB order of the
    fused pass ("direct", the default, or "spectral"; see below).  With
    ``split_edges`` (the default, where the shape allows it: osamp 64, ntap
    <= 4) :meth:`correlate_into` takes the re-PFB blocks that straddle two
    iPFB tiles in two halves, one from each tile's column pass, added in
    stage 2: no edge-block launch and no time-stream buffer per lane.
    """

    def __init__(self, ninput, chan_numbers, acclen, cut, osamp, nblock, repfb_chanstart, repfb_chanend,
                 filt_thresh=0.45, ntap=4, filt=None, win=None, tile_batch=4, nchan=2049, row_pass="auto",
                 lanes=2, stage2_overlap=False, ipfb_form="auto", split_edges=True):
        t = _lib.torch()
        self.ninput = ninput
        self.acclen = int(acclen)
        self.cut = int(cut)
        self.nslice = self.acclen + 2 * self.cut
        self.osamp = int(osamp)
        self.nblock = int(nblock)
        self.ntap = int(ntap)
        self.lblock = 4096 * self.osamp
        self.c_lo = int(repfb_chanstart)
        self.c_hi = int(repfb_chanend)
        self.nfine = self.c_hi - self.c_lo
        self.tile_batch = int(tile_batch)
        # inputs whose stage 1 runs concurrently in correlate_into (one HIP
        # stream, tile buffer and stream buffer each)
        self.lanes = max(1, int(lanes))
        self.stage2_overlap = bool(stage2_overlap)
        self.split_edges = bool(split_edges)
        if nchan != 2049:
            raise NotImplementedError("the fused iPFB handles 2049-channel spectra (lblock 4096)")
        chan_numbers = np.asarray(chan_numbers)
        if chan_numbers.size == 0 or chan_numbers.min() < 0 or chan_numbers.max() > 2048:
            raise ValueError("coarse channel numbers must lie in [0, 2048]")
        # a repeated channel number keeps its last column (numpy's fancy
        # assignment, baseband_data_classes.py:72): earlier copies are -1, which
        # the row passes skip
        chan_use = chan_numbers.astype(np.int32).copy()
        seen = set()
        for k in range(len(chan_use) - 1, -1, -1):
            if int(chan_use[k]) in seen:
                chan_use[k] = -1
            else:
                seen.add(int(chan_use[k]))
        self.chan = _lib.to_device(chan_use, t.int32)
        self.ncols = int(self.chan.numel())
        self.filt = _lib.to_device(filt, t.complex64) if filt is not None else _pfb.get_filter(self.nslice, filt_thresh,
                                                                                                  ntap=ntap)
        self.win = _lib.to_device(win, t.float32) if win is not None else _pfb.window(ntap, self.lblock)
        # optional per-launch hook timer(name, fn) (bench.py times each kernel
        # with HIP events on the stream it is launched on)
        self.timer = None
        self._tiles = {}
        self._ws = None
        self._xin_f = None
        self._lane_streams = None
        self._s2_stream = None   # stage-2 stream and per-input events (stage2_overlap)
        self._s2_events = []
        # row pass: "fft" (2048-point FFT per spectrum) or "mfma" (matrix-core
        # synthesis from the ncols non-zero channels); with a power-of-two
        # nslice both write the blocked tile layout, which the column pass reads
        # as one contiguous strip.  "auto" takes the FFT form (faster on MI355X
        # at 100 channels: 0.107 vs 0.164 ms for 16 352 spectra).
        if row_pass not in ("auto", "fft", "mfma"):
            raise ValueError("row_pass must be 'auto', 'fft' or 'mfma'")
        self.pow2 = self.nslice & (self.nslice - 1) == 0
        if row_pass == "mfma" and not self.pow2:
            raise ValueError("the matrix-core row pass needs a power-of-two nslice")
        if row_pass == "mfma" and self.ncols > 512:
            raise ValueError("the matrix-core row pass takes at most 512 channels")
        self.row_pass = "mfma" if row_pass == "mfma" else "fft"
        # iPFB order of the fused pass: "direct" is the reference's order
        # (row pass, then both transforms along time in the column pass);
        # "spectral" runs the transform along time first, on the channel
        # columns only (alb_ipfb_chan_fft_packed4 + alb_ipfb_rows_spectral +
        # alb_ipfb_columns_repfb_spectral), leaving one 4096-point transform
        # per column pair in the column pass.  Both give the same result to
        # float32 rounding.  "auto" takes the direct order: on MI355X the
        # spectral one's column pass is 10% faster but its two launches before
        # it cost more than the row pass (iPFB launches 1.18 against 1.07 ms per
        # 16 tiles on the batched workload); :meth:`autotune` tries both where
        # both apply.
        if ipfb_form not in ("auto", "spectral", "direct"):
            raise ValueError("ipfb_form must be 'auto', 'spectral' or 'direct'")
        self.spectral_ok = self.nslice == 4096 and self.row_pass == "fft"
        if ipfb_form == "spectral" and not self.spectral_ok:
            raise ValueError("the spectral-order iPFB needs nslice 4096 and the FFT row pass")
        self.ipfb_form = "spectral" if ipfb_form == "spectral" else "direct"
        self._chan = {}
        self.coeff = None
        if self.row_pass == "mfma":
            nbytes = int(_lib.load().alb_ipfb_rows_coeff_bytes(self.ncols))
            self.coeff = t.empty(nbytes, dtype=t.uint8, device=_lib.device())
            _lib.call("alb_ipfb_rows_coeff", _lib.ptr(self.chan), self.ncols, _lib.ptr(self.coeff), _lib.stream())

    def _launch(self, name, entry, *args):
        """Call library entry point ``entry``, through the timer hook when set."""
        if self.timer is None:
            _lib.call(entry, *args)
        else:
            self.timer(name, lambda: _lib.call(entry, *args))

    def _rows(self, packed, rowmap, t0, nt, y):
        """Row pass of tiles t0 .. t0+nt-1 into ``y``; returns the tile row
        stride argument of the column passes (4096 row-major, 8 blocked)."""
        s = _lib.stream()
        if self.row_pass == "mfma":
            _lib.call("alb_ipfb_rows_packed4_mfma", _lib.ptr(packed), _lib.ptr(rowmap), _lib.ptr(self.chan),
                      self.ncols, self.acclen, self.cut, t0, nt, _lib.ptr(self.coeff), _lib.ptr(y), s)
            return 8
        if self.pow2:
            self._launch("ipfb_rows_packed4", "alb_ipfb_rows_packed4_blocked", _lib.ptr(packed), _lib.ptr(rowmap),
                         _lib.ptr(self.chan), self.ncols, self.acclen, self.cut, t0, nt, _lib.ptr(y), s)
            return 8
        _lib.call("alb_ipfb_rows_packed4", _lib.ptr(packed), _lib.ptr(rowmap), _lib.ptr(self.chan), self.ncols,
                  self.acclen, self.cut, t0, nt, _lib.ptr(y), s)
        return 4096

    # ---------------------------------------------------------------- plan
    def plan(self, nchunks):
        lchunk = 4096 * self.acclen
        jobs = get_chunkier(nchunks, lchunk, self.nblock, self.lblock, self.ntap)
        nblocks_total = len(jobs) * self.nblock
        rows_needed = (nblocks_total + self.ntap - 1) * self.osamp
        rows = max(rows_needed, nchunks * self.acclen)
        return jobs, nblocks_total, rows

    def _tile_buffer(self, ntiles, lane=0):
        """float32 buffer of ``ntiles`` iPFB tiles: the row pass's blocked tiles,
        or the spectral form's strips ([1024][nslice/2+1][4] complex64 per tile)."""
        t = _lib.torch()
        need = ntiles * max(self.nslice * 4096, 1024 * (self.nslice // 2 + 1) * 8)
        if lane not in self._tiles or self._tiles[lane].numel() < need:
            self._tiles[lane] = t.empty(need, dtype=t.float32, device=_lib.device())
        return self._tiles[lane]

    def _chan_buffer(self, ntiles, lane=0):
        """Spectral form: the channel columns' transforms along time of
        ``ntiles`` tiles ([ntiles][nslice][ncols] complex64)."""
        t = _lib.torch()
        need = ntiles * self.nslice * self.ncols
        if lane not in self._chan or self._chan[lane].numel() < need:
            self._chan[lane] = t.empty(need, dtype=t.complex64, device=_lib.device())
        return self._chan[lane]

    def _workspace(self, nblk, ninput=1):
        """Stage-1 workspace of ``ninput`` inputs of ``nblk`` blocks each
        (input m at element m * per); returns (tensor, per)."""
        t = _lib.torch()
        per = int(_lib.load().alb_repfb_workspace_bytes(nblk, self.osamp)) // 8
        need = per * ninput
        if self._ws is None or self._ws.numel() < need:
            self._ws = t.empty(need, dtype=t.complex64, device=_lib.device())
        return self._ws, per

    # --------------------------------------------------------------- stages
    def ipfb_stream(self, packed, rowmap, nchunks, stream_out, tiles=None):
        """Fill ``stream_out`` ([rows, 4096] float32) with the iPFB time stream
        of one input: ``packed`` uint8 [nchunks, acclen, ncols] (present rows
        first), ``rowmap`` int32 [nchunks, acclen].  ``tiles = (c_lo, c_hi)``
        (default all ``nchunks``): only those tiles, tile t's kept rows at
        rows ``(t - c_lo) * acclen`` of ``stream_out``; rows past them are
        zeroed."""
        c_lo, c_hi = tiles if tiles is not None else (0, nchunks)
        s = _lib.stream()
        rows_valid = (c_hi - c_lo) * self.acclen
        total = stream_out.shape[0]
        if total > rows_valid:
            tail = stream_out[rows_valid:]
            self._launch("fill_zero", "alb_fill_zero", _lib.ptr(tail), tail.numel() * 4, s)
        tb = self.tile_batch
        y = self._tile_buffer(min(tb, c_hi - c_lo))
        for t0 in range(c_lo, c_hi, tb):
            nt = min(tb, c_hi - t0)
            ldy = self._rows(packed, rowmap, t0, nt, y)
            out = stream_out[(t0 - c_lo) * self.acclen:]
            _pfb.ipfb_columns(y, ldy, self.nslice * 4096, nt, self.nslice, 4096, self.filt, self.cut, out, 4096,
                              self.acclen * 4096)

    def block_batch(self, bpj, limit=512):
        """Blocks per unfused ``alb_repfb`` call: a multiple of the job length
        ``bpj``, at most ``limit`` blocks and (at least one job) whose
        workspace stays within ``BATCH_WORKSPACE_LIMIT`` (one re-PFB block is
        1 GiB of workspace at osamp 65536)."""
        per = int(_lib.load().alb_repfb_workspace_bytes(1, self.osamp))
        nb = min(int(limit), max(1, BATCH_WORKSPACE_LIMIT // per))
        return max(1, nb // bpj) * bpj

    def repfb(self, stream_rows, nblocks_total, out, sb, sc, bpj, sj, block_batch=None):
        """Re-PFB all ``nblocks_total`` spectra of one input's stream into
        ``out`` (element addressing of ``alb_repfb``), ``block_batch`` blocks
        per call (default :meth:`block_batch`)."""
        if block_batch is None:
            block_batch = self.block_batch(bpj)
        s = _lib.stream()
        for b0 in range(0, nblocks_total, block_batch):
            nb = min(block_batch, nblocks_total - b0)
            ws, _ = self._workspace(nb)
            ts = stream_rows[b0 * self.osamp:]
            # the batch's first spectrum is global block b0: shift the output base
            j0, r0 = divmod(b0, bpj)
            base = out.view(-1)[j0 * sj + r0 * sb:] if b0 else out.view(-1)
            if r0 != 0:
                raise ValueError("block_batch must be a multiple of the job length")
            _lib.call("alb_repfb", _lib.ptr(ts), _lib.ptr(self.win), self.ntap, nb, self.osamp, _lib.ptr(ws),
                      self.c_lo, self.c_hi, _lib.ptr(base), sb, sc, bpj, sj, s)

    # ----------------------------------------------------- fused iPFB/re-PFB
    def fused_supported(self, nblocks_total):
        """True when the column pass can carry re-PFB stage 1 (nslice 4096,
        osamp 32 or 64) and the all-blocks workspace stays within
        ``FUSED_WORKSPACE_LIMIT``."""
        lib = _lib.load()
        rc = lib.alb_ipfb_columns_repfb(None, 4096, 0, 0, self.nslice, None, self.nslice // 2 + 1, self.cut, None, 0,
                                        self.acclen, None, self.ntap, self.osamp, 0, None, None)
        ws = int(lib.alb_repfb_workspace_bytes(nblocks_total, self.osamp))
        return rc == 0 and ws <= FUSED_WORKSPACE_LIMIT

    def interior(self, g0, nblocks_total):
        """Re-PFB blocks ``[lo, hi)`` whose rows all lie in the tile whose kept
        rows start at stream row ``g0``."""
        lo, hi = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("alb_repfb_interior", int(g0), self.acclen, self.osamp, self.ntap, int(nblocks_total),
                  ctypes.byref(lo), ctypes.byref(hi))
        return lo.value, hi.value

    def edge_blocks(self, nchunks, nblocks_total, tiles=None, row_lo=0):
        """Blocks not interior to any tile (sorted int32): tiles ``tiles =
        (c_lo, c_hi)`` (default the ``nchunks`` tiles) of a stream window
        whose row 0 is stream row ``row_lo`` (tile t's kept rows start at
        window row ``t * acclen - row_lo``)."""
        c_lo, c_hi = tiles if tiles is not None else (0, nchunks)
        mask = np.ones(nblocks_total, dtype=bool)
        for c in range(c_lo, c_hi):
            lo, hi = self.interior(c * self.acclen - row_lo, nblocks_total)
            mask[lo:hi] = False
        return np.nonzero(mask)[0].astype(np.int32)

    def stream_repfb(self, packed, rowmap, nchunks, stream_out, nblocks_total, out, sb, sc, bpj, sj):
        """unpack -> iPFB -> re-PFB of one input into ``out`` (addressing of
        ``alb_repfb``).  Fused form: the column pass produces re-PFB stage 1 of
        every block inside one tile and writes only the stream rows the other
        (edge) blocks read; those run from the stream buffer once all tiles are
        done, then stage 2 runs over all blocks.  Other shapes: the separate
        column pass and re-PFB."""
        if not self.fused_supported(nblocks_total):
            self.ipfb_stream(packed, rowmap, nchunks, stream_out)
            self.repfb(stream_out, nblocks_total, out, sb, sc, bpj, sj)
            return
        ws, _ = self._workspace(nblocks_total)
        self.stream_stage1(packed, rowmap, nchunks, stream_out, nblocks_total, ws)
        _lib.call("alb_repfb_stage2", _lib.ptr(ws), nblocks_total, self.osamp, self.c_lo, self.c_hi, _lib.ptr(out),
                  sb, sc, bpj, sj, _lib.stream())

    def zero_from(self, rows_valid, nblocks_total):
        """First block (of ``nblocks_total``) whose rows all lie at or past
        window row ``rows_valid``: it and every later block read only the
        zero-filled tail of the stream (helper_gpu_stream.py:215-219), so
        their re-PFB output is exactly zero."""
        return int(min(nblocks_total, max(0, -(-int(rows_valid) // self.osamp))))

    def stream_stage1(self, packed, rowmap, nchunks, stream_out, nblocks_total, ws, lane=0, tiles=None, row_lo=0,
                      skip_zero_tail=False, halves=None, p_slots=0):
        """Fused form up to re-PFB stage 1: every block's stage-1 result of one
        input in the workspace ``ws`` (row pass, column pass with stage 1 of
        the interior blocks, then the edge blocks from the stream buffer; in
        the spectral form the channel transform along time and the spectra
        along the samples replace the row pass), on
        the current stream with tile buffer ``lane``.  ``tiles = (c_lo,
        c_hi)`` and ``row_lo`` select a window of the stream (a group of
        jobs): ``stream_out`` row 0 and block 0 start at stream row
        ``row_lo``; rows of the window past the tiles are zeroed.  With
        ``skip_zero_tail`` the blocks from :meth:`zero_from` on are left out
        (their workspace is not written) and only the zero rows that earlier
        blocks read are cleared.  With ``halves`` (a complex64 device tensor
        of ``p_slots`` blocks) the column passes take the split form: the
        second halves of the tile-boundary blocks go to ``halves`` (added by
        :meth:`correlate_into`'s stage 2), ``stream_out`` is not written and
        there is no edge-block launch."""
        c_lo, c_hi = tiles if tiles is not None else (0, nchunks)
        if halves is not None:
            s = _lib.stream()
            tb = self.tile_batch
            y = self._tile_buffer(min(tb, c_hi - c_lo), lane)
            if self.ipfb_form == "spectral":
                E = self._chan_buffer(min(tb, c_hi - c_lo), lane)
            for t0 in range(c_lo, c_hi, tb):
                nt = min(tb, c_hi - t0)
                geo = (self.cut, t0 * self.acclen - row_lo, self.acclen, _lib.ptr(self.win), self.ntap, self.osamp,
                       nblocks_total, _lib.ptr(ws), _lib.ptr(halves), t0 - c_lo, int(p_slots), s)
                if self.ipfb_form == "spectral":
                    self._launch("ipfb_chan_fft", "alb_ipfb_chan_fft_packed4", _lib.ptr(packed), _lib.ptr(rowmap),
                                 _lib.ptr(self.chan), self.ncols, self.acclen, self.cut, t0, nt, _lib.ptr(E), s)
                    self._launch("ipfb_rows_spectral", "alb_ipfb_rows_spectral", _lib.ptr(E), _lib.ptr(self.chan),
                                 self.ncols, self.nslice, nt, _lib.ptr(y), s)
                    self._launch("ipfb_columns_repfb", "alb_ipfb_columns_repfb_spectral_split", _lib.ptr(y), nt,
                                 self.nslice, _lib.ptr(self.filt), self.filt.shape[1], *geo)
                else:
                    ldy = self._rows(packed, rowmap, t0, nt, y)
                    self._launch("ipfb_columns_repfb", "alb_ipfb_columns_repfb_split", _lib.ptr(y), ldy,
                                 self.nslice * 4096, nt, self.nslice, _lib.ptr(self.filt), self.filt.shape[1], *geo)
            return
        rows_valid = max(0, c_hi * self.acclen - row_lo)
        if stream_out.shape[0] < max((nblocks_total + self.ntap - 1) * self.osamp, rows_valid):
            raise ValueError("stream buffer too short for the re-PFB blocks")
        s = _lib.stream()
        total = stream_out.shape[0]
        nb_used = nblocks_total
        if skip_zero_tail:
            nb_used = self.zero_from(rows_valid, nblocks_total)
            total = min(total, max(rows_valid, (nb_used + self.ntap - 1) * self.osamp))
        if total > rows_valid:
            tail = stream_out[rows_valid:total]
            self._launch("fill_zero", "alb_fill_zero", _lib.ptr(tail), tail.numel() * 4, s)
        tb = self.tile_batch
        y = self._tile_buffer(min(tb, c_hi - c_lo), lane)
        if self.ipfb_form == "spectral":
            E = self._chan_buffer(min(tb, c_hi - c_lo), lane)
            for t0 in range(c_lo, c_hi, tb):
                nt = min(tb, c_hi - t0)
                self._launch("ipfb_chan_fft", "alb_ipfb_chan_fft_packed4", _lib.ptr(packed), _lib.ptr(rowmap),
                             _lib.ptr(self.chan), self.ncols, self.acclen, self.cut, t0, nt, _lib.ptr(E), s)
                self._launch("ipfb_rows_spectral", "alb_ipfb_rows_spectral", _lib.ptr(E), _lib.ptr(self.chan),
                             self.ncols, self.nslice, nt, _lib.ptr(y), s)
                self._launch("ipfb_columns_repfb", "alb_ipfb_columns_repfb_spectral", _lib.ptr(y), nt, self.nslice,
                             _lib.ptr(self.filt), self.filt.shape[1], self.cut, _lib.ptr(stream_out),
                             t0 * self.acclen - row_lo, self.acclen, _lib.ptr(self.win), self.ntap, self.osamp,
                             nblocks_total, _lib.ptr(ws), s)
        else:
            for t0 in range(c_lo, c_hi, tb):
                nt = min(tb, c_hi - t0)
                ldy = self._rows(packed, rowmap, t0, nt, y)
                self._launch("ipfb_columns_repfb", "alb_ipfb_columns_repfb", _lib.ptr(y), ldy, self.nslice * 4096,
                             nt, self.nslice, _lib.ptr(self.filt), self.filt.shape[1], self.cut, _lib.ptr(stream_out),
                             t0 * self.acclen - row_lo, self.acclen, _lib.ptr(self.win), self.ntap, self.osamp,
                             nblocks_total, _lib.ptr(ws), s)
        edges = self.edge_blocks(nchunks, nblocks_total, (c_lo, c_hi), row_lo)
        edges = np.ascontiguousarray(edges[edges < nb_used])
        self._launch("repfb_stage1_edges", "alb_repfb_stage1_list", _lib.ptr(stream_out), _lib.ptr(self.win),
                     self.ntap, edges.ctypes.data_as(ctypes.c_void_p), len(edges), self.osamp, _lib.ptr(ws), s)

    # ------------------------------------------------------------ pipelines
    def spectra(self, packed, rowmap, nchunks, fused=True):
        """Single input (rechannelizer): complex64 ``[nblocks_total, nfine]`` of
        ``cupy_pfb(stream)[:, c_lo:c_hi]`` over all jobs."""
        t = _lib.torch()
        jobs, nbt, rows = self.plan(nchunks)
        stream = t.empty((rows, 4096), dtype=t.float32, device=_lib.device())
        out = t.empty((nbt, self.nfine), dtype=t.complex64, device=stream.device)
        if fused:
            self.stream_repfb(packed, rowmap, nchunks, stream, nbt, out, self.nfine, 1, nbt, 0)
        else:
            self.ipfb_stream(packed, rowmap, nchunks, stream)
            self.repfb(stream, nbt, out, self.nfine, 1, nbt, 0, block_batch=self.block_batch(nbt, nbt))
        return out

    def fused_shape(self):
        """True when the column pass can carry re-PFB stage 1 (nslice 4096,
        osamp 32 or 64)."""
        lib = _lib.load()
        return lib.alb_ipfb_columns_repfb(None, 4096, 0, 0, self.nslice, None, self.nslice // 2 + 1, self.cut, None, 0,
                                          self.acclen, None, self.ntap, self.osamp, 0, None, None) == 0

    def split_shape(self):
        """True when the fused column pass can take the tile-boundary blocks
        in two halves (nslice 4096, osamp 64, ntap <= 4)."""
        lib = _lib.load()
        return self.fused_shape() and lib.alb_ipfb_columns_repfb_split(
            None, 4096, 0, 0, self.nslice, None, self.nslice // 2 + 1, self.cut, 0, self.acclen, None, self.ntap,
            self.osamp, 0, None, None, 0, 0, None) == 0

    def groups(self, njobs, group_jobs=None):
        """Job groups ``[s0, s1)`` of one pass: ``group_jobs`` jobs each
        (default all at once), in the fused form at most as many as keep every
        input's stage-1 workspace within ``INPUTS_WORKSPACE_LIMIT``."""
        if njobs <= 0:
            return []
        G = njobs if group_jobs is None else int(group_jobs)
        if self.fused_shape():
            per_job = int(_lib.load().alb_repfb_workspace_bytes(self.nblock, self.osamp)) * self.ninput
            G = min(G, max(1, INPUTS_WORKSPACE_LIMIT // per_job))
        G = max(1, min(G, njobs))
        return [(s0, min(njobs, s0 + G)) for s0 in range(0, njobs, G)]

    def window(self, s0, s1, nvalid):
        """Stream rows and tiles of the jobs ``[s0, s1)``: their blocks read
        stream rows ``[row_lo, row_hi)``, produced by tiles ``[c_lo, c_hi)``
        of the ``nvalid`` tiles with data (helper_gpu_stream.py:148-214: each
        job starts ``nblock`` blocks after the previous one and reads
        ``ntap - 1`` blocks beyond its own)."""
        K, R, a = self.nblock, self.osamp, self.acclen
        row_lo = s0 * K * R
        row_hi = s1 * K * R + (self.ntap - 1) * R
        c_lo = min(nvalid, row_lo // a)
        c_hi = min(nvalid, -(-row_hi // a))
        return row_lo, row_hi, c_lo, max(c_lo, c_hi)

    def _window_rows(self, s0, s1, nvalid):
        """Rows of a window buffer that serves the jobs ``[s0, s1)`` both from
        stream row ``row_lo`` (fused form) and from tile ``c_lo`` (unfused)."""
        row_lo, row_hi, c_lo, c_hi = self.window(s0, s1, nvalid)
        a = self.acclen
        return max(row_hi - row_lo, c_hi * a - row_lo, row_hi - c_lo * a, (c_hi - c_lo) * a)

    def group_jobs_for(self, window_bytes=STREAM_WINDOW_BYTES):
        """Jobs per group whose time-stream window fits in ``window_bytes``
        (at least one job)."""
        rows = int(window_bytes) // (4096 * 4)
        per = self.nblock * self.osamp
        return max(1, (rows - (self.ntap - 1) * self.osamp - 2 * self.acclen) // per)

    def buffers(self, nchunks, group_jobs=None):
        """Device buffers of one correlate pass over ``nchunks`` chunks,
        allocated once and reused (a captured hipGraph may not allocate).
        ``group_jobs`` (default: all jobs at once) bounds the memory: the time
        stream window, the stage-1 workspace and the correlator input then
        hold one group of jobs, whatever the length of the stream.  Also sizes
        the tile and re-PFB work buffers."""
        t = _lib.torch()
        jobs, nbt, _ = self.plan(nchunks)
        grp = self.groups(len(jobs), group_jobs)
        G = max([s1 - s0 for s0, s1 in grp] or [0])
        rows = max([self._window_rows(s0, s1, nchunks) for s0, s1 in grp] or [0])
        M, K, F = self.ninput, self.nblock, self.nfine
        dev = _lib.device()
        bufs = {"jobs": jobs, "nbt": nbt, "rows": rows, "groups": grp,
                # correlator input [S][K][F][M] of one group: a re-PFB block's
                # channels of all inputs are one contiguous run (whole-line
                # stage-2 stores); the correlator takes it through its strided form
                "xin": t.empty((G, K, F, M), dtype=t.complex64, device=dev),
                "stream": t.empty((rows, 4096), dtype=t.float32, device=dev),
                "vis": t.empty((len(jobs), F, M, M), dtype=t.complex64, device=dev)}
        self._tile_buffer(min(self.tile_batch, nchunks))
        if self.ipfb_form == "spectral":
            self._chan_buffer(min(self.tile_batch, nchunks))
        self.prepare()
        if self.fused_shape():
            self._workspace(G * K, M)
            nl = min(self.lanes, M)
            bufs["split"] = self.split_edges and self.split_shape()
            if bufs["split"]:
                # the second halves of the tile-boundary blocks: ntap slots per
                # tile of the largest window, per input; no lane writes a stream
                # buffer, so every lane shares the (unwritten) one
                ntw = max([c_hi - c_lo for c_lo, c_hi in (self.window(s0, s1, nchunks)[2:] for s0, s1 in grp)] or [1])
                slot = int(_lib.load().alb_repfb_split_slot_bytes(self.osamp)) // 8
                bufs["p_slots"] = self.ntap * max(1, ntw)
                bufs["halves"] = t.empty(M * bufs["p_slots"] * slot, dtype=t.complex64, device=dev)
                bufs["lane_stream_bufs"] = [bufs["stream"]] * nl
            else:
                # concurrent stage 1 of up to ``lanes`` inputs: a stream buffer
                # and a tile buffer per lane
                bufs["lane_stream_bufs"] = [bufs["stream"]] + [t.empty((rows, 4096), dtype=t.float32, device=dev)
                                                               for _ in range(nl - 1)]
            for lane in range(1, nl):
                self._tile_buffer(min(self.tile_batch, nchunks), lane)
                if self.ipfb_form == "spectral":
                    self._chan_buffer(min(self.tile_batch, nchunks), lane)
            if nl > 1 and (self._lane_streams is None or len(self._lane_streams) < nl):
                self._lane_streams = [t.cuda.Stream() for _ in range(nl)]
            # made here, before any capture: stream and event creation are not captured work
            if self._s2_stream is None:
                self._s2_stream = t.cuda.Stream()
            while len(self._s2_events) < M:
                self._s2_events.append(t.cuda.Event())
        else:
            self._workspace(self.block_batch(K))
        return bufs

    @staticmethod
    def _valid(nchunks, nvalid):
        """Chunks holding data: ``nvalid`` (at most ``nchunks``) or all."""
        if nvalid is None:
            return nchunks
        if not 0 <= nvalid <= nchunks:
            raise ValueError("nvalid must lie in [0, nchunks]")
        return int(nvalid)

    def prepare(self):
        """Build every transform table this pipeline's kernels use (4096-point
        rows, nslice-point or chirp-z columns, the re-PFB stage-2 phases) so
        that a pass captured into a hipGraph finds them built: a first build
        synchronises, which the library refuses on a capturing stream."""
        _lib.call("alb_prepare_transform", 4096)
        if self.osamp <= 64:
            _lib.call("alb_prepare_repfb", self.osamp)
        if self.pow2:
            _lib.call("alb_prepare_transform", self.nslice)
        else:
            _lib.call("alb_prepare_transform", _pfb.bluestein_tables(self.nslice)[2])

    def check_precision(self, precision):
        """Raise before any launch if the correlator cannot run ``precision``
        for this pipeline's input count (the reduced-precision matrix-core
        forms take at most 8 inputs)."""
        from .correlations import PRECISION_MODES
        if precision not in PRECISION_MODES:
            raise ValueError("precision must be one of %s" % sorted(PRECISION_MODES))
        if precision != "fp32" and self.ninput > 8:
            raise NotImplementedError("precision %r takes at most 8 correlator inputs (nant*npol = %d); use 'fp32'"
                                      % (precision, self.ninput))

    def correlate_into(self, packed_list, rowmaps, nchunks, bufs, fused=True, precision="fp32", nvalid=None):
        """Launch the whole pass into ``bufs`` (no allocation, no host sync, so
        it can be captured in a hipGraph once :meth:`buffers` has built the
        transform tables).  Jobs are planned for ``nchunks`` chunks of which
        the first ``nvalid`` (default all) hold data; the time stream past them
        is zero, as the reference's zero-filled incomplete buffer
        (helper_gpu_stream.py:215-219).  The jobs go through in the groups of
        ``bufs`` (one group unless :meth:`buffers` was given ``group_jobs``).
        Returns the ``(M, M, nfine, njobs)`` F-order view of ``bufs["vis"]``."""
        from .correlations import PRECISION_MODES
        self.check_precision(precision)
        nv = self._valid(nchunks, nvalid)
        M, K, F = self.ninput, self.nblock, self.nfine
        xin, stream, vis = bufs["xin"], bufs["stream"], bufs["vis"]
        use_fused = fused and self.fused_shape()
        bb = self.block_batch(K)
        scale = np.float32(1.0) / np.float32(K)
        a = self.acclen
        for s0, s1 in bufs["groups"]:
            row_lo, row_hi, c_lo, c_hi = self.window(s0, s1, nv)
            ng = s1 - s0
            nbg = ng * K
            if use_fused:
                # blocks from nbz on read only the zero-filled stream tail:
                # their channels are exactly zero, written by one fill instead
                # of stage 1 and stage 2
                nbz = self.zero_from(max(0, c_hi * a - row_lo), nbg)
                if nbz < nbg:
                    ztail = xin.view(-1)[nbz * F * M:nbg * F * M]
                    self._launch("fill_zero", "alb_fill_zero", _lib.ptr(ztail), ztail.numel() * 8, _lib.stream())
                # stage 1 of every input into its own workspace, then one
                # stage-2 launch writes all inputs' channels into the
                # correlator layout.  Inputs go round-robin to ``lanes`` HIP
                # streams (default priority), each with its own tile buffer
                # (and, without the split form, its own stream window), so one
                # input's launches overlap the next input's row and column
                # passes; the current stream waits for all of them.
                ws, per = self._workspace(nbg, M)
                lane_bufs = bufs.get("lane_stream_bufs", [stream])[:self.lanes]
                nl = len(lane_bufs)
                s2 = self.stage2_overlap and M > 1 and self._s2_stream is not None and len(self._s2_events) >= M
                # split form: each input's tile-boundary halves, and stage 2 adds them
                split = bool(bufs.get("split"))
                p_slots = bufs.get("p_slots", 0)
                pper = p_slots * (int(_lib.load().alb_repfb_split_slot_bytes(self.osamp)) // 8) if split else 0
                halves = (lambda m: bufs["halves"][m * pper:]) if split else (lambda m: None)

                def stage2(ws_ptr, half_ptr, nin, out_ptr):
                    if split:
                        self._launch("repfb_stage2", "alb_repfb_stage2_split", ws_ptr, per * 8, half_ptr, pper * 8, nin,
                                     nbz, self.osamp, self.c_lo, self.c_hi, out_ptr, 1, F * M, M, K, K * F * M,
                                     c_lo * a - row_lo, a, c_hi - c_lo, self.ntap, _lib.stream())
                    else:
                        self._launch("repfb_stage2", "alb_repfb_stage2_inputs", ws_ptr, per * 8, nin, nbz, self.osamp,
                                     self.c_lo, self.c_hi, out_ptr, 1, F * M, M, K, K * F * M, _lib.stream())

                if nl == 1 and not s2:
                    for m in range(M):
                        self.stream_stage1(packed_list[m], rowmaps[m], nchunks, stream, nbg, ws[m * per:],
                                           tiles=(c_lo, c_hi), row_lo=row_lo, skip_zero_tail=True, halves=halves(m),
                                           p_slots=p_slots)
                else:
                    t = _lib.torch()
                    main = t.cuda.current_stream()
                    lanes = self._lane_streams[:nl] if nl > 1 else [main]
                    if nl > 1:
                        for st in lanes:
                            st.wait_stream(main)
                    if s2:
                        self._s2_stream.wait_stream(main)
                    for m in range(M):
                        lane = m % nl
                        with t.cuda.stream(lanes[lane]):
                            self.stream_stage1(packed_list[m], rowmaps[m], nchunks, lane_bufs[lane], nbg,
                                               ws[m * per:], lane=lane, tiles=(c_lo, c_hi), row_lo=row_lo,
                                               skip_zero_tail=True, halves=halves(m), p_slots=p_slots)
                        if s2:
                            # input m's stage 2 (its own workspace, its own channels of
                            # xin) once its stage 1 is done, beside the next inputs' stage 1
                            ev = self._s2_events[m]
                            ev.record(lanes[lane])
                            self._s2_stream.wait_event(ev)
                            with t.cuda.stream(self._s2_stream):
                                stage2(_lib._vp(ws.data_ptr() + m * per * 8),
                                       _lib._vp(bufs["halves"].data_ptr() + m * pper * 8) if split else None, 1,
                                       _lib._vp(xin.data_ptr() + m * 8))
                    if nl > 1:
                        for st in lanes:
                            main.wait_stream(st)
                    if s2:
                        main.wait_stream(self._s2_stream)
                if not s2:
                    stage2(_lib.ptr(ws), _lib.ptr(bufs["halves"]) if split else None, M, _lib.ptr(xin))
            else:
                # the window starts at tile c_lo; the group's blocks at row_lo
                for m in range(M):
                    self.ipfb_stream(packed_list[m], rowmaps[m], nchunks, stream, (c_lo, c_hi))
                    self.repfb(stream[row_lo - c_lo * a:], nbg, xin.view(-1)[m:], F * M, M, K, K * F * M,
                               block_batch=bb)
            self._launch("xcorr_all_ant_strided", "alb_xcorr_all_ant_strided", _lib.ptr(xin), M, K, F, ng, F * M, M,
                         K * F * M, _lib.ptr(vis[s0]), float(scale), PRECISION_MODES[precision], _lib.stream())
        return vis.permute(3, 2, 1, 0)

    # launch-shape knobs of the library tried by autotune (0 = built-in rule)
    TUNE_KNOBS = (("ipfb_cols_width", (0, 2, 8)), ("ipfb_cols_group", (0, 2, 8, 64)), ("ipfb_rows_group", (0, 2, 4)))

    def autotune(self, packed_list, rowmaps, nchunks, bufs, precision="fp32", reps=3, tile_batches=(4, 8, 16, 32)):
        """Choose the chunk schedule and launch shapes of :meth:`correlate_into`
        for this pipeline's shapes by timing whole passes on the given inputs
        (no counterpart in the reference, whose launch shapes are cuFFT's and
        cuBLAS's own).

        Coordinate descent, one setting at a time with the others at their
        best so far: iPFB tiles per row / column launch (``tile_batch``, up
        to ``nchunks``), concurrent stage-1 lanes (up to the lane buffers
        ``bufs`` holds), stage 2 beside stage 1 or after it
        (``stage2_overlap``), the iPFB order (``ipfb_form``, where the
        spectral order applies), then the library's launch-shape knobs in
        ``TUNE_KNOBS`` (``alb_tune_set``; process-wide, so they also apply
        to other pipelines in this process).  Each candidate runs one untimed
        pass and ``reps`` timed ones (HIP events on the current stream, which
        joins every lane stream); its time is the median.  The best settings
        stay in force.  Returns ``{"tile_batch", "lanes", <knob>: value ...,
        "ms": best pass time, "tried": {candidate: ms}}``.  Larger tile
        batches allocate larger tile buffers (kept for later passes).
        """
        t = _lib.torch()

        def timed():
            self.correlate_into(packed_list, rowmaps, nchunks, bufs, precision=precision)
            ms = []
            for _ in range(int(reps)):
                a = t.cuda.Event(enable_timing=True)
                b = t.cuda.Event(enable_timing=True)
                a.record()
                self.correlate_into(packed_list, rowmaps, nchunks, bufs, precision=precision)
                b.record()
                b.synchronize()
                ms.append(a.elapsed_time(b))
            return float(np.median(ms))

        lib = _lib.load()
        tried = {}
        nl_max = len(bufs.get("lane_stream_bufs", [None]))

        def set_tile_batch(v):
            self.tile_batch = v
            for lane in range(nl_max):
                self._tile_buffer(min(v, nchunks), lane)
                if self.ipfb_form == "spectral":
                    self._chan_buffer(min(v, nchunks), lane)

        def set_lanes(v):
            self.lanes = v

        def set_overlap(v):
            self.stage2_overlap = bool(v)

        def set_form(v):
            self.ipfb_form = v
            if v == "spectral":
                for lane in range(nl_max):
                    self._chan_buffer(min(self.tile_batch, nchunks), lane)

        def knob_setter(key):
            return lambda v: _lib.call("alb_tune_set", key.encode(), int(v))

        space = [("tile_batch", sorted({min(int(v), nchunks) for v in tile_batches} | {min(self.tile_batch, nchunks)}),
                  self.tile_batch, set_tile_batch),
                 ("lanes", sorted({1, nl_max}), min(self.lanes, nl_max), set_lanes),
                 ("stage2_overlap", [0, 1], int(self.stage2_overlap), set_overlap)]
        if self.spectral_ok and self.fused_shape():
            space.append(("ipfb_form", ["direct", "spectral"], self.ipfb_form, set_form))
        space += [(key, vals, int(lib.alb_tune_get(key.encode())), knob_setter(key)) for key, vals in self.TUNE_KNOBS]
        best = {}
        for name, vals, cur, setter in space:
            setter(cur)
            times = {}
            for v in vals if cur in vals else [cur] + list(vals):
                setter(v)
                times[v] = timed()
                tried["%s=%s" % (name, v)] = round(times[v], 4)
            pick = min(times, key=times.get)
            setter(pick)
            best[name] = pick
        best["ms"] = timed()
        best["tried"] = tried
        return best

    def correlate_per_chunk(self, packed_list, rowmaps, nchunks, bufs, nvalid=None):
        """The per-chunk launch form in plain FP32, following the reference's
        loop structure (helper_gpu_stream.py:157-236): per input and chunk one
        row pass and one column pass, per job and input one re-PFB launch, per
        job one correlator launch.  Same result as :meth:`correlate_into`."""
        nchunks_planned = nchunks
        nchunks = self._valid(nchunks, nvalid)
        M, K, F = self.ninput, self.nblock, self.nfine
        stream, vis = bufs["stream"], bufs["vis"]
        if stream.shape[0] < self.plan(nchunks_planned)[2]:
            raise ValueError("the per-chunk form needs whole-stream buffers (buffers() without group_jobs)")
        njobs = len(bufs["jobs"])
        # the reference's per-job correlator input: F-order (M, K, F) per job
        t = _lib.torch()
        if self._xin_f is None or tuple(self._xin_f.shape) != (njobs, F, K, M):
            self._xin_f = t.empty((njobs, F, K, M), dtype=t.complex64, device=_lib.device())
        xin = self._xin_f
        s = _lib.stream()
        y = self._tile_buffer(1)
        ws, _ = self._workspace(K)
        rows_valid = nchunks * self.acclen
        scale = np.float32(1.0) / np.float32(K)
        for m in range(M):
            if stream.shape[0] > rows_valid:
                tail = stream[rows_valid:]
                _lib.call("alb_fill_zero", _lib.ptr(tail), tail.numel() * 4, s)
            for c in range(nchunks):
                _lib.call("alb_ipfb_rows_packed4", _lib.ptr(packed_list[m]), _lib.ptr(rowmaps[m]), _lib.ptr(self.chan),
                          self.ncols, self.acclen, self.cut, c, 1, _lib.ptr(y), s)
                _pfb.ipfb_columns(y, 4096, self.nslice * 4096, 1, self.nslice, 4096, self.filt, self.cut,
                                  stream[c * self.acclen:], 4096, self.acclen * 4096)
            for j in range(njobs):
                ts = stream[j * K * self.osamp:]
                base = xin.view(-1)[j * M * K * F + m:]
                _lib.call("alb_repfb", _lib.ptr(ts), _lib.ptr(self.win), self.ntap, K, self.osamp, _lib.ptr(ws),
                          self.c_lo, self.c_hi, _lib.ptr(base), M, M * K, K, M * K * F, s)
        for j in range(njobs):
            _lib.call("alb_xcorr_all_ant", _lib.ptr(xin[j]), M, K, F, _lib.ptr(vis[j]), float(scale), s)
        return vis.permute(3, 2, 1, 0)

    def correlate(self, packed_list, rowmaps, nchunks, return_xin=False, fused=True, precision="fp32", nvalid=None,
                  group_jobs=None):
        """All inputs: ``vis`` complex64 F-order ``(M, M, nfine, njobs)`` of
        ``avg_xcorr_all_ant_gpu`` per job (helper_gpu_stream.py:236).
        ``return_xin`` (the last group's correlator input) needs one group."""
        bufs = self.buffers(nchunks, group_jobs)
        res = self.correlate_into(packed_list, rowmaps, nchunks, bufs, fused=fused, precision=precision, nvalid=nvalid)
        if return_xin:
            return res, bufs["xin"].permute(3, 1, 2, 0)
        return res


def rowmaps_from_specnums(specnums_list, start_specnum, acclen, row_bases=None):
    """Per chunk row maps ``(specnum - start) % acclen -> packed row`` on the
    device (baseband_data_classes.py:53).  Chunk c's present spectra are
    packed rows ``row_bases[c] + p`` (default ``c*acclen``: the stacked
    BasebandFileIterator layout)."""
    t = _lib.torch()
    n = len(specnums_list)
    rm = t.empty((n, acclen), dtype=t.int32, device=_lib.device())
    s = _lib.stream()
    for c, sp in enumerate(specnums_list):
        sp = np.asarray(sp, dtype=np.int64)
        d = _lib.to_device(sp if len(sp) else np.zeros(1, dtype=np.int64), t.int64)
        base = c * acclen if row_bases is None else int(row_bases[c])
        _lib.call("alb_build_rowmap", _lib.ptr(d), len(sp), int(start_specnum), int(acclen), base, _lib.ptr(rm[c]), s)
    return rm


class StreamRowmap:
    """Row map of a stream whose present spectra are stored consecutively,
    rebuilt on the device from its spec_nums by one ``alb_build_rowmap``
    launch (so a captured pass can include it): the spec_nums of
    ``[start, start + nchunks*acclen)`` are located once on the host, the
    map ``[nchunks, acclen]`` int32 is filled by :meth:`build`."""

    def __init__(self, specnums, start_specnum, nchunks, acclen):
        t = _lib.torch()
        sp = np.asarray(specnums if _lib.is_numpy(specnums) else specnums.cpu().numpy(), dtype=np.int64)
        self.start, self.total = int(start_specnum), int(nchunks) * int(acclen)
        self.lo = int(np.searchsorted(sp, self.start, side="left"))
        self.hi = int(np.searchsorted(sp, self.start + self.total, side="left"))
        self.spec = _lib.to_device(sp[self.lo:self.hi] if self.hi > self.lo else np.zeros(1, dtype=np.int64), t.int64)
        self.rowmap = t.empty((nchunks, acclen), dtype=t.int32, device=_lib.device())

    def build(self):
        _lib.call("alb_build_rowmap", _lib.ptr(self.spec), self.hi - self.lo, self.start, self.total, self.lo,
                  _lib.ptr(self.rowmap), _lib.stream())
        return self.rowmap


def rowmap_stream(specnums, start_specnum, nchunks, acclen):
    """Row map of a whole stream whose present spectra are stored
    consecutively (packed row p holds spec_num ``specnums[p]``): re-blocks
    arrival chunks of any length into ``nchunks`` fixed chunks of ``acclen``
    spectra, missing spec_nums zero-filled.  ``specnums``: sorted int64
    (numpy or device).  Returns int32 ``[nchunks, acclen]`` on the device."""
    t = _lib.torch()
    sp = np.asarray(specnums if _lib.is_numpy(specnums) else specnums.cpu().numpy(), dtype=np.int64)
    total = nchunks * acclen
    lo = np.searchsorted(sp, start_specnum, side="left")
    hi = np.searchsorted(sp, start_specnum + total, side="left")
    rm = t.empty((nchunks, acclen), dtype=t.int32, device=_lib.device())
    d = _lib.to_device(sp[lo:hi] if hi > lo else np.zeros(1, dtype=np.int64), t.int64)
    _lib.call("alb_build_rowmap", _lib.ptr(d), int(hi - lo), int(start_specnum), int(total), int(lo), _lib.ptr(rm),
              _lib.stream())
    return rm


def stack_packed(chunks, key, acclen, ncols):
    """Stack per-chunk packed rows (present rows first) into [nchunks, acclen, ncols]."""
    out = np.zeros((len(chunks), acclen, ncols), dtype=np.uint8)
    for c, ch in enumerate(chunks):
        a = np.asarray(ch[key] if not hasattr(ch[key], "cpu") else ch[key].cpu().numpy(), dtype=np.uint8)
        out[c, : a.shape[0]] = a
    return out


def repfb_xcorr_avg_chunks(chunks, start_specnums, chan_numbers, repfb_chanstart, repfb_chanend, acclen, nchunks,
                           nblock, osamp, cut=10, filt_thresh=0.45, win=None, filt=None, npol=2, return_xin=False):
    """helper_gpu_stream.py:62-254 on in-memory chunks: ``chunks[j]`` is the
    list of packed-4-bit BasebandFileIterator outputs of antenna j (dicts with
    uint8 ``pol0``/``pol1`` and int64 ``specnums``).  Returns
    ``(vis, missing_fraction, freqs, win, filt)``; vis is numpy complex64
    F-order ``(nant*npol, nant*npol, nchan, njobs)``."""
    nant = len(chunks)
    ncols = len(chan_numbers)
    pipe = RePFBPipeline(nant * npol, chan_numbers, acclen, cut, osamp, nblock, repfb_chanstart, repfb_chanend,
                         filt_thresh=filt_thresh, win=win, filt=filt)
    packed, rmaps = [], []
    for j in range(nant):
        rm = rowmaps_from_specnums([c["specnums"] for c in chunks[j][:nchunks]], start_specnums[j], acclen)
        for k in range(npol):
            packed.append(_lib.to_device(stack_packed(chunks[j][:nchunks], "pol%d" % k, acclen, ncols)))
            rmaps.append(rm)
    out = pipe.correlate(packed, rmaps, nchunks, return_xin=return_xin)
    vis = out[0] if return_xin else out
    jobs, _, _ = pipe.plan(nchunks)
    lblock = 4096 * osamp
    npres = [[len(c["specnums"]) for c in chunks[j][:nchunks]] for j in range(nant)]
    miss = missing_fraction(npres, acclen, nblock, lblock, 4096 * acclen, 4, jobs)
    visn = np.asfortranarray(_lib.host(vis))
    freqs = np.arange(repfb_chanstart, repfb_chanend)
    res = (np.ma.masked_invalid(visn), miss, freqs, pipe.win, pipe.filt)
    if return_xin:
        return res + (out[1],)
    return res


def repfb_xcorr_avg(idxs, files, acclen, nchunks, nblock, chanstart, chanend, osamp, cut=10, filt_thresh=0.45,
                    cupy_win_big=None, filt=None, verbose=False, precision="fp32", group_jobs=None,
                    window_bytes=STREAM_WINDOW_BYTES, **pipeline_kw):
    """helper_gpu_stream.py:62-254: re-PFB + cross-correlate ``nchunks`` iPFB
    chunks of ``acclen`` spectra read from baseband files.

    ``idxs[j]`` is antenna j's start offset (spectra past the first spectrum
    of ``files[j][0]``), ``files[j]`` its list of 4-bit ``.raw`` files;
    coarse channel indices ``[chanstart, chanend)`` of the files' channel
    list are selected, and the kept fine channels are
    ``[channels[chanstart]*osamp, channels[chanend]*osamp)`` in channel
    numbers (helper_gpu_stream.py:104-109).  Returns
    ``(vis, missing_fraction, freqs, win, filt)`` as the reference:
    ``vis`` a masked complex64 F-order ``(nant*npol, nant*npol, nchan, njobs)``
    array (input row ``ant*npol + pol``, helper_gpu_stream_clean.py:375),
    ``missing_fraction`` float64 ``(nant, njobs)`` percentages,
    ``freqs`` the kept fine channel numbers, ``win`` and ``filt`` the window
    and Wiener filter used (pass them back to the next call, as
    xcorr_gpu_list.py:70-82 does).

    The files are read through :class:`BasebandFileIterator` (packed form)
    straight into one stacked device buffer per input; if an antenna runs
    out of files, every antenna stops at that chunk (the reference's
    ``zip``) and the stream past it is zero.  The jobs go through in groups
    of ``group_jobs`` (default: as many as keep one input's time-stream
    window within ``window_bytes``), so device memory does not grow with the
    length of the run, as the reference's rolling ``pfb_buf``
    (helper_gpu_stream.py:88-89, 140-214).  ``precision`` selects the
    correlator arithmetic (``correlations.PRECISION_MODES``); other keywords
    go to :class:`RePFBPipeline`.
    """
    import time
    from . import baseband_data_classes as bdc
    t = _lib.torch()
    t0 = time.perf_counter()
    nant, npol, ntap = len(idxs), 2, 4
    its = [bdc.BasebandFileIterator(files[j], 0, idxs[j], acclen, nchunks=nchunks, chanstart=chanstart,
                                    chanend=chanend, type="packed") for j in range(nant)]
    obj = its[-1].obj
    if obj.bit_mode != 4:
        raise NotImplementedError("repfb_xcorr_avg reads 4-bit baseband files (bit mode %d)" % obj.bit_mode)
    channels = np.asarray(obj.channels, dtype=np.int64)
    nchan = int(obj.chanend - obj.chanstart) * osamp
    if obj.chanend >= len(channels):
        raise ValueError("chanend must index a channel of the file (the kept band ends at channels[chanend]*osamp)")
    repfb_chanstart = int(channels[obj.chanstart]) * osamp
    repfb_chanend = int(channels[obj.chanend]) * osamp
    if repfb_chanend - repfb_chanstart != nchan:
        raise ValueError("the selected coarse channels must have consecutive channel numbers")
    chan_numbers = channels[its[-1].channel_idxs]
    ncols = its[0].ncols
    start_specnums = [it.spec_num_start for it in its]
    dev = _lib.device()
    packed = [t.empty((nchunks, acclen, ncols), dtype=t.uint8, device=dev) for _ in range(nant * npol)]
    present = [[] for _ in range(nant)]
    nread = 0
    for c in range(nchunks):
        try:
            got = [it.read_into(packed[npol * j][c], packed[npol * j + 1][c]) for j, it in enumerate(its)]
        except StopIteration:
            break
        for j in range(nant):
            present[j].append(got[j])
        nread += 1
    t1 = time.perf_counter()
    pipe = RePFBPipeline(nant * npol, chan_numbers, acclen, cut, osamp, nblock, repfb_chanstart, repfb_chanend,
                         filt_thresh=filt_thresh, ntap=ntap, filt=filt, win=cupy_win_big, **pipeline_kw)
    rmaps = []
    for j in range(nant):
        rm = rowmaps_from_specnums(present[j], start_specnums[j], acclen) if nread else \
            t.full((nchunks, acclen), -1, dtype=t.int32, device=dev)
        rmaps += [rm] * npol
    G = group_jobs if group_jobs is not None else pipe.group_jobs_for(window_bytes)
    vis_dev = pipe.correlate(packed, rmaps, nchunks, precision=precision, nvalid=nread, group_jobs=G)
    jobs, _, _ = pipe.plan(nchunks)
    lblock = 4096 * osamp
    npres = [[len(s) for s in present[j]] for j in range(nant)]
    miss = missing_fraction(npres, acclen, nblock, lblock, 4096 * acclen, ntap, jobs)
    vis = np.asfortranarray(_lib.host(vis_dev))
    if verbose:
        print("repfb_xcorr_avg: %d antennas, %d of %d chunks read in %.3f s, %d jobs correlated in %.3f s"
              % (nant, nread, nchunks, t1 - t0, len(jobs), time.perf_counter() - t1))
    return (np.ma.masked_invalid(vis), np.asfortranarray(miss), np.arange(repfb_chanstart, repfb_chanend), pipe.win,
            pipe.filt)


def result_filename(init_t, acclen, osamp, nchunks, chanstart, chanend):
    """xcorr_gpu_list.py:90: name of one processed interval's result file."""
    return "stream_xcorr_all_ant_4bit_%s_%s_%s_%s_%s_%s.npz" % (init_t, acclen, osamp, nchunks, chanstart, chanend)


def write_result(fpath, vis, missing_fraction, chans):
    """xcorr_gpu_list.py:92: ``np.savez`` of ``data`` (the visibilities),
    ``mask`` (their invalid-value mask), ``missing_fraction`` and ``chans``
    (the fine channel numbers), as the reference's driver writes them."""
    v = np.ma.masked_invalid(vis) if not isinstance(vis, np.ma.MaskedArray) else vis
    np.savez(fpath, data=v.data, mask=v.mask, missing_fraction=np.asarray(missing_fraction), chans=np.asarray(chans))
    return fpath
