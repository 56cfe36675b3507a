"""Baseband file reader and chunk iterator -- the call surface of
src/correlations/baseband_data_classes.py.

A ``.raw`` file is a header followed by fixed-size packets
``[>u4 spec_num][bytes_per_packet - 4 payload bytes]``
(baseband_data_classes.py:156-214).  Here a file is parsed with one
structured read of the header and a memory map of the packet block; the
spec_num column stays on the host (int64), the payload goes to the GPU in one
copy (baseband_data_classes.py:217-218), and unpacking, repacking and the
spec_num zero fill run in libalbatros_hip.

``BasebandFileIterator`` is planned in spectrum-number space: step c covers
the window ``[S + c*acclen, S + (c+1)*acclen)`` of absolute spectrum numbers
(S = first spectrum of the start file + idxstart), and file k owns the
spectra after the last spectrum of file k-1.  A step's present spectra are
found with one ``searchsorted`` per file the window touches and unpacked
straight into the step's output rows.  This gives the reference's
per-step outputs (baseband_data_classes.py:522-715: present rows packed at
the top of zeroed ``(acclen, ncols)`` arrays, their spectrum numbers in
``specnums``, files advanced once a window reaches a file's last spectrum,
``StopIteration`` when the files run out, 2^32 wraps carried into later
files).  Where the reference's loop would step its cursor backwards (a start
index past the end of the start file, or a file wholly inside the previous
one), the window here stays fixed and such spectra are skipped.
"""
import os

import numpy as np

from . import _lib
from . import unpacking as unpk

WRAP = 1 << 32


def fill_arr(specnum, spec_per_packet):
    """baseband_data_classes.py:15-46: per-packet spec_num -> per-spectrum numbers."""
    s = np.asarray(specnum)
    return (s[:, None] + np.arange(spec_per_packet, dtype=s.dtype)).ravel()


fill_arr_cpu = fill_arr


def add_constant(arr, const):
    """baseband_data_classes.py:75-98 (in place)."""
    arr += const


def make_continuous_gpu(spec, chan_specnum, start_specnum, channels, nspec, nchans=2049, out=None):
    """baseband_data_classes.py:48-73: the spec_num zero fill.

    Rows present in ``spec`` (packed at the top) go to row
    ``(specnum - start) % nspec``, columns to channel numbers ``channels``;
    everything else is zero.  Returns ``spec`` itself when nothing is missing
    and all channels are present (as the reference does).
    """
    t = _lib.torch()
    rows = np.asarray(chan_specnum, dtype=np.int64)
    chans = np.asarray(channels, dtype=np.int64)
    if rows.shape[0] == nspec and len(chans) == nchans:
        if out is not None:
            out[:] = spec
            return out
        return spec
    isnp = _lib.is_numpy(spec)
    sp = _lib.to_device(spec, t.complex64)
    if out is None:
        o = t.empty((nspec, nchans), dtype=t.complex64, device=sp.device)
    else:
        o = out
    # a repeated channel number keeps its last column (numpy fancy assignment,
    # baseband_data_classes.py:72); earlier copies are marked -1 and skipped
    if len(chans) and len(np.unique(chans)) != len(chans):
        chans = chans.copy()
        _, last_rev = np.unique(chans[::-1], return_index=True)
        keep = np.zeros(len(chans), dtype=bool)
        keep[len(chans) - 1 - last_rev] = True
        chans[~keep] = -1
    # spectra landing on one row (a repeated spec_num, or spec_nums nspec
    # apart) keep the last one, as numpy's fancy assignment does: only the
    # last spectrum of each row is scattered
    slot = (rows - int(start_specnum)) % int(nspec)
    if len(slot) != len(np.unique(slot)):
        _, first_rev = np.unique(slot[::-1], return_index=True)
        keep = np.sort(len(slot) - 1 - first_rev)
        sp = sp[_lib.to_device(keep, t.int64)]
        rows = rows[keep]
    sn = _lib.to_device(rows if len(rows) else np.zeros(1, dtype=np.int64), t.int64)
    ch = _lib.to_device(chans if len(chans) else np.zeros(1, dtype=np.int64), t.int64)
    _lib.call("alb_make_continuous", _lib.ptr(sp), sp.shape[1] if sp.ndim == 2 else 1, _lib.ptr(sn), len(rows),
              int(start_specnum), int(nspec), _lib.ptr(ch), len(chans), _lib.ptr(o), int(nchans), _lib.stream())
    return _lib.host(o) if isnp and out is None else o


def get_rows_from_specnum(stidx, endidx, spec_arr):
    """baseband_data_classes.py:493-519: rows of the sorted ``spec_arr`` whose
    numbers lie in ``[stidx, endidx)``."""
    lo, hi = np.searchsorted(spec_arr, [stidx, endidx], side="left")
    return int(lo), int(hi)


# ------------------------------------------------------------------ header
_HEAD_FIELDS = [("bytes_per_packet", ">u8"), ("length_channels", ">u8"), ("spectra_per_packet", ">u8"),
                ("bit_mode", ">u8"), ("have_trimble", ">u8")]
_GPS_FIELDS = [("gps_week", ">u8"), ("gps_timestamp", ">u8"), ("gps_latitude", ">f8"), ("gps_longitude", ">f8"),
               ("gps_elevation", ">f8")]


def _read_header(f):
    """Parse the header at the start of open file ``f``
    (baseband_data_classes.py:156-174): a big-endian uint64 byte count n of
    the rest of the header, then five uint64 fields, (n - 80) / 8 uint64
    channel numbers, the GPS week and timestamp (uint64) and latitude,
    longitude, elevation (float64).  Returns (total header bytes, record)."""
    first = f.read(8)
    if len(first) != 8:
        raise ValueError("not a baseband file: no header")
    n = int(np.frombuffer(first, dtype=">u8")[0])
    nch = (n - 80) // 8
    if nch < 0 or 80 + 8 * nch != n:
        raise ValueError("not a baseband file: header length %d" % n)
    dt = np.dtype(_HEAD_FIELDS + [("channels", ">u8", (nch,))] + _GPS_FIELDS)
    body = f.read(n)
    if len(body) != n:
        raise ValueError("not a baseband file: truncated header")
    return 8 + n, np.frombuffer(body, dtype=dt, count=1)[0]


class Baseband:
    """baseband_data_classes.py:100-229: header and packets of one .raw file.

    Attributes as the reference: the header fields, ``channels`` (4-bit files
    list each channel's two pols, so every second number is kept and
    ``length_channels`` halves; 1-bit files list channel pairs, so each
    number c becomes c, c+1 and ``length_channels`` doubles), ``num_packets``,
    ``read_packets`` (``readlen`` >= 1: that many packets, in (0, 1): that
    fraction, 0: none, otherwise all), ``raw_data`` (device uint8
    ``[packets, bytes_per_packet - 4]``), ``spec_num`` (host int64; one 2^32
    wrap inside the file is undone, a second raises ValueError) and the lazy
    per-spectrum ``spec_idx``.
    """

    def __init__(self, file_name, readlen=-1):
        with open(file_name, "rb") as f:
            self.header_bytes, h = _read_header(f)
            size = os.fstat(f.fileno()).st_size
        for name, _ in _HEAD_FIELDS + _GPS_FIELDS:
            v = h[name]
            setattr(self, name, float(v) if name in ("gps_latitude", "gps_longitude", "gps_elevation") else int(v))
        ch = np.asarray(h["channels"], dtype=np.uint64)
        if self.bit_mode == 1:
            ch = np.stack([ch, ch + np.uint64(1)], axis=1).ravel()
            self.length_channels = int(self.length_channels * 2)
        elif self.bit_mode == 4:
            ch = ch[::2]
            self.length_channels = int(self.length_channels / 2)
        self.channels = ch
        self._overflowed = False
        self._spec_idx = None
        self.raw_data = None
        self.spec_num = None
        self.num_packets = (size - self.header_bytes) // self.bytes_per_packet
        if readlen >= 1:
            self.read_packets = int(readlen)
        elif 0 < readlen < 1:
            self.read_packets = int(self.num_packets * readlen)
        elif readlen == 0:
            self.read_packets = 0
        else:
            self.read_packets = -1
        if self.read_packets != 0:
            count = self.num_packets if self.read_packets < 0 else min(self.read_packets, self.num_packets)
            self._load_packets(file_name, count)

    def _load_packets(self, file_name, count):
        """Map the first ``count`` packets; spec_num to the host, payload to the GPU."""
        bpp = self.bytes_per_packet
        if count > 0:
            mm = np.memmap(file_name, dtype=np.uint8, mode="r", offset=self.header_bytes, shape=(count, bpp))
            spec = np.ascontiguousarray(mm[:, :4]).view(">u4").ravel().astype(np.int64)
            payload = np.ascontiguousarray(mm[:, 4:])
            del mm
        else:
            spec = np.zeros(0, dtype=np.int64)
            payload = np.zeros((0, bpp - 4), dtype=np.uint8)
        t = _lib.torch()
        self.raw_data = _lib.to_device(payload, t.uint8)
        down = np.flatnonzero(np.diff(spec) < 0)
        if len(down) > 1:
            raise ValueError("%s: spec_num decreases %d times (at most one 2^32 wrap per file)"
                             % (file_name, len(down)))
        if len(down) == 1:
            spec[down[0] + 1:] += WRAP
            self._overflowed = True
        self.spec_num = spec

    @property
    def spec_idx(self):
        """baseband_data_classes.py:231-244: per-spectrum numbers, with the
        missing-packet locations recorded as ``_missing_idx`` (packet before
        each jump), ``_missing_loc`` (first missing spectrum, relative to the
        file's first) and ``_missing_num`` (missing spectra)."""
        if self._spec_idx is None:
            spp = self.spectra_per_packet
            self._spec_idx = fill_arr(self.spec_num, spp)
            jump = np.diff(self.spec_num)
            at = np.flatnonzero(jump != spp)
            self._missing_idx = at
            self._missing_loc = (self.spec_num[at] + spp - self.spec_num[0]).astype(np.int64)
            self._missing_num = (jump[at] - spp).astype(np.int64)
        return self._spec_idx

    def _shift_spec_num(self, offset):
        """Add ``offset`` to every spectrum number (wraps of earlier files)."""
        if offset:
            self.spec_num += offset
            if self._spec_idx is not None:
                self._spec_idx += offset

    def print_header(self):
        print("Header Bytes = %s. Bytes per packet = %s. Channel length = %s. Spectra per packet: %s. Bit mode: %s. "
              "Total packets = %s. Read packets = %s. Have trimble = %s. Channels: %s GPS week = %s. GPS timestamp = "
              "%s. GPS latitude = %s. GPS longitude = %s. GPS elevation = %s." % (
                  self.header_bytes, self.bytes_per_packet, self.length_channels, self.spectra_per_packet,
                  self.bit_mode, self.num_packets, self.read_packets, self.have_trimble, self.channels, self.gps_week,
                  self.gps_timestamp, self.gps_latitude, self.gps_longitude, self.gps_elevation))

    def __str__(self):
        self.print_header()
        return ""

    def get_hist(self, mode=-1):
        """baseband_data_classes.py:287-305: raw-level histogram of every spectrum."""
        rowend = self.spec_num.shape[0] * self.spectra_per_packet
        return unpk.hist(self.raw_data, 0, rowend, self.length_channels, self.bit_mode, mode)

    def _assign_channels(self, channels=None, chanstart=None, chanend=None):
        """baseband_data_classes.py:307-342: the selected channel indices
        ``channel_idxs`` and the span ``[chanstart, chanend)`` (for a list,
        ``chanend = chanstart + count``); 1-bit files need an even, continuous
        span."""
        if isinstance(channels, np.ndarray):
            if channels.dtype != np.int64:
                raise ValueError("channels should be integers!")
            sel = np.sort(channels)
        elif isinstance(channels, (list, tuple)):
            sel = np.asarray(sorted(channels), dtype=np.int64)
        else:
            lo = 0 if chanstart is None else chanstart
            hi = self.length_channels if chanend is None else chanend
            sel = np.arange(lo, hi, dtype=np.int64)
            self.chanstart, self.chanend = lo, hi
        if channels is not None:
            self.chanstart = sel[0]
            self.chanend = sel[0] + len(sel)
        self.channel_idxs = sel
        if self.bit_mode == 1:
            if self.chanstart % 2 != 0 or self.chanend % 2 != 0:
                raise ValueError("Start and end channel indices must be divisible by 2 in 1-bit mode.")
            if np.any(np.diff(self.channel_idxs) != 1):
                raise ValueError("A continuous range of channels must be provided in 1-bit mode.")


def get_header(file_name, verbose=False):
    """baseband_data_classes.py:344-365: the public attributes of a header-only read."""
    obj = Baseband(file_name, readlen=0)
    if verbose:
        obj.print_header()
    return {k: v for k, v in obj.__dict__.items() if not k.startswith("_")}


class _Selected(Baseband):
    """Common constructor of the float and packed readers."""

    def __init__(self, file_name, readlen=-1, rowstart=None, rowend=None, channels=None, chanstart=0, chanend=None,
                 unpack=True):
        super().__init__(file_name, readlen)
        self._assign_channels(channels=channels, chanstart=chanstart, chanend=chanend)
        if unpack:
            # baseband_data_classes.py:414-418: a falsy rowstart or rowend means all rows
            if rowstart and rowend:
                self.pol0, self.pol1 = self._unpack(rowstart, rowend)
            else:
                self.pol0, self.pol1 = self._unpack(0, len(self.spec_idx))

    def _unpack(self, rowstart, rowend, out=None):
        raise NotImplementedError

    def _columns(self):
        """Output columns per spectrum of ``_unpack``."""
        raise NotImplementedError


class BasebandFloat(_Selected):
    """baseband_data_classes.py:368-433: complex64 pols of the selected channels."""

    def _unpack(self, rowstart, rowend, out=None):
        if self.bit_mode == 4:
            return unpk.unpack_4bit(self.raw_data, rowstart, rowend, self.channel_idxs, self.length_channels, out=out)
        if self.bit_mode == 1:
            return unpk.unpack_1bit(self.raw_data, rowstart, rowend, self.channel_idxs, self.length_channels, out=out)
        raise ValueError("bit mode %d not supported" % self.bit_mode)

    def _columns(self):
        return len(self.channel_idxs)


class BasebandPacked(_Selected):
    """baseband_data_classes.py:435-490: per-pol packed bytes (sortpols layout)."""

    def _unpack(self, rowstart, rowend, out=None):
        return unpk.sortpols(self.raw_data, self.length_channels, self.bit_mode, rowstart, rowend, self.chanstart,
                             self.chanend, out=out)

    def _columns(self):
        if self.bit_mode == 1:
            return int(np.ceil((self.chanend - self.chanstart) / 4))
        return int(self.chanend - self.chanstart)


class BasebandFileIterator:
    """baseband_data_classes.py:522-715: ``acclen`` spectra per step across
    files, gaps and 2^32 wraps.

    Each step returns ``{"pol0", "pol1", "specnums"}``: zeroed device arrays
    ``(acclen, ncols)`` whose first ``len(specnums)`` rows hold the present
    spectra of the step's window in order, and their absolute spectrum
    numbers (int64).  ``spec_num_start`` is the first spectrum of the next
    window.  :meth:`read_into` fills caller buffers instead (the streaming
    re-PFB reads every chunk of an input into one stacked buffer).
    """

    def __init__(self, file_paths, fileidx, idxstart, acclen, nchunks=None, channels=None, chanstart=0, chanend=None,
                 type="packed"):
        if type == "float":
            self._reader, self.dtype = BasebandFloat, "complex64"
        elif type == "packed":
            self._reader, self.dtype = BasebandPacked, "uint8"
        else:
            raise ValueError("type must be 'float' or 'packed'")
        self.type = type
        self.acclen = int(acclen)
        self.file_paths = file_paths
        self.fileidx = fileidx
        self.nchunks = nchunks
        self.chunksread = 0
        self._wraps = 0           # files with an internal 2^32 wrap opened so far
        self._done = False
        self.obj = self._open(fileidx, channels=channels, chanstart=chanstart, chanend=chanend)
        self.channel_idxs = self.obj.channel_idxs
        self.spec_num_start = int(idxstart + self.obj.spec_idx[0])
        # first spectrum the current file owns (the start file: all from the first window)
        self._owned_from = self.spec_num_start
        if self.obj.bit_mode == 1 and type == "packed" and self.obj.chanstart % 2:
            raise ValueError("ERROR: Start channel index must be even.")
        self.ncols = self.obj._columns()

    @property
    def _OVERFLOW_CTR(self):
        """Wraps carried into later files (the reference's counter name)."""
        return self._wraps

    def _open(self, k, **sel):
        """Reader of file k, spectrum numbers shifted by the wraps of the
        files opened before it (baseband_data_classes.py:601-608)."""
        obj = self._reader(self.file_paths[k], unpack=False, **sel)
        obj._shift_spec_num(self._wraps * WRAP)
        if obj._overflowed:
            self._wraps += 1
        return obj

    def _plan(self, lo, hi):
        """Pieces (reader, rowstart, rowend) holding the present spectra of the
        window ``[lo, hi)``, in order.  A window that reaches the current
        file's last spectrum opens the next file (StopIteration when there is
        none, as the reference raises)."""
        pieces = []
        while True:
            obj = self.obj
            idx = obj.spec_idx
            last = int(idx[-1])
            r0, r1 = np.searchsorted(idx, [max(lo, self._owned_from), min(hi, last + 1)], side="left")
            if r1 > r0:
                pieces.append((obj, int(r0), int(r1)))
            if hi <= last:
                return pieces
            self.fileidx += 1
            if self.fileidx >= len(self.file_paths):
                self._done = True
                raise StopIteration("BFI ran out of files")
            self.obj = self._open(self.fileidx, channels=self.channel_idxs)
            self._owned_from = last + 1
            if hi <= last + 1:
                return pieces

    def _step(self):
        if self._done or (self.nchunks and self.chunksread == self.nchunks):
            raise StopIteration("%s chunks read" % self.chunksread)
        lo = self.spec_num_start
        hi = lo + self.acclen
        try:
            pieces = self._plan(lo, hi)
        finally:
            self.spec_num_start = hi
        self.chunksread += 1
        return pieces

    def read_into(self, pol0, pol1):
        """Read the next step into device views ``pol0``, ``pol1``
        (``(acclen, ncols)``, row-contiguous): present rows first, the rest
        zeroed.  Returns the step's ``specnums``."""
        pieces = self._step()
        i = 0
        spec = []
        for obj, r0, r1 in pieces:
            n = r1 - r0
            obj._unpack(r0, r1, out=(pol0[i:i + n], pol1[i:i + n]))
            spec.append(obj.spec_idx[r0:r1])
            i += n
        if i < self.acclen:
            tail0, tail1 = pol0[i:], pol1[i:]
            s = _lib.stream()
            _lib.call("alb_fill_zero", _lib.ptr(tail0), tail0.numel() * tail0.element_size(), s)
            _lib.call("alb_fill_zero", _lib.ptr(tail1), tail1.numel() * tail1.element_size(), s)
        return np.concatenate(spec) if spec else np.zeros(0, dtype=np.int64)

    def __iter__(self):
        return self

    def __next__(self):
        t = _lib.torch()
        tdt = t.complex64 if self.dtype == "complex64" else t.uint8
        pol0 = t.empty((self.acclen, self.ncols), dtype=tdt, device=_lib.device())
        pol1 = t.empty((self.acclen, self.ncols), dtype=tdt, device=_lib.device())
        specnums = self.read_into(pol0, pol1)
        return {"pol0": pol0, "pol1": pol1, "specnums": specnums}
