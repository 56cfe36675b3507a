"""Baseband file reader and chunk iterator -- the call surface of
src/correlations/baseband_data_classes.py.

Headers and spec_num stay on the host (int64 numpy); raw packet bytes go to
the GPU as a torch uint8 tensor (baseband_data_classes.py:217-218); unpacking,
repacking and the spec_num zero fill run in libalbatros_hip.
"""
import os
import struct
import time

import numpy as np

from . import _lib
from . import unpacking as unpk


def fill_arr(specnum, spec_per_packet):
    """baseband_data_classes.py:15-46: expand per-packet spec_num to per spectrum."""
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
        seen = set()
        for k in range(len(chans) - 1, -1, -1):
            if chans[k] in seen:
                chans[k] = -1
            else:
                seen.add(chans[k])
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
    """baseband_data_classes.py:493-519: rows of ``spec_arr`` in [stidx, endidx)."""
    l = np.searchsorted(spec_arr, stidx, side="left")
    r = np.searchsorted(spec_arr, endidx, side="left")
    return l, r


class Baseband:
    """baseband_data_classes.py:100-229: header + packets of one .raw file.

    Header: big-endian uint64 header_bytes, bytes_per_packet, length_channels,
    spectra_per_packet, bit_mode, have_trimble, channels[...], gps_week,
    gps_timestamp, then float64 latitude, longitude, elevation.  Packets are
    ``[>u4 spec_num][bytes_per_packet-4 bytes]``; a single 2^32 wrap inside a
    file is undone, two raise ValueError.
    """

    def __init__(self, file_name, readlen=-1):
        with open(file_name, "rb") as f:
            header_bytes = struct.unpack(">Q", f.read(8))[0]
            self.header_bytes = 8 + header_bytes
            self.bytes_per_packet = struct.unpack(">Q", f.read(8))[0]
            self.length_channels = struct.unpack(">Q", f.read(8))[0]
            self.spectra_per_packet = struct.unpack(">Q", f.read(8))[0]
            self.bit_mode = struct.unpack(">Q", f.read(8))[0]
            self.have_trimble = struct.unpack(">Q", f.read(8))[0]
            nch = int((header_bytes - 8 * 10) / 8)
            self.channels = np.frombuffer(f.read(self.header_bytes - 88), ">%dQ" % nch)[0]
            self.gps_week = struct.unpack(">Q", f.read(8))[0]
            self.gps_timestamp = struct.unpack(">Q", f.read(8))[0]
            self.gps_latitude = struct.unpack(">d", f.read(8))[0]
            self.gps_longitude = struct.unpack(">d", f.read(8))[0]
            self.gps_elevation = struct.unpack(">d", f.read(8))[0]
            self._overflowed = False
            self._spec_idx = None
            self.raw_data = None
            self.spec_num = None
            if self.bit_mode == 1:
                self.channels = np.ravel(np.column_stack((self.channels, self.channels + 1)))
                self.length_channels = int(self.length_channels * 2)
            if self.bit_mode == 4:
                self.channels = self.channels[::2]
                self.length_channels = int(self.length_channels / 2)
            self.num_packets = (os.fstat(f.fileno()).st_size - self.header_bytes) // self.bytes_per_packet
            if readlen >= 1:
                self.read_packets = int(readlen)
            elif 0 < readlen < 1:
                self.read_packets = int(self.num_packets * readlen)
            elif readlen == 0:
                self.read_packets = 0
            else:
                self.read_packets = -1
            if self.read_packets != 0:
                f.seek(self.header_bytes)
                data = np.fromfile(f, count=self.read_packets,
                                   dtype=[("spec_num", ">I"), ("spectra", "%dB" % (self.bytes_per_packet - 4))])
                t = _lib.torch()
                self.raw_data = _lib.to_device(np.ascontiguousarray(data["spectra"]).reshape(len(data), -1), t.uint8)
                self.spec_num = np.array(data["spec_num"], dtype="int64", order="c")
                self._wrap_loc = np.where(np.diff(self.spec_num) < 0)[0]
                if len(self._wrap_loc) == 1:
                    self.spec_num[self._wrap_loc[0] + 1:] += 2 ** 32
                    self._overflowed = True
                elif len(self._wrap_loc) > 1:
                    raise ValueError("Why are there two -ve diffs in specnum? Investigate this file")

    @property
    def spec_idx(self):
        """baseband_data_classes.py:231-244 (missing-packet locations recorded too)."""
        if self._spec_idx is None:
            self._spec_idx = fill_arr(self.spec_num, self.spectra_per_packet)
            specdiff = np.diff(self.spec_num)
            idx = np.where(specdiff != self.spectra_per_packet)[0]
            self._missing_idx = idx
            self._missing_loc = (self.spec_num[idx] + self.spectra_per_packet - self.spec_num[0]).astype("int64")
            self._missing_num = (specdiff[idx] - self.spectra_per_packet).astype("int64")
        return self._spec_idx

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
        """baseband_data_classes.py:287-305."""
        rowend = self.spec_num.shape[0] * self.spectra_per_packet
        return unpk.hist(self.raw_data, 0, rowend, self.length_channels, self.bit_mode, mode)

    def _assign_channels(self, channels=None, chanstart=None, chanend=None):
        """baseband_data_classes.py:307-342."""
        if isinstance(channels, np.ndarray):
            channels.sort()
            if channels.dtype == "int64":
                self.channel_idxs = channels
                self.chanstart = channels[0]
                self.chanend = len(channels) + channels[0]
            else:
                raise ValueError("channels should be integers!")
        elif isinstance(channels, (list, tuple)):
            channels = sorted(channels)
            self.channel_idxs = np.asarray(channels, dtype="int64")
            self.chanstart = channels[0]
            self.chanend = len(channels) + channels[0]
        else:
            self.chanstart = 0 if chanstart is None else chanstart
            self.chanend = self.length_channels if chanend is None else chanend
            self.channel_idxs = np.arange(self.chanstart, self.chanend, dtype="int64")
        if self.bit_mode == 1:
            if self.chanstart % 2 != 0 or self.chanend % 2 != 0:
                raise ValueError("Start and end channel indices must be divisible by 2 in 1-bit mode.")
            if np.any(np.diff(self.channel_idxs) - 1):
                raise ValueError("A continuous range of channels must be provided in 1-bit mode.")


def get_header(file_name, verbose=False):
    """baseband_data_classes.py:344-365."""
    obj = Baseband(file_name, readlen=0)
    if verbose:
        obj.print_header()
    return {k: v for k, v in obj.__dict__.items() if not k.startswith("_")}


class BasebandFloat(Baseband):
    """baseband_data_classes.py:368-433: complex64 unpacked pols."""

    def __init__(self, file_name, readlen=-1, rowstart=None, rowend=None, channels=None, chanstart=0, chanend=None,
                 unpack=True):
        super().__init__(file_name, readlen)
        self._assign_channels(channels=channels, chanstart=chanstart, chanend=chanend)
        if unpack:
            if rowstart and rowend:
                self.pol0, self.pol1 = self._unpack(rowstart, rowend)
            else:
                self.pol0, self.pol1 = self._unpack(0, len(self.spec_idx))

    def _unpack(self, rowstart, rowend):
        if self.bit_mode == 4:
            return unpk.unpack_4bit(self.raw_data, rowstart, rowend, self.channel_idxs, self.length_channels)
        elif self.bit_mode == 1:
            return unpk.unpack_1bit(self.raw_data, rowstart, rowend, self.channel_idxs, self.length_channels)


class BasebandPacked(Baseband):
    """baseband_data_classes.py:435-490: per-pol packed bytes (sortpols)."""

    def __init__(self, file_name, readlen=-1, rowstart=None, rowend=None, channels=None, chanstart=0, chanend=None,
                 unpack=True):
        super().__init__(file_name, readlen)
        self._assign_channels(channels=channels, chanstart=chanstart, chanend=chanend)
        if unpack:
            if rowstart and rowend:
                self.pol0, self.pol1 = self._unpack(rowstart, rowend)
            else:
                self.pol0, self.pol1 = self._unpack(0, len(self.spec_idx))

    def _unpack(self, rowstart, rowend):
        return unpk.sortpols(self.raw_data, self.length_channels, self.bit_mode, rowstart, rowend, self.chanstart,
                             self.chanend)


class BasebandFileIterator:
    """baseband_data_classes.py:522-715: ``acclen`` spectra per step across
    files, gaps and 2^32 wraps; present rows packed at the top of zeroed
    ``(acclen, ncols)`` device arrays, ``specnums`` = their spectrum numbers."""

    def __init__(self, file_paths, fileidx, idxstart, acclen, nchunks=None, channels=None, chanstart=0, chanend=None,
                 type="packed"):
        self._OVERFLOW_CTR = 0
        self.acclen = acclen
        self.file_paths = file_paths
        self.fileidx = fileidx
        self.nchunks = nchunks
        self.chunksread = 0
        self.type = type
        self.file_loader = self.get_file_loader()
        self.obj = self.file_loader(file_paths[fileidx], channels=channels, chanstart=chanstart, chanend=chanend,
                                    unpack=False)
        self.channel_idxs = self.obj.channel_idxs
        self.spec_num_start = idxstart + self.obj.spec_idx[0]
        if self.obj.bit_mode == 1 and self.type == "packed":
            if self.obj.chanstart % 2 > 0:
                raise ValueError("ERROR: Start channel index must be even.")
            self.ncols = int(np.ceil((self.obj.chanend - self.obj.chanstart) / 4))
        else:
            self.ncols = len(self.channel_idxs)

    def get_file_loader(self):
        if self.type == "float":
            myclass = BasebandFloat
            self.dtype = "complex64"
        elif self.type == "packed":
            myclass = BasebandPacked
            self.dtype = "uint8"
        else:
            raise ValueError("type must be 'float' or 'packed'")

        def file_loader(*args, **kwargs):
            obj = myclass(*args, **kwargs)
            if self._OVERFLOW_CTR > 0:
                add_constant(obj.spec_num, self._OVERFLOW_CTR * 2 ** 32)
            if obj._overflowed:
                self._OVERFLOW_CTR += 1
            return obj
        return file_loader

    def __iter__(self):
        return self

    def __next__(self):
        t = _lib.torch()
        if self.nchunks and self.chunksread == self.nchunks:
            raise StopIteration(f"{self.nchunks} chunks read!")
        tdt = t.complex64 if self.dtype == "complex64" else t.uint8
        pol0 = t.zeros((self.acclen, self.ncols), dtype=tdt, device=_lib.device())
        pol1 = t.zeros((self.acclen, self.ncols), dtype=tdt, device=_lib.device())
        specnums = np.array([], dtype="int64")
        rem = self.acclen
        i = 0
        while rem:
            if self.spec_num_start < self.obj.spec_num[0]:
                step = min(self.obj.spec_num[0] - self.spec_num_start, rem)
                rem -= step
                self.spec_num_start += step
            else:
                l = self.obj.spec_idx[-1] - self.spec_num_start + 1
                if rem >= l:
                    rowstart, rowend = get_rows_from_specnum(self.spec_num_start, self.spec_num_start + l,
                                                             self.obj.spec_idx)
                    specnums = np.append(specnums, self.obj.spec_idx[rowstart:rowend])
                    rem -= l
                    if rowend > rowstart:
                        a, b = self.obj._unpack(rowstart, rowend)
                        pol0[i:i + rowend - rowstart] = a
                        pol1[i:i + rowend - rowstart] = b
                    i += rowend - rowstart
                    self.spec_num_start += l
                    self.fileidx += 1
                    if len(self.file_paths) == self.fileidx:
                        raise StopIteration("BFI Ran out of files!")
                    self.obj = self.file_loader(self.file_paths[self.fileidx], channels=self.channel_idxs,
                                                unpack=False)
                else:
                    rowstart, rowend = get_rows_from_specnum(self.spec_num_start, self.spec_num_start + rem,
                                                             self.obj.spec_idx)
                    specnums = np.append(specnums, self.obj.spec_idx[rowstart:rowend])
                    if rowend > rowstart:
                        a, b = self.obj._unpack(rowstart, rowend)
                        pol0[i:i + rowend - rowstart] = a
                        pol1[i:i + rowend - rowstart] = b
                    self.spec_num_start += rem
                    rem = 0
                    i += rowend - rowstart
        self.chunksread += 1
        return {"pol0": pol0, "pol1": pol1, "specnums": specnums}
