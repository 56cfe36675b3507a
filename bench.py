#!/usr/bin/env python3
"""Benchmark of the ALBATROS hot path on one MI355X.

Headline workload (one "step"): the batched streaming pass -- 2 antennas x 2
pols of 4-bit baseband, 300 coarse channels, 64 arrival chunks of uneven
length (2^10..2^14 spectra) with runs of lost spectra, re-blocked by
spec_num into 4088-spectrum iPFB tiles -> unpack + spec_num zero fill ->
Wiener inverse PFB (thresh 0.45) -> re-PFB 64x (ntap 4, 19200 fine channels)
-> 4x4 visibilities per 1024-spectrum job, as one hipGraph replay that also
builds the row map from the spec_nums.  Inputs are resident in HBM before the
timed region.  value = ADC-equivalent samples per second (inputs x coarse
spectra x 4096 / s / 1e6).

Also reported: the same workload in the plain-FP32 per-chunk launch form and
the speedup over it, live per-kernel HIP-event times, the roofline of the
dominant kernel, the pipeline-level roofline, a CPU baseline (the reference
CPU path restated, timed on a bounded sample of the same stream, all cores
and one core) and secondary workloads (the single-input rechannelizer with
its own roofline and CPU baseline, 4-bit / 1-bit unpack and sortpols, the
direct 4-bit and 1-bit correlators with the reference's own OpenMP C timed
beside them, the precision correlator).
"""
import argparse
import ctypes
import itertools
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP32_MFMA_TFLOPS = 157.3
HERE = os.path.dirname(os.path.abspath(__file__))
METRIC = BASE_METRIC = "Msamples/s unpack→iPFB→PFB→xcorr, % HBM/MFMA roofline; vs plain FP32"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE /
    WRITE_SIZE, scaled by the copy-kernel calibration), kept as key=value
    text under profiles/.  None when not measured."""
    p = os.path.join(HERE, "profiles", "traffic_%s.txt" % name)
    if not os.path.exists(p):
        return None
    kv = {}
    for line in open(p):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            kv[k.strip()] = v.strip()
    try:
        return float(kv["bytes_per_launch"])
    except (KeyError, ValueError):
        return None


class Timer:
    """HIP events around each launch on torch's current stream (the stream
    every entry point is given)."""

    def __init__(self, torch):
        self.t = torch
        self.ev = {}
        self.on = False

    def __call__(self, name, fn):
        if not self.on:
            fn()
            return
        a = self.t.cuda.Event(enable_timing=True)
        b = self.t.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        self.ev.setdefault(name, []).append((a, b))

    def summary(self):
        self.t.cuda.synchronize()
        return {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in self.ev.items()}

    def counts(self):
        return {k: len(v) for k, v in self.ev.items()}


def rechannelizer(args, torch):
    from albatros_axion_detection_amd import _lib, pfb, repfb, synth
    acclen, cut, osamp, nchunks = 4088, 4, 64, args.nchunks
    chan0, nkeep = 64, 100
    n_coarse = acclen * nchunks
    log("[bench] generating rechannelizer input: %d coarse spectra" % n_coarse)
    packed_np, chans = synth.rechannelizer_input(n_coarse=n_coarse, chan0=chan0, nkeep=nkeep, seed=1)
    nslice = acclen + 2 * cut
    nblk = (n_coarse // osamp) - 3  # every complete re-PFB block of the stream
    c_lo, c_hi = chan0 * osamp, (chan0 + nkeep) * osamp
    pipe = repfb.RePFBPipeline(1, chans, acclen, cut, osamp, nblk, c_lo, c_hi, filt_thresh=0.45,
                               tile_batch=nchunks, row_pass=args.row_pass, ipfb_form=args.ipfb_form)
    ldy = 8 if pipe.pow2 else 4096
    dev = _lib.device()
    packed = _lib.to_device(packed_np.reshape(nchunks, acclen, nkeep))
    # spec_num of every present spectrum of the nchunks chunks, in arrival order
    specnums = _lib.to_device(np.arange(n_coarse, dtype=np.int64))
    rowmap = torch.empty((nchunks, acclen), dtype=torch.int32, device=dev)
    rows = max((nblk + 3) * osamp, n_coarse)
    stream = torch.zeros((rows, 4096), dtype=torch.float32, device=dev)
    # tiles: the row pass's blocked tiles, or the spectral form's strips
    # ([1024][nslice/2+1][4] complex64 per tile); chan_t: the spectral form's
    # channel transforms along time
    tiles = torch.empty(nchunks * max(nslice * 4096, 1024 * (nslice // 2 + 1) * 8), dtype=torch.float32, device=dev)
    spectral = pipe.ipfb_form == "spectral" and not args.unfused
    chan_t = torch.empty((nchunks, nslice, nkeep), dtype=torch.complex64, device=dev) if spectral else None
    ws = torch.empty(int(_lib.load().alb_repfb_workspace_bytes(nblk, osamp)) // 8, dtype=torch.complex64,
                     device=dev)
    out = torch.empty((nblk, c_hi - c_lo), dtype=torch.complex64, device=dev)
    timer = Timer(torch)
    s = _lib.stream

    fused = not args.unfused
    if fused and not pipe.fused_supported(nblk):
        raise SystemExit("fused column pass unavailable for this shape")
    edges = pipe.edge_blocks(nchunks, nblk)
    edges_p = edges.ctypes.data_as(ctypes.c_void_p)

    def step():
        # the chunks are consecutive acclen-spectrum windows of one spec_num
        # range, so one call maps all of them: row = spec_num - start
        timer("rowmap", lambda: _lib.call(
            "alb_build_rowmap", _lib.ptr(specnums), n_coarse, 0, n_coarse, 0, _lib.ptr(rowmap), s()))
        if spectral:
            timer("ipfb_chan_fft", lambda: _lib.call(
                "alb_ipfb_chan_fft_packed4", _lib.ptr(packed), _lib.ptr(rowmap), _lib.ptr(pipe.chan), nkeep, acclen,
                cut, 0, nchunks, _lib.ptr(chan_t), s()))
            timer("ipfb_rows_spectral", lambda: _lib.call(
                "alb_ipfb_rows_spectral", _lib.ptr(chan_t), _lib.ptr(pipe.chan), nkeep, nslice, nchunks,
                _lib.ptr(tiles), s()))
        elif pipe.row_pass == "mfma":
            timer("ipfb_rows_mfma", lambda: _lib.call(
                "alb_ipfb_rows_packed4_mfma", _lib.ptr(packed), _lib.ptr(rowmap), _lib.ptr(pipe.chan), nkeep, acclen,
                cut, 0, nchunks, _lib.ptr(pipe.coeff), _lib.ptr(tiles), s()))
        else:
            timer("ipfb_rows_packed4", lambda: _lib.call(
                "alb_ipfb_rows_packed4_blocked" if pipe.pow2 else "alb_ipfb_rows_packed4", _lib.ptr(packed),
                _lib.ptr(rowmap), _lib.ptr(pipe.chan), nkeep, acclen, cut, 0, nchunks, _lib.ptr(tiles), s()))
        if spectral:
            timer("ipfb_columns_repfb", lambda: _lib.call(
                "alb_ipfb_columns_repfb_spectral", _lib.ptr(tiles), nchunks, nslice, _lib.ptr(pipe.filt),
                pipe.filt.shape[1], cut, _lib.ptr(stream), 0, acclen, _lib.ptr(pipe.win), 4, osamp, nblk,
                _lib.ptr(ws), s()))
        if fused and not spectral:
            timer("ipfb_columns_repfb", lambda: _lib.call(
                "alb_ipfb_columns_repfb", _lib.ptr(tiles), ldy, nslice * 4096, nchunks, nslice, _lib.ptr(pipe.filt),
                pipe.filt.shape[1], cut, _lib.ptr(stream), 0, acclen, _lib.ptr(pipe.win), 4, osamp, nblk,
                _lib.ptr(ws), s()))
        if fused:
            timer("repfb_stage1_edges", lambda: _lib.call(
                "alb_repfb_stage1_list", _lib.ptr(stream), _lib.ptr(pipe.win), 4, edges_p, len(edges), osamp,
                _lib.ptr(ws), s()))
        else:
            timer("ipfb_columns", lambda: _lib.call(
                "alb_ipfb_columns", _lib.ptr(tiles), ldy, nslice * 4096, nchunks, nslice, 4096, _lib.ptr(pipe.filt),
                pipe.filt.shape[1], cut, _lib.ptr(stream), 4096, acclen * 4096, s()))
            timer("repfb_stage1", lambda: _lib.call(
                "alb_repfb_stage1", _lib.ptr(stream), _lib.ptr(pipe.win), 4, nblk, osamp, _lib.ptr(ws), s()))
        timer("repfb_stage2", lambda: _lib.call(
            "alb_repfb_stage2", _lib.ptr(ws), nblk, osamp, c_lo, c_hi, _lib.ptr(out), c_hi - c_lo, 1, nblk, 0, s()))

    # Launch shapes: the library defaults (column-pass strip width 4, XCD
    # group 2 and row-pass group 2 at this tile count), the shapes the committed
    # rocprofv3 summaries were taken with, unless a flag sets one.  --sweep
    # times every strip width x row group first (untimed) and keeps the
    # fastest; they lie within a few per cent of each other.
    fixed = (("ipfb_cols_width", args.cols_width), ("ipfb_cols_group", args.cols_group),
             ("ipfb_rows_group", args.rows_group))
    for key, v in fixed:
        _lib.call("alb_tune_set", key.encode(), v)
    sweep = {}
    if args.sweep:
        knobs = (("ipfb_cols_width", "w", [args.cols_width] if args.cols_width else [8, 4, 2]),
                 ("ipfb_rows_group", "r", [args.rows_group] if args.rows_group else [4, 2]))
        combos = {}
        for vals in itertools.product(*(k[2] for k in knobs)):
            for (key, _, _), v in zip(knobs, vals):
                _lib.call("alb_tune_set", key.encode(), v)
            step()
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(3):
                step()
            b.record()
            torch.cuda.synchronize()
            name = "_".join("%s%d" % (k[1], v) for k, v in zip(knobs, vals))
            sweep[name] = a.elapsed_time(b) / 3
            combos[name] = vals
        best = min(sweep, key=sweep.get)
        for (key, _, _), v in zip(knobs, combos[best]):
            _lib.call("alb_tune_set", key.encode(), v)
        log("[bench] launch-shape sweep (ms/step): %s -> %s" % (
            json.dumps({k: round(v, 4) for k, v in sweep.items()}), best))
    lib = _lib.load()
    tuned = {key: int(lib.alb_tune_get(key.encode())) for key, _ in fixed}   # 0: library default

    # Clock settle (untimed, before the warm-up): a GPU that has been idle
    # runs its first ~10 ms of work at lower clocks, so the step is repeated
    # for about 100 ms first (the launch-shape sweep, when asked for, does
    # the same).
    t_settle = time.perf_counter()
    while time.perf_counter() - t_settle < 0.1:
        for _ in range(20):
            step()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region 1: K steps back to back (the throughput figure)
    t0e = torch.cuda.Event(enable_timing=True)
    t1e = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t0e.record()
    for _ in range(args.steps):
        step()
    t1e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0e.elapsed_time(t1e) / args.steps
    # timed region 2: the same K steps with a HIP event pair around every
    # launch on its stream (per-kernel durations; the events add gaps between
    # launches, so the step time above is taken without them)
    timer.on = True
    for _ in range(args.steps):
        step()
    kern = timer.summary()
    timer.on = False
    log("[bench] rechannelizer: %.3f ms/step (wall %.3f ms/step); kernels %s" % (
        ms, 1e3 * wall / args.steps, json.dumps({k: round(v, 4) for k, v in kern.items()})))

    # algorithmic bytes per launch of each kernel: the per-unit figures of the
    # rechannelizer (per coarse spectrum: raw 100 B in, time stream 16384 B
    # written and 16384 B read, 800 B of fine channels out) times the units
    # the launch covers; the fused column pass produces and consumes the time
    # stream, so it carries both stream terms.  The edge blocks' re-reads of
    # the stream are design overhead, counted as what that launch reads.
    alg = {
        "ipfb_rows_packed4": n_coarse * nkeep,                          # raw bytes in
        "ipfb_rows_mfma": n_coarse * nkeep,                             # raw bytes in
        "ipfb_chan_fft": n_coarse * nkeep,                              # raw bytes in
        "ipfb_columns": n_coarse * 4096 * 4,                            # time stream out
        "repfb_stage1": (nblk + 3) * osamp * 4096 * 4,                  # time stream in
        "ipfb_columns_repfb": 2 * n_coarse * 4096 * 4,                  # time stream out + in
        "repfb_stage1_edges": len(edges) * osamp * 4 * 4096 * 4,        # stream rows of the edge blocks
        "repfb_stage2": nblk * (c_hi - c_lo) * 8,                       # fine channels out
    }
    dom = max((k for k in kern if k in alg), key=lambda k: kern[k])
    achieved = alg[dom] / (kern[dom] * 1e-3) / 1e9
    traffic = read_traffic(dom + "_rechannelizer")
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": alg[dom], "avg_launch_ms": round(kern[dom], 5)}
    pipe_bytes = 33668 * n_coarse
    pipe_roof = {"bound": "hbm", "algorithmic_bytes_per_step": pipe_bytes,
                 "achieved": round(pipe_bytes / (ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(pipe_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    value = n_coarse * 4096 / (ms * 1e-3) / 1e6
    res = {
        "value": value, "ms_per_step": ms, "kernels_ms": {k: round(v, 5) for k, v in kern.items()},
        "roofline": roof, "pipeline_roofline": pipe_roof,
        "config": {"workload": "rechannelizer: 1 input, %d coarse 4-bit spectra (%d chunks x %d, nslice %d, cut %d), "
                               "coarse ch %d..%d -> %d fine ch, osamp %d, ntap 4, Wiener thresh 0.45, %d re-PFB "
                               "spectra" % (n_coarse, nchunks, acclen, nslice, cut, chan0, chan0 + nkeep - 1,
                                            c_hi - c_lo, osamp, nblk),
                   "n_coarse": n_coarse, "acclen": acclen, "cut": cut, "osamp": osamp, "nchan_fine": c_hi - c_lo,
                   "nblock": nblk, "ipfb_form": "spectral" if spectral else "direct"},
        "ref_unit_msps": n_coarse * 2049 * 8 / (ms * 1e-3) / 1e6,
        "tuning": dict(tuned, sweep_ms_per_step={k: round(v, 4) for k, v in sweep.items()}),
    }
    ctx = dict(packed_np=packed_np, acclen=acclen, cut=cut, osamp=osamp, chan0=chan0, nkeep=nkeep,
               filt=pipe.filt, win=pipe.win)
    return res, ctx


def cpu_baseline(ctx, budget_s):
    """The reference CPU rechannelizer (pfb_cpu_utils.py cpu_ipfb / cpu_pfb with
    scipy.fft workers = cores) on one chunk of the same workload."""
    from albatros_axion_detection_amd import _lib
    from oracle import cpu_port
    acclen, cut, osamp = ctx["acclen"], ctx["cut"], ctx["osamp"]
    nkeep, chan0 = ctx["nkeep"], ctx["chan0"]
    sample_chunks = 1
    raw = ctx["packed_np"][: sample_chunks * acclen]
    filt = _lib.host(ctx["filt"]).astype(np.complex64)
    win = _lib.host(ctx["win"]).astype(np.float32)
    cores = cpu_port.ncores()

    def run():
        hi = (raw >> 4).astype(np.int8)
        lo = (raw & 15).astype(np.int8)
        hi = np.where(hi > 8, hi - 16, hi)
        lo = np.where(lo > 8, lo - 16, lo)
        full = np.zeros((sample_chunks * acclen, 2049), np.complex64)  # make_continuous zero fill
        full[:, chan0:chan0 + nkeep] = hi + 1j * lo
        chunks = [full[c * acclen:(c + 1) * acclen] for c in range(sample_chunks)]
        return cpu_port.rechannelize(chunks, filt, win, acclen, cut, osamp, chan0 * osamp,
                                     (chan0 + nkeep) * osamp)

    run()
    times = []
    t_all = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_all) < budget_s:
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    nspec = sample_chunks * acclen
    return {"value": nspec * 4096 / t / 1e6, "unit": "Msamples/s", "cores": cores, "kind": "port",
            "sample": "1 chunk of the same workload (%d coarse spectra): unpack + zero fill, cpu_ipfb, cpu_pfb "
                      "(scipy.fft float32, workers=%d), median of %d runs" % (nspec, cores, len(times))}


def secondary(args, torch):
    """Unpack family (2^18 spectra x 1024 channels), direct 4-bit (2 ant x 2
    pol, 128 ch, 2^17 spectra, 4096-blocks), precision correlator (8 x 1024 x
    6400, FP32 on MFMA) and direct 1-bit correlator."""
    from albatros_axion_detection_amd import _lib, synth
    out = {}
    dev = _lib.device()

    def timeit(fn, n=10):
        # clock settle: the call repeated for about 100 ms, untimed, so a GPU
        # that was idle (this leg run on its own) is timed at its working
        # clocks, as it is after the batched pass
        t_settle = time.perf_counter()
        while time.perf_counter() - t_settle < 0.1:
            for _ in range(5):
                fn()
            torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / n

    # unpack family on 2^18 spectra x 1024 channels of random raw bytes:
    # algorithmic bytes = raw bytes read + both pols written
    T, nch = 2 ** 18, 1024
    torch.cuda.manual_seed(5)
    raw4 = torch.randint(0, 256, (T, 2 * nch), dtype=torch.uint8, device=dev)
    raw1 = raw4[:, : nch // 2].contiguous()
    chans = torch.arange(nch, dtype=torch.int64, device=dev)
    c0 = torch.empty((T, nch), dtype=torch.complex64, device=dev)
    c1 = torch.empty_like(c0)
    s0 = torch.empty((T, nch), dtype=torch.uint8, device=dev)
    s1 = torch.empty_like(s0)
    legs = {
        "unpack_4bit_float": (lambda: _lib.call("alb_unpack_4bit_float", _lib.ptr(raw4), _lib.ptr(c0), _lib.ptr(c1), 0,
                                                T, _lib.ptr(chans), nch, nch, _lib.stream()),
                              T * 2 * nch + 2 * T * nch * 8),
        "unpack_1bit_float": (lambda: _lib.call("alb_unpack_1bit_float", _lib.ptr(raw1), _lib.ptr(c0), _lib.ptr(c1), 0,
                                                T, 0, nch, nch, _lib.stream()),
                              T * nch // 2 + 2 * T * nch * 8),
        "sortpols_4bit": (lambda: _lib.call("alb_sortpols", _lib.ptr(raw4), _lib.ptr(s0), _lib.ptr(s1), 0, T, nch, nch,
                                            4, 0, nch, _lib.stream()),
                          T * 2 * nch + 2 * T * nch),
        "sortpols_1bit": (lambda: _lib.call("alb_sortpols", _lib.ptr(raw1), _lib.ptr(s0), _lib.ptr(s1), 0, T, nch // 4,
                                            nch, 1, 0, nch, _lib.stream()),
                          T * nch // 2 + 2 * T * nch // 4),
    }
    forms = {}
    for name, (fn, nbytes) in legs.items():
        ms = timeit(fn)
        forms[name] = {"ms": round(ms, 5), "GB_s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                       "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out["unpack"] = {"workload": "2^18 spectra x 1024 channels, raw bytes -> complex64 pols (unpack) or per-pol "
                                 "bytes (sortpols); GB/s of raw bytes read + pols written", "forms": forms}
    del raw4, raw1, chans, c0, c1, s0, s1
    # direct 4-bit: 4 inputs x 2^17 spectra x 128 ch, 10 pairs, 4096-blocks
    lib = _lib.load()
    inp_np = synth.direct4(T=2 ** 17, nchan=128, seed=0)
    inp = _lib.to_device(inp_np)
    ptrs = (_lib._vp * 4)(*[inp[i].data_ptr() for i in range(4)])
    b4 = 68091904
    forms, res4 = {}, {}
    for name, m in (("mfma_i8", 2), ("valu", 1)):
        o4 = torch.empty((32, 10, 128, 2), dtype=torch.int64, device=dev)
        lib.alb_tune_set(b"xcorr_int_method", m)
        ms = timeit(lambda: _lib.call("alb_xcorr_4bit_allpairs", ptrs, 4, 2 ** 17, 128, 4096, _lib.ptr(o4),
                                      _lib.stream()), n=50)
        res4[name] = o4
        forms[name] = {"ms": round(ms, 5), "Msamples_s": round(4 * 2 ** 17 * 128 / (ms * 1e-3) / 1e6, 1),
                       "hbm_frac": round(b4 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # two channels per 16 x 16 x 64 int8 product (4 inputs): 32 blocks x 64 channel pairs x 64 steps
    forms["mfma_i8"]["mfma_frac"] = round(32 * 64 * 64 * 32768 / (forms["mfma_i8"]["ms"] * 1e-3) / 5e15, 5)
    lib.alb_tune_set(b"xcorr_int_method", 0)
    out["direct4"] = {"workload": "4 inputs x 2^17 spectra x 128 ch, 10 pairs, 4096-blocks, int64 sums",
                      "forms": forms, "forms_bit_identical": bool(torch.equal(res4["mfma_i8"], res4["valu"]))}
    if not args.no_cpu:
        log("[bench] CPU reference: direct 4-bit ...")
        out["direct4"]["cpu_baseline"] = cpu_reference_direct(4, inp_np, 128, 4096, 32)
    del inp, res4
    # direct 1-bit: 16 inputs x T spectra x 1024 ch (256 B), 136 pairs, 8192-blocks
    T1 = args.onebit_T
    inp = synth.direct1(ninput=16, T=T1, nchan=1024, seed=3, device=dev)
    ptrs = (_lib._vp * 16)(*[inp[i].data_ptr() for i in range(16)])
    nb1 = T1 // 8192
    b1 = 16 * T1 * 256 + nb1 * 136 * 1024 * 2 * 8
    forms, res1 = {}, {}
    # the matrix-core forms: one FP4 tile per channel (the default here) and
    # three int8 products per channel; then the XOR form
    for name, m, form, n in (("mfma_fp4", 2, 1, 10), ("mfma_i8", 2, 2, 5), ("xor_valu", 1, 0, 2)):
        o1 = torch.empty((nb1, 136, 1024, 2), dtype=torch.int64, device=dev)
        lib.alb_tune_set(b"xcorr_int_method", m)
        lib.alb_tune_set(b"xcorr1_form", form)
        ms = timeit(lambda: _lib.call("alb_xcorr_1bit_allpairs", ptrs, 16, T1, 256, 1024, 8192, _lib.ptr(o1),
                                      _lib.stream()), n=n)
        res1[name] = o1
        forms[name] = {"ms": round(ms, 4), "Msamples_s": round(16 * T1 * 1024 / (ms * 1e-3) / 1e6, 1),
                       "hbm_frac": round(b1 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    lib.alb_tune_set(b"xcorr1_form", 0)
    # FP4: two 16 x 16 x 128 products per channel per 64 spectra, against the dense FP4 peak
    forms["mfma_fp4"]["mfma_frac"] = round(2 * 1024 * (T1 // 64) * 65536 / (forms["mfma_fp4"]["ms"] * 1e-3) / 10.1e15, 4)
    forms["mfma_i8"]["mfma_frac"] = round(3 * 1024 * (T1 // 64) * 32768 / (forms["mfma_i8"]["ms"] * 1e-3) / 5e15, 4)
    lib.alb_tune_set(b"xcorr_int_method", 0)
    out["direct1"] = {"workload": "16 inputs x %d spectra x 1024 ch, 136 pairs, 8192-blocks, int64 sums" % T1,
                      "forms": forms,
                      "forms_bit_identical": bool(torch.equal(res1["mfma_fp4"], res1["xor_valu"]) and
                                                  torch.equal(res1["mfma_i8"], res1["xor_valu"]))}
    if not args.no_cpu:
        log("[bench] CPU reference: direct 1-bit ...")
        out["direct1"]["cpu_baseline"] = cpu_reference_direct(1, inp[:, :8192].cpu().numpy(), 1024, 8192, 1)
    del inp, res1
    # precision correlator: 8 inputs x 1024 fine spectra x 6400 channels, four
    # arithmetic forms on the matrix cores, each against a float64 product
    x = synth.precision_xcorr(M=8, T=1024, F=6400, seed=2, device=dev)
    xd = x.permute(2, 1, 0)                     # C-order [F][T][M]
    x64 = xd.to(torch.complex128)
    ref = torch.matmul(x64.transpose(1, 2).contiguous(), x64.conj().contiguous()) / 1024.0   # [f, i, j]
    dg = torch.sqrt(torch.diagonal(ref, dim1=1, dim2=2).real)
    norm = dg[:, :, None] * dg[:, None, :]
    del x64
    vis = torch.empty((6400, 8, 8), dtype=torch.complex64, device=dev)
    b = 422707200
    flops = 16 * 16 * 1024 * 2 * 6400           # one realified 16x16 Gram product per channel
    peaks = {"fp32": FP32_MFMA_TFLOPS, "bf16x3": 2500.0, "fp16": 2500.0, "fp16x3": 2500.0}
    prec = {}
    for mode, name in enumerate(("fp32", "bf16x3", "fp16", "fp16x3")):
        ms = timeit(lambda: _lib.call("alb_xcorr_all_ant_prec", _lib.ptr(xd), 8, 1024, 6400, _lib.ptr(vis),
                                      float(np.float32(1) / np.float32(1024)), mode, _lib.stream()))
        c = vis.transpose(1, 2).to(torch.complex128)   # vis[f, j, i] = C_ij
        err = float(torch.max(torch.abs(c - ref) / norm))
        nprod = 3 if name.endswith("x3") else 1
        prec[name] = {"ms": round(ms, 5), "Msamples_s": round(8 * 1024 * 6400 / (ms * 1e-3) / 1e6, 1),
                      "max_rel_err_vs_fp64": err,
                      "hbm_frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "mfma_frac": round(nprod * flops / (ms * 1e-3) / 1e12 / peaks[name], 4)}
    out["xcorr_precision"] = {"workload": "8 inputs x 1024 x 6400 ch complex64 (F-order), full 8x8 per channel; "
                                          "error = max |C - C64| / sqrt(C64_ii C64_jj)", "forms": prec}
    if not args.no_cpu:
        log("[bench] CPU port: all-antenna correlator ...")
        out["xcorr_precision"]["cpu_baseline"] = cpu_port_xcorr(x.cpu().numpy(), args.cpu_budget)
    del x, xd, vis, ref, norm
    return out


def cpu_port_xcorr(x, budget_s):
    """The reference's CPU correlator (avg_xcorr_all_ant_cpu,
    correlations.py:115-151: upper-triangle loop, complex64 products summed in
    complex128, parallel over channels) restated in oracle.cpu_port, on the
    whole (M, K, F) workload: all cores (median of up to 5 runs within the
    budget) and one core (one run)."""
    from oracle import cpu_port
    model, cores = cpu_info()
    M, K, F = x.shape
    cpu_port.avg_xcorr_all_ant(x, cores)
    times = []
    t_all = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_all) < budget_s:
        t0 = time.perf_counter()
        cpu_port.avg_xcorr_all_ant(x, cores)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    t1c0 = time.perf_counter()
    cpu_port.avg_xcorr_all_ant(x, 1)
    t1 = time.perf_counter() - t1c0
    nsamp = M * K * F
    return {"value": round(nsamp / t / 1e6, 2), "unit": "Msamples/s", "cores": cores, "kind": "port",
            "cpu_model": model, "one_core_value": round(nsamp / t1 / 1e6, 2),
            "sample": "the whole workload (%d inputs x %d spectra x %d channels), numpy threads over channel "
                      "chunks; median of %d runs (one-core figure: one thread, one run)" % (M, K, F, len(times))}


def cpu_info():
    """CPU model (/proc/cpuinfo) and the cores this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def batched(args, torch):
    """Headline: the batched streaming workload.  2 antennas x 2 pols, 300
    coarse channels of 4-bit data, 64 arrival chunks of uneven length (uniform
    in [2^10, 2^14] spectra) with runs of 5 lost spectra (~0.5%), re-blocked by
    spec_num into iPFB tiles of 4088 (+8 carried) spectra -> iPFB (Wiener
    0.45) -> re-PFB 64x (ntap 4, 19200 fine channels) -> 4x4 visibilities per
    1024-spectrum job.  One step = one hipGraph replay of the row-map build
    (alb_build_rowmap from the stream's spec_nums) and the batched pass
    (fused column pass + re-PFB stage 1, one stage 2 for every input, every
    job correlated at once), inputs resident in HBM.  The plain-FP32 form
    (per-chunk / per-job launches, FP32 correlator) is timed on the same
    inputs.  With --autotune, the pipeline's autotuner first picks tile
    batch, lanes and launch shapes on these inputs."""
    from albatros_axion_detection_amd import _lib, repfb, synth
    dev = _lib.device()
    acclen, cut, osamp, nblock, nchan, c0 = 4088, 4, 64, 1024, 300, 1000
    log("[bench] generating the batched streaming input ...")
    packed, specnums, lengths = synth.batched_stream(nant=2, nchan=nchan, nchunks=64, lmin=2 ** 10, lmax=2 ** 14,
                                                     loss=0.005, run=5, seed=4, device=dev)
    total = int(lengths.sum())
    nfixed = total // acclen
    chan = np.arange(c0, c0 + nchan)
    # two stage-1 lanes: inputs alternate between two HIP streams, so one
    # input's row, column and edge-block launches run beside the next one's
    # (1.5% faster than one lane; four lanes measured the same as two).  The
    # per-kernel times below come from passes with one lane, where every
    # launch runs alone.
    pipe = repfb.RePFBPipeline(4, chan, acclen, cut, osamp, nblock, c0 * osamp, (c0 + nchan) * osamp, tile_batch=16,
                               lanes=args.lanes, stage2_overlap=args.stage2_overlap, ipfb_form=args.ipfb_form)
    srm = repfb.StreamRowmap(specnums, 0, nfixed, acclen)
    rm = srm.build()
    plist = [packed[m] for m in range(4)]
    rms = [rm] * 4
    bufs = pipe.buffers(nfixed)
    njobs = len(bufs["jobs"])
    lib = _lib.load()
    knob_keys = [k for k, _ in repfb.RePFBPipeline.TUNE_KNOBS]
    knobs_before = {k: int(lib.alb_tune_get(k.encode())) for k in knob_keys}
    tune = None
    if args.autotune:
        log("[bench] autotuning the batched pass ...")
        tune = pipe.autotune(plist, rms, nfixed, bufs, precision=args.precision, reps=2)
        log("[bench] autotune: %s" % json.dumps({k: v for k, v in tune.items() if k != "tried"}))
    timer = Timer(torch)

    def step():
        timer("rowmap", srm.build)
        return pipe.correlate_into(plist, rms, nfixed, bufs, precision=args.precision)

    # plain-FP32 per-chunk form: the reference check of the batched result
    pipe.correlate_per_chunk(plist, rms, nfixed, bufs)
    torch.cuda.synchronize()
    ref = bufs["vis"].clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    if args.no_graph:
        replay = step
    else:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            step()
        replay = g.replay
    replay()
    torch.cuda.synchronize()
    vis = bufs["vis"]
    # max |difference| / rms over all visibilities (edge fine channels carry
    # only leakage, so per-channel ratios there measure FP32 rounding noise)
    err = float(torch.max(torch.abs(vis - ref)) / torch.sqrt(torch.mean(torch.abs(ref) ** 2)))
    # clock settle (~100 ms of replays), warm-up, then K replays between two events
    t_settle = time.perf_counter()
    while time.perf_counter() - t_settle < 0.1:
        replay()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        replay()
    t0e = torch.cuda.Event(enable_timing=True)
    t1e = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t0e.record()
    for _ in range(args.steps):
        replay()
    t1e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0e.elapsed_time(t1e) / args.steps
    # per-kernel durations: eager passes with a HIP event pair around every
    # launch on the stream it goes to, with one stage-1 lane so that no two
    # launches overlap (on two lanes an event pair would also time the other
    # lane's kernels); the step time above is taken without them
    nk = max(1, min(args.steps, 5))
    lanes, overlap = pipe.lanes, pipe.stage2_overlap
    pipe.lanes, pipe.stage2_overlap = 1, False
    timer.on = True
    pipe.timer = timer
    for _ in range(nk):
        step()
    kern = timer.summary()
    counts = {k: v // nk for k, v in timer.counts().items()}
    timer.on = False
    pipe.timer = None
    pipe.lanes, pipe.stage2_overlap = lanes, overlap
    eager0 = torch.cuda.Event(enable_timing=True)
    eager1 = torch.cuda.Event(enable_timing=True)
    eager0.record()
    for _ in range(3):
        step()
    eager1.record()
    torch.cuda.synchronize()
    ms_eager = eager0.elapsed_time(eager1) / 3
    log("[bench] batched: %.3f ms/step graph (wall %.3f), %.3f ms eager; kernels %s" % (
        ms, 1e3 * wall / args.steps, ms_eager, json.dumps({k: round(v, 4) for k, v in kern.items()})))

    def timed(fn, n):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(n):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / n

    # the plain-FP32 form at its own best column-pass width (it runs the plain
    # column pass, whose fastest width differs from the fused pass's)
    run_p = lambda: (srm.build(), pipe.correlate_per_chunk(plist, rms, nfixed, bufs))
    tuned_w = int(lib.alb_tune_get(b"ipfb_cols_width"))
    chunk_ms = {}
    for w in ((2, 4, 8) if not args.no_plain else ()):
        _lib.call("alb_tune_set", b"ipfb_cols_width", w)
        run_p()
        chunk_ms[w] = timed(run_p, 2)
    _lib.call("alb_tune_set", b"ipfb_cols_width", tuned_w)
    ms_chunk = min(chunk_ms.values()) if chunk_ms else None
    for k, v in knobs_before.items():
        _lib.call("alb_tune_set", k.encode(), v)
    del replay
    n_coarse = nfixed * acclen                      # coarse spectra per input
    nsamp = 4 * n_coarse * 4096                     # ADC-equivalent samples per step
    # algorithmic bytes per coarse spectrum per input: raw 300 B in, time
    # stream 16384 B written + 16384 B read, kept fine channels 2400 B written
    # + 2400 B read (19200 channels x 8 B / 64 spectra per re-PFB block)
    per_spec = 300 + 2 * 16384 + 2 * 19200 * 8 // 64
    pipe_bytes = 4 * n_coarse * per_spec
    # per-launch algorithmic bytes of each kernel (its terms of per_spec x the
    # coarse spectra the launch covers): the fused column pass writes and reads
    # the time stream; the row pass reads the raw bytes; stage 2 writes the
    # kept fine channels and the correlator reads them; the edge blocks'
    # stream re-reads are counted as what that launch reads
    nedge = len(pipe.edge_blocks(nfixed, bufs["nbt"]))
    per_pass = {
        "ipfb_rows_packed4": 4 * n_coarse * 300,
        "ipfb_chan_fft": 4 * n_coarse * 300,
        "ipfb_columns_repfb": 4 * n_coarse * 2 * 16384,
        "repfb_stage1_edges": 4 * nedge * osamp * 4 * 4096 * 4,
        "repfb_stage2": 4 * n_coarse * 2400,
        "xcorr_all_ant_strided": 4 * n_coarse * 2400,
    }
    tot = {k: kern[k] * counts.get(k, 1) for k in kern}
    dom = max((k for k in tot if k in per_pass), key=lambda k: tot[k])
    alg = per_pass[dom] / max(1, counts.get(dom, 1))
    achieved = alg / (kern[dom] * 1e-3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(dom + "_batched"),
            "algorithmic_bytes_per_launch": int(alg), "avg_launch_ms": round(kern[dom], 5),
            "launches_per_step": counts.get(dom, 1)}
    pipe_roof = {"bound": "hbm", "algorithmic_bytes_per_step": pipe_bytes,
                 "achieved": round(pipe_bytes / (ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(pipe_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    res = {
        "value": nsamp / (ms * 1e-3) / 1e6, "ms_per_step": ms, "roofline": roof, "pipeline_roofline": pipe_roof,
        "kernels_ms": {k: round(v, 5) for k, v in kern.items()}, "launches_per_step": counts,
        "config": {"workload": "batched streaming: 2 ant x 2 pol, 300 coarse ch 4-bit, 64 arrival chunks "
                               "U[2^10, 2^14] spectra with ~0.5%% loss in runs of 5 (%d spectra) -> %d iPFB tiles "
                               "of 4088 (+8 carried), Wiener 0.45 -> re-PFB 64x ntap 4 (19200 fine ch) -> 4x4 vis "
                               "over %d jobs of 1024 spectra; %s per step" % (
                                   total, nfixed, njobs, "one eager pass" if args.no_graph else "one hipGraph replay"),
                   "inputs": 4, "coarse_channels": nchan, "arrival_chunks": 64, "spectra_per_input": n_coarse,
                   "acclen": acclen, "cut": cut, "osamp": osamp, "nblock": nblock, "jobs": njobs,
                   "correlator": args.precision, "stage1_lanes": args.lanes, "ipfb_form": pipe.ipfb_form,
                   "stage2_overlap": args.stage2_overlap},
        "plain_fp32": {"form": "per-chunk / per-job launches, FP32 correlator, fastest column-pass width",
                       "ms_per_step": round(ms_chunk, 3),
                       "Msamples_s": round(nsamp / (ms_chunk * 1e-3) / 1e6, 1),
                       "ms_by_cols_width": {str(w): round(v, 3) for w, v in chunk_ms.items()}} if ms_chunk else None,
        "speedup_vs_plain_fp32": round(ms_chunk / ms, 3) if ms_chunk else None,
        "eager_ms_per_step": round(ms_eager, 3),
        "max_abs_diff_over_rms_vs_plain_fp32": err,
        "ref_unit_msps": 4 * n_coarse * 2049 * 8 / (ms * 1e-3) / 1e6,
        "autotune": tune,
    }
    ctx = dict(packed=packed, specnums=specnums, acclen=acclen, cut=cut, osamp=osamp, c0=c0, nchan=nchan,
               filt=pipe.filt, win=pipe.win)
    return res, ctx


def cpu_baseline_batched(ctx, budget_s, sample_chunks=2):
    """The reference CPU path on a bounded sample of the same stream: the
    first ``sample_chunks`` iPFB chunks of each of the 4 inputs, spec_num zero
    filled, through cpu_ipfb (2*cut carried rows) and cpu_pfb at 64x
    (pfb_cpu_utils.py:90-172, scipy.fft float32), then the visibilities of
    the re-PFB blocks those chunks fill with the complex128 all-antenna
    correlator (correlations.py:115-151).  All cores (median of up to 5 runs
    within the budget) and one core (one run)."""
    from albatros_axion_detection_amd import _lib
    from oracle import cpu_port
    acclen, cut, osamp, c0, nchan = ctx["acclen"], ctx["cut"], ctx["osamp"], ctx["c0"], ctx["nchan"]
    sp = _lib.host(ctx["specnums"]) if not isinstance(ctx["specnums"], np.ndarray) else ctx["specnums"]
    nrow = int(np.searchsorted(sp, sample_chunks * acclen))
    raw = ctx["packed"][:, :nrow].cpu().numpy()          # [4][rows][300]
    filt = _lib.host(ctx["filt"]).astype(np.complex64)
    win = _lib.host(ctx["win"]).astype(np.float32)
    model, cores = cpu_info()

    def run(workers):
        xin = []
        for m in range(4):
            b = raw[m]
            hi = (b >> 4).astype(np.int8)
            lo = (b & 15).astype(np.int8)
            hi = np.where(hi > 8, hi - 16, hi)
            lo = np.where(lo > 8, lo - 16, lo)
            chunks = []
            for c in range(sample_chunks):
                sel = (sp[:nrow] >= c * acclen) & (sp[:nrow] < (c + 1) * acclen)
                full = np.zeros((acclen, 2049), np.complex64)       # make_continuous_gpu zero fill
                full[np.ix_(sp[:nrow][sel] - c * acclen, np.arange(c0, c0 + nchan))] = hi[sel] + 1j * lo[sel]
                chunks.append(full)
            xin.append(cpu_port.rechannelize(chunks, filt, win, acclen, cut, osamp, c0 * osamp,
                                             (c0 + nchan) * osamp, workers))
        x = np.stack(xin)                                             # (M, K, F)
        K = x.shape[1]
        return np.einsum("ikf,jkf->ijf", x.astype(np.complex128), np.conj(x.astype(np.complex128))) / K

    run(cores)
    times = []
    t_all = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_all) < budget_s:
        t0 = time.perf_counter()
        run(cores)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    t1c0 = time.perf_counter()
    run(1)
    t1 = time.perf_counter() - t1c0
    nsamp = 4 * sample_chunks * acclen * 4096
    return {"value": round(nsamp / t / 1e6, 2), "unit": "Msamples/s", "cores": cores, "kind": "port",
            "cpu_model": model, "one_core_value": round(nsamp / t1 / 1e6, 2),
            "sample": "%d iPFB chunks of 4088 spectra x 4 inputs of the same stream (%d coarse spectra): unpack + "
                      "zero fill, cpu_ipfb, cpu_pfb 64x (scipy.fft float32, workers=%d), complex128 4x4 correlator "
                      "of the %d re-PFB blocks; median of %d runs (one-core figure: workers=1, one run)"
                      % (sample_chunks, 4 * sample_chunks * acclen, cores, sample_chunks * acclen // osamp - 3,
                         len(times))}


class _StdoutToStderr:
    """Send file descriptor 1 to stderr while reference C code runs (its
    printf output must not mix with the bench's one JSON line)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)


def _omp_threads(n):
    import ctypes.util
    name = ctypes.util.find_library("gomp") or "libgomp.so.1"
    try:
        ctypes.CDLL(name).omp_set_num_threads(int(n))
        return True
    except OSError:
        return False


def cpu_reference_direct(kind, inputs, nchan, blocklen, nblocks):
    """The reference's own OpenMP correlators (oracle/_ref, built from
    correlations_cpu.c) over ``nblocks`` blocks of every pair: 4-bit
    avg_autocorr_4bit / avg_xcorr_4bit (:39, :128) or 1-bit avg_xcorr_1bit
    (:273), one call per pair per block as the reference's callers make
    them.  All cores and one core."""
    from oracle import ref as oref
    if not oref.available():
        return None
    model, cores = cpu_info()
    n = inputs.shape[0]
    pairs = [(i, j) for i in range(n) for j in range(i, n)]

    def run():
        for b in range(nblocks):
            blk = inputs[:, b * blocklen:(b + 1) * blocklen]
            for i, j in pairs:
                if kind == 4 and i == j:
                    oref.avg_autocorr_4bit(blk[i], blocklen)
                elif kind == 4:
                    oref.avg_xcorr_4bit(blk[i], blk[j], blocklen)
                else:
                    oref.avg_xcorr_1bit(blk[i], blk[j], nchan, blocklen)

    out = {}
    with _StdoutToStderr():
        for label, thr in (("all", cores), ("one", 1)):
            _omp_threads(thr)
            run()
            ts = []
            for _ in range(3 if label == "all" else 1):
                t0 = time.perf_counter()
                run()
                ts.append(time.perf_counter() - t0)
            out[label] = float(np.median(ts))
        _omp_threads(cores)
    nsamp = n * nblocks * blocklen * nchan
    return {"value": round(nsamp / out["all"] / 1e6, 1), "unit": "Msamples/s", "cores": cores, "kind": "reference",
            "cpu_model": model, "one_core_value": round(nsamp / out["one"] / 1e6, 1),
            "sample": "%d of the workload's blocks of %d spectra, every pair (%d), one reference call per pair per "
                      "block (OpenMP threads = %d; median of 3 runs, one-core: 1 run)"
                      % (nblocks, blocklen, len(pairs), cores)}


def copy_calibration(torch, nbytes=1 << 30, reps=5):
    """Plain device-to-device copies of a known size: the FETCH_SIZE /
    WRITE_SIZE counters read for these dispatches give the factor that turns
    the counters into HBM bytes (expected nbytes read and nbytes written)."""
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    print(json.dumps({"copy_bytes": nbytes, "reps": reps, "ms": round(ms, 5),
                      "GBps_read_plus_write": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nchunks", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--onebit-T", type=int, default=2 ** 20)
    ap.add_argument("--cols-width", type=int, default=0, choices=[0, 2, 4, 8],
                    help="column-pass strip width (0: library default, 4)")
    ap.add_argument("--cols-group", type=int, default=0, choices=[0, 2, 4, 8, 16, 32, 64],
                    help="column-pass strips per XCD group (0: library default, 2 up to 4 tiles, 64 above)")
    ap.add_argument("--rows-group", type=int, default=0, choices=[0, 2, 4],
                    help="spectra per workgroup of the FFT row pass (0: library default, 2 up to 4 tiles, 4 above)")
    ap.add_argument("--autotune", action="store_true",
                    help="run the pipeline autotuner on the batched pass's inputs first and time its choice "
                         "(default: the library's launch shapes, which the committed profiles were taken with)")
    ap.add_argument("--no-autotune", action="store_true", help="accepted for compatibility (the default)")
    ap.add_argument("--ipfb-form", default="auto", choices=["auto", "spectral", "direct"],
                    help="iPFB order of the fused pass: direct (the reference's order; auto takes it) or "
                         "spectral (transform along time first, on the channel columns)")
    ap.add_argument("--stage2-overlap", action="store_true",
                    help="run each input's re-PFB stage 2 on a second stream beside the next inputs' stage 1 "
                         "(default: one stage-2 launch for all inputs after their stage 1, which measured faster)")
    ap.add_argument("--lanes", type=int, default=2, choices=[1, 2],
                    help="concurrent stage-1 lanes of the batched pass (2: two inputs' stage 1 side by side, "
                         "1: every kernel runs alone; the per-kernel times are taken with one lane either way)")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager passes instead of hipGraph replays (the counter-collection runs)")
    ap.add_argument("--no-plain", action="store_true",
                    help="do not time the plain-FP32 per-chunk form (it still runs once as the result check)")
    ap.add_argument("--precision", default="fp16x3", choices=["fp32", "bf16x3", "fp16", "fp16x3"],
                    help="correlator arithmetic of the batched pass (FP16 hi/lo split by default)")
    ap.add_argument("--sweep", action="store_true",
                    help="time every column-pass strip width x row-pass group before timing and keep the fastest")
    ap.add_argument("--row-pass", default="auto", choices=["auto", "fft", "mfma"],
                    help="iPFB row pass: FFT per spectrum or matrix-core synthesis from the kept channels")
    ap.add_argument("--unfused", action="store_true",
                    help="time the separate column pass + re-PFB stage 1 instead of the fused kernel")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 plain device copies of 1 GiB (HBM counter calibration) and exit")
    ap.add_argument("--rechannelizer-only", action="store_true",
                    help="only run the single-input rechannelizer workload and print its JSON (its profiles)")
    ap.add_argument("--secondary-only", action="store_true",
                    help="only run the secondary workloads (unpack, direct correlators, precision correlator) "
                         "and print their JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (asked for %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    sys.path.insert(0, HERE)
    if args.copy_calibration:
        copy_calibration(torch)
        return
    if args.secondary_only:
        print(json.dumps(secondary(args, torch)), flush=True)
        return
    if args.rechannelizer_only:
        rres, rctx = rechannelizer(args, torch)
        if not args.no_cpu:
            rres["cpu_baseline"] = cpu_baseline(rctx, args.cpu_budget)
        print(json.dumps(rres), flush=True)
        return
    res, ctx = batched(args, torch)
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (batched stream sample) ...")
        cpu = cpu_baseline_batched(ctx, args.cpu_budget)
        log("[bench] CPU baseline: %.1f Msamples/s on %d cores, %.1f on one" % (cpu["value"], cpu["cores"],
                                                                                cpu["one_core_value"]))
    del ctx
    sec = None
    if not args.no_secondary:
        log("[bench] rechannelizer ...")
        rres, rctx = rechannelizer(args, torch)
        if not args.no_cpu:
            rres["cpu_baseline"] = cpu_baseline(rctx, args.cpu_budget)
        rres["value_Msamples_s"] = round(rres.pop("value"), 2)
        log("[bench] secondary workloads ...")
        sec = {"rechannelizer": rres}
        sec.update(secondary(args, torch))
    line = {
        "metric": METRIC,
        "value": round(res["value"], 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 4-bit Gaussian baseband, correlated antennas, runs of lost spectra)",
        "config": res["config"],
        "roofline": res["roofline"],
        "pipeline_roofline": res["pipeline_roofline"],
        "cpu_baseline": cpu,
        "speedup_vs_plain_fp32": res["speedup_vs_plain_fp32"],
        "plain_fp32": res["plain_fp32"],
        "max_abs_diff_over_rms_vs_plain_fp32": res["max_abs_diff_over_rms_vs_plain_fp32"],
        "eager_ms_per_step": res["eager_ms_per_step"],
        "kernels_ms": res["kernels_ms"],
        "launches_per_step": res["launches_per_step"],
        "ref_unit_msps": round(res["ref_unit_msps"], 1),
        "autotune": res["autotune"],
        "secondary": sec,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
