"""CPU oracle for the ALBATROS baseband hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may use it, and only as the reference
results tests compare against or as the timed CPU restatement of the
reference algorithm -- never as the
thing measured on the GPU and never as a fallback.

Modules
-------
unpack   4-bit / 1-bit decode, sortpols repack, histograms (integer, exact)
corr     integer correlators (4-bit, 1-bit, van Vleck, 2-antenna merge) and
         the all-antenna complex correlator (complex128 accumulation)
pfb      sinc-Hamming window, window spectrum (matft), Wiener filter,
         inverse PFB and forward PFB, all in float64
stream   the streaming re-PFB + correlate driver restated on in-memory
         chunks (chunk planner, zero fill, overlap bookkeeping)
orbcomm  the FFT lag correlation and delay phase of the ORBCOMM delay search
ref      ctypes bindings of the reference's own C sources compiled into
         ``oracle/_ref`` (built only where the reference tree is present)

Every function cites the reference file:line it restates (paths relative
to ``albatros_analysis/``).  How each function is pinned (golden vectors,
closed forms, brute force, the compiled reference) is listed in
``PINNING`` below; a function listed with ``None`` would be
"parity unpinned" (there is none at present).
"""

PINNING = {
    "unpack.unpack_4bit": "reference fixtures test_4bit_float_unpack.py:28-125, "
                          "exhaustive 256-byte table vs the mars2019 rule, oracle/_ref",
    "unpack.unpack_1bit": "reference fixtures test_1bit_unpack_gpu.py:186-218, "
                          "exhaustive 256-byte table, oracle/_ref",
    "unpack.sortpols": "reference asserts test_1bit_unpack.py:81-107, "
                       "test_1bit_xcorr_2ant.py:76-94, oracle/_ref (all tails)",
    "unpack.hist_4bit/hist_1bit": "reference asserts test_4bit_float_unpack.py:128-170, "
                                  "test_1bit_unpack.py:110-134, oracle/_ref",
    "corr.xcorr_4bit_sums": "brute force complex128 on decoded values, oracle/_ref",
    "corr.xcorr_4bit_2ant_sums": "brute-force set intersection, oracle/_ref",
    "corr.xcorr_1bit_sums": "+-1 algebra vs decoded complex product, oracle/_ref (all tails)",
    "corr.vanvleck_sums": "brute force on decoded bits, test_1bit_xcorr_2ant.py:119-169 "
                          "property, oracle/_ref",
    "corr.avg_xcorr_all_ant": "reference constants test_gpu_all_ant_complex_xcorr.py:35-68, "
                              "numpy einsum in complex128",
    "corr.avg_xcorr_all_ant_c64_products": "reference constants test_gpu_all_ant_complex_xcorr.py:35-68 "
                                           "exactly, complex128 form to product rounding, Hermitian",
    "corr.van_vleck_correction": "arcsine law of 1-bit Gaussian correlations (recovers the correlation "
                                 "coefficient), closed-form special values",
    "stream.missing_fraction": "hand-traced worked cases of helper_gpu_stream.py:157-233 "
                               "(uneven chunks, a chunk spanning jobs, carried remainders past a job)",
    "orbcomm.coarse_xcorr": "direct lag sums (lag_sums_brute) on small inputs; the lag of a "
                            "delayed copy",
    "orbcomm.apply_delay": "unit modulus, zero delay is the identity, linear phase in frequency",
    "pfb.sinc_hamming": "numpy.hamming * numpy.sinc closed form (library), symmetry",
    "pfb.matft_fp64": "float64 numpy.fft.rfft of the window matrix (the reference's "
                      "own construction) vs the 4-term closed form",
    "pfb.wiener_filter": "1/conj(W) identity at thresh=0, stable-form identity",
    "pfb.ipfb": "exact inversion of the forward PFB at thresh=0 on a circulant "
                "signal, brute-force DFT at tiny sizes",
    "pfb.pfb": "brute-force DFT at tiny sizes, tone at a channel centre",
    "stream.repfb_xcorr_avg": "stream continuity invariant, chunk planner vs "
                              "helper_gpu_stream.py:256-284 worked cases",
}
