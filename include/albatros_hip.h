/*
 * albatros_hip.h -- C ABI of libalbatros_hip.so, the MI355X (gfx950) hot path of
 * the ALBATROS baseband pipeline: 4/1-bit unpack, Wiener-filtered inverse PFB,
 * forward re-PFB (upchannelization) and per-channel cross-antenna correlation.
 *
 * Conventions (all entry points):
 *   - every data pointer is a DEVICE pointer (hipMalloc / torch CUDA storage);
 *   - sizes and strides are int64_t element counts; complex data is interleaved
 *     float32 (re, im) -- numpy complex64 / torch.complex64;
 *   - `stream` is a hipStream_t (NULL = default stream); work is stream ordered,
 *     nothing synchronises the host unless stated;
 *   - return 0 on success, 1 invalid argument, 2 HIP runtime error, 3 size not
 *     supported; alb_last_error() gives the message of the calling thread;
 *   - no entry point keeps a pointer to caller memory after it returns; the
 *     library's own device memory is read-only transform tables built in float64
 *     on first use of a transform size (one per device and size) in a fixed 2 MiB
 *     device array (hipMalloc only for tables beyond it: transforms above 2^24
 *     points), or ahead of time by alb_prepare_transform / alb_prepare_repfb --
 *     which code captured into a hipGraph must do, since a build synchronises and
 *     a first use on a capturing stream is refused (ALB_EINVAL).
 *
 * Each declaration cites the reference interface it replaces
 * (paths relative to albatros_analysis/ of PhilJ0ly/albatros_axion_detection).
 */
#ifndef ALBATROS_HIP_H
#define ALBATROS_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- housekeeping ------------------------------------------------------ */
const char* alb_last_error(void);
int alb_version(void);
int alb_stream_synchronize(void* stream);
/* Zero nbytes of device memory with a kernel (safe inside graph capture). */
int alb_fill_zero(void* p, int64_t nbytes, void* stream);
/* Build the twiddle table of an N-point transform (N a power of two) on the current device
 * (no counterpart in the reference, whose cuFFT plans, pycufft.py:35-76, hold their own tables):
 * the row passes use 4096, the column passes nslice (or the chirp-z length P), the fused column
 * pass nslice and 4096.  Synchronises the device; call it before capturing a pass into a graph. */
int alb_prepare_transform(int64_t N);
/* Build the transform tables of the re-PFB at osamp (a power of two 1..64; stage 2's 4096-point
 * twiddles) on the current device (no counterpart in the reference, whose cupy_pfb,
 * pfb_gpu_utils.py:105-121, hands the FFT to cuFFT).  Synchronises the device; allocates nothing. */
int alb_prepare_repfb(int64_t osamp);

/* ---- unpack (src/correlations/unpacking.c, unpacking.py) --------------- */
/* unpacking.c:155 unpack_4bit_float / unpacking.py:140 _unpack_4bit_float_gpu.
 * data: raw spectra of 2*nchan bytes; pol0/pol1: complex64 [rowend-rowstart][ncols];
 * channels: int64 column indices (device). */
int alb_unpack_4bit_float(const uint8_t* data, void* pol0, void* pol1, int64_t rowstart, int64_t rowend,
                          const int64_t* channels, int64_t ncols, int64_t nchan, void* stream);
/* unpacking.c:207 unpack_1bit_float / unpacking.py:165 _unpack_1bit_float_gpu. */
int alb_unpack_1bit_float(const uint8_t* data, void* pol0, void* pol1, int64_t rowstart, int64_t rowend,
                          int64_t chanstart, int64_t chanend, int64_t nchan, void* stream);
/* unpacking.c:248 sortpols (bit_depth 4 or 1). */
int alb_sortpols(const uint8_t* data, uint8_t* pol0, uint8_t* pol1, int64_t rowstart, int64_t rowend, int64_t ncols,
                 int64_t nchan, int bit_depth, int64_t chanstart, int64_t chanend, void* stream);
/* unpacking.c:7 hist_4bit / :76 hist_1bit -> uint64 [2^bit_depth][nchan] (no write past the end). */
int alb_hist(const uint8_t* data, uint64_t* hist, int64_t rowstart, int64_t rowend, int64_t nchan, int bit_depth,
             int mode, void* stream);
/* baseband_data_classes.py:48 make_continuous_gpu: zero-filled [nspec][nchans] complex64 with
 * out[(specnum[p]-start) % nspec][channels[k]] = spec[p*spec_ld + k]; columns with channels[k] < 0
 * (callers mark all but the last copy of a repeated channel so) or >= nchans are skipped. */
int alb_make_continuous(const void* spec, int64_t spec_ld, const int64_t* specnum, int64_t npresent, int64_t start,
                        int64_t nspec, const int64_t* channels, int64_t ncols, void* out, int64_t nchans,
                        void* stream);
/* unpacking_gpu.cu:24 dropped_packets: zero packet i when spec_num[i+1]-spec_num[i] != spectra_per_packet;
 * *num_dropped (device uint64) receives the count. */
int alb_dropped_packets(uint8_t* data, const int64_t* spec_num, int64_t num_packets, int64_t spectra_per_packet,
                        int64_t nchan, int bit_depth, uint64_t* num_dropped, void* stream);

/* ---- integer correlators (src/correlations/correlations_cpu.c) --------- */
/* correlations_cpu.c:39 avg_autocorr_4bit (autocorr=1) and :128 avg_xcorr_4bit (autocorr=0);
 * with match != NULL (from alb_match_rows) :188 avg_xcorr_4bit_2ant.  out: int64 [ncol][2] exact sums. */
int alb_xcorr_4bit_pair(const uint8_t* data0, const uint8_t* data1, int64_t nrows, int64_t ncol, const int* match,
                        int autocorr, int64_t* out, void* stream);
/* correlations_cpu.c:201-214 spec_num merge: match[i] = row of list 1 or -1, *count = matches (device int). */
int alb_match_rows(const int64_t* specnum0, const int64_t* specnum1, int64_t n0, int64_t n1, int64_t idxstart0,
                   int64_t idxstart1, int* match, int* count, void* stream);
/* correlations_cpu.c:273 avg_xcorr_1bit (vanvleck=0 -> int64 [nchan][2]) and :435 / :533
 * avg_xcorr_1bit_vanvleck(_2ant) (vanvleck=1 -> int64 [4][nchan] = R0, R1, IM0, IM1). */
int alb_xcorr_1bit_pair(const uint8_t* data0, const uint8_t* data1, int64_t nrows, int64_t ncol, int64_t nchan,
                        const int* match, int vanvleck, int64_t* out, void* stream);
/* All baselines i <= j of ninput per-pol 4-bit inputs [nrows][ncol] in one pass (the per-pair
 * correlations_cpu.c:128 applied to every pair, as scripts/autocorravg.py:114-116 does pair by pair);
 * out int64 [nrows/blocklen][npair][ncol][2].  inputs: HOST array of ninput device pointers. */
int alb_xcorr_4bit_allpairs(const uint8_t* const* inputs, int ninput, int64_t nrows, int64_t ncol, int64_t blocklen,
                            int64_t* out, void* stream);
/* Same for sortpols 1-bit inputs (correlations_cpu.c:273 on every pair); out int64 [nblock][npair][nchan][2]. */
int alb_xcorr_1bit_allpairs(const uint8_t* const* inputs, int ninput, int64_t nrows, int64_t ncol, int64_t nchan,
                            int64_t blocklen, int64_t* out, void* stream);

/* ---- float correlator (src/correlations/correlations_gpu.py, cgemm_batch.cu) -- */
/* correlations_gpu.py:95-121 avg_xcorr_all_ant_gpu, whose cgemm_strided_batched call (:114-119) passes
 * x as both operands: x complex64 F-order (M, K, F); out complex64 F-order (M, M, F) = x x^H * scale
 * (any M: the matrix cores for M <= 8, a VALU form above). */
int alb_xcorr_all_ant(const void* x, int64_t M, int64_t K, int64_t F, void* out, float scale, void* stream);
/* The same product in a chosen arithmetic on the matrix cores, FP32 accumulation:
 * mode 0 FP32 (= alb_xcorr_all_ant), 1 BF16 hi/lo split (hi hi^T + hi lo^T + lo hi^T),
 * 2 FP16, 3 FP16 hi/lo split; modes 1-3 take M <= 8 (ENOTSUP otherwise). */
int alb_xcorr_all_ant_prec(const void* x, int64_t M, int64_t K, int64_t F, void* out, float scale, int mode,
                           void* stream);
/* The matrix-core correlator on a strided input (no counterpart in the reference, whose
 * cgemm_strided_batched takes the F-order (M, K, F) block per job): S jobs of F channels,
 * element (input m, time k, channel f, job s) at x[s*ss + f*sf + k*sk + m] (complex64),
 * out as alb_xcorr_all_ant over the F*S channels (job-major).  Mode as alb_xcorr_all_ant_prec
 * (modes 1-3 take M <= 8; mode 0 above 8 runs the VALU form).  The batched re-PFB writes [S][K][F][M] (sk = F*M, sf = M). */
int alb_xcorr_all_ant_strided(const void* x, int64_t M, int64_t K, int64_t F, int64_t S, int64_t sk, int64_t sf,
                              int64_t ss, void* out, float scale, int mode, void* stream);

/* correlations.py:115-151 avg_xcorr_all_ant_cpu: out(i, j, f) = sum_k x(i,k,f) conj(x(j,k,f)) / K with every
 * product rounded to complex64 (FP32 multiplies and adds, no fused multiply-add) and the sum and the
 * division by K in float64, as the reference's Numba loop does; lower triangle = conjugate.
 * Element (m, k, f) of x at x[m*sm + k*sk + f*sf], output (i, j, f) at out[i*oi + j*oj + f*of]
 * (complex64 element strides; the reference's C-order output is oi = M*F, oj = F, of = 1). */
int alb_xcorr_all_ant_f64acc(const void* x, int64_t M, int64_t K, int64_t F, int64_t sm, int64_t sk, int64_t sf,
                             void* out, int64_t oi, int64_t oj, int64_t of, void* stream);
/* cgemm_batch.cu:8-54 cgemm_strided_batched: C = A B^H for batch_count column-major batches,
 * A (M, K), B (N, K), C (M, N) complex64 with batch strides M*K, N*K, M*N (alpha 1, beta 0). */
int alb_cgemm_strided_batched(const void* A, const void* B, void* C, int64_t M, int64_t N, int64_t K,
                              int64_t batch_count, void* stream);

/* ---- PFB / inverse PFB (src/utils/pfb_gpu_utils.py, pycufft.py/.cu) ---- */
/* pfb_gpu_utils.py:123 get_matft + :29 calculate_filter in float64: win float32 [ntap*lblock];
 * matft / filt complex64 [lblock][nslice/2+1] (either may be NULL). */
int alb_matft_filter(const float* win, int ntap, int64_t lblock, int64_t nslice, double thresh, void* matft,
                     void* filt, void* stream);
/* pfb_gpu_utils.py:29 calculate_filter from a given complex64 matft (float64 arithmetic). */
int alb_filter_from_matft(const void* matft, int64_t n, double thresh, void* filt, void* stream);
/* pycufft.py:139 irfft / pycufft.cu:159 cufft_c2r_mohan (+ the /n of pycufft.py:153), along rows. */
int alb_irfft_rows(const void* in, int64_t ldi, float* out, int64_t ldo, int64_t nrows, int64_t n, void* stream);
/* pycufft.py:120 rfft / pycufft.cu:68 cufft_r2c_mohan along rows; bins [c0, c1) written. */
int alb_rfft_rows(const float* in, int64_t ldi, void* out, int64_t ldo, int64_t nrows, int64_t n, int64_t c0,
                  int64_t c1, void* stream);
/* pycufft.py:92 _fft / pycufft.cu:122 cufft_c2c along rows; direction -1 forward, +1 inverse (with 1/n). */
int alb_fft_rows(const void* in, int64_t ldi, void* out, int64_t ldo, int64_t nrows, int64_t n, int direction,
                 void* stream);
/* pycufft.py:92-131 _fft / pycufft.cu:122 cufft_c2c along rows of n = 2^15..2^24 points (longer than
 * alb_fft_rows takes): x and out complex64 [nbatch][n] (they may be the same buffer), work a scratch
 * buffer of the same size (aliasing neither); direction -1 forward, +1 inverse (with 1/n). */
int alb_fft_large(const void* x, void* work, void* out, int64_t n, int64_t nbatch, int direction, void* stream);
/* src/utils/orbcomm_utils_gpu.py:39-52 coarse_xcorr: FFT lag correlation per channel of f1, f2 complex64
 * [N][nchan] (C order), zero-padded to P points (a power of two, N + dN <= P <= 2^24; the reference pads
 * to 2N, and any P >= N + dN gives the same lags): out complex64 [nchan][2 dN] holds lag -dN + i at
 * column i, sum_t f1[t + lag] conj(f2[t]) / (N - |lag|).  work: 3 * nchan * P complex64. */
int alb_lag_xcorr(const void* f1, const void* f2, int64_t N, int64_t nchan, int64_t dN, int64_t P, void* work,
                  void* out, void* stream);
/* src/utils/orbcomm_utils_gpu.py:13-37 apply_delay: out[t][c] = arr[t][c] exp(-2 pi i freqs[c] delay[t])
 * for arr complex64 [nspec][nchan], delay float64 [nspec], freqs float64 [nchan]: the phase and its
 * cosine and sine in float64, the product rounded to complex64 (out may be arr). */
int alb_apply_delay(const void* arr, const double* delay, const double* freqs, int64_t nspec, int64_t nchan, void* out,
                    void* stream);
/* pfb_gpu_utils.py:105 cupy_pfb for lblock <= 16384 (coarse PFB): channels [c0, c1) of
 * (nblock, lblock/2+1) written with row stride ldo. */
int alb_pfb_small(const float* ts, int64_t nblock, int64_t lblock, const float* win, int ntap, int64_t c0, int64_t c1,
                  void* out, int64_t ldo, void* stream);
/* pfb_gpu_utils.py:84-91 (the column half of cupy_ipfb) over ntiles tiles; kept rows
 * [cut, nslice-cut) written (helper_gpu_stream.py:199).  Tiles are row-major with row stride
 * ldy >= ncols, or (ldy = 8) in the blocked layout of alb_ipfb_rows_packed4_mfma. */
int alb_ipfb_columns(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int64_t nslice, int64_t ncols,
                     const void* filt, int64_t ldf, int cut, float* out, int64_t ldo, int64_t out_tile_stride,
                     void* stream);
/* The same column pass for any even nslice (chirp-z form with P-point FFTs, P the power of two
 * >= 2*nslice-1, P <= 16384): chirp complex64 [nslice] = exp(-i pi m^2/nslice), bhat complex64 [P]
 * = FFT_P of b (b[m] = conj(chirp[m]) for m < nslice, b[P-m] = conj(chirp[m]) for 0 < m < nslice). */
int alb_ipfb_columns_any(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int64_t nslice,
                         int64_t ncols, const void* filt, int64_t ldf, int cut, float* out, int64_t ldo,
                         int64_t out_tile_stride, const void* chirp, const void* bhat, int64_t P, void* stream);
/* Row half of the streaming iPFB with the 4-bit unpack and spec_num zero fill fused in
 * (unpacking.py:140 + baseband_data_classes.py:48 + pfb_gpu_utils.py:80, helper_gpu_stream.py:195-199):
 * bytes [rows][ncols] packed 4-bit spectra of one pol, rowmap [nchunk][acclen] from alb_build_rowmap,
 * chan[ncols] coarse channel numbers (a negative number drops the column, which is how a caller
 * keeps only the last copy of a repeated channel; numbers above 2048 are dropped too); tiles tile0 .. tile0+ntiles-1 of
 * nslice = acclen+2*cut rows each -> out [ntiles][nslice][4096] float32. */
int alb_ipfb_rows_packed4(const uint8_t* bytes, const int* rowmap, const int* chan, int64_t ncols, int64_t acclen,
                          int cut, int64_t tile0, int64_t ntiles, float* out, void* stream);
/* The same row pass written in the blocked tile layout out[ntiles][4096/8][nslice][8], read by
 * the column passes with ldy = 8 (each column strip one contiguous span). */
int alb_ipfb_rows_packed4_blocked(const uint8_t* bytes, const int* rowmap, const int* chan, int64_t ncols,
                                  int64_t acclen, int cut, int64_t tile0, int64_t ntiles, float* out, void* stream);
/* The same row pass as a matrix-core synthesis from the ncols non-zero channels
 * (irfft of pycufft.py:139-161 as y = X T with a float64-built FP16 hi/lo cos/sin table T,
 * FP32 accumulation), written in the blocked tile layout out[ntiles][4096/8][nslice][8]
 * that the column passes take with ldy = 8.  coeff: alb_ipfb_rows_coeff_bytes(ncols) bytes
 * filled once by alb_ipfb_rows_coeff from the channel numbers. */
int alb_ipfb_rows_packed4_mfma(const uint8_t* bytes, const int* rowmap, const int* chan, int64_t ncols, int64_t acclen,
                               int cut, int64_t tile0, int64_t ntiles, const void* coeff, float* out, void* stream);
int64_t alb_ipfb_rows_coeff_bytes(int64_t ncols);
int alb_ipfb_rows_coeff(const int* chan, int64_t ncols, void* coeff, void* stream);
/* pfb_gpu_utils.py:105 cupy_pfb with lblock = 4096*osamp (re-PFB, helper_gpu_stream.py:222), osamp a power
 * of two up to 65536 (the reference's production runs use 65536, scripts/xcorr/config_axion_gpu.json:13;
 * above 64 stage 2 sums each kept channel directly, for the few channels per bin such runs keep):
 * fine channels [c_lo, c_hi) of nblock spectra; spectrum b, channel c goes to
 * out[(b/bpj)*sj + (b%bpj)*sb + (c-c_lo)*sc] (bpj = blocks per re-PFB job, so all jobs of a
 * stream land in the correlator's F-order (M, nblock, nchan, njobs) layout); A is workspace of
 * alb_repfb_workspace_bytes(nblock, osamp) bytes. */
int alb_repfb(const float* ts, const float* win, int ntap, int64_t nblock, int64_t osamp, void* A, int64_t c_lo,
              int64_t c_hi, void* out, int64_t sb, int64_t sc, int64_t bpj, int64_t sj, void* stream);
int64_t alb_repfb_workspace_bytes(int64_t nblock, int64_t osamp);
/* The two halves of alb_repfb (same arguments): polyphase window + osamp-point DFT down each
 * column into A, then twiddle + 4096-point FFT along the columns into the fine channels. */
int alb_repfb_stage1(const float* ts, const float* win, int ntap, int64_t nblock, int64_t osamp, void* A,
                     void* stream);
int alb_repfb_stage2(const void* A, int64_t nblock, int64_t osamp, int64_t c_lo, int64_t c_hi, void* out, int64_t sb,
                     int64_t sc, int64_t bpj, int64_t sj, void* stream);
/* alb_repfb_stage2 for ninput inputs in one launch (no counterpart in the reference, which runs
 * cupy_pfb per input, helper_gpu_stream.py:222): input m's workspace at A + m*a_stride bytes
 * (a_stride >= alb_repfb_workspace_bytes(nblock, osamp)), its channels at out + m*sm complex
 * elements plus the alb_repfb_stage2 addressing.  With the correlator layout (sm = 1) the inputs'
 * neighbouring elements are written by neighbouring workgroups. */
int alb_repfb_stage2_inputs(const void* A, int64_t a_stride, int64_t ninput, int64_t nblock, int64_t osamp,
                            int64_t c_lo, int64_t c_hi, void* out, int64_t sm, int64_t sb, int64_t sc, int64_t bpj,
                            int64_t sj, void* stream);
/* The iPFB column pass (alb_ipfb_columns, ncols = 4096, row stride 4096) fused with re-PFB stage 1:
 * the time stream is the input of pfb_gpu_utils.py:105 cupy_pfb at lblock = 4096*osamp
 * (helper_gpu_stream.py:199-222).  Kept row i of tile c is stream row
 * g = row0 + c*rows_per_tile + (i - cut).  For every re-PFB block b < nblocks whose stream rows
 * [osamp*b, osamp*(b+ntap)) all lie in one tile (alb_repfb_interior) A[b] is produced in the
 * same launch; only stream rows read by the other blocks are written to ts ([rows][4096]).
 * Those other blocks go through alb_repfb_stage1_list once every tile that feeds them is done.
 * Tiles row-major (ldy >= 4096) or blocked (ldy = 8, alb_ipfb_rows_packed4_mfma).  row0 may be negative:
 * ts is then a window of the stream whose row 0 is the first row of block 0, rows before it are not
 * written, and ts must hold every row >= 0 of the tiles (a stream processed a group of jobs at a time).
 * nslice 4096 with osamp 32 or 64 (ENOTSUP otherwise: use alb_ipfb_columns + alb_repfb_stage1).
 * With ntiles = 0 only the shape (nslice, cut, ntap, osamp) is checked; pointers may be null. */
int alb_ipfb_columns_repfb(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int64_t nslice,
                           const void* filt, int64_t ldf, int cut, float* ts, int64_t row0, int64_t rows_per_tile,
                           const float* win, int ntap, int64_t osamp, int64_t nblocks, void* A, void* stream);
/* Spectral-order streaming iPFB (pfb_gpu_utils.py:56-91 cupy_ipfb: dd = irfft(dat, axis=1), then
 * rfft of dd along time).  The two transforms act on different axes, so the one along time runs
 * first, on the ncols non-zero channel columns only; three launches replace the row pass and the
 * forward half of the column pass (alb_ipfb_rows_packed4_blocked + alb_ipfb_columns_repfb), nslice 4096.
 * Step 1: E [ntiles][nslice][ncols] complex64 = FFT along the nslice rows of each tile (rows, bytes,
 * rowmap, chan and zero fill exactly as alb_ipfb_rows_packed4; channels 0 and 2048 keep the real part,
 * which is all irfft reads of them, pycufft.py:139-161). */
int alb_ipfb_chan_fft_packed4(const uint8_t* bytes, const int* rowmap, const int* chan, int64_t ncols, int64_t acclen,
                              int cut, int64_t tile0, int64_t ntiles, void* E, void* stream);
/* Step 2: D [ntiles][1024][nslice/2+1][4] complex64, D[t][s][f][c] = 4096 * ddft[4s+c][f] where
 * ddft = rfft(irfft(dat, axis=1), axis=0) of tile t (the reference's ddft before the filter,
 * pfb_gpu_utils.py:84-87), from step 1's E: one inverse 4096-point transform per f of the bins
 * X_f[chan] = E[f], X_f[4096 - chan] = conj(E[-f]).  chan must hold distinct numbers (repeats -1). */
int alb_ipfb_rows_spectral(const void* E, const int* chan, int64_t ncols, int64_t nslice, int64_t ntiles, void* D,
                           void* stream);
/* Step 3: alb_ipfb_columns_repfb (same arguments and outputs) on step 2's D in place of the tiles:
 * Wiener filter (pfb_gpu_utils.py:87), inverse transform along time, re-PFB stage 1 of the interior
 * blocks, the stream rows of the others.  nslice 4096 with osamp 32 or 64. */
int alb_ipfb_columns_repfb_spectral(const void* D, int64_t ntiles, int64_t nslice, const void* filt, int64_t ldf,
                                    int cut, float* ts, int64_t row0, int64_t rows_per_tile, const float* win,
                                    int ntap, int64_t osamp, int64_t nblocks, void* A, void* stream);
/* Split form of alb_ipfb_columns_repfb and its spectral step 3 (no counterpart call in the reference;
 * it computes the same cupy_pfb input, helper_gpu_stream.py:183-222, with no stream buffer): each
 * tile computes re-PFB stage 1 of every block that touches its kept rows, over those rows only.  The
 * tile that holds block b's first row writes A[b]; the next tile writes its half to P slot
 * ntap*(tile0 + t - 1) + (b - first block straddling tile t's first row), t the tile within this
 * launch and tile0 this launch's first tile's index in the stream window, whose tile 0 (first kept
 * row row0 - tile0*rows_per_tile) must start at or before window row 0 (ALB_EINVAL otherwise).  Consecutive tiles must be rows_per_tile = nslice - 2*cut rows apart.
 * P holds p_slots slots of alb_repfb_split_slot_bytes(osamp) bytes; launches that would write past
 * them fail with ALB_EINVAL.  nslice 4096, osamp 64, ntap <= 4, strip width 4 (ENOTSUP otherwise;
 * with ntiles = 0 only the shape is checked). */
int alb_ipfb_columns_repfb_split(const float* y, int64_t ldy, int64_t tile_stride, int64_t ntiles, int64_t nslice,
                                 const void* filt, int64_t ldf, int cut, int64_t row0, int64_t rows_per_tile,
                                 const float* win, int ntap, int64_t osamp, int64_t nblocks, void* A, void* P,
                                 int64_t tile0, int64_t p_slots, void* stream);
int alb_ipfb_columns_repfb_spectral_split(const void* D, int64_t ntiles, int64_t nslice, const void* filt, int64_t ldf,
                                          int cut, int64_t row0, int64_t rows_per_tile, const float* win, int ntap,
                                          int64_t osamp, int64_t nblocks, void* A, void* P, int64_t tile0,
                                          int64_t p_slots, void* stream);
/* alb_repfb_stage2_inputs (osamp 64) after the split column pass: each block that straddles two of
 * the window's ntiles tiles gets its second half from P (input m at P + m*p_stride bytes) added to
 * A before the transform.  The window's tile c starts at window row g_first + c*rows_per_tile
 * (g_first <= 0); every block's first row must lie in its tiles. */
int alb_repfb_stage2_split(const void* A, int64_t a_stride, const void* P, int64_t p_stride, int64_t ninput,
                           int64_t nblock, int64_t osamp, int64_t c_lo, int64_t c_hi, void* out, int64_t sm, int64_t sb,
                           int64_t sc, int64_t bpj, int64_t sj, int64_t g_first, int64_t rows_per_tile, int64_t ntiles,
                           int ntap, void* stream);
/* Bytes of one P slot (one block of stage-1 workspace). */
int64_t alb_repfb_split_slot_bytes(int64_t osamp);
/* The split form's index functions as the kernels use them (host-only, no GPU work): the blocks
 * [*lo, *hi) of nblocks that touch a tile's kept rows [g0, g0 + nkeep) and *e0, the first block
 * straddling row g0; and the P slot stage 2 adds to block b (-1: none; -2: arguments out of range). */
int alb_repfb_split_blocks(int64_t g0, int64_t nkeep, int64_t osamp, int ntap, int64_t nblocks, int64_t* lo,
                           int64_t* hi, int64_t* e0);
int64_t alb_repfb_split_slot(int64_t b, int64_t g_first, int64_t rows_per_tile, int64_t ntiles, int ntap,
                             int64_t osamp);
/* alb_repfb with its stage-1 workspace in a fixed device array of the library instead of A: allocates
 * nothing (no counterpart in the reference; for callers that cannot call hipMalloc -- code under hipGraph
 * capture, tight serving loops, harnesses that forbid allocation).  Requires osamp <= 64 and
 * alb_repfb_workspace_bytes(nblock, osamp) <= alb_repfb_noalloc_max_bytes() (64 MiB: 64 blocks at
 * osamp 64), else ALB_EINVAL before any launch.  Each device holds its own array until the process
 * ends; calls that overlap on one device share it and race. */
int alb_repfb_noalloc(const float* ts, const float* win, int ntap, int64_t nblock, int64_t osamp, int64_t c_lo,
                      int64_t c_hi, void* out, int64_t sb, int64_t sc, int64_t bpj, int64_t sj, void* stream);
int64_t alb_repfb_noalloc_max_bytes(void);
/* Interior blocks [*lo, *hi) of a tile whose kept rows are stream rows [g0, g0 + nkeep) (g0 may be
 * negative; blocks start at 0). */
int alb_repfb_interior(int64_t g0, int64_t nkeep, int64_t osamp, int ntap, int64_t nblocks, int64_t* lo, int64_t* hi);
/* alb_repfb_stage1 for the n blocks listed in the host array blocks (block b reads ts rows
 * [osamp*b, osamp*(b+ntap)), writes A[b]); osamp 16, 32 or 64. */
int alb_repfb_stage1_list(const float* ts, const float* win, int ntap, const int* blocks, int64_t n, int64_t osamp,
                          void* A, void* stream);

/* ---- streaming helpers (scripts/xcorr/helper_gpu_stream.py) ------------ */
/* Row map of one chunk for the fused iPFB: rowmap[(specnum[p]-start) % acclen] = row_base + p
 * (the packed row holding spectrum p), -1 elsewhere (baseband_data_classes.py:53 modular row of
 * make_continuous_gpu); of several spectra on one row the last (largest p) is kept. */
int alb_build_rowmap(const int64_t* specnum, int64_t npresent, int64_t start, int64_t acclen, int64_t row_base,
                     int* rowmap, void* stream);

/* ---- tuning (no counterpart in the reference) ---------------------------- */
/* Set a launch-shape knob; 0 restores the built-in choice.  Keys:
 *   "ipfb_cols_width"  columns per workgroup of the nslice = 4096 column pass (2, 4 or 8, default 4);
 *   "ipfb_cols_group"  column strips per XCD group of the column-pass order (2..64; unset: 2 for launches of up to 4 tiles, 64 above);
 *   "ipfb_rows_group"  spectra per workgroup of the packed 4-bit row pass (2 or 4; unset: 2 for launches of up to 4 tiles, 4 above);
 *   "xcorr_int_method" all-pairs integer correlators: 1 VALU (XOR form for 1-bit),
 *                      2 int8 matrix cores (blocklen % 64 == 0; 1-bit also ncol % 4 == 0),
 *                      0 matrix cores whenever the shape allows. */
int alb_tune_set(const char* key, int64_t value);
int64_t alb_tune_get(const char* key);

#ifdef __cplusplus
}
#endif
#endif /* ALBATROS_HIP_H */
