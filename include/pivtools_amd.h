/*
 * pivtools_amd.h -- C ABI of the MI355X correlation hot path.
 *
 * Two layers, both plain C (no HIP or torch types in any signature):
 *
 *  1. Drop-in host-pointer entry points with the exact names, arguments,
 *     ownership and error codes of the reference's ctypes-bound libraries
 *     libbulkxcorr2d / libfusedwarp (built here as libbulkxcorr2d.so and
 *     libfusedwarp.so).  Each call is synchronous: host -> HBM, kernels,
 *     HBM -> host.  Reference declarations cited per function.
 *
 *  2. The device-pointer API (pivamd_*) underneath, exported by
 *     libpivtools_amd.so: every pointer is device memory, work is queued on
 *     the given stream (a hipStream_t passed as void*), nothing is copied.
 *     The Python host layer and the benchmark drive this layer.
 *
 * Error codes of layer 1 are the reference's (pivtools_cli/lib/common.h:31-36):
 *   0 none, 1 NOMEM, 2 NOPLAN_FWD, 4 NOPLAN_BWD (unsupported window size),
 *   8 NOPLAN (selector misuse / device failure), 9 OUT_OF_BOUNDS.
 * Layer 2 returns 0 on success, 1 bad argument, 2 unsupported size,
 *   3 launch failure, 4 out of device memory, 5 runtime error.
 */
#ifndef PIVTOOLS_AMD_H
#define PIVTOOLS_AMD_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================= */
/* Layer 1: libbulkxcorr2d drop-in                                          */
/* ======================================================================= */

/* Replaces bulkxcorr2d, pivtools_cli/lib/PIV_2d_cross_correlate.h:18-21
 * (ctypes binding pivtools_cli/piv/piv_backend/cpu_instantaneous.py:118-165).
 * Outputs (N, nPeaks, nWy, nWx) at n*nPeaks*nWin + p*nWin + w; fPkRatio
 * (N, nWy, nWx); masked or out-of-bounds windows are left untouched;
 * fCorrelPlane_Out (nullable) receives the raw planes (N*nWin*h*w). */
unsigned char bulkxcorr2d(const float *fImageA, const float *fImageB, const float *fMask,
                          const int *nImageSize, int N_images, const float *fWinCtrsX,
                          const float *fWinCtrsY, const int *nWindows, float *fWindowWeightA,
                          bool bEnsemble, const float *fWindowWeightB, const int *nWindowSize,
                          int nPeaks, int iPeakFinder, float *fPkLocX, float *fPkLocY,
                          float *fPkHeight, float *fSx, float *fSy, float *fSxy, float *fPkRatio,
                          float *fCorrelPlane_Out);

/* Replaces bulkxcorr2d_accumulate, PIV_2d_cross_correlate.h:37-51.
 * Adds (+=) the central nFitWindowSize crop of every pair's AB plane into
 * fCorrelPlane_Sum (nWin * fh * fw); the caller zeroes it. */
unsigned char bulkxcorr2d_accumulate(const float *fImageA_stack, const float *fImageB_stack,
                                     const float *fMask, const int *nImageSize, int N_images,
                                     const float *fWinCtrsX, const float *fWinCtrsY,
                                     const int *nWindows, const float *fWindowWeightA,
                                     const float *fWindowWeightB, const int *nWindowSize,
                                     const int *nFitWindowSize, float *fCorrelPlane_Sum);

/* Replaces bulkxcorr2d_accumulate_triple, PIV_2d_cross_correlate.h:66-86
 * (ctypes binding unit-tests/test_ensemble_options.py:43-65). */
unsigned char bulkxcorr2d_accumulate_triple(
    const float *fImageA_stack, const float *fImageB_stack, const float *fMask,
    const int *nImageSize, int N_images, const float *fWinCtrsX, const float *fWinCtrsY,
    const int *nWindows, const float *fWindowWeightA_AB, const float *fWindowWeightB_AB,
    const float *fAutoWeightA, const float *fAutoWeightB, const int *nWindowSize,
    const int *nFitWindowSize, int bMeanSubtract, int bPerPairNorm, float *fCorrAB_Sum,
    float *fCorrAA_Sum, float *fCorrBB_Sum);

/* PIV_2d_cross_correlate.h:88-89 */
float fminvec(const float *fVec, int n);
float fmaxvec(const float *fVec, int n);

/* PIV_2d_cross_correlate.h:97-99.  fft_s accumulates device seconds of the
 * fused correlate-and-fit kernel; fit_s stays 0 (the fit is fused in). */
void bulkxcorr2d_set_timing_enabled(int on);
void bulkxcorr2d_reset_timing(void);
void bulkxcorr2d_get_timing(double *fft_s, double *fit_s);

/* PIV_2d_cross_correlate.h:107-109.  0 = scalar fitter flavour (libm
 * exponential), 1 = batch flavour (Cephes exponential); both run on the
 * device, so batch is always available. */
int bulkxcorr2d_set_peakfit_impl(int impl);
int bulkxcorr2d_get_peakfit_impl(void);
int bulkxcorr2d_peakfit_batch_available(void);

/* peak_locate_lm.h:39-51.  One plane (any N[0] x N[1] >= 8 x 8). */
void lsqpeaklocate_lm(const float *xcorr, const int *N, float *peak_loc, int nPeaks,
                      int iFitType, float *std_dev, const float *fPlaneWeight);
float peak_detectability(const float *xcorr, const int *N);
/* Returns 1 when a HIP device is present (the loader's guard,
 * pivtools_cli/piv/piv_backend/lib_guard.py:6-26). */
int pivtools_cpu_supported(void);

/* peak_locate_lm_batch.h:39-51.  Lane width W = peakfit_batch_lanes() = 16,
 * the reference's PEAKFIT_MAX_LANES (:26-28): fits min(L_real, 16) planes and
 * writes exactly 3*W floats to each of peak_loc and std_dev. */
void lsqpeaklocate_lm_batch(const float *planes, int L_real, const int *N, int iFitType,
                            float *peak_loc, float *std_dev, const float *fPlaneWeight);
int peakfit_batch_available(void);
int peakfit_batch_lanes(void);

/* xcorr.h:26-27.  convolve: centre crop of the 2N zero-padded correlation
 * (any size).  xcorr: shift(IFFT(FFT(a) .* conj(FFT(b))))/numel for built
 * sizes up to 128 (192/256 return 4). */
unsigned convolve(const float *w1, const float *w2, float *conv, const int *N);
unsigned xcorr(const float *w1, const float *w2, float *corr, const int *N);
/* BUILD-SIMD.md verify step.  Returns 16, equal to peakfit_batch_lanes(),
 * which the reference's batch gate requires (PIV_2d_cross_correlate.c:231-239). */
int codelet_lanes(void);

/* ======================================================================= */
/* Layer 1: libfusedwarp drop-in                                            */
/* ======================================================================= */

/* fused_warp.h:21-22.  Both values run the same device kernel. */
void fused_warp_set_impl(int impl);
int fused_warp_get_impl(void);

/* fused_warp.h:55-66 */
int fused_symmetric_warp(const float *img_a, const float *img_b, float *out_a, float *out_b,
                         const float *pred_dy, const float *pred_dx, int H, int W, int nPY,
                         int nPX, const float *ctrs_y, const float *ctrs_x, int interp_mode,
                         int round_shifts);

/* fused_warp.h:78-93 */
int fused_symmetric_warp_batch(const float *imgs_a, const float *imgs_b, float *outs_a,
                               float *outs_b, const float *pred_dy, const float *pred_dx, int N,
                               int H, int W, int nPY, int nPX, const float *ctrs_y,
                               const float *ctrs_x, int interp_mode, int shared_predictor,
                               int round_shifts);

/* ======================================================================= */
/* Layer 1: libkspacefit drop-in                                            */
/* ======================================================================= */

/* Replaces kspace_lm_fit_batch, pivtools_cli/lib/kspace_lm_fit.h:54-61
 * (caller pivtools_cli/piv/piv_backend/kspace_lm_fitting.py:877-900).
 * Same layouts and contract: planes (n_windows, corr_h, corr_w) FP64,
 * P_win centred or NULL, every output written.  n_threads is accepted and
 * ignored (one workgroup per window).  Returns 0, -1 (axis length outside
 * 8/12/16/24/32/48/64/96/128) or -2 (device failure). */
int kspace_lm_fit_batch(const double *R_AA, const double *R_BB, const double *R_AB,
                        const unsigned char *mask_flat, const double *P_win, int n_windows, int corr_h,
                        int corr_w, int use_kx4, int use_ky4, int n_threads, double *gauss_flat,
                        int *status_flat, double *initial_guess_flat, double *diag_gain, double *diag_N0,
                        double *diag_b4x, double *diag_b4y, double *diag_cost_per_pt, int *diag_n_valid,
                        int *diag_iter, unsigned char *diag_conv);
/* kspace_lm_fit.h:65.  Loader probe: compute units of the current device. */
int kspace_lm_fit_max_threads(void);
/* kspace_lm_fit.h:79-82.  D in natural order and the coloured N0 seed. */
int kspace_debug_prep(const double *R_AA, const double *R_BB, const double *R_AB, const double *P_centred,
                      int H, int W, double *out_Dv, double *out_n0_seed);
/* kspace_lm_fit.h:86-87.  fftshift(fft2(ifftshift(plane))) in natural order. */
int kspace_fft2_centered(const double *plane, int H, int W, double *out_re, double *out_im);

/* ======================================================================= */
/* Layer 2: device-pointer API (libpivtools_amd.so)                          */
/* ======================================================================= */

/* 1 if an NY x NX window has a correlation kernel. */
int pivamd_size_supported(int ny, int nx);

/* Transform of pivamd_xcorr: 0 = FP32 FFT (default), 1 = split-precision
 * MFMA DFT (fp16 hi/lo operands, FP32 accumulation) for 16/32/48/64-point
 * windows.  Returns the previous mode. */
int pivamd_set_dft_mode(int mode);

/* Launch variants of the correlation kernel per square window size N (no
 * counterpart in the reference, whose C path has one build per size):
 * R = 1, 2, 4, 8 is the FP32 FFT with each line transform shared by R
 * lanes, 16 the split-precision MFMA DFT, 36 the 4-lane FFT as a persistent
 * kernel that prefetches the next window (N >= 64).  _variants writes up to max_out codes
 * and returns how many are built; _set_variant selects one for N
 * (0 = default, process-wide, like bulkxcorr2d_set_peakfit_impl) and
 * returns the previous selection, or -1 if v is not built for N;
 * _default_variant and _get_variant return -1 for an unsupported N.  The
 * autotuner (pivtools_amd/autotune.py) times them and keeps the fastest. */
int pivamd_xcorr_variants(int N, int *out, int max_out);
int pivamd_xcorr_default_variant(int N);
int pivamd_xcorr_set_variant(int N, int v);
int pivamd_xcorr_get_variant(int N);

/* d_out = centre crop of the 2N zero-padded correlation of a and b (FP64). */
int pivamd_convolve(const float *d_a, const float *d_b, int ny, int nx, float *d_out, void *stream);

/* d_out[ny*nx] = numel / (wB (*) wB), the plane weight of bulkxcorr2d. */
int pivamd_plane_weight(const float *d_wB, int ny, int nx, float *d_out, void *stream);

/* One bulkxcorr2d over device buffers.  d_mask may be NULL (no mask);
 * d_planes may be NULL.  impl: 0 scalar, 1 batch fitter flavour.  d_work is
 * caller-owned device scratch of pivamd_xcorr_workspace() bytes. */
int pivamd_xcorr(const float *d_imgA, const float *d_imgB, const float *d_mask, int H, int W,
                 int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                 const float *d_wA, const float *d_wB, const float *d_planeWeight, int ensemble,
                 int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX, float *d_pkY,
                 float *d_pkH, float *d_sx, float *d_sy, float *d_sxy, float *d_ratio,
                 float *d_planes, void *d_work, unsigned long work_bytes, void *stream);

/* No counterpart in the reference: pivamd_xcorr without the workspace
 * argument, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * makes no allocation: the peak records (pivamd_xcorr_workspace() bytes) live
 * in a file-scope device array of pivamd_xcorr_noalloc_max_bytes() = 64 MiB
 * per GPU, held until the process ends; a larger call returns PIVAMD_ERR_SIZE
 * (2) with a message.  All calls on one GPU share that array, so calls that
 * overlap on one GPU (other streams or threads) race. */
int pivamd_xcorr_noalloc(const float *d_imgA, const float *d_imgB, const float *d_mask, int H, int W,
                         int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                         const float *d_wA, const float *d_wB, const float *d_planeWeight, int ensemble,
                         int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX,
                         float *d_pkY, float *d_pkH, float *d_sx, float *d_sy, float *d_sxy,
                         float *d_ratio, float *d_planes, void *stream);
unsigned long pivamd_xcorr_noalloc_max_bytes(void);

/* No counterpart in the reference (its caller makes one bulkxcorr2d call per
 * batch, pivtools_cli/piv/piv_backend/cpu_instantaneous.py:505-528): one
 * ragged launch of the correlation for up to 8 jobs -- pivamd_xcorr calls of
 * one square window size over different batches (image sizes, window grids,
 * buffers) -- whose windows share one grid of workgroups, then each job's
 * fit.  Per job the results equal pivamd_xcorr's.  FP32 FFT forms only (the
 * split DFT mode is not used); returns PIVAMD_ERR_ARG for more than 8 jobs
 * or a job's missing buffer, PIVAMD_ERR_SIZE for a non-square or unbuilt
 * size.  pivamd_xcorr_job_size() is sizeof(PivamdXcorrJob). */
typedef struct PivamdXcorrJob {
    const void *imgA, *imgB;        /* frames (n_images, H, W): float32, or 16-bit counts when u16 != 0 */
    int u16;
    int H, W, n_images;
    const float *mask, *ctrX, *ctrY; /* mask nullable */
    int nWy, nWx;
    const float *wA, *wB, *planeWeight;
    float *pkX, *pkY, *pkH, *sx, *sy, *sxy, *ratio;
    float *planes;                   /* nullable */
    void *work;                      /* pivamd_xcorr_workspace() bytes */
    unsigned long work_bytes;
} PivamdXcorrJob;
int pivamd_xcorr_ragged(const PivamdXcorrJob *jobs, int n_jobs, int winY, int winX, int nPeaks,
                        int finder, int impl, void *stream);
int pivamd_xcorr_job_size(void);

/* Device workspace bytes pivamd_xcorr needs (0 in ensemble mode). */
unsigned long pivamd_xcorr_workspace(int n_images, int nWy, int nWx, int winY, int winX, int nPeaks,
                                     int ensemble);

/* One fused_symmetric_warp_batch over device buffers.  interp_mode 0:
 * bicubic and 1: Lanczos-3, both bit-identical to fused_warp.c; 2: the same
 * bicubic evaluated separably with fused multiply-adds (last-bit differences
 * only, about half the instructions).  Other values: PIVAMD_ERR_ARG. */
int pivamd_warp(const float *d_imgA, const float *d_imgB, float *d_outA, float *d_outB,
                const float *d_predDy, const float *d_predDx, int N, int H, int W, int nPY,
                int nPX, const float *d_ctrY, const float *d_ctrX, int interp_mode,
                int shared_predictor, int round_shifts, void *stream);

/* pivamd_xcorr and pivamd_warp over 16-bit camera counts (uint16 frames,
 * converted exactly to float on load): the same results as the float entry
 * points on the converted frames, without a float copy of the frames in
 * HBM.  The reference takes float frames only (pivtools_cli/lib/
 * PIV_2d_cross_correlate.h:18-21, fused_warp.h:55-93); its caller converts
 * the camera frames to float first. */
int pivamd_xcorr_u16(const unsigned short *d_imgA, const unsigned short *d_imgB, const float *d_mask,
                     int H, int W, int n_images, const float *d_ctrX, const float *d_ctrY, int nWy,
                     int nWx, const float *d_wA, const float *d_wB, const float *d_planeWeight,
                     int ensemble, int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX,
                     float *d_pkY, float *d_pkH, float *d_sx, float *d_sy, float *d_sxy,
                     float *d_ratio, float *d_planes, void *d_work, unsigned long work_bytes,
                     void *stream);
int pivamd_warp_u16(const unsigned short *d_imgA, const unsigned short *d_imgB, float *d_outA,
                    float *d_outB, const float *d_predDy, const float *d_predDx, int N, int H, int W,
                    int nPY, int nPX, const float *d_ctrY, const float *d_ctrX, int interp_mode,
                    int shared_predictor, int round_shifts, void *stream);

/* Form of the interp_mode 2 warp, process-global (no counterpart in the
 * reference beyond its own form selector, fused_warp.h:21-22): 1 (default)
 * stages each tile's source rectangles of A and B in LDS as floats, 0 reads
 * every tap from global memory.  Both give the same values (signed zeros
 * aside).  set returns PIVAMD_ERR_ARG for other values. */
int pivamd_warp_set_variant(int v);
int pivamd_warp_get_variant(void);

/* Ensemble sums over device buffers.  triple = 1: bulkxcorr2d_accumulate_triple
 * (AB, AA, BB with optional mean removal and per-pair normalisation; sumW*
 * are the weight sums of the mean removal); triple = 0:
 * bulkxcorr2d_accumulate (AB only, weights d_wAab / d_wBab).  Sums are
 * added (+=) to the fitY x fitX central crops in d_out*; d_work is
 * pivamd_ensemble_workspace() bytes of device scratch (FP64 spectra). */
unsigned long pivamd_ensemble_workspace(int nWy, int nWx, int winY, int winX);
int pivamd_accumulate(const float *d_A, const float *d_B, const float *d_mask, int H, int W,
                      int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                      const float *d_wAab, const float *d_wBab, const float *d_wAA,
                      const float *d_wBB, int winY, int winX, int fitY, int fitX, int meanSub,
                      int perPairNorm, int triple, float sumWBab, float sumWAA, float sumWBB,
                      float *d_outAB, float *d_outAA, float *d_outBB, void *d_work,
                      unsigned long work_bytes, void *stream);

/* Camera payload ingest (device buffers).  The reference decodes frames on
 * the host (pivtools_core/image_handling/readers/set_reader.py:1296-1350
 * decode_mono12p, pivraw_reader.py:896-961 _sensor_pixels/_decode_into);
 * these expand the payload in HBM into uint16 frames for pivamd_xcorr_u16 /
 * pivamd_warp_u16.  Frame f's payload starts at d_raw + f * frame_stride
 * (e.g. the .pivraw container's 4096-byte-aligned stride).
 *  - Mono12p: H*W pixels (H*W even) in H*W*3/2 bytes, PFNC LSB-first packing;
 *    d_out (n_frames, H, W).
 *  - BayerRG8: an H x W mosaic (H, W even) summed over 2x2 superpixels;
 *    d_out (n_frames, H/2, W/2).
 * d_dark is null or one dark frame in the sensor type (uint16 H x W for
 * Mono12p, uint8 H x W for BayerRG8) subtracted per photosite as
 * max(p, d) - d from every frame of the call (the caller passes the dark of
 * the frames' A/B role).  Returns PIVAMD_ERR_SIZE for an odd pixel count /
 * odd mosaic, PIVAMD_ERR_ARG for a stride shorter than one payload. */
int pivamd_unpack_mono12p(const void *d_raw, unsigned long frame_stride, int n_frames, int H, int W,
                          const unsigned short *d_dark, unsigned short *d_out, void *stream);
int pivamd_bayer_rg8_superpixel(const void *d_raw, unsigned long frame_stride, int n_frames, int H, int W,
                                const unsigned char *d_dark, unsigned short *d_out, void *stream);

/* DaVis .set stream decode (device buffers): the reference decodes a frame
 * stream's pixels, applies its file-backed corrections and its geometry on
 * the host (pivtools_core/image_handling/readers/set_reader.py:1196-1480
 * pixel encodings, :477-585 corrections and geometries, :1907-2000 the read
 * path); this does all three in one pass.  encoding: 0 mono-10p,
 * 1 mono-10pmsb, 2 mono-12p, 3 mono-12pmsb, 4 mono-12packed, 5 raw-16-bit,
 * 6 raw-8-bit.  d_scmos, d_dark, d_bgnd: null or one H x W uint16 image in
 * sensor coordinates each, applied in that order as clip(x - S + 300, 0,
 * 65535), max(x - D, 0), max(x - B, 0).  geometry: 0 none, 1 rotate-180,
 * 2 vertical-mirror, 3 horizontal mirror, 4 rotate-cw-90, 5 rotate-ccw-90,
 * 6 rotate-cw-90 + vertical mirror, 7 rotate-ccw-90 + vertical mirror; 4-7
 * write W x H frames.  Frame f's payload starts at d_raw + f * frame_stride;
 * d_out holds n_frames output frames.  PIVAMD_ERR_SIZE when H x W does not
 * fill whole groups of the packing.  pivamd_davis_payload_bytes gives one
 * frame's payload bytes (0 for such a size or an unknown encoding). */
unsigned long pivamd_davis_payload_bytes(int encoding, int H, int W);
int pivamd_davis_decode(const void *d_raw, unsigned long frame_stride, int n_frames, int H, int W,
                        int encoding, const unsigned short *d_scmos, const unsigned short *d_dark,
                        const unsigned short *d_bgnd, int geometry, unsigned short *d_out, void *stream);

/* The serial decompressors of the DaVis containers, host memory (no device
 * work).  pivamd_davis_tokens decodes nvals values of the DaVis token stream
 * (.im7 pack 1, IMX .set entries; davis_tokens.py:46-130) from data[*off_io]
 * with running value *prev_io, updating both and *written; returns 0, or
 * 1 truncated, 2 a 0x7f run past nvals, 3 a value outside 0..65535.
 * pivamd_lz4_block decompresses one LZ4 block (.im7 pack 20, LZ4 .set
 * entries; im7_reader.py:270-291) into dst and returns its size, or -1. */
int pivamd_davis_tokens(const unsigned char *data, unsigned long n, unsigned long *off_io,
                        unsigned short *out, unsigned long nvals, long long *prev_io, unsigned long *written);
long long pivamd_lz4_block(const unsigned char *src, unsigned long n, unsigned char *dst, unsigned long cap);

/* TIFF frames (generic_readers.py:81-122, read_tiff through tifffile, and its
 * _luma, :33-51).  Host memory: pivamd_tiff_lzw and pivamd_packbits decode one
 * strip or tile (compression 5, 32773) into dst and return its size (LZW: -1
 * for a malformed or old-style stream).  Device: pivamd_gray_from_samples
 * turns n frames of H x W pixels of S samples (interleaved, or planar S x H x
 * W) into float32 frames: the sample (S 1), OpenCV's BT.601 RGB2GRAY of the
 * first three (S 3 or 4; 8/16-bit (9798 R + 19235 G + 3735 B + 2^14) >> 15,
 * float 0.299 R + 0.587 G + 0.114 B), or the samples' mean (other S).
 * sample_type: 0 u8, 1 u16, 2 i8, 3 i16, 4 u32, 5 i32, 6 f32; 1 is returned
 * for luma of another type. */
long long pivamd_tiff_lzw(const unsigned char *src, unsigned long n, unsigned char *dst, unsigned long cap);
long long pivamd_packbits(const unsigned char *src, unsigned long n, unsigned char *dst, unsigned long cap);
/* PNG frames (generic_readers.py:125-176, read_png_jpeg for PNG): undo the
 * scanline filters of rows rows of rowbytes bytes (src: filter byte + row
 * each), bpp bytes per pixel; 0, or -1 for a filter type above 4.  The gray
 * conversion is pivamd_gray_from_samples, as for TIFF. */
int pivamd_png_unfilter(const unsigned char *src, unsigned long rows, unsigned long rowbytes, int bpp,
                        unsigned char *dst);
/* JPEG frames (generic_readers.py:125-176, read_png_jpeg for JPEG, which
 * decodes with libjpeg-turbo's defaults): baseline and extended sequential
 * Huffman 8-bit files, host memory.  pivamd_jpeg_info gives width, height and
 * channels (1 gray, 3 RGB) in info[0..2]; pivamd_jpeg_decode writes height x
 * width x channels bytes.  Status 0, or 1 malformed, 2 progressive, 3
 * arithmetic coding, 4 lossless / hierarchical, 5 12-bit, 6 two or four
 * components, 7 unsupported sampling factors, 8 out too small. */
int pivamd_jpeg_info(const unsigned char *data, unsigned long n, int *info);
int pivamd_jpeg_decode(const unsigned char *data, unsigned long n, unsigned char *out, unsigned long cap);
int pivamd_gray_from_samples(const void *d_src, int n_frames, int H, int W, int S, int planar, int sample_type,
                             float *d_out, void *stream);

/* Multipass glue (device buffers, planar fields [dy plane, dx plane]):
 * uint16 ingest, displacement gate + predictor add-back, normalised-median
 * test, nearest infill, edge padding, one axis of the Gaussian predictor
 * smoothing, bicubic predictor remap, and a plain copy (counter calibration). */
int pivamd_u16_to_f32(const void *d_in, float *d_out, unsigned long n, void *stream);
int pivamd_copy_f32(const float *d_in, float *d_out, unsigned long n, void *stream);
int pivamd_gate(const float *pkx, const float *pky, const float *pkh, const float *pred,
                const float *mask, int N, int nWin, float tx, float ty, float *u,
                unsigned char *nanm, signed char *reason, float *pkh_out, void *stream);
int pivamd_median(const float *u, const float *pkh, const float *mask, int N, int ny, int nx,
                  float eps, float thr, int size, unsigned char *nanm, signed char *reason,
                  float *uo, float *pkh_out, void *stream);
/* The outlier stage in general: the OR of the median test (use_median),
 * the peak-magnitude test pkh < peak_thr or NaN (use_peak,
 * outlier_detection.py:47-84) and precomputed flags d_extra (nullable),
 * never at masked windows, followed by the median stage's bookkeeping.
 * pivamd_median is this stage with the median test alone. */
int pivamd_outlier_stage(const float *u, const float *pkh, const float *mask, int N, int ny, int nx,
                         int use_median, float eps, float thr, int size, int use_peak, float peak_thr,
                         const unsigned char *d_extra, unsigned char *nanm, signed char *reason, float *uo,
                         float *pkh_out, unsigned char *choice, void *stream);
/* Several peaks per window (nPeaks > 1) and secondary-peak substitution
 * (cpu_instantaneous.py:586-760).  pivamd_gate_peaks gates every peak of the
 * (N, P, nWin) bulkxcorr2d outputs and adds the predictor (pxa, pya, pha),
 * and takes the primary peak as pivamd_gate does (choice = 1).
 * pivamd_outlier_detect is the outlier test without the final bookkeeping:
 * d_active null = the first round (reason 10), else re-judging the windows
 * of a substitution round (reason 12).  pivamd_peak_select is one round
 * (windows still invalid move to their next peak; usel / pk_sel gather every
 * window's current peak); pivamd_peak_final closes the stage. */
int pivamd_gate_peaks(const float *pkx, const float *pky, const float *pkh, const float *pred,
                      const float *mask, int N, int P, int nWin, float tx, float ty, float *pxa, float *pya,
                      float *pha, float *u, unsigned char *nanm, signed char *reason, float *pkh0,
                      unsigned char *choice, void *stream);
int pivamd_outlier_detect(const float *u, const float *pkh, const float *mask, int N, int ny, int nx,
                          int use_median, float eps, float thr, int size, int use_peak, float peak_thr,
                          const unsigned char *d_extra, const unsigned char *d_active, unsigned char *nanm,
                          signed char *reason, void *stream);
int pivamd_peak_select(const float *pxa, const float *pya, const float *pha, int N, int P, int nWin,
                       unsigned char *nanm, signed char *reason, unsigned char *choice, float *usel,
                       float *pk_sel, unsigned char *active, void *stream);
int pivamd_peak_final(const float *usel, const float *pk_sel, const float *mask, int N, int nWin,
                      unsigned char *nanm, signed char *reason, unsigned char *choice, float *uo,
                      float *pkh_out, void *stream);
/* The sigma (outlier_detection.py:244-313; field 0 = vector norm, 1 = ux,
 * 2 = uy) and divergence/vorticity (:316-373; NaN thresholds = 6 * 1.4826 *
 * MAD of the image's finite values) outlier methods on the gated planar
 * field u, masked windows read as NaN; flags are OR-ed into d_extra (one
 * byte per window) for pivamd_outlier_stage.  d_work is
 * pivamd_outlier_methods_workspace() bytes. */
unsigned long pivamd_outlier_methods_workspace(int N, int ny, int nx);
int pivamd_outlier_sigma(const float *u, const float *d_mask, int N, int ny, int nx, float sigma_thr,
                         int field, unsigned char *d_extra, void *d_work, unsigned long work_bytes,
                         void *stream);
int pivamd_outlier_div_vort(const float *u, const float *d_mask, int N, int ny, int nx, float div_thr,
                            float vort_thr, unsigned char *d_extra, void *d_work, unsigned long work_bytes,
                            void *stream);
int pivamd_infill_nearest(const float *u, const float *mask, int N, int ny, int nx, float py,
                          float px, float *uo, int *d_err, void *stream);
int pivamd_pad_edge(const float *u, int N, int ny, int nx, int pre_y, int pre_x, int nyp, int nxp,
                    float *out, void *stream);
/* Least-squares infill of the NaN vectors of each image
 * (infilling.py:144-260 _fill_least_squares via infill_harmonic (order 1)
 * and infill_biharmonic (order 2)); d_mask (nullable) marks excluded cells
 * (masked windows), neither donors nor filled.  FP64 normal equations solved
 * by conjugate gradients to a relative residual of 1e-13, one workgroup per
 * image; d_work is pivamd_infill_ls_workspace() bytes.  *d_err |= 2 if a
 * solve misses the reference's 1e-6 acceptance, |= 4 for fill cells no
 * difference row ties to a valid cell (left at 0; the reference raises). */
unsigned long pivamd_infill_ls_workspace(int N, int ny, int nx);
int pivamd_infill_ls(const float *u, const float *d_mask, int N, int ny, int nx, float py, float px,
                     int order, float *uo, void *d_work, unsigned long work_bytes, int *d_err, void *stream);
/* Masked windows of a predictor field before it is padded
 * (cpu_instantaneous.py:843-854 -> infilling.py:808-870 fill_masked_predictor):
 * 'nearest' fills every masked cell (d_mask != 0) from the nearest unmasked
 * cell in physical distance; 'harmonic' applies its least-squares solution
 * as a linear map u[p][fill_idx[i]] = sum_j gt[j*nf + i] * u[p][bnd_idx[j]]
 * (FP64, in place, every plane p < planes of plane_stride floats); gt is the
 * mask geometry's map, computed once on the host. */
/* infilling.py:609-695 (infill_knn): NaN vectors (not excluded) become the
 * k-nearest-neighbour regression of the valid vectors in physical
 * coordinates, weights 1/distance (uniform != 0: equal), k = min(n_neighbors,
 * donors) with n_neighbors <= 128, FP64 sums stored as float32; ties at equal
 * distance in row-major order.  d_work: pivamd_infill_knn_workspace(N)
 * bytes; *d_err |= 1 when an image with holes has no donor. */
unsigned long pivamd_infill_knn_workspace(int N);
int pivamd_infill_knn(const float *u, const float *d_excl, int N, int ny, int nx, float py, float px,
                      int n_neighbors, int uniform, float *out, void *d_work, unsigned long work_bytes,
                      int *d_err, void *stream);
/* infilling.py:505-597 (infill_rbf_local, scipy RBFInterpolator with
 * `neighbors`): each NaN vector (not excluded) is interpolated from its k =
 * min(neighbors, donors) nearest valid vectors (neighbors <= 128; physical
 * coordinates; ties at the k-th distance in row-major order) by the kernel
 * phi (0 linear, 1 thin_plate_spline, 2 cubic, 3 quintic, 4 multiquadric,
 * 5 inverse_multiquadric, 6 inverse_quadratic, 7 gaussian; epsilon = 1 for
 * 0-3 as scipy's default) plus scipy's default polynomial degree, smoothing
 * on the kernel diagonal, solved in FP64 by LU with partial pivoting, stored
 * as float32.  d_work: pivamd_infill_rbf_workspace(N, ny, nx) bytes;
 * *d_err |= 1 (fewer than 4 donors in an image with holes), 2 (singular
 * system), 4 (candidate buffer exceeded by a very anisotropic pitch). */
unsigned long pivamd_infill_rbf_workspace(int N, int ny, int nx);
int pivamd_infill_rbf(const float *u, const float *d_excl, int N, int ny, int nx, float py, float px,
                      int neighbors, int kernel, double epsilon, double smoothing, float *out, void *d_work,
                      unsigned long work_bytes, int *d_err, void *stream);

/* Linear griddata infill (infilling.py:440-503 infill_griddata_linear:
 * scipy griddata 'linear' over the valid vectors in physical coordinates,
 * 'nearest' outside their convex hull).  Each NaN vector (not excluded) is
 * the barycentric interpolation in a Delaunay triangle of the donors: the
 * first triangle, in row-major order of its vertices, among the k nearest
 * donors (k = 24, 48, 96, 192 in turn) that holds the hole, whose
 * circumcircle holds no donor and lies within the k-th donor's distance;
 * the nearest donor when no triangle holds the hole.  FP64, stored as
 * float32; where cocircular donors leave several Delaunay triangles the
 * first is taken (scipy's Qhull takes one by its own processing order).
 * d_work: pivamd_infill_griddata_workspace(N, ny, nx) bytes.  *d_err bits:
 * 1 an image with holes has no donor, 2 a hole's triangle reaches beyond 192
 * donors, 4 the candidate search outgrew its buffer. */
unsigned long pivamd_infill_griddata_workspace(int N, int ny, int nx);
int pivamd_infill_griddata(const float *u, const float *d_excl, int N, int ny, int nx, float py, float px,
                           float *out, void *d_work, unsigned long work_bytes, int *d_err, void *stream);
int pivamd_fill_masked_nearest(const float *u, const float *d_mask, int N, int ny, int nx, float py,
                               float px, float *uo, int *d_err, void *stream);
int pivamd_fill_linear(float *u, int planes, long plane_stride, int nf, const int *d_fill_idx, int nb,
                       const int *d_bnd_idx, const double *d_gt, void *stream);
int pivamd_smooth_axis(const float *in, int planes, int ny, int nx, int axis, const double *d_wts,
                       int radius, float *out, void *stream);
int pivamd_remap_cubic(const float *in, int N, int nyo, int nxo, const float *map_x,
                       const float *map_y, int ny, int nx, float *out, void *stream);

/* k-space fit over device buffers (same arguments as kspace_lm_fit_batch,
 * plus pwin_shared = 1 when d_pwin holds one shape for every window, a
 * workspace of pivamd_kspace_workspace bytes and a stream). */
unsigned long pivamd_kspace_workspace(int n_windows, int corr_h, int corr_w);
int pivamd_kspace_fit(const double *d_aa, const double *d_bb, const double *d_ab,
                      const unsigned char *d_mask, const double *d_pwin, int n_windows, int corr_h,
                      int corr_w, int use_kx4, int use_ky4, double *d_gauss, int *d_status, double *d_init,
                      double *d_gain, double *d_n0, double *d_b4x, double *d_b4y, double *d_cpp,
                      int *d_nvalid, int *d_iter, unsigned char *d_conv, int pwin_shared, void *d_work,
                      unsigned long work_bytes, void *stream);
int pivamd_kspace_fft2_centered(const double *d_plane, int corr_h, int corr_w, double *d_re, double *d_im,
                                void *d_work, unsigned long work_bytes, void *stream);
/* d_work: pivamd_kspace_workspace(1, H, W) + 512 bytes. */
int pivamd_kspace_debug_prep(const double *d_aa, const double *d_bb, const double *d_ab,
                             const double *d_pcentred, int corr_h, int corr_w, double *d_dv,
                             double *d_n0seed, void *d_work, unsigned long work_bytes, void *stream);

/* Ensemble pass on the device (single_pass_accumulator.py:555-795,
 * 1067-1290): FP64 image sums, mean images, plane finalize (sums / N minus
 * the background planes (NULL: none), optional plane-mean removal and
 * envelope divide (NULL: none), AB x ab_scale, division by
 * sqrt(max(AA(0) BB(0), 1e-12)), FP64 out), and displacement / stress
 * extraction from the k-space rows (peak height x peak_scale in FP64 before
 * the clip to [0, 1]: sqrt(sum area / A area) for the single run type,
 * single_pass_accumulator.py:1191-1210, else 1). */
int pivamd_sum_images(const float *d_imgs, int n, unsigned long hw, double *d_sum, void *stream);
int pivamd_mean_image(const double *d_sum, double n, unsigned long hw, float *d_out, void *stream);
int pivamd_ensemble_finalize(const float *d_sAA, const float *d_sBB, const float *d_sAB, const float *d_bAA,
                             const float *d_bBB, const float *d_bAB, const float *d_envAuto,
                             const float *d_envAB, int n_windows, int corr_h, int corr_w, float n_images,
                             float ab_scale, int dc_zero, double *d_oAA, double *d_oBB, double *d_oAB,
                             void *stream);
int pivamd_ensemble_extract(const double *d_gauss, const int *d_status, const unsigned char *d_mask,
                            int n_windows, int corr_h, int corr_w, double peak_scale, float *d_ux, float *d_uy,
                            float *d_uu, float *d_vv, float *d_uv, float *d_peak, int *d_reason, void *stream);
/* Single run type image padding (window_utils.py:447-520, apply_single_mode_padding
 * with pad_value 0): n frames of H x W (16-bit counts when src_u16, else
 * float32) into zero-bordered float32 planes of Hp x Wp at (pad_top, pad_left);
 * 2 (size) for padded frames of 2^31 pixels or more. */
int pivamd_pad_frames(const void *d_src, int src_u16, int n, int H, int W, int pad_top, int pad_left, int Hp,
                      int Wp, float *d_dst, void *stream);

/* Peak search + fit + detectability on n planes of ny x nx in device
 * memory (window-major); outputs peak_loc[3*n] (row, col, height) and
 * std_dev[3*n] per plane, component-major with stride n, and ratio[n]. */
int pivamd_peakfit(const float *d_planes, int n, int ny, int nx, int nPeaks, int finder, int impl,
                   const float *d_planeWeight, float *d_loc, float *d_std, float *d_ratio,
                   void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PIVTOOLS_AMD_H */
