// pivamd_dev.h -- internal declarations shared by the HIP translation units.
// The public C ABI (host-pointer drop-ins and the device-pointer API) is
// declared in include/pivtools_amd.h.
#pragma once

#include <hip/hip_runtime.h>

// Error codes of the device API (the host-pointer drop-ins translate them to
// the reference's codes, pivtools_cli/lib/common.h:31-36).
#define PIVAMD_OK 0
#define PIVAMD_ERR_ARG 1
#define PIVAMD_ERR_SIZE 2
#define PIVAMD_ERR_LAUNCH 3
#define PIVAMD_ERR_NOMEM 4
#define PIVAMD_ERR_RUNTIME 5

// One correlation call of a ragged launch (pivamd_xcorr_ragged); the same
// layout as the PivamdXcorrJob of the public header.
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

extern "C" {
int pivamd_size_supported(int ny, int nx);
int pivamd_set_dft_mode(int mode);
int pivamd_xcorr_variants(int N, int *out, int max_out);
int pivamd_xcorr_default_variant(int N);
int pivamd_xcorr_set_variant(int N, int v);
int pivamd_xcorr_get_variant(int N);
int pivamd_convolve(const float *d_a, const float *d_b, int ny, int nx, float *d_out, hipStream_t st);
int pivamd_plane_weight(const float *d_wB, int ny, int nx, float *d_out, hipStream_t st);
int pivamd_xcorr(const float *d_imgA, const float *d_imgB, const float *d_mask, int H, int W,
                 int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                 const float *d_wA, const float *d_wB, const float *d_planeWeight, int ensemble,
                 int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX, float *d_pkY,
                 float *d_pkH, float *d_sx, float *d_sy, float *d_sxy, float *d_ratio,
                 float *d_planes, void *d_work, size_t work_bytes, hipStream_t st);
size_t pivamd_xcorr_workspace(int n_images, int nWy, int nWx, int winY, int winX, int nPeaks,
                              int ensemble);
int pivamd_warp(const float *d_imgA, const float *d_imgB, float *d_outA, float *d_outB,
                const float *d_predDy, const float *d_predDx, int N, int H, int W, int nPY, int nPX,
                const float *d_ctrY, const float *d_ctrX, int interp_mode, int shared_predictor,
                int round_shifts, hipStream_t st);
int pivamd_xcorr_u16(const unsigned short *d_imgA, const unsigned short *d_imgB, const float *d_mask,
                     int H, int W, int n_images, const float *d_ctrX, const float *d_ctrY, int nWy,
                     int nWx, const float *d_wA, const float *d_wB, const float *d_planeWeight,
                     int ensemble, int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX,
                     float *d_pkY, float *d_pkH, float *d_sx, float *d_sy, float *d_sxy,
                     float *d_ratio, float *d_planes, void *d_work, size_t work_bytes, hipStream_t st);
int pivamd_warp_u16(const unsigned short *d_imgA, const unsigned short *d_imgB, float *d_outA,
                    float *d_outB, const float *d_predDy, const float *d_predDx, int N, int H, int W,
                    int nPY, int nPX, const float *d_ctrY, const float *d_ctrX, int interp_mode,
                    int shared_predictor, int round_shifts, hipStream_t st);
size_t pivamd_ensemble_workspace(int nWy, int nWx, int winY, int winX);
int pivamd_accumulate(const float *d_A, const float *d_B, const float *d_mask, int H, int W, int n_images,
                      const float *d_ctrX, const float *d_ctrY, int nWy, int nWx, const float *d_wAab,
                      const float *d_wBab, const float *d_wAA, const float *d_wBB, int winY, int winX,
                      int fitY, int fitX, int meanSub, int perPairNorm, int triple, float sumWBab,
                      float sumWAA, float sumWBB, float *d_outAB, float *d_outAA, float *d_outBB,
                      void *d_work, size_t work_bytes, hipStream_t st);
size_t pivamd_kspace_workspace(int n_windows, int corr_h, int corr_w);
int pivamd_kspace_fit(const double *d_aa, const double *d_bb, const double *d_ab, const unsigned char *d_mask,
                      const double *d_pwin, int n_windows, int corr_h, int corr_w, int use_kx4, int use_ky4,
                      double *d_gauss, int *d_status, double *d_init, double *d_gain, double *d_n0,
                      double *d_b4x, double *d_b4y, double *d_cpp, int *d_nvalid, int *d_iter,
                      unsigned char *d_conv, int pwin_shared, void *d_work, size_t work_bytes, hipStream_t st);
int pivamd_sum_images(const float *d_imgs, int n, unsigned long hw, double *d_sum, hipStream_t st);
int pivamd_mean_image(const double *d_sum, double n, unsigned long hw, float *d_out, hipStream_t st);
int pivamd_ensemble_finalize(const float *d_sAA, const float *d_sBB, const float *d_sAB, const float *d_bAA,
                             const float *d_bBB, const float *d_bAB, const float *d_envAuto, const float *d_envAB,
                             int n_windows, int corr_h, int corr_w, float n_images, float ab_scale, int dc_zero,
                             double *d_oAA, double *d_oBB, double *d_oAB, hipStream_t st);
int pivamd_ensemble_extract(const double *d_gauss, const int *d_status, const unsigned char *d_mask, int n_windows,
                            int corr_h, int corr_w, double peak_scale, float *d_ux, float *d_uy, float *d_uu,
                            float *d_vv, float *d_uv, float *d_peak, int *d_reason, hipStream_t st);
int pivamd_pad_frames(const void *d_src, int src_u16, int n, int H, int W, int pad_top, int pad_left, int Hp, int Wp,
                      float *d_dst, hipStream_t st);
int pivamd_kspace_fft2_centered(const double *d_plane, int corr_h, int corr_w, double *d_re, double *d_im,
                                void *d_work, size_t work_bytes, hipStream_t st);
int pivamd_kspace_debug_prep(const double *d_aa, const double *d_bb, const double *d_ab,
                             const double *d_pcentred, int corr_h, int corr_w, double *d_dv, double *d_n0seed,
                             void *d_work, size_t work_bytes, hipStream_t st);
int pivamd_peakfit(const float *d_planes, int n, int ny, int nx, int nPeaks, int finder, int impl,
                   const float *d_planeWeight, float *d_loc, float *d_std, float *d_ratio,
                   hipStream_t st);
}
