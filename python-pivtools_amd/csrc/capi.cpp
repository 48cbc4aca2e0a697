// capi.cpp -- host-pointer drop-ins for the reference's ctypes libraries.
//
// Built three times: with PIVAMD_CAPI_XCORR into libbulkxcorr2d.so (the
// symbols of pivtools_cli/lib/PIV_2d_cross_correlate.h, peak_locate_lm.h,
// peak_locate_lm_batch.h and xcorr.h), with PIVAMD_CAPI_WARP into
// libfusedwarp.so (pivtools_cli/lib/fused_warp.h) and with
// PIVAMD_CAPI_KSPACE into libkspacefit.so (pivtools_cli/lib/kspace_lm_fit.h).  Every call is
// synchronous: the caller's host arrays go to HBM, the device kernels of
// libpivtools_amd.so run on this library's stream, results come back.
// Ownership is the reference's: the caller allocates every input and
// output; device buffers are this library's grow-only cache.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "pivamd_dev.h"

namespace {

// Reference error bits (pivtools_cli/lib/common.h:31-36).
constexpr unsigned char kErrNone = 0, kErrNomem = 1, kErrNoplanBwd = 4, kErrNoplan = 8;

std::mutex g_mu;

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Ctx {
    int dev = -1;
    hipStream_t st = nullptr;
    Buf bufs[24];
};

Ctx g_ctx;

bool ctx_ready()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (g_ctx.st && g_ctx.dev == dev) return true;
    if (g_ctx.st) return false;   // one device per process for the drop-ins
    if (hipStreamCreateWithFlags(&g_ctx.st, hipStreamNonBlocking) != hipSuccess) return false;
    g_ctx.dev = dev;
    return true;
}

// Device buffer slot k with at least `bytes` (grow-only cache).
void *dbuf(int k, size_t bytes)
{
    Buf &b = g_ctx.bufs[k];
    if (bytes == 0) bytes = 4;
    if (b.cap >= bytes) return b.p;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    if (hipMalloc(&b.p, bytes) != hipSuccess) return nullptr;
    b.cap = bytes;
    return b.p;
}

bool h2d(void *d, const void *h, size_t bytes)
{
    return bytes == 0 || hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g_ctx.st) == hipSuccess;
}

bool d2h(void *h, const void *d, size_t bytes)
{
    return bytes == 0 || hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g_ctx.st) == hipSuccess;
}

bool sync() { return hipStreamSynchronize(g_ctx.st) == hipSuccess; }

}  // namespace

#ifdef PIVAMD_CAPI_XCORR

static volatile int g_kernel_timing = 0;
// last plane weight computed by bulkxcorr2d: its wB contents, size and the
// device buffer holding it (guarded by g_mu)
static std::vector<float> g_pw_key;
static int g_pw_wy = 0, g_pw_wx = 0;
static const float *g_pw_dev = nullptr;
static double g_t_fft = 0.0, g_t_fit = 0.0;
static volatile int g_peakfit_impl = 0;
// Lane width of the batch fitter's output arrays.  The reference publishes
// PEAKFIT_MAX_LANES = 16 as the bound callers size their buffers by
// (peak_locate_lm_batch.h:26-28, PIV_2d_cross_correlate.c:149), so the
// drop-in reports 16 and writes exactly 3 x 16 floats per output array.
static const int kBatchLanes = 16;

extern "C" {

void bulkxcorr2d_set_timing_enabled(int on) { g_kernel_timing = on ? 1 : 0; }
void bulkxcorr2d_reset_timing(void) { g_t_fft = 0.0; g_t_fit = 0.0; }
void bulkxcorr2d_get_timing(double *fft_s, double *fit_s)
{
    if (fft_s) *fft_s = g_t_fft;
    if (fit_s) *fit_s = g_t_fit;
}

int peakfit_batch_available(void) { return 1; }
int peakfit_batch_lanes(void) { return kBatchLanes; }
int codelet_lanes(void) { return kBatchLanes; }

int bulkxcorr2d_set_peakfit_impl(int impl)
{
    g_peakfit_impl = impl ? 1 : 0;
    return 0;
}
int bulkxcorr2d_get_peakfit_impl(void) { return g_peakfit_impl; }
int bulkxcorr2d_peakfit_batch_available(void) { return 1; }

int pivtools_cpu_supported(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

// MIN/MAX of common.h: MIN(A,B) = A < B ? A : B
float fminvec(const float *v, int n)
{
    float r = v[0];
    for (int i = 1; i < n; ++i) r = (r < v[i]) ? r : v[i];
    return r;
}
float fmaxvec(const float *v, int n)
{
    float r = v[0];
    for (int i = 1; i < n; ++i) r = (r > v[i]) ? r : v[i];
    return r;
}

unsigned char bulkxcorr2d(const float *fImageA, const float *fImageB, const float *fMask,
                          const int *nImageSize, int N_images, const float *fWinCtrsX,
                          const float *fWinCtrsY, const int *nWindows, float *fWindowWeightA,
                          bool bEnsemble, const float *fWindowWeightB, const int *nWindowSize,
                          int nPeaks, int iPeakFinder, float *fPkLocX, float *fPkLocY,
                          float *fPkHeight, float *fSx, float *fSy, float *fSxy, float *fPkRatio,
                          float *fCorrelPlane_Out)
{
    const int H = nImageSize[0], W = nImageSize[1];
    const int nWy = nWindows[0], nWx = nWindows[1];
    const int wy = nWindowSize[0], wx = nWindowSize[1];
    const int impl = g_peakfit_impl;   // read once per call
    if (!pivamd_size_supported(wy, wx)) {
        fprintf(stderr, "bulkxcorr2d: window %dx%d has no device kernel\n", wy, wx);
        return kErrNoplanBwd;
    }
    if (N_images <= 0 || nWy <= 0 || nWx <= 0) return kErrNone;
    if (nPeaks < 1) return kErrNoplan;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return kErrNoplan;
    const size_t nWin = (size_t)nWy * nWx, numel = (size_t)wy * wx;
    const size_t img = (size_t)N_images * H * W * sizeof(float);
    const size_t outn = (size_t)N_images * nPeaks * nWin * sizeof(float);
    const size_t ratn = (size_t)N_images * nWin * sizeof(float);
    const size_t pln = (size_t)N_images * nWin * numel * sizeof(float);
    const size_t work = pivamd_xcorr_workspace(N_images, nWy, nWx, wy, wx, nPeaks, bEnsemble ? 1 : 0);
    float *dA = (float *)dbuf(0, img), *dB = (float *)dbuf(1, img);
    float *dM = (float *)dbuf(2, nWin * sizeof(float));
    float *dCx = (float *)dbuf(3, nWx * sizeof(float)), *dCy = (float *)dbuf(4, nWy * sizeof(float));
    float *dWa = (float *)dbuf(5, numel * sizeof(float)), *dWb = (float *)dbuf(6, numel * sizeof(float));
    float *dPw = (float *)dbuf(22, numel * sizeof(float));   // slot of the cached plane weight only
    float *dO[6];
    for (int k = 0; k < 6; ++k) dO[k] = (float *)dbuf(8 + k, outn);
    float *dR = (float *)dbuf(14, ratn);
    float *dP = fCorrelPlane_Out ? (float *)dbuf(15, pln) : nullptr;
    void *dWk = dbuf(16, work);
    if (!dA || !dB || !dM || !dCx || !dCy || !dWa || !dWb || !dPw || !dR || !dWk ||
        (fCorrelPlane_Out && !dP))
        return kErrNomem;
    for (int k = 0; k < 6; ++k)
        if (!dO[k]) return kErrNomem;
    float *hO[6] = {fPkLocX, fPkLocY, fPkHeight, fSx, fSy, fSxy};
    // one weight for A and B (the instantaneous path's call) goes up once and
    // reaches the kernel as one pointer, which then reads it once per sample
    const bool same_w = fWindowWeightA == fWindowWeightB ||
                        std::memcmp(fWindowWeightA, fWindowWeightB, numel * sizeof(float)) == 0;
    if (same_w) dWb = dWa;
    bool ok = h2d(dA, fImageA, img) && h2d(dB, fImageB, img) && h2d(dM, fMask, nWin * sizeof(float)) &&
              h2d(dCx, fWinCtrsX, nWx * sizeof(float)) && h2d(dCy, fWinCtrsY, nWy * sizeof(float)) &&
              h2d(dWa, fWindowWeightA, numel * sizeof(float)) &&
              (same_w || h2d(dWb, fWindowWeightB, numel * sizeof(float)));
    // the library leaves masked / out-of-bounds windows untouched: start
    // from the caller's contents
    if (!bEnsemble) {
        for (int k = 0; k < 6; ++k) ok = ok && h2d(dO[k], hO[k], outn);
        ok = ok && h2d(dR, fPkRatio, ratn);
    }
    if (dP) ok = ok && h2d(dP, fCorrelPlane_Out, pln);
    if (!ok) return kErrNoplan;
    // The plane weight depends on wB alone: it is recomputed only when the
    // caller passes a different weight (or size) than the previous call.
    const bool pw_hit = g_pw_dev == dPw && g_pw_wy == wy && g_pw_wx == wx &&
                        g_pw_key.size() == numel &&
                        std::memcmp(g_pw_key.data(), fWindowWeightB, numel * sizeof(float)) == 0;
    if (!pw_hit) {
        g_pw_dev = nullptr;
        if (pivamd_plane_weight(dWb, wy, wx, dPw, g_ctx.st) != 0) return kErrNoplan;
        g_pw_key.assign(fWindowWeightB, fWindowWeightB + numel);
        g_pw_wy = wy;
        g_pw_wx = wx;
        g_pw_dev = dPw;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timing = g_kernel_timing != 0;
    if (timing) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, g_ctx.st);
    }
    const int rc = pivamd_xcorr(dA, dB, dM, H, W, N_images, dCx, dCy, nWy, nWx, dWa, dWb, dPw,
                                bEnsemble ? 1 : 0, wy, wx, nPeaks, iPeakFinder, impl, dO[0], dO[1],
                                dO[2], dO[3], dO[4], dO[5], dR, dP, dWk, work, g_ctx.st);
    if (timing) (void)hipEventRecord(e1, g_ctx.st);
    if (rc != 0) return rc == PIVAMD_ERR_SIZE ? kErrNoplanBwd : kErrNoplan;
    if (!bEnsemble) {
        for (int k = 0; k < 6; ++k) ok = ok && d2h(hO[k], dO[k], outn);
        ok = ok && d2h(fPkRatio, dR, ratn);
    }
    if (dP) ok = ok && d2h(fCorrelPlane_Out, dP, pln);
    ok = ok && sync();
    if (timing) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) g_t_fft += ms * 1e-3;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return ok ? kErrNone : kErrNoplan;
}

// Shared host side of bulkxcorr2d_accumulate(_triple): inputs to HBM, the
// caller's sums to HBM (the library adds into them), one device call, sums
// back.  An unsupported window size reports NOMEM, as the reference's
// failed batched plan does (PIV_2d_cross_correlate.c:545-546, 718-721).
static unsigned char run_accumulate(const float *A, const float *B, const float *mask, const int *nImageSize,
                                    int N_images, const float *ctrX, const float *ctrY, const int *nWindows,
                                    const float *wAab, const float *wBab, const float *wAA, const float *wBB,
                                    const int *nWindowSize, const int *nFit, int meanSub, int perPairNorm,
                                    int triple, float *outAB, float *outAA, float *outBB)
{
    const int H = nImageSize[0], W = nImageSize[1];
    const int nWy = nWindows[0], nWx = nWindows[1];
    const int wy = nWindowSize[0], wx = nWindowSize[1];
    const int fy = nFit ? nFit[0] : wy, fx = nFit ? nFit[1] : wx;
    // the device ensemble sums are built for square windows
    if (wy != wx || !pivamd_size_supported(wy, wx)) return kErrNomem;
    if (nWy <= 0 || nWx <= 0) return kErrNone;
    const size_t numel = (size_t)wy * wx, nWin = (size_t)nWy * nWx;
    float sWB = 0.0f, sWAA = 0.0f, sWBB = 0.0f;
    if (meanSub && triple)
        for (size_t i = 0; i < numel; ++i) { sWB += wBab[i]; sWAA += wAA[i]; sWBB += wBB[i]; }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return kErrNoplan;
    const size_t img = (size_t)(N_images > 0 ? N_images : 1) * H * W * sizeof(float);
    const size_t outb = nWin * fy * fx * sizeof(float);
    const size_t work = pivamd_ensemble_workspace(nWy, nWx, wy, wx);
    float *dA = (float *)dbuf(0, img), *dB = (float *)dbuf(1, img);
    float *dM = (float *)dbuf(2, nWin * sizeof(float));
    float *dCx = (float *)dbuf(3, nWx * sizeof(float)), *dCy = (float *)dbuf(4, nWy * sizeof(float));
    float *dw[4];
    for (int k = 0; k < 4; ++k) dw[k] = (float *)dbuf(5 + k, numel * sizeof(float));
    float *dO[3];
    for (int k = 0; k < 3; ++k) dO[k] = (float *)dbuf(9 + k, outb);
    void *dWk = dbuf(16, work);
    if (!dA || !dB || !dM || !dCx || !dCy || !dWk) return kErrNomem;
    for (int k = 0; k < 4; ++k) if (!dw[k]) return kErrNomem;
    for (int k = 0; k < 3; ++k) if (!dO[k]) return kErrNomem;
    const float *hw[4] = {wAab, wBab, triple ? wAA : wAab, triple ? wBB : wBab};
    float *ho[3] = {outAB, outAA, outBB};
    const size_t imgn = (size_t)N_images * H * W * sizeof(float);
    bool ok = h2d(dA, A, imgn) && h2d(dB, B, imgn) && h2d(dM, mask, nWin * sizeof(float)) &&
              h2d(dCx, ctrX, nWx * sizeof(float)) && h2d(dCy, ctrY, nWy * sizeof(float));
    // identical weight contents share one device buffer, which lets the
    // kernel take the auto spectra from the AB transform (wAab == wAA and
    // wBab == wBB, the reference's default call)
    const float *dwk[4];
    for (int k = 0; k < 4; ++k) {
        dwk[k] = dw[k];
        for (int j = 0; j < k; ++j)
            if (std::memcmp(hw[j], hw[k], numel * sizeof(float)) == 0) {
                dwk[k] = dwk[j];
                break;
            }
        if (dwk[k] == dw[k]) ok = ok && h2d(dw[k], hw[k], numel * sizeof(float));
    }
    for (int k = 0; k < (triple ? 3 : 1); ++k) ok = ok && h2d(dO[k], ho[k], outb);
    if (!ok) return kErrNoplan;
    const int rc = pivamd_accumulate(dA, dB, dM, H, W, N_images, dCx, dCy, nWy, nWx, dwk[0], dwk[1], dwk[2], dwk[3],
                                     wy, wx, fy, fx, meanSub, perPairNorm, triple, sWB, sWAA, sWBB, dO[0],
                                     triple ? dO[1] : nullptr, triple ? dO[2] : nullptr, dWk, work, g_ctx.st);
    if (rc != 0) return rc == PIVAMD_ERR_SIZE ? kErrNomem : kErrNoplan;
    for (int k = 0; k < (triple ? 3 : 1); ++k) ok = ok && d2h(ho[k], dO[k], outb);
    ok = ok && sync();
    return ok ? kErrNone : kErrNoplan;
}

unsigned char bulkxcorr2d_accumulate(const float *fImageA_stack, const float *fImageB_stack, const float *fMask,
                                     const int *nImageSize, int N_images, const float *fWinCtrsX,
                                     const float *fWinCtrsY, const int *nWindows, const float *fWindowWeightA,
                                     const float *fWindowWeightB, const int *nWindowSize,
                                     const int *nFitWindowSize, float *fCorrelPlane_Sum)
{
    return run_accumulate(fImageA_stack, fImageB_stack, fMask, nImageSize, N_images, fWinCtrsX, fWinCtrsY,
                          nWindows, fWindowWeightA, fWindowWeightB, nullptr, nullptr, nWindowSize,
                          nFitWindowSize, 0, 0, 0, fCorrelPlane_Sum, nullptr, nullptr);
}

unsigned char bulkxcorr2d_accumulate_triple(const float *fImageA_stack, const float *fImageB_stack,
                                            const float *fMask, const int *nImageSize, int N_images,
                                            const float *fWinCtrsX, const float *fWinCtrsY,
                                            const int *nWindows, const float *fWindowWeightA_AB,
                                            const float *fWindowWeightB_AB, const float *fAutoWeightA,
                                            const float *fAutoWeightB, const int *nWindowSize,
                                            const int *nFitWindowSize, int bMeanSubtract, int bPerPairNorm,
                                            float *fCorrAB_Sum, float *fCorrAA_Sum, float *fCorrBB_Sum)
{
    return run_accumulate(fImageA_stack, fImageB_stack, fMask, nImageSize, N_images, fWinCtrsX, fWinCtrsY,
                          nWindows, fWindowWeightA_AB, fWindowWeightB_AB, fAutoWeightA, fAutoWeightB,
                          nWindowSize, nFitWindowSize, bMeanSubtract, bPerPairNorm, 1, fCorrAB_Sum,
                          fCorrAA_Sum, fCorrBB_Sum);
}

static int run_peakfit(const float *planes, int n, int ny, int nx, int nPeaks, int finder, int impl,
                       const float *w, float *loc, float *std, float *ratio)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return -1;
    const size_t pl = (size_t)n * ny * nx * sizeof(float);
    const size_t out = (size_t)3 * nPeaks * n * sizeof(float);
    float *dP = (float *)dbuf(17, pl), *dW = w ? (float *)dbuf(18, (size_t)ny * nx * sizeof(float)) : nullptr;
    float *dL = (float *)dbuf(19, out), *dS = (float *)dbuf(20, out);
    float *dR = (float *)dbuf(21, (size_t)n * sizeof(float));
    if (!dP || !dL || !dS || !dR || (w && !dW)) return -1;
    bool ok = h2d(dP, planes, pl) && (!w || h2d(dW, w, (size_t)ny * nx * sizeof(float)));
    if (!ok) return -1;
    if (pivamd_peakfit(dP, n, ny, nx, nPeaks, finder, impl, dW, dL, dS, dR, g_ctx.st) != 0) return -1;
    ok = (!loc || d2h(loc, dL, out)) && (!std || d2h(std, dS, out)) &&
         (!ratio || d2h(ratio, dR, (size_t)n * sizeof(float))) && sync();
    return ok ? 0 : -1;
}

void lsqpeaklocate_lm(const float *xcorr, const int *N, float *peak_loc, int nPeaks, int iFitType,
                      float *std_dev, const float *fPlaneWeight)
{
    if (run_peakfit(xcorr, 1, N[0], N[1], nPeaks, iFitType, 0, fPlaneWeight, peak_loc, std_dev,
                    nullptr) != 0) {
        fprintf(stderr, "lsqpeaklocate_lm: device call failed\n");
        for (int p = 0; p < nPeaks; ++p) {
            peak_loc[p] = NAN; peak_loc[nPeaks + p] = NAN; peak_loc[2 * nPeaks + p] = 0;
            std_dev[p] = 0; std_dev[nPeaks + p] = 0; std_dev[2 * nPeaks + p] = 0;
        }
    }
}

float peak_detectability(const float *xcorr, const int *N)
{
    float r = 0.0f;
    if (run_peakfit(xcorr, 1, N[0], N[1], 1, 3, 0, nullptr, nullptr, nullptr, &r) != 0) {
        fprintf(stderr, "peak_detectability: device call failed\n");
        return 0.0f;
    }
    return r;
}

void lsqpeaklocate_lm_batch(const float *planes, int L_real, const int *N, int iFitType, float *peak_loc,
                            float *std_dev, const float *fPlaneWeight)
{
    const int Wl = kBatchLanes;
    for (int l = 0; l < Wl; ++l) {
        peak_loc[l] = NAN; peak_loc[Wl + l] = NAN; peak_loc[2 * Wl + l] = 0;
        std_dev[l] = 0; std_dev[Wl + l] = 0; std_dev[2 * Wl + l] = 0;
    }
    if (iFitType < 4 || iFitType > 6) {
        fprintf(stderr, "lsqpeaklocate_lm_batch: unsupported iFitType %d "
                        "(caller must route 3-point fits to the scalar path)\n", iFitType);
        return;
    }
    const int L = L_real < Wl ? L_real : Wl;
    if (L <= 0) return;
    std::vector<float> loc(3 * (size_t)L), sd(3 * (size_t)L);
    if (run_peakfit(planes, L, N[0], N[1], 1, iFitType, 1, fPlaneWeight, loc.data(), sd.data(),
                    nullptr) != 0) {
        fprintf(stderr, "lsqpeaklocate_lm_batch: device call failed\n");
        return;
    }
    for (int c = 0; c < 3; ++c)
        for (int l = 0; l < L; ++l) {
            peak_loc[c * Wl + l] = loc[(size_t)c * L + l];
            std_dev[c * Wl + l] = sd[(size_t)c * L + l];
        }
}

static bool built_size(int n)
{
    switch (n) {
    case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256:
        return true;
    default:
        return false;
    }
}

unsigned convolve(const float *w1, const float *w2, float *conv, const int *N)
{
    if (!built_size(2 * N[0]) || !built_size(2 * N[1])) return kErrNoplanBwd;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return kErrNoplan;
    const size_t b = (size_t)N[0] * N[1] * sizeof(float);
    float *d1 = (float *)dbuf(17, b), *d2 = (float *)dbuf(18, b), *dc = (float *)dbuf(19, b);
    if (!d1 || !d2 || !dc) return kErrNomem;
    bool ok = h2d(d1, w1, b) && h2d(d2, w2, b);
    ok = ok && pivamd_convolve(d1, d2, N[0], N[1], dc, g_ctx.st) == 0;
    ok = ok && d2h(conv, dc, b) && sync();
    return ok ? kErrNone : kErrNoplan;
}

unsigned xcorr(const float *w1, const float *w2, float *corr, const int *N)
{
    const int ny = N[0], nx = N[1];
    if (!pivamd_size_supported(ny, nx)) return kErrNoplanBwd;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return kErrNoplan;
    const size_t b = (size_t)ny * nx * sizeof(float);
    std::vector<float> ones((size_t)ny * nx, 1.0f);
    const float cy = (ny - 1) / 2.0f, cx = (nx - 1) / 2.0f;
    float *d1 = (float *)dbuf(17, b), *d2 = (float *)dbuf(18, b), *dp = (float *)dbuf(19, b);
    float *dw = (float *)dbuf(20, b), *dc = (float *)dbuf(21, 2 * sizeof(float));
    if (!d1 || !d2 || !dp || !dw || !dc) return kErrNomem;
    const float ctr[2] = {cx, cy};
    bool ok = h2d(d1, w1, b) && h2d(d2, w2, b) && h2d(dw, ones.data(), b) && h2d(dc, ctr, sizeof(ctr));
    // slot0 (the "B" image of bulkxcorr2d) = w1, slot1 ("A") = w2; ensemble
    // mode skips the mean removal, the weights are ones
    ok = ok && pivamd_xcorr(d2, d1, nullptr, ny, nx, 1, dc, dc + 1, 1, 1, dw, dw, nullptr, 1, ny, nx, 1, 3,
                            0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, dp,
                            nullptr, 0, g_ctx.st) == 0;
    ok = ok && d2h(corr, dp, b) && sync();
    return ok ? kErrNone : kErrNoplan;
}

}  // extern "C"

#endif  // PIVAMD_CAPI_XCORR

#ifdef PIVAMD_CAPI_WARP

static volatile int g_warp_impl = 1;

extern "C" {

void fused_warp_set_impl(int impl) { g_warp_impl = impl ? 1 : 0; }
int fused_warp_get_impl(void) { return g_warp_impl; }

int fused_symmetric_warp_batch(const float *imgs_a, const float *imgs_b, float *outs_a, float *outs_b,
                               const float *pred_dy, const float *pred_dx, int N, int H, int W, int nPY,
                               int nPX, const float *ctrs_y, const float *ctrs_x, int interp_mode,
                               int shared_predictor, int round_shifts)
{
    // input validation of fused_warp.c:491-493 (ERROR_NOMEM on bad input)
    if (N <= 0 || H <= 0 || W <= 0 || nPY <= 0 || nPX <= 0) return 1;
    if (!imgs_a || !imgs_b || !outs_a || !outs_b) return 1;
    if (!pred_dy || !pred_dx || !ctrs_y || !ctrs_x) return 1;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return 1;
    const size_t img = (size_t)N * H * W * sizeof(float);
    const size_t pred = (size_t)(shared_predictor ? 1 : N) * nPY * nPX * sizeof(float);
    float *dA = (float *)dbuf(0, img), *dB = (float *)dbuf(1, img);
    float *oA = (float *)dbuf(2, img), *oB = (float *)dbuf(3, img);
    float *dy = (float *)dbuf(4, pred), *dx = (float *)dbuf(5, pred);
    float *cy = (float *)dbuf(6, nPY * sizeof(float)), *cx = (float *)dbuf(7, nPX * sizeof(float));
    if (!dA || !dB || !oA || !oB || !dy || !dx || !cy || !cx) return 1;
    bool ok = h2d(dA, imgs_a, img) && h2d(dB, imgs_b, img) && h2d(dy, pred_dy, pred) &&
              h2d(dx, pred_dx, pred) && h2d(cy, ctrs_y, nPY * sizeof(float)) &&
              h2d(cx, ctrs_x, nPX * sizeof(float));
    // fused_warp.c:514: 0 is bicubic, anything else Lanczos-3 (both in the
    // reference's exact arithmetic)
    ok = ok && pivamd_warp(dA, dB, oA, oB, dy, dx, N, H, W, nPY, nPX, cy, cx, interp_mode == 0 ? 0 : 1,
                           shared_predictor, round_shifts, g_ctx.st) == 0;
    ok = ok && d2h(outs_a, oA, img) && d2h(outs_b, oB, img) && sync();
    return ok ? 0 : 1;
}

int fused_symmetric_warp(const float *img_a, const float *img_b, float *out_a, float *out_b,
                         const float *pred_dy, const float *pred_dx, int H, int W, int nPY, int nPX,
                         const float *ctrs_y, const float *ctrs_x, int interp_mode, int round_shifts)
{
    return fused_symmetric_warp_batch(img_a, img_b, out_a, out_b, pred_dy, pred_dx, 1, H, W, nPY, nPX,
                                      ctrs_y, ctrs_x, interp_mode, 1, round_shifts);
}

}  // extern "C"

#endif  // PIVAMD_CAPI_WARP

#ifdef PIVAMD_CAPI_KSPACE

namespace {

// kspace_lm_fit.h:48-53 return codes
int ks_code(int e) { return e == PIVAMD_OK ? 0 : (e == PIVAMD_ERR_SIZE ? -1 : -2); }

bool ks_axis(int n)
{
    switch (n) {
    case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96: case 128: return true;
    default: return false;
    }
}

}  // namespace

extern "C" {

// kspace_lm_fit.h:54-61.  n_threads is the reference's OpenMP width and has
// no meaning on the device; every window is fitted by its own workgroup.
int kspace_lm_fit_batch(const double *R_AA, const double *R_BB, const double *R_AB,
                        const unsigned char *mask_flat, const double *P_win, int n_windows, int corr_h,
                        int corr_w, int use_kx4, int use_ky4, int n_threads, double *gauss_flat,
                        int *status_flat, double *initial_guess_flat, double *diag_gain, double *diag_N0,
                        double *diag_b4x, double *diag_b4y, double *diag_cost_per_pt, int *diag_n_valid,
                        int *diag_iter, unsigned char *diag_conv)
{
    (void)n_threads;
    if (!ks_axis(corr_h) || !ks_axis(corr_w)) return -1;
    if (n_windows <= 0) return 0;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return -2;
    const size_t n = (size_t)n_windows, P = (size_t)corr_h * corr_w, pb = n * P * sizeof(double);
    double *aa = (double *)dbuf(0, pb), *bb = (double *)dbuf(1, pb), *ab = (double *)dbuf(2, pb);
    unsigned char *mask = (unsigned char *)dbuf(3, n);
    double *pw = P_win ? (double *)dbuf(4, pb) : nullptr;
    // outputs: gauss, init (16 n each), five diagnostics (n each) in doubles, then 3 int arrays, conv
    const size_t nd = 32 * n + 5 * n;
    char *out = (char *)dbuf(5, nd * sizeof(double) + 3 * n * sizeof(int) + n);
    const size_t wb = pivamd_kspace_workspace(n_windows, corr_h, corr_w);
    void *work = dbuf(6, wb);
    if (!aa || !bb || !ab || !mask || (P_win && !pw) || !out || !work) return -2;
    double *g = (double *)out, *ig = g + 16 * n, *dgn = ig + 16 * n, *dn0 = dgn + n, *dbx = dn0 + n,
           *dby = dbx + n, *dcp = dby + n;
    int *st = (int *)(dcp + n), *nv = st + n, *itr = nv + n;
    unsigned char *cv = (unsigned char *)(itr + n);
    if (!h2d(aa, R_AA, pb) || !h2d(bb, R_BB, pb) || !h2d(ab, R_AB, pb) || !h2d(mask, mask_flat, n) ||
        (P_win && !h2d(pw, P_win, pb)))
        return -2;
    const int e = pivamd_kspace_fit(aa, bb, ab, mask, pw, n_windows, corr_h, corr_w, use_kx4, use_ky4, g, st, ig,
                                    dgn, dn0, dbx, dby, dcp, nv, itr, cv, 0, work, wb, g_ctx.st);
    if (e != PIVAMD_OK) {
        sync();
        return ks_code(e);
    }
    const size_t d16 = 16 * n * sizeof(double), d1 = n * sizeof(double), i1 = n * sizeof(int);
    const bool ok = d2h(gauss_flat, g, d16) && d2h(initial_guess_flat, ig, d16) && d2h(diag_gain, dgn, d1) &&
                    d2h(diag_N0, dn0, d1) && d2h(diag_b4x, dbx, d1) && d2h(diag_b4y, dby, d1) &&
                    d2h(diag_cost_per_pt, dcp, d1) && d2h(status_flat, st, i1) && d2h(diag_n_valid, nv, i1) &&
                    d2h(diag_iter, itr, i1) && d2h(diag_conv, cv, n) && sync();
    return ok ? 0 : -2;
}

// kspace_lm_fit.h:65.  A liveness probe for the loader: the number of
// compute units of the current device (the device-side width), 0 without one.
int kspace_lm_fit_max_threads(void)
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    return prop.multiProcessorCount;
}

// kspace_lm_fit.h:79-82
int kspace_debug_prep(const double *R_AA, const double *R_BB, const double *R_AB, const double *P_centred, int H,
                      int W, double *out_Dv, double *out_n0_seed)
{
    if (!ks_axis(H) || !ks_axis(W)) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return -2;
    const size_t pb = (size_t)H * W * sizeof(double);
    double *aa = (double *)dbuf(0, pb), *bb = (double *)dbuf(1, pb), *ab = (double *)dbuf(2, pb);
    double *pc = P_centred ? (double *)dbuf(4, pb) : nullptr;
    double *dv = (double *)dbuf(5, pb + 64);
    const size_t wb = pivamd_kspace_workspace(1, H, W) + 512;
    void *work = dbuf(6, wb);
    if (!aa || !bb || !ab || (P_centred && !pc) || !dv || !work) return -2;
    if (!h2d(aa, R_AA, pb) || !h2d(bb, R_BB, pb) || !h2d(ab, R_AB, pb) || (P_centred && !h2d(pc, P_centred, pb)))
        return -2;
    const int e = pivamd_kspace_debug_prep(aa, bb, ab, pc, H, W, dv, dv + (size_t)H * W, work, wb, g_ctx.st);
    if (e != PIVAMD_OK) {
        sync();
        return ks_code(e);
    }
    return d2h(out_Dv, dv, pb) && d2h(out_n0_seed, dv + (size_t)H * W, sizeof(double)) && sync() ? 0 : -2;
}

// kspace_lm_fit.h:86-87
int kspace_fft2_centered(const double *plane, int H, int W, double *out_re, double *out_im)
{
    if (!ks_axis(H) || !ks_axis(W)) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!ctx_ready()) return -2;
    const size_t pb = (size_t)H * W * sizeof(double);
    double *pl = (double *)dbuf(0, pb), *re = (double *)dbuf(1, pb), *im = (double *)dbuf(2, pb);
    void *work = dbuf(6, 5 * pb);
    if (!pl || !re || !im || !work) return -2;
    if (!h2d(pl, plane, pb)) return -2;
    const int e = pivamd_kspace_fft2_centered(pl, H, W, re, im, work, 5 * pb, g_ctx.st);
    if (e != PIVAMD_OK) {
        sync();
        return ks_code(e);
    }
    return d2h(out_re, re, pb) && d2h(out_im, im, pb) && sync() ? 0 : -2;
}

// The reference's loader also requires the CPU-support probe in every library
// (lib_guard.py:6-26); here it reports whether a HIP device is present.
int pivtools_cpu_supported(void)
{
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0 ? 1 : 0;
}

}  // extern "C"

#endif  // PIVAMD_CAPI_KSPACE
