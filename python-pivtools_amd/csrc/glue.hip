// glue.hip -- device-resident multipass glue between correlation passes.
//
// The reference runs these steps in host NumPy/cv2/scipy between its two C
// calls (pivtools_cli/piv/piv_backend/cpu_instantaneous.py:555-865); here
// they are small kernels so a whole pass stays on the device:
//   gate_kernel       displacement gate N/2 (first pass) or N/4, NaN sentinel,
//                     predictor add-back, nan_mask / nan_reason (:576-635)
//   median_kernel     normalised-median test (any odd size), vector norm, 3x3 guard
//                     (outlier_detection.py:87-243) and the reason codes (:726-757)
//   infill_kernel     nearest-donor fill in physical distance, ties to the
//                     smallest column then row, as scipy's EDT resolves them
//                     (infilling.py:114-141)
//   infill_kernel     (second mode) nearest fill of masked windows of a
//                     predictor field (infilling.py:808-870, 'nearest')
//   fill_linear_kernel the same field's harmonic fill ('harmonic'), applied
//                     as a precomputed linear map of the unmasked neighbours
//   pad_kernel        edge padding into the predictor grid (:836-851)
//   smooth_kernel     scipy gaussian_filter, mode 'nearest', axis 0 then 1 with
//                     float32 rounding between axes (:1135-1147)
//   remap_kernel      predictor onto the new window grid, bicubic a=-0.75 on a
//                     1/32-pixel fixed-point grid with replicate border, the
//                     interpolation cv2.remap(INTER_CUBIC) performs (:1050-1071)
//   u16_to_f32_kernel 12-bit ingest
// Fields are planar: [component][image][row][col] with component 0 = dy
// (uy) and 1 = dx (ux), which is what the warp kernel consumes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pivamd_dev.h"

namespace pivamd {

// The glue restates NumPy float32 expressions: no fused multiply-adds.
#pragma clang fp contract(off)

__global__ void u16_to_f32_kernel(const uint16_t *in, float *out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += stride) out[k] = (float)in[k];
}

// pk fields (N, ny, nx) of a single-peak bulkxcorr2d; pred planar (2, N, ny, nx)
// or null; u planar (2, N, ny, nx) out; nan_mask u8, reason i8 out.
__global__ void gate_kernel(const float *pkx, const float *pky, const float *pkh, const float *pred,
                            const float *mask, int N, int nWin, float tx, float ty, float *u,
                            uint8_t *nanm, int8_t *reason, float *pkh_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = N * nWin;
    if (i >= total) return;
    const int w = i % nWin;
    float x = pkx[i], y = pky[i], h = pkh[i];
    const bool masked = mask && mask[w] != 0.0f;
    const bool fit_fail = (x != x) || (y != y);
    const bool large = (fabsf(x) > tx) || (fabsf(y) > ty);
    if (masked || large) {
        x = __builtin_nanf("");
        y = __builtin_nanf("");
        h = __builtin_nanf("");
    }
    if (pred) {
        y += pred[i];
        x += pred[(size_t)total + i];
    }
    const bool nm = (x != x) || (y != y) || masked;
    u[i] = nm ? 0.0f : y;
    u[(size_t)total + i] = nm ? 0.0f : x;
    nanm[i] = nm ? 1 : 0;
    int8_t r = 0;
    if (fit_fail) r = 1;
    if (large) r = 6;
    reason[i] = r;
    pkh_out[i] = h;
}

// Multi-peak gate (nPeaks > 1, cpu_instantaneous.py:586-635): pk fields
// (N, P, nWin) as bulkxcorr2d lays them out.  Every peak is NaN when its
// window is masked or the peak lies outside the gate, then the predictor is
// added (pxa, pya, pha keep all peaks); the primary peak becomes u (0 at
// invalid vectors), nan_mask and reasons as in gate_kernel; choice = 1.
__global__ void gate_peaks_kernel(const float *pkx, const float *pky, const float *pkh, const float *pred,
                                  const float *mask, int N, int P, int nWin, float tx, float ty, float *pxa,
                                  float *pya, float *pha, float *u, uint8_t *nanm, int8_t *reason, float *pkh0,
                                  uint8_t *choice)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = N * nWin;
    if (i >= total) return;
    const int n = i / nWin, w = i % nWin;
    const bool masked = mask && mask[w] != 0.0f;
    const float nan = __builtin_nanf("");
    bool fit_fail0 = false, large0 = false;
    for (int p = 0; p < P; ++p) {
        const size_t k = ((size_t)n * P + p) * nWin + w;
        float x = pkx[k], y = pky[k], h = pkh[k];
        const bool large = (fabsf(x) > tx) || (fabsf(y) > ty);
        if (p == 0) {
            fit_fail0 = (x != x) || (y != y);
            large0 = large;
        }
        if (masked || large) {
            x = nan;
            y = nan;
            h = nan;
        }
        if (pred) {
            y += pred[i];
            x += pred[(size_t)total + i];
        }
        pxa[k] = x;
        pya[k] = y;
        pha[k] = h;
        if (p == 0) {
            const bool nm = (x != x) || (y != y) || masked;
            u[i] = nm ? 0.0f : y;
            u[(size_t)total + i] = nm ? 0.0f : x;
            nanm[i] = nm ? 1 : 0;
            pkh0[i] = h;
        }
    }
    int8_t r = 0;
    if (fit_fail0) r = 1;
    if (large0) r = 6;
    reason[i] = r;
    choice[i] = 1;
}

// One substitution round (cpu_instantaneous.py:677-699): windows still
// invalid with peaks left move to their next peak; every window's current
// peak is gathered into usel / pk_sel (NaN where that peak is invalid), and
// the moved windows are re-judged from scratch (reason 13 if the new peak is
// NaN, else valid).
__global__ void peak_select_kernel(const float *pxa, const float *pya, const float *pha, int N, int P, int nWin,
                                   uint8_t *nanm, int8_t *reason, uint8_t *choice, float *usel, float *pk_sel,
                                   uint8_t *active)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = N * nWin;
    if (i >= total) return;
    const int n = i / nWin, w = i % nWin;
    const bool act = nanm[i] != 0 && choice[i] < P;
    int ch = choice[i];
    if (act) ch += 1;
    choice[i] = (uint8_t)ch;
    const size_t k = ((size_t)n * P + (ch - 1)) * nWin + w;
    const float x = pxa[k], y = pya[k];
    usel[i] = y;
    usel[(size_t)total + i] = x;
    pk_sel[i] = pha[k];
    active[i] = act ? 1 : 0;
    if (act) {
        const bool sub_nan = (x != x) || (y != y);
        nanm[i] = sub_nan ? 1 : 0;
        reason[i] = sub_nan ? 13 : 0;
    }
}

// After the substitution rounds (cpu_instantaneous.py:724-764): masked
// windows stay invalid, a NaN height of the chosen peak invalidates, the
// remaining reasons are filled in, and the outputs take NaN (0 at masked
// windows, choice 0) at invalid vectors.
__global__ void peak_final_kernel(const float *usel, const float *pk_sel, const float *mask, int N, int nWin,
                                  uint8_t *nanm, int8_t *reason, uint8_t *choice, float *uo, float *pkh_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = N * nWin;
    if (i >= total) return;
    const int w = i % nWin;
    const bool masked = mask && mask[w] != 0.0f;
    bool nm = nanm[i] != 0 || masked;
    int8_t rs = reason[i];
    const float hv = pk_sel[i];
    if (hv != hv) {
        nm = true;
        if (rs == 0) rs = 13;
    }
    if (nm && rs == 0) rs = 9;
    if (masked) rs = -1;
    nanm[i] = nm ? 1 : 0;
    reason[i] = rs;
    const float nanv = __builtin_nanf("");
    uo[i] = masked ? 0.0f : (nm ? nanv : usel[i]);
    uo[(size_t)total + i] = masked ? 0.0f : (nm ? nanv : usel[(size_t)total + i]);
    pkh_out[i] = nm ? nanv : hv;
    if (nm) choice[i] = 0;
}

// Batcher's odd-even merge sort of K registers (any K; 132 compare-exchanges
// for K = 24, 19 for K = 8), fully unrolled so every index is a constant.
template <int K>
__device__ __forceinline__ void sort_net(float (&w)[K])
{
#pragma unroll
    for (int p = 1; p < K; p += p)
#pragma unroll
        for (int k = p; k > 0; k /= 2)
#pragma unroll
            for (int j = k % p; j + k < K; j += k + k)
#pragma unroll
                for (int i = 0; i < k && i < K - j - k; ++i)
                    if ((j + i) / (p + p) == (j + i + k) / (p + p)) {
                        const float a = w[j + i], b = w[j + i + k];
                        w[j + i] = fminf(a, b);
                        w[j + i + k] = fmaxf(a, b);
                    }
}

// nanmedian of K values held in registers (bottleneck semantics: NaNs
// dropped, the middle element for an odd count, the mean of the middle pair
// for an even one, NaN when nothing is left).  NaNs become +inf and sort
// last; the n values in front are the finite ones (and any real +inf, equal
// to the stand-ins), so the middle of those is the median.
// ROOT: v holds squares of non-negative values; the median of their square
// roots is returned.  sqrtf is monotone (non-decreasing, NaN to NaN), so the
// order statistics of the roots are the roots of the order statistics: only
// the one or two selected values need a root, the same bits as taking every
// root first.
template <int K, bool ROOT = false>
__device__ __forceinline__ float nanmedian_reg(const float (&v)[K])
{
    float w[K];
    int n = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const bool ok = v[i] == v[i];
        n += ok ? 1 : 0;
        w[i] = ok ? v[i] : __builtin_inff();
    }
    if (n == 0) return __builtin_nanf("");
    sort_net<K>(w);
    const int klo = (n - 1) >> 1, khi = n >> 1;
    float lo = w[0], hi = w[0];
#pragma unroll
    for (int i = 1; i < K; ++i) {
        lo = i == klo ? w[i] : lo;
        hi = i == khi ? w[i] : hi;
    }
    if constexpr (ROOT) {
        lo = sqrtf(lo);
        hi = sqrtf(hi);
    }
    return (n & 1) ? lo : 0.5f * (lo + hi);
}

// Normalised-median test of one node with a SIZE x SIZE neighbourhood.
template <int SIZE>
__device__ __forceinline__ bool median_outlier(const float *uy, const float *ux, const float *mask, int r,
                                               int c, int ny, int nx, float eps, float thr)
{
    constexpr int K = SIZE * SIZE - 1, h = SIZE / 2;
    float vx[K], vy[K];
    int k = 0;
#pragma unroll
    for (int dr = -h; dr <= h; ++dr)
#pragma unroll
        for (int dc = -h; dc <= h; ++dc) {
            if (dr == 0 && dc == 0) continue;
            const int rr = r + dr, cc = c + dc;
            float a = __builtin_nanf(""), b = __builtin_nanf("");
            if (rr >= 0 && rr < ny && cc >= 0 && cc < nx) {
                const int q = rr * nx + cc;
                if (!(mask && mask[q] != 0.0f)) {
                    a = ux[q];
                    b = uy[q];
                }
            }
            vx[k] = a;
            vy[k] = b;
            ++k;
        }
    const float mu = nanmedian_reg<K>(vx), mv = nanmedian_reg<K>(vy);
    const float x0 = ux[r * nx + c], y0 = uy[r * nx + c];
    const float ddx = x0 - mu, ddy = y0 - mv;
    const float r0 = sqrtf(ddx * ddx + ddy * ddy);
    float ri2[K];    // squared residual norms; the median takes the roots
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const float ex = vx[i] - mu, ey = vy[i] - mv;
        ri2[i] = ex * ex + ey * ey;
    }
    const float rm = nanmedian_reg<K, true>(ri2);
    const float nr = r0 / (rm + eps);
    const bool exceeded = (nr > thr) || !isfinite(nr);
    int present = 0;
#pragma unroll
    for (int dr = -1; dr <= 1; ++dr)
#pragma unroll
        for (int dc = -1; dc <= 1; ++dc) {
            const int rr = r + dr, cc = c + dc;
            if (rr < 0 || rr >= ny || cc < 0 || cc >= nx) continue;
            const int q = rr * nx + cc;
            const bool m = mask && mask[q] != 0.0f;
            const float a = m ? __builtin_nanf("") : ux[q], b = m ? __builtin_nanf("") : uy[q];
            if ((isfinite(a) && isfinite(b)) || m) ++present;
        }
    return exceeded || present < 6;
}

// The k-th smallest (0-based) of the values f(0..K-1) that are not NaN,
// found by counting (O(K^2) compares, nothing stored): v_i is the k-th when
// fewer than k + 1 values are below it and more than k are at or below it.
template <typename F>
__device__ __forceinline__ float kth_finite(F f, int K, int k)
{
    float best = __builtin_nanf("");
    for (int i = 0; i < K; ++i) {
        const float vi = f(i);
        if (vi != vi) continue;
        int lt = 0, le = 0;
        for (int j = 0; j < K; ++j) {
            const float vj = f(j);
            lt += vj < vi ? 1 : 0;
            le += vj <= vi ? 1 : 0;
        }
        if (lt <= k && k < le) best = vi;
    }
    return best;
}

// nanmedian (bottleneck semantics, as nanmedian_reg) of f(0..K-1).
template <typename F>
__device__ __forceinline__ float nanmedian_count(F f, int K)
{
    int n = 0;
    for (int i = 0; i < K; ++i) {
        const float v = f(i);
        n += v == v ? 1 : 0;
    }
    if (n == 0) return __builtin_nanf("");
    const float lo = kth_finite(f, K, (n - 1) >> 1);
    if (n & 1) return lo;
    const float hi = kth_finite(f, K, n >> 1);
    return 0.5f * (lo + hi);
}

// The normalised-median test for any odd neighbourhood size (the reference
// accepts any odd size >= 3, outlier_detection.py:170-171): the same
// arithmetic as median_outlier<SIZE>, with the neighbours read by index and
// the medians selected by counting instead of a sorting network.
__device__ bool median_outlier_any(const float *uy, const float *ux, const float *mask, int r, int c, int ny,
                                   int nx, float eps, float thr, int size)
{
    const int h = size / 2, K = size * size - 1;
    // neighbour j (row-major, the centre skipped), NaN outside or masked
    auto nb = [&](int j, const float *f) {
        const int e = j < K / 2 ? j : j + 1;
        const int rr = r + e / size - h, cc = c + e % size - h;
        if (rr < 0 || rr >= ny || cc < 0 || cc >= nx) return __builtin_nanf("");
        const int q = rr * nx + cc;
        if (mask && mask[q] != 0.0f) return __builtin_nanf("");
        return f[q];
    };
    const float mu = nanmedian_count([&](int j) { return nb(j, ux); }, K);
    const float mv = nanmedian_count([&](int j) { return nb(j, uy); }, K);
    const float x0 = ux[r * nx + c], y0 = uy[r * nx + c];
    const float ddx = x0 - mu, ddy = y0 - mv;
    const float r0 = sqrtf(ddx * ddx + ddy * ddy);
    const float rm = nanmedian_count([&](int j) {
        const float ex = nb(j, ux) - mu, ey = nb(j, uy) - mv;
        return sqrtf(ex * ex + ey * ey);
    }, K);
    const float nr = r0 / (rm + eps);
    const bool exceeded = (nr > thr) || !isfinite(nr);
    int present = 0;
    for (int dr = -1; dr <= 1; ++dr)
        for (int dc = -1; dc <= 1; ++dc) {
            const int rr = r + dr, cc = c + dc;
            if (rr < 0 || rr >= ny || cc < 0 || cc >= nx) continue;
            const int q = rr * nx + cc;
            const bool m = mask && mask[q] != 0.0f;
            const float a = m ? __builtin_nanf("") : ux[q], b = m ? __builtin_nanf("") : uy[q];
            if ((isfinite(a) && isfinite(b)) || m) ++present;
        }
    return exceeded || present < 6;
}

// Normalised-median outlier test on u (planar, masked cells read as NaN),
// then the end-of-stage bookkeeping of cpu_instantaneous.py:726-757.
// The outlier methods combine by OR (outlier_detection.py:420-500
// apply_outlier_detection): the median test (use_median), the peak
// magnitude test pkh < peak_thr or NaN (use_peak, :47-84) and flags other
// kernels computed into extra (nullable); masked windows are never flagged.
//
// mode 0: detection and the final bookkeeping (choice, nullable, gets 0 at
// invalid vectors and 1 elsewhere); mode 1: detection only (the first round
// before secondary peaks); mode 2: re-judging a substitution round, only the
// windows flagged in active whose new peak is not NaN (reason 12).
__global__ void median_kernel(const float *u, const float *pkh, const float *mask, int N, int ny, int nx,
                              int use_median, float eps, float thr, int size, int use_peak, float peak_thr,
                              const uint8_t *extra, int mode, const uint8_t *active, uint8_t *nanm, int8_t *reason,
                              float *uo, float *pkh_out, uint8_t *choice)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nWin = ny * nx, total = N * nWin;
    if (i >= total) return;
    const int n = i / nWin, w = i % nWin, r = w / nx, c = w % nx;
    const float *uy = u + (size_t)n * nWin;
    const float *ux = u + (size_t)total + (size_t)n * nWin;
    const bool self_masked = mask && mask[w] != 0.0f;
    bool outlier = false;
    const float hv = pkh[i];
    if (mode == 2 && !(active[i] && uy[w] == uy[w] && ux[w] == ux[w])) return;
    if (!self_masked) {
        if (use_median)
            outlier = size == 3   ? median_outlier<3>(uy, ux, mask, r, c, ny, nx, eps, thr)
                      : size == 5 ? median_outlier<5>(uy, ux, mask, r, c, ny, nx, eps, thr)
                      : size == 7 ? median_outlier<7>(uy, ux, mask, r, c, ny, nx, eps, thr)
                                  : median_outlier_any(uy, ux, mask, r, c, ny, nx, eps, thr, size);
        if (use_peak) outlier = outlier || !(hv >= peak_thr);
        if (extra) outlier = outlier || extra[i] != 0;
    }
    bool nm = nanm[i] != 0;
    int8_t rs = reason[i];
    if (mode == 2) {
        if (outlier) {
            nanm[i] = 1;
            reason[i] = 12;
        }
        return;
    }
    if (outlier) {
        nm = true;
        if (rs == 0) rs = 10;
    }
    if (mode == 1) {
        nanm[i] = nm ? 1 : 0;
        reason[i] = rs;
        return;
    }
    if (hv != hv) {
        nm = true;
        if (rs == 0) rs = 13;
    }
    if (nm && rs == 0) rs = 9;
    if (self_masked) rs = -1;
    nanm[i] = nm ? 1 : 0;
    reason[i] = rs;
    const float nanv = __builtin_nanf("");
    uo[i] = self_masked ? 0.0f : (nm ? nanv : uy[w]);
    uo[(size_t)total + i] = self_masked ? 0.0f : (nm ? nanv : ux[w]);
    pkh_out[i] = nm ? nanv : hv;
    if (choice) choice[i] = nm ? 0 : 1;
}

// Nearest-donor fill of NaN cells (both components from the same donor).
// Donors: finite in both components and not excluded (masked).  Distance
// (dy*py)^2 + (dx*px)^2, ties to the smallest column, then the smallest row.
// One lane per cell; a ring search stops once the ring is farther than the
// best distance found.  A cell with no donor anywhere sets *err.
__global__ void infill_kernel(const float *u, const float *mask, int N, int ny, int nx, float py, float px,
                              int fill_masked, float *uo, int *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nWin = ny * nx, total = N * nWin;
    if (i >= total) return;
    const int n = i / nWin, w = i % nWin, r = w / nx, c = w % nx;
    const float *uy = u + (size_t)n * nWin;
    const float *ux = u + (size_t)total + (size_t)n * nWin;
    const float y0 = uy[w], x0 = ux[w];
    uo[i] = y0;
    uo[(size_t)total + i] = x0;
    // fill_masked = 0: fill NaN cells, masked cells are no donors (the
    // infill of a pass); 1: fill the masked cells themselves from the
    // unmasked ones (the masked-window fill of a predictor field)
    const bool fill = fill_masked ? (mask && mask[w] != 0.0f) : ((y0 != y0) || (x0 != x0));
    if (!fill) return;
    double best = 1e300;
    int br = -1, bc = -1;
    const int maxring = ny > nx ? ny : nx;
    for (int ring = 1; ring <= maxring; ++ring) {
        // every cell of this ring is at least ring*min(py,px) away
        const double lo = (double)ring * (double)(py < px ? py : px);
        if (br >= 0 && lo * lo > best) break;
        for (int dr = -ring; dr <= ring; ++dr) {
            const int rr = r + dr;
            if (rr < 0 || rr >= ny) continue;
            const bool edge_row = (dr == -ring || dr == ring);
            for (int dc = -ring; dc <= ring; dc += (edge_row ? 1 : 2 * ring)) {
                const int cc = c + dc;
                if (cc < 0 || cc >= nx) continue;
                const int k = rr * nx + cc;
                if (mask && mask[k] != 0.0f) continue;
                const float a = uy[k], b = ux[k];
                if (!(isfinite(a) && isfinite(b))) continue;
                const double d = ((double)dr * py) * ((double)dr * py) + ((double)dc * px) * ((double)dc * px);
                if (d < best || (d == best && (cc < bc || (cc == bc && rr < br)))) {
                    best = d;
                    br = rr;
                    bc = cc;
                }
            }
        }
    }
    if (br < 0) {
        // no donor anywhere: flag it (the caller raises, as the reference
        // does) and leave a zero displacement so later passes stay defined
        atomicOr(err, 1);
        uo[i] = 0.0f;
        uo[(size_t)total + i] = 0.0f;
        return;
    }
    uo[i] = uy[br * nx + bc];
    uo[(size_t)total + i] = ux[br * nx + bc];
}

// A linear fill of chosen cells from chosen donor cells, the same in every
// plane: u[p][fill_idx[i]] = sum_j gt[j][i] * u[p][bnd_idx[j]], in FP64,
// rounded once.  The masked-window harmonic fill of a predictor field has
// this form: its least-squares solution is a fixed linear map (the mask's
// geometry only) of the unmasked cells next to the mask.  One lane per fill
// cell (gt rows are read coalesced across lanes), one grid row per plane;
// fill and donor cells are disjoint, so the plane is updated in place.
__global__ void fill_linear_kernel(float *u, long plane_stride, int nf, const int *__restrict__ fill_idx, int nb,
                                   const int *__restrict__ bnd_idx, const double *__restrict__ gt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    float *pl = u + (size_t)blockIdx.y * plane_stride;
    double acc = 0.0;
    for (int j = 0; j < nb; ++j) acc += gt[(size_t)j * nf + i] * (double)pl[bnd_idx[j]];
    pl[fill_idx[i]] = (float)acc;
}

// np.pad(mode='edge') of the planar field (2, N, ny, nx) into (2, N, nyp, nxp).
__global__ void pad_kernel(const float *u, int N, int ny, int nx, int pre_y, int pre_x, int nyp, int nxp,
                           float *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = nyp * nxp, total = 2 * N * per;
    if (i >= total) return;
    const int plane = i / per, k = i % per;
    const int r = min(max(k / nxp - pre_y, 0), ny - 1);
    const int c = min(max(k % nxp - pre_x, 0), nx - 1);
    out[i] = u[(size_t)plane * ny * nx + r * nx + c];
}

// One axis of scipy.ndimage.gaussian_filter1d (mode 'nearest'): double
// accumulation of the symmetric kernel in the order of scipy's symmetric
// correlate1d (centre tap, then the paired taps from the outermost pair
// inward), rounded to float32.
__global__ void smooth_kernel(const float *in, int planes, int ny, int nx, int axis, const double *wts,
                              int radius, float *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = ny * nx, total = planes * per;
    if (i >= total) return;
    const int p = i / per, k = i % per, r = k / nx, c = k % nx;
    const float *src = in + (size_t)p * per;
    double s;
    if (axis == 0) {
        s = (double)src[r * nx + c] * wts[0];
        for (int j = radius; j >= 1; --j) {
            const int ra = min(max(r + j, 0), ny - 1), rb = min(max(r - j, 0), ny - 1);
            s += ((double)src[ra * nx + c] + (double)src[rb * nx + c]) * wts[j];
        }
    } else {
        s = (double)src[r * nx + c] * wts[0];
        for (int j = radius; j >= 1; --j) {
            const int ca = min(max(c + j, 0), nx - 1), cb = min(max(c - j, 0), nx - 1);
            s += ((double)src[r * nx + ca] + (double)src[r * nx + cb]) * wts[j];
        }
    }
    out[i] = (float)s;
}

__device__ __forceinline__ void cubic_coeffs(float x, float *w)
{
    const float A = -0.75f;
    w[0] = ((A * (x + 1) - 5 * A) * (x + 1) + 8 * A) * (x + 1) - 4 * A;
    w[1] = ((A + 2) * x - (A + 3)) * x * x + 1;
    w[2] = ((A + 2) * (1 - x) - (A + 3)) * (1 - x) * (1 - x) + 1;
    w[3] = 1.f - w[0] - w[1] - w[2];
}

// Bicubic remap of the planar field (2, N, nyo, nxo) at fractional indices
// (map_y, map_x) of the new grid (ny, nx): coordinates rounded to 1/32 px
// (round half to even), per-axis cubic weights on that lattice, replicate
// border, with the summation order of cv2.remap(INTER_CUBIC) for stencils
// inside the field and for those that reach a border.
__global__ void remap_kernel(const float *in, int N, int nyo, int nxo, const float *map_x, const float *map_y,
                             int ny, int nx, float *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = ny * nx, total = 2 * N * per;
    if (i >= total) return;
    const int plane = i / per, k = i % per;
    const float *src = in + (size_t)plane * nyo * nxo;
    const int X = __float2int_rn(map_x[k] * 32.0f);
    const int Y = __float2int_rn(map_y[k] * 32.0f);
    const int sx = (X >> 5) - 1, sy = (Y >> 5) - 1;
    float wx[4], wy[4];
    cubic_coeffs((float)(X & 31) * (1.0f / 32.0f), wx);
    cubic_coeffs((float)(Y & 31) * (1.0f / 32.0f), wy);
    // a stencil inside the field is summed row by row (four row sums added
    // in order), one reaching a border adds its 16 clamped-index products
    // one at a time in row-major order
    const bool inside = sx >= 0 && sy >= 0 && sx <= nxo - 4 && sy <= nyo - 4;
    float sum = 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int rr = min(max(sy + a, 0), nyo - 1);
        const float *row = src + (size_t)rr * nxo;
        float rs = 0.0f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int cc = min(max(sx + b, 0), nxo - 1);
            const float p = row[cc] * (wy[a] * wx[b]);
            if (inside) rs += p;
            else sum += p;
        }
        if (inside) sum += rs;
    }
    out[i] = sum;
}

// Plain copy of n floats, the calibration reference for HBM byte counters.
__global__ void copy_kernel(const float *in, float *out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += stride) out[k] = in[k];
}

static inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace pivamd

using namespace pivamd;

#define PIVAMD_LAUNCHED() (hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH)

extern "C" int pivamd_u16_to_f32(const void *d_in, float *d_out, size_t n, hipStream_t st)
{
    if (n == 0) return 0;
    unsigned g = nblk(n);
    if (g > 65536u * 8) g = 65536u * 8;
    hipLaunchKernelGGL(u16_to_f32_kernel, dim3(g), dim3(256), 0, st, (const uint16_t *)d_in, d_out, n);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_copy_f32(const float *d_in, float *d_out, size_t n, hipStream_t st)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(copy_kernel, dim3(2048), dim3(256), 0, st, d_in, d_out, n);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_gate(const float *pkx, const float *pky, const float *pkh, const float *pred,
                           const float *mask, int N, int nWin, float tx, float ty, float *u, uint8_t *nanm,
                           int8_t *reason, float *pkh_out, hipStream_t st)
{
    const size_t n = (size_t)N * nWin;
    if (n == 0) return 0;
    hipLaunchKernelGGL(gate_kernel, dim3(nblk(n)), dim3(256), 0, st, pkx, pky, pkh, pred, mask, N, nWin, tx, ty,
                       u, nanm, reason, pkh_out);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_median(const float *u, const float *pkh, const float *mask, int N, int ny, int nx,
                             float eps, float thr, int size, uint8_t *nanm, int8_t *reason, float *uo,
                             float *pkh_out, hipStream_t st)
{
    if (size < 3 || size % 2 == 0) return PIVAMD_ERR_ARG;
    const size_t n = (size_t)N * ny * nx;
    if (n == 0) return 0;
    hipLaunchKernelGGL(median_kernel, dim3(nblk(n)), dim3(256), 0, st, u, pkh, mask, N, ny, nx, 1, eps, thr, size,
                       0, 0.0f, (const uint8_t *)nullptr, 0, (const uint8_t *)nullptr, nanm, reason, uo, pkh_out,
                       (uint8_t *)nullptr);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_outlier_stage(const float *u, const float *pkh, const float *mask, int N, int ny, int nx,
                                    int use_median, float eps, float thr, int size, int use_peak, float peak_thr,
                                    const uint8_t *extra, uint8_t *nanm, int8_t *reason, float *uo,
                                    float *pkh_out, uint8_t *choice, hipStream_t st)
{
    if (use_median && (size < 3 || size % 2 == 0)) return PIVAMD_ERR_ARG;
    const size_t n = (size_t)N * ny * nx;
    if (n == 0) return 0;
    hipLaunchKernelGGL(median_kernel, dim3(nblk(n)), dim3(256), 0, st, u, pkh, mask, N, ny, nx, use_median, eps, thr,
                       size, use_peak, peak_thr, extra, 0, (const uint8_t *)nullptr, nanm, reason, uo, pkh_out,
                       choice);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_outlier_detect(const float *u, const float *pkh, const float *mask, int N, int ny, int nx,
                                     int use_median, float eps, float thr, int size, int use_peak, float peak_thr,
                                     const uint8_t *extra, const uint8_t *active, uint8_t *nanm, int8_t *reason,
                                     hipStream_t st)
{
    if (use_median && (size < 3 || size % 2 == 0)) return PIVAMD_ERR_ARG;
    const size_t n = (size_t)N * ny * nx;
    if (n == 0) return 0;
    hipLaunchKernelGGL(median_kernel, dim3(nblk(n)), dim3(256), 0, st, u, pkh, mask, N, ny, nx, use_median, eps, thr,
                       size, use_peak, peak_thr, extra, active ? 2 : 1, active, nanm, reason, (float *)nullptr,
                       (float *)nullptr, (uint8_t *)nullptr);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_gate_peaks(const float *pkx, const float *pky, const float *pkh, const float *pred,
                                 const float *mask, int N, int P, int nWin, float tx, float ty, float *pxa, float *pya,
                                 float *pha, float *u, uint8_t *nanm, int8_t *reason, float *pkh0, uint8_t *choice,
                                 hipStream_t st)
{
    const size_t n = (size_t)N * nWin;
    if (n == 0) return 0;
    if (P < 1 || P > 255) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL(gate_peaks_kernel, dim3(nblk(n)), dim3(256), 0, st, pkx, pky, pkh, pred, mask, N, P, nWin, tx,
                       ty, pxa, pya, pha, u, nanm, reason, pkh0, choice);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_peak_select(const float *pxa, const float *pya, const float *pha, int N, int P, int nWin,
                                  uint8_t *nanm, int8_t *reason, uint8_t *choice, float *usel, float *pk_sel,
                                  uint8_t *active, hipStream_t st)
{
    const size_t n = (size_t)N * nWin;
    if (n == 0) return 0;
    if (P < 1 || P > 255) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL(peak_select_kernel, dim3(nblk(n)), dim3(256), 0, st, pxa, pya, pha, N, P, nWin, nanm, reason,
                       choice, usel, pk_sel, active);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_peak_final(const float *usel, const float *pk_sel, const float *mask, int N, int nWin,
                                 uint8_t *nanm, int8_t *reason, uint8_t *choice, float *uo, float *pkh_out,
                                 hipStream_t st)
{
    const size_t n = (size_t)N * nWin;
    if (n == 0) return 0;
    hipLaunchKernelGGL(peak_final_kernel, dim3(nblk(n)), dim3(256), 0, st, usel, pk_sel, mask, N, nWin, nanm, reason,
                       choice, uo, pkh_out);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_infill_nearest(const float *u, const float *mask, int N, int ny, int nx, float py,
                                     float px, float *uo, int *d_err, hipStream_t st)
{
    const size_t n = (size_t)N * ny * nx;
    if (n == 0) return 0;
    if (!(py > 0.0f) || !(px > 0.0f)) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL(infill_kernel, dim3(nblk(n)), dim3(256), 0, st, u, mask, N, ny, nx, py, px, 0, uo, d_err);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_fill_masked_nearest(const float *u, const float *mask, int N, int ny, int nx, float py,
                                          float px, float *uo, int *d_err, hipStream_t st)
{
    const size_t n = (size_t)N * ny * nx;
    if (n == 0) return 0;
    if (!mask || !(py > 0.0f) || !(px > 0.0f)) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL(infill_kernel, dim3(nblk(n)), dim3(256), 0, st, u, mask, N, ny, nx, py, px, 1, uo, d_err);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_fill_linear(float *u, int planes, long plane_stride, int nf, const int *fill_idx, int nb,
                                  const int *bnd_idx, const double *gt, hipStream_t st)
{
    if (planes <= 0 || nf <= 0) return 0;
    if (nb <= 0 || !fill_idx || !bnd_idx || !gt || planes > 65535) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL(fill_linear_kernel, dim3((unsigned)((nf + 255) / 256), (unsigned)planes), dim3(256), 0, st, u,
                       plane_stride, nf, fill_idx, nb, bnd_idx, gt);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_pad_edge(const float *u, int N, int ny, int nx, int pre_y, int pre_x, int nyp, int nxp,
                               float *out, hipStream_t st)
{
    const size_t n = (size_t)2 * N * nyp * nxp;
    if (n == 0) return 0;
    hipLaunchKernelGGL(pad_kernel, dim3(nblk(n)), dim3(256), 0, st, u, N, ny, nx, pre_y, pre_x, nyp, nxp, out);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_smooth_axis(const float *in, int planes, int ny, int nx, int axis, const double *d_wts,
                                  int radius, float *out, hipStream_t st)
{
    const size_t n = (size_t)planes * ny * nx;
    if (n == 0) return 0;
    hipLaunchKernelGGL(smooth_kernel, dim3(nblk(n)), dim3(256), 0, st, in, planes, ny, nx, axis, d_wts, radius,
                       out);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_remap_cubic(const float *in, int N, int nyo, int nxo, const float *map_x,
                                  const float *map_y, int ny, int nx, float *out, hipStream_t st)
{
    const size_t n = (size_t)2 * N * ny * nx;
    if (n == 0) return 0;
    hipLaunchKernelGGL(remap_kernel, dim3(nblk(n)), dim3(256), 0, st, in, N, nyo, nxo, map_x, map_y, ny, nx, out);
    return PIVAMD_LAUNCHED();
}
