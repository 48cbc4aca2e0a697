// outliers.hip -- the sigma and divergence/vorticity outlier methods.
//
// Two of the reference's outlier methods besides the median and
// peak-magnitude tests (pivtools_cli/piv/piv_backend/outlier_detection.py):
//
//   sigma     (:244-313)  the 8-neighbour mean and standard deviation of a
//             scalar field (the vector norm, or each component), from
//             scipy.ndimage.uniform_filter(size 3, mode 'constant') of the
//             value, its square and the finite count; a node is an outlier
//             when |v - mean8| > k std8, or v, mean8 are not finite, or
//             std8 == 0, or no neighbour is finite.
//   div_vort  (:316-373)  central differences with periodic wrap (np.roll),
//             divergence and vorticity, and thresholds 6 * 1.4826 * MAD about
//             the median of the finite values unless given.
//
// Masked windows read as NaN (apply_outlier_detection, :450-455).  Every
// float32 operation is the reference's, in its order, with contraction off,
// so the flags are the reference's bit for bit:
//   * uniform_filter runs axis 0 then axis 1; along each line it keeps a
//     running sum in double (the first window summed in order, then
//     tmp += new - old) and stores tmp / 3 rounded to float32;
//   * np.median of float32 takes the middle element, or (a + b) / 2 in
//     float32 for an even count; the order statistics come from a radix
//     select over the finite values, one workgroup per image and field.
// The flags are OR-ed into a byte per window for pivamd_outlier_stage.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pivamd_dev.h"

namespace pivamd {

// 0: |u| (combine 'norm'); 1: ux; 2: uy
__device__ inline float sigma_value(const float *uy, const float *ux, const float *mask, int w, int field)
{
#pragma clang fp contract(off)
    const float nan = __builtin_nanf("");
    const bool m = mask && mask[w] != 0.0f;
    const float a = m ? nan : uy[w], b = m ? nan : ux[w];
    if (field == 1) return b;
    if (field == 2) return a;
    const float s = b * b + a * a;
    return sqrtf(s);
}

// uniform_filter1d(size 3, constant 0) of one line in the reference's
// running-sum form; `get(k)` is element k of the line, `put(k, v)` stores.
template <class Get, class Put>
__device__ inline void uniform3_line(int n, Get get, Put put)
{
    double tmp = 0.0;
    tmp += 0.0;                                   // the zero before the line
    tmp += (double)get(0);
    tmp += n > 1 ? (double)get(1) : 0.0;
    put(0, (float)(tmp / 3.0));
    for (int k = 1; k < n; ++k) {
        const double add = k + 1 < n ? (double)get(k + 1) : 0.0;
        const double sub = k - 2 >= 0 ? (double)get(k - 2) : 0.0;
        tmp += add - sub;
        put(k, (float)(tmp / 3.0));
    }
}

// axis 0: one lane per (image, column); the three filtered fields go to t0..t2
__global__ void sigma_axis0_kernel(const float *u, const float *mask, int N, int ny, int nx, int field,
                                   float *t0, float *t1, float *t2)
{
#pragma clang fp contract(off)
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N * nx) return;
    const int img = id / nx, c = id % nx;
    const size_t G = (size_t)ny * nx, total = (size_t)N * G;
    const float *uy = u + img * G, *ux = u + total + img * G;
    auto v0 = [&](int r) {
        const float v = sigma_value(uy, ux, mask, r * nx + c, field);
        return isfinite(v) ? v : 0.0f;
    };
    float *o0 = t0 + img * G, *o1 = t1 + img * G, *o2 = t2 + img * G;
    uniform3_line(ny, [&](int r) { return v0(r); }, [&](int r, float x) { o0[r * nx + c] = x; });
    uniform3_line(ny, [&](int r) { const float x = v0(r); return x * x; }, [&](int r, float x) { o1[r * nx + c] = x; });
    uniform3_line(ny,
                  [&](int r) { return isfinite(sigma_value(uy, ux, mask, r * nx + c, field)) ? 1.0f : 0.0f; },
                  [&](int r, float x) { o2[r * nx + c] = x; });
}

// axis 1 and the test: one lane per (image, row)
__global__ void sigma_axis1_kernel(const float *u, const float *mask, int N, int ny, int nx, int field, float thr,
                                   const float *t0, const float *t1, const float *t2, float *s0, float *s1,
                                   float *s2, uint8_t *extra)
{
#pragma clang fp contract(off)
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= N * ny) return;
    const int img = id / ny, r = id % ny;
    const size_t G = (size_t)ny * nx, total = (size_t)N * G;
    const float *uy = u + img * G, *ux = u + total + img * G;
    const size_t row = img * G + (size_t)r * nx;
    uniform3_line(nx, [&](int c) { return t0[row + c]; }, [&](int c, float x) { s0[row + c] = x; });
    uniform3_line(nx, [&](int c) { return t1[row + c]; }, [&](int c, float x) { s1[row + c] = x; });
    uniform3_line(nx, [&](int c) { return t2[row + c]; }, [&](int c, float x) { s2[row + c] = x; });
    for (int c = 0; c < nx; ++c) {
        const float v = sigma_value(uy, ux, mask, r * nx + c, field);
        const bool fin = isfinite(v);
        const float cv = fin ? v : 0.0f, cc = fin ? 1.0f : 0.0f;
        const float sum9 = s0[row + c] * 9.0f, sumsq9 = s1[row + c] * 9.0f, cnt9 = s2[row + c] * 9.0f;
        const float sum8 = sum9 - cv;
        const float sq = cv * cv;
        const float sumsq8 = sumsq9 - sq;
        const float cnt8 = cnt9 - cc;
        const float mean8 = sum8 / cnt8;
        const float q = sumsq8 / cnt8;
        const float mm = mean8 * mean8;
        float var8 = q - mm;
        var8 = (var8 != var8) ? var8 : (var8 > 0.0f ? var8 : 0.0f);   // np.maximum keeps NaN
        const float std8 = sqrtf(var8);
        const float lim = thr * std8;
        const bool out = (fabsf(v - mean8) > lim) || !fin || !isfinite(mean8) || (std8 == 0.0f) || (cnt8 < 1.0f);
        if (out) extra[row + c] = 1;
    }
}

// divergence and vorticity with periodic central differences
__global__ void divvort_kernel(const float *u, const float *mask, int N, int ny, int nx, float *dv, float *vo)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t G = (size_t)ny * nx, total = (size_t)N * G;
    if ((size_t)i >= total) return;
    const int img = i / (int)G, w = i % (int)G, r = w / nx, c = w % nx;
    const float *uy = u + img * G, *ux = u + total + img * G;
    const float nan = __builtin_nanf("");
    auto X = [&](int rr, int cc) { const int k = rr * nx + cc; return (mask && mask[k] != 0.0f) ? nan : ux[k]; };
    auto Y = [&](int rr, int cc) { const int k = rr * nx + cc; return (mask && mask[k] != 0.0f) ? nan : uy[k]; };
    const int cp = c + 1 == nx ? 0 : c + 1, cm = c == 0 ? nx - 1 : c - 1;
    const int rp = r + 1 == ny ? 0 : r + 1, rm = r == 0 ? ny - 1 : r - 1;
    const float dudx = 0.5f * (X(r, cp) - X(r, cm));
    const float dudy = 0.5f * (X(rp, c) - X(rm, c));
    const float dvdx = 0.5f * (Y(r, cp) - Y(r, cm));
    const float dvdy = 0.5f * (Y(rp, c) - Y(rm, c));
    dv[i] = dudx + dvdy;
    vo[i] = dvdx - dudy;
}

__device__ inline uint32_t ord_key(float x)
{
    const uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ inline float key_val(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

constexpr int kSelThreads = 1024;

// k-th smallest (0-based) of the values val(i), i < n, that are finite, by
// an 8-bit radix select on the order-preserving keys.  Block-uniform.
template <class Val>
__device__ float block_select(int n, int k, Val val, uint32_t *hist, uint32_t *s_bc)
{
    uint32_t prefix = 0u, pmask = 0u;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0u;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const float x = val(i);
            if (!isfinite(x)) continue;
            const uint32_t key = ord_key(x);
            if ((key & pmask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t cum = 0u, b = 0u;
            for (; b < 256u; ++b) {
                if (cum + hist[b] > (uint32_t)k) break;
                cum += hist[b];
            }
            s_bc[0] = b;
            s_bc[1] = cum;
        }
        __syncthreads();
        const uint32_t b = s_bc[0];
        k -= (int)s_bc[1];
        prefix |= b << shift;
        pmask |= 255u << shift;
        __syncthreads();
    }
    return key_val(prefix);
}

// np.median of the finite values of val(0..n-1); NaN if none.
template <class Val>
__device__ float block_median(int n, Val val, uint32_t *hist, uint32_t *s_bc, int *s_cnt)
{
    __syncthreads();
    if (threadIdx.x == 0) *s_cnt = 0;
    __syncthreads();
    int cnt = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) cnt += isfinite(val(i)) ? 1 : 0;
    atomicAdd(s_cnt, cnt);
    __syncthreads();
    const int m = *s_cnt;
    if (m == 0) return __builtin_nanf("");
    const int h = m / 2;
    if (m & 1) return block_select(n, h, val, hist, s_bc);
    const float lo = block_select(n, h - 1, val, hist, s_bc);
    const float hi = block_select(n, h, val, hist, s_bc);
    const float s = lo + hi;
    return s / 2.0f;
}

// thresholds: blockIdx.x = image, blockIdx.y = 0 divergence / 1 vorticity
__global__ void __launch_bounds__(kSelThreads) divvort_thresh_kernel(const float *dv, const float *vo, int G,
                                                                     float *thr)
{
#pragma clang fp contract(off)
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_bc[2];
    __shared__ int s_cnt;
    const float *x = (blockIdx.y == 0 ? dv : vo) + (size_t)blockIdx.x * G;
    const float med = block_median(G, [&](int i) { return x[i]; }, hist, s_bc, &s_cnt);
    float t;
    if (med != med) {
        t = __builtin_inff();                    // no finite value
    } else {
        const float dmed = block_median(G, [&](int i) { return fabsf(x[i] - med); }, hist, s_bc, &s_cnt);
        const float mad = 1.4826f * dmed;
        t = 6.0f * mad;
    }
    if (threadIdx.x == 0) thr[2 * blockIdx.x + blockIdx.y] = t;
}

__global__ void divvort_flag_kernel(const float *dv, const float *vo, int N, int G, const float *thr, float div_t,
                                    float vort_t, uint8_t *extra)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * G) return;
    const int img = i / G;
    const float td = div_t == div_t ? div_t : thr[2 * img];
    const float tv = vort_t == vort_t ? vort_t : thr[2 * img + 1];
    if ((fabsf(dv[i]) > td) || (fabsf(vo[i]) > tv)) extra[i] = 1;
}

}  // namespace pivamd

using namespace pivamd;

#define PIVAMD_LAUNCHED() (hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH)

extern "C" unsigned long pivamd_outlier_methods_workspace(int N, int ny, int nx)
{
    if (N <= 0 || ny <= 0 || nx <= 0) return 0;
    const size_t G = (size_t)ny * nx;
    return (unsigned long)(6 * (size_t)N * G * sizeof(float) + 2 * (size_t)N * sizeof(float) + 256);
}

extern "C" int pivamd_outlier_sigma(const float *u, const float *mask, int N, int ny, int nx, float sigma_thr,
                                    int field, unsigned char *extra, void *d_work, unsigned long work_bytes,
                                    hipStream_t st)
{
    if (N <= 0 || ny <= 0 || nx <= 0) return 0;
    if (field < 0 || field > 2 || !u || !extra) return PIVAMD_ERR_ARG;
    if (!d_work || work_bytes < pivamd_outlier_methods_workspace(N, ny, nx)) return PIVAMD_ERR_NOMEM;
    const size_t NG = (size_t)N * ny * nx;
    float *t = (float *)d_work;
    hipLaunchKernelGGL(sigma_axis0_kernel, dim3((N * nx + 127) / 128), dim3(128), 0, st, u, mask, N, ny, nx, field,
                       t, t + NG, t + 2 * NG);
    hipLaunchKernelGGL(sigma_axis1_kernel, dim3((N * ny + 127) / 128), dim3(128), 0, st, u, mask, N, ny, nx, field,
                       sigma_thr, t, t + NG, t + 2 * NG, t + 3 * NG, t + 4 * NG, t + 5 * NG, extra);
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_outlier_div_vort(const float *u, const float *mask, int N, int ny, int nx, float div_thr,
                                       float vort_thr, unsigned char *extra, void *d_work, unsigned long work_bytes,
                                       hipStream_t st)
{
    if (N <= 0 || ny <= 0 || nx <= 0) return 0;
    if (!u || !extra || N > 65535) return PIVAMD_ERR_ARG;
    if (!d_work || work_bytes < pivamd_outlier_methods_workspace(N, ny, nx)) return PIVAMD_ERR_NOMEM;
    const int G = ny * nx;
    const size_t NG = (size_t)N * G;
    float *dv = (float *)d_work, *vo = dv + NG, *thr = dv + 6 * NG;
    hipLaunchKernelGGL(divvort_kernel, dim3((unsigned)((NG + 255) / 256)), dim3(256), 0, st, u, mask, N, ny, nx, dv,
                       vo);
    if (div_thr != div_thr || vort_thr != vort_thr)
        hipLaunchKernelGGL(divvort_thresh_kernel, dim3(N, 2), dim3(kSelThreads), 0, st, dv, vo, G, thr);
    hipLaunchKernelGGL(divvort_flag_kernel, dim3((unsigned)((NG + 255) / 256)), dim3(256), 0, st, dv, vo, N, G, thr,
                       div_thr, vort_thr, extra);
    return PIVAMD_LAUNCHED();
}
