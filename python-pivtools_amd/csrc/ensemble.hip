// ensemble.hip -- ensemble correlation accumulation (gfx950).
//
// Replaces bulkxcorr2d_accumulate_triple and bulkxcorr2d_accumulate
// (pivtools_cli/lib/PIV_2d_cross_correlate.c:432-484, 502-833).  For every
// window and every image pair it forms
//     AB = xcorr(slot0 = (B - mB) wB_AB, slot1 = (A - mA) wA_AB)
//     AA = power((A - mA) wAutoA),   BB = power((B - mB) wAutoB)
// (optional weighted means, optional per-pair normalisation by the zero-lag
// auto energies, pairs with a non-positive energy skipped) and adds the
// central crop of each plane to the caller's sums.
//
// By linearity of the inverse FFT and of the crop, the sums are taken in
// the frequency domain: each window keeps the half-spectrum sums of AB, AA
// and BB in FP64 (device workspace, owned lane by lane, no atomics), and is
// inverse-transformed once after the last pair.  Per pair and window the
// work is two complex 2D FFTs: z1 = pAB_B + i pAB_A gives F(pAB_B) and
// F(pAB_A), z2 = pAA + i pBB gives F(pAA) and F(pBB).
#include <hip/hip_runtime.h>

#include "fft_reg.h"
#include "pivamd_dev.h"

namespace pivamd {

struct EnsArgs {
    const float *imgA, *imgB, *mask, *ctrX, *ctrY, *wAab, *wBab, *wAA, *wBB;
    float *outAB, *outAA, *outBB;
    double *acc;            // nWin * 4 * NY * (NX/2+1)
    int H, W, nImages, nWy, nWx, fitY, fitX, meanSub, perPairNorm, triple, nblocks;
    int sharedW;            // wAab == wAA and wBab == wBB: AA, BB come from the AB transform
    float sumWBab, sumWAA, sumWBB;
};

constexpr int kEnsFlush = 64;   // images per FP32 register partial before the FP64 flush

constexpr int kEnsMaxLds = 150 * 1024;

template <int N>
struct EnsCfg {
    static constexpr int T = N;
    static constexpr int PX = N + 1;
    static constexpr int tile = N * PX;
    static constexpr int HW = N / 2 + 1;
    static constexpr int lds_win = 2 * tile * 4 + 6 * T * 4;
    static constexpr int wpb_thr = (256 / T) > 0 ? (256 / T) : 1;
    static constexpr int wpb_lds = kEnsMaxLds / lds_win;
    static constexpr int WPB = wpb_thr < wpb_lds ? wpb_thr : wpb_lds;
    static constexpr int BLOCK = WPB * T;
};

// Sum over the T lanes of one window via LDS, every lane summing in lane
// order (identical bits in every lane).
__device__ __forceinline__ void ens_sum3(float &a, float &b, float &c, float *ra, float *rb, float *rc,
                                         int t, int T)
{
    ra[t] = a;
    rb[t] = b;
    rc[t] = c;
    __syncthreads();
    float sa = 0.0f, sb = 0.0f, sc = 0.0f;
    for (int k = 0; k < T; ++k) {
        sa += ra[k];
        sb += rb[k];
        sc += rc[k];
    }
    __syncthreads();
    a = sa;
    b = sb;
    c = sc;
}

// Forward 2D FFT of the packed tile (column pass from registers, row pass
// through LDS).  On return the tile holds Z[ky][kx].
template <int N>
__device__ __forceinline__ void fwd2d(float (&zr)[N], float (&zi)[N], float *Sr, float *Si, int t, bool on)
{
    constexpr int PX = N + 1;
    if (on) {
        fft_reg<N, false>(zr, zi);
#pragma unroll
        for (int y = 0; y < N; ++y) {
            Sr[y * PX + t] = zr[y];
            Si[y * PX + t] = zi[y];
        }
    }
    __syncthreads();
    if (on) {
        float xr[N], xi[N];
#pragma unroll
        for (int x = 0; x < N; ++x) {
            xr[x] = Sr[t * PX + x];
            xi[x] = Si[t * PX + x];
        }
        fft_reg<N, false>(xr, xi);
#pragma unroll
        for (int x = 0; x < N; ++x) {
            Sr[t * PX + x] = xr[x];
            Si[t * PX + x] = xi[x];
        }
    }
    __syncthreads();
}

template <int N>
__global__ void __launch_bounds__((EnsCfg<N>::BLOCK)) ens_kernel(EnsArgs g)
{
    using C = EnsCfg<N>;
    constexpr int T = C::T, PX = C::PX, WPB = C::WPB, HW = C::HW;
    constexpr int numel = N * N;
    __shared__ float s_re[WPB * C::tile];
    __shared__ float s_im[WPB * C::tile];
    __shared__ float s_red[6][WPB * T];

    const int per = (int)gridDim.x >> 3;
    const int bb = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
    if (bb >= g.nblocks) return;
    const int wloc = (int)threadIdx.x / T;
    const int t = (int)threadIdx.x - wloc * T;
    float *Sr = s_re + wloc * C::tile;
    float *Si = s_im + wloc * C::tile;
    float *r0 = &s_red[0][wloc * T], *r1 = &s_red[1][wloc * T], *r2 = &s_red[2][wloc * T];

    const int nWin = g.nWy * g.nWx;
    const int w = bb * WPB + wloc;
    bool valid = w < nWin;
    int row0 = 0, col0 = 0;
    if (valid) {
        const int ii = w % g.nWx, jj = w / g.nWx;
        if (g.mask && g.mask[w] == 1.0f) valid = false;
        // float arithmetic, as PIV_2d_cross_correlate.c:555-556 / 743-744
        row0 = (int)floorf(g.ctrY[jj] - ((float)N - 1.0f) / 2.0f + 0.5f);
        col0 = (int)floorf(g.ctrX[ii] - ((float)N - 1.0f) / 2.0f + 0.5f);
        if (row0 < 0 || col0 < 0 || row0 + N > g.H || col0 + N > g.W) valid = false;
    }
    double *acc = g.acc + (size_t)(valid ? w : 0) * 4 * N * HW;
    const bool specLane = valid && t < HW;
    if (specLane) {
        for (int k = 0; k < 4 * N; ++k) acc[(size_t)k * HW + t] = 0.0;
    }

    const size_t plane = (size_t)g.H * g.W;
    for (int n = 0; n < g.nImages; ++n) {
        float ra[N], rb[N];
        float sAB_B = 0.0f, sAA = 0.0f, sBB = 0.0f;
        if (valid) {
            const float *pa = g.imgA + (size_t)n * plane + (size_t)row0 * g.W + col0 + t;
            const float *pb = g.imgB + (size_t)n * plane + (size_t)row0 * g.W + col0 + t;
#pragma unroll
            for (int y = 0; y < N; ++y) {
                ra[y] = pa[(size_t)y * g.W];
                rb[y] = pb[(size_t)y * g.W];
                if (g.meanSub) {
                    sAB_B += rb[y] * g.wBab[y * N + t];
                    sAA += ra[y] * g.wAA[y * N + t];
                    sBB += rb[y] * g.wBB[y * N + t];
                }
            }
        } else {
#pragma unroll
            for (int y = 0; y < N; ++y) { ra[y] = 0.0f; rb[y] = 0.0f; }
        }
        float mB = 0.0f, mAA = 0.0f, mBB = 0.0f;
        if (g.meanSub) {
            ens_sum3(sAB_B, sAA, sBB, r0, r1, r2, t, T);
            mB = (g.sumWBab > 0.0f) ? sAB_B / g.sumWBab : 0.0f;
            mAA = (g.sumWAA > 0.0f) ? sAA / g.sumWAA : 0.0f;
            mBB = (g.sumWBB > 0.0f) ? sBB / g.sumWBB : 0.0f;
        }
        float z1r[N], z1i[N], z2r[N], z2i[N];
        float eAA = 0.0f, eBB = 0.0f, dummy = 0.0f;
#pragma unroll
        for (int y = 0; y < N; ++y) {
            z1r[y] = (rb[y] - mB) * g.wBab[y * N + t];
            z1i[y] = (ra[y] - mAA) * g.wAab[y * N + t];
            z2r[y] = (ra[y] - mAA) * g.wAA[y * N + t];
            z2i[y] = (rb[y] - mBB) * g.wBB[y * N + t];
            eAA += z2r[y] * z2r[y];
            eBB += z2i[y] * z2i[y];
        }
        double sAB = 1.0, sA = 1.0, sB = 1.0;
        bool use = valid;
        if (g.perPairNorm) {
            ens_sum3(eAA, eBB, dummy, r0, r1, r2, t, T);
            if (eAA <= 0.0f || eBB <= 0.0f) use = false;
            sAB = (double)(1.0f / sqrtf(eAA * eBB));
            sA = (double)(1.0f / eAA);
            sB = (double)(1.0f / eBB);
        }
        // ---- AB: z1 = pAB_B + i pAB_A, C = F(pAB_B) conj(F(pAB_A)) ----
        fwd2d<N>(z1r, z1i, Sr, Si, t, use);
        if (use && t < HW) {
            const int mx = (N - t) % N;
#pragma unroll 4
            for (int ky = 0; ky < N; ++ky) {
                const int my = (N - ky) % N;
                const float zR = Sr[ky * PX + t], zI = Si[ky * PX + t];
                const float mR = Sr[my * PX + mx], mI = Si[my * PX + mx];
                const float xr_ = 0.5f * (zR + mR), xi_ = 0.5f * (zI - mI);      // F(pAB_B)
                const float yr_ = 0.5f * (zI + mI), yi_ = -0.5f * (zR - mR);     // F(pAB_A)
                const double cr = (double)(xr_ * yr_ + xi_ * yi_);
                const double ci = (double)(xi_ * yr_ - xr_ * yi_);
                acc[(size_t)(0 * N + ky) * HW + t] += sAB * cr;
                acc[(size_t)(1 * N + ky) * HW + t] += sAB * ci;
            }
        }
        if (g.triple) {
            __syncthreads();
            // ---- AA, BB: z2 = pAA + i pBB -------------------------------
            fwd2d<N>(z2r, z2i, Sr, Si, t, use);
            if (use && t < HW) {
                const int mx = (N - t) % N;
#pragma unroll 4
                for (int ky = 0; ky < N; ++ky) {
                    const int my = (N - ky) % N;
                    const float zR = Sr[ky * PX + t], zI = Si[ky * PX + t];
                    const float mR = Sr[my * PX + mx], mI = Si[my * PX + mx];
                    const float ar = 0.5f * (zR + mR), ai = 0.5f * (zI - mI);
                    const float br = 0.5f * (zI + mI), bi = -0.5f * (zR - mR);
                    acc[(size_t)(2 * N + ky) * HW + t] += sA * (double)(ar * ar + ai * ai);
                    acc[(size_t)(3 * N + ky) * HW + t] += sB * (double)(br * br + bi * bi);
                }
            }
        }
        __syncthreads();
    }

    // ---- readout: one c2r per accumulated spectrum, shift, crop, += -------
    const int oh = g.fitY, ow = g.fitX;
    const int sy = (N - oh) / 2, sx = (N - ow) / 2;
    for (int which = 0; which < (g.triple ? 3 : 1); ++which) {
        if (valid && t < HW) {
            float cr[N], ci[N];
#pragma unroll
            for (int ky = 0; ky < N; ++ky) {
                if (which == 0) {
                    cr[ky] = (float)acc[(size_t)(0 * N + ky) * HW + t];
                    ci[ky] = (float)acc[(size_t)(1 * N + ky) * HW + t];
                } else {
                    cr[ky] = (float)acc[(size_t)((which + 1) * N + ky) * HW + t];
                    ci[ky] = 0.0f;
                }
            }
            fft_reg<N, true>(cr, ci);
#pragma unroll
            for (int y = 0; y < N; ++y) {
                Sr[y * PX + t] = cr[y];
                Si[y * PX + t] = ci[y];
            }
        }
        __syncthreads();
        if (valid && t < N / 2) {
            const int ya = t, yb = t + N / 2;
            float xr[N], xi[N];
#pragma unroll
            for (int k = 0; k <= N / 2; ++k) {
                float g1r = Sr[ya * PX + k], g1i = Si[ya * PX + k];
                float g2r = Sr[yb * PX + k], g2i = Si[yb * PX + k];
                if (k == 0 || k == N / 2) { g1i = 0.0f; g2i = 0.0f; }
                xr[k] = g1r - g2i;
                xi[k] = g1i + g2r;
                if (k != 0 && k != N / 2) {
                    xr[N - k] = g1r + g2i;
                    xi[N - k] = -g1i + g2r;
                }
            }
            fft_reg<N, true>(xr, xi);
            const float mul = 1.0f / (float)numel;
            float *out = which == 0 ? g.outAB : (which == 1 ? g.outAA : g.outBB);
            out += (size_t)w * oh * ow;
            // plane row (y + N/2) % N holds c[y]; crop rows sy..sy+oh-1
#pragma unroll
            for (int x = 0; x < N; ++x) {
                const int xs = (x + N / 2) % N;
                if (xs >= sx && xs < sx + ow) {
                    if (yb >= sy && yb < sy + oh) out[(yb - sy) * ow + (xs - sx)] += xr[x] * mul;
                    if (ya >= sy && ya < sy + oh) out[(ya - sy) * ow + (xs - sx)] += xi[x] * mul;
                }
            }
        }
        __syncthreads();
    }
}

// Register-accumulating variant for N <= 32.  Lane t of a window owns the
// spectral row ky = t (all N lanes busy) and keeps the AB, AA and BB sums of
// that row in FP32 registers across the call's images, flushing them into the
// FP64 workspace every kEnsFlush images; the workspace is touched once per
// flush instead of once per image.  With shared weights (wAab == wAA,
// wBab == wBB) pAA = pAB_A and pBB = pAB_B, so one complex transform per pair
// gives all three spectra.
template <int N, bool SHARED>
__global__ void __launch_bounds__((EnsCfg<N>::BLOCK)) ens_kernel_reg(EnsArgs g)
{
    using C = EnsCfg<N>;
    constexpr int T = C::T, PX = C::PX, WPB = C::WPB, HW = C::HW;
    constexpr int numel = N * N;
    __shared__ float s_re[WPB * C::tile];
    __shared__ float s_im[WPB * C::tile];
    __shared__ float s_red[6][WPB * T];

    const int per = (int)gridDim.x >> 3;
    const int bb = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
    if (bb >= g.nblocks) return;
    const int wloc = (int)threadIdx.x / T;
    const int t = (int)threadIdx.x - wloc * T;
    float *Sr = s_re + wloc * C::tile;
    float *Si = s_im + wloc * C::tile;
    float *r0 = &s_red[0][wloc * T], *r1 = &s_red[1][wloc * T], *r2 = &s_red[2][wloc * T];

    const int nWin = g.nWy * g.nWx;
    const int w = bb * WPB + wloc;
    bool valid = w < nWin;
    int row0 = 0, col0 = 0;
    if (valid) {
        const int ii = w % g.nWx, jj = w / g.nWx;
        if (g.mask && g.mask[w] == 1.0f) valid = false;
        // float arithmetic, as PIV_2d_cross_correlate.c:555-556 / 743-744
        row0 = (int)floorf(g.ctrY[jj] - ((float)N - 1.0f) / 2.0f + 0.5f);
        col0 = (int)floorf(g.ctrX[ii] - ((float)N - 1.0f) / 2.0f + 0.5f);
        if (row0 < 0 || col0 < 0 || row0 + N > g.H || col0 + N > g.W) valid = false;
    }
    double *acc = g.acc + (size_t)(valid ? w : 0) * 4 * N * HW;
    const int ky = t, my = (N - ky) % N;

    float aRe[HW], aIm[HW], aAA[HW], aBB[HW];
#pragma unroll
    for (int k = 0; k < HW; ++k) { aRe[k] = 0.0f; aIm[k] = 0.0f; aAA[k] = 0.0f; aBB[k] = 0.0f; }
    bool first = true;
    auto flush = [&]() {
        if (valid) {
            double *row = acc + (size_t)ky * HW;
            // an opaque base keeps the 4 x HW bin addresses out of the image
            // loop: hoisted, they held about 140 registers for the whole call
            asm volatile("" : "+v"(row));
#pragma unroll
            for (int k = 0; k < HW; ++k) {
                double *p0 = row + k, *p1 = row + N * HW + k, *p2 = row + 2 * N * HW + k, *p3 = row + 3 * N * HW + k;
                *p0 = (first ? 0.0 : *p0) + (double)aRe[k];
                *p1 = (first ? 0.0 : *p1) + (double)aIm[k];
                *p2 = (first ? 0.0 : *p2) + (double)aAA[k];
                *p3 = (first ? 0.0 : *p3) + (double)aBB[k];
                aRe[k] = 0.0f; aIm[k] = 0.0f; aAA[k] = 0.0f; aBB[k] = 0.0f;
                // one bin at a time: hoisting every bin's FP64 loads would
                // raise the kernel's register peak for a once-per-flush step
                asm volatile("" ::: "memory");
            }
        }
        first = false;
    };

    const size_t plane = (size_t)g.H * g.W;
    for (int n = 0; n < g.nImages; ++n) {
        // keep the weight columns out of the loop-carried register set: they
        // are re-read from the cache each image instead of pinned in VGPRs
        asm volatile("" ::: "memory");
        float z1r[N], z1i[N];
        // The frames as buffer resources based at image n: one 32-bit lane
        // offset serves all N rows (the row step is a uniform offset), where
        // 64-bit row addresses cost two registers per row and load.
        const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(g.imgA + (size_t)n * plane), (short)0, (int)(plane * sizeof(float)), 0x00020000);
        const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(g.imgB + (size_t)n * plane), (short)0, (int)(plane * sizeof(float)), 0x00020000);
        const int voff = (row0 * g.W + col0 + t) * (int)sizeof(float);
        auto pa = [&](int y) {
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rA, voff, y * g.W * 4, 0));
        };
        auto pb = [&](int y) {
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rB, voff, y * g.W * 4, 0));
        };
        float mB = 0.0f, mAA = 0.0f, mBB = 0.0f;
        float eAA = 0.0f, eBB = 0.0f, dummy = 0.0f;
        if constexpr (SHARED) {
            // wBab is wBB and wAab is wAA (and their sums agree), so the AB
            // operands are the auto operands: two weight columns per image
            // instead of four, the same values bit for bit
            float sAA = 0.0f, sBB = 0.0f;
            if (valid) {
#pragma unroll
                for (int y = 0; y < N; ++y) {
                    z1r[y] = pb(y);
                    z1i[y] = pa(y);
                    if (g.meanSub) {
                        sAA += z1i[y] * g.wAA[y * N + t];
                        sBB += z1r[y] * g.wBB[y * N + t];
                    }
                }
            } else {
#pragma unroll
                for (int y = 0; y < N; ++y) { z1r[y] = 0.0f; z1i[y] = 0.0f; }
            }
            if (g.meanSub) {
                ens_sum3(sBB, sAA, dummy, r0, r1, r2, t, T);
                mAA = (g.sumWAA > 0.0f) ? sAA / g.sumWAA : 0.0f;
                mBB = (g.sumWBB > 0.0f) ? sBB / g.sumWBB : 0.0f;
            }
#pragma unroll
            for (int y = 0; y < N; ++y) {
                z1r[y] = (z1r[y] - mBB) * g.wBB[y * N + t];
                z1i[y] = (z1i[y] - mAA) * g.wAA[y * N + t];
                eAA += z1i[y] * z1i[y];
                eBB += z1r[y] * z1r[y];
            }
        } else {
            float sAB_B = 0.0f, sAA = 0.0f, sBB = 0.0f;
            if (valid) {
#pragma unroll
                for (int y = 0; y < N; ++y) {
                    z1r[y] = pb(y);
                    z1i[y] = pa(y);
                    if (g.meanSub) {
                        sAB_B += z1r[y] * g.wBab[y * N + t];
                        sAA += z1i[y] * g.wAA[y * N + t];
                        sBB += z1r[y] * g.wBB[y * N + t];
                    }
                }
            } else {
#pragma unroll
                for (int y = 0; y < N; ++y) { z1r[y] = 0.0f; z1i[y] = 0.0f; }
            }
            if (g.meanSub) {
                ens_sum3(sAB_B, sAA, sBB, r0, r1, r2, t, T);
                mB = (g.sumWBab > 0.0f) ? sAB_B / g.sumWBab : 0.0f;
                mAA = (g.sumWAA > 0.0f) ? sAA / g.sumWAA : 0.0f;
                mBB = (g.sumWBB > 0.0f) ? sBB / g.sumWBB : 0.0f;
            }
#pragma unroll
            for (int y = 0; y < N; ++y) {
                const float b = z1r[y], a = z1i[y];
                z1r[y] = (b - mB) * g.wBab[y * N + t];
                z1i[y] = (a - mAA) * g.wAab[y * N + t];
                const float pA = (a - mAA) * g.wAA[y * N + t], pB = (b - mBB) * g.wBB[y * N + t];
                eAA += pA * pA;
                eBB += pB * pB;
            }
        }
        float sAB = 1.0f, sA = 1.0f, sB = 1.0f;
        bool use = valid;
        if (g.perPairNorm) {
            ens_sum3(eAA, eBB, dummy, r0, r1, r2, t, T);
            if (eAA <= 0.0f || eBB <= 0.0f) use = false;
            sAB = 1.0f / sqrtf(eAA * eBB);
            sA = 1.0f / eAA;
            sB = 1.0f / eBB;
        }
        fwd2d<N>(z1r, z1i, Sr, Si, t, use);
        if (use) {
#pragma unroll
            for (int kx = 0; kx < HW; ++kx) {
                const int mx = (N - kx) % N;
                const float zR = Sr[ky * PX + kx], zI = Si[ky * PX + kx];
                const float mR = Sr[my * PX + mx], mI = Si[my * PX + mx];
                const float xr_ = 0.5f * (zR + mR), xi_ = 0.5f * (zI - mI);      // F(pAB_B)
                const float yr_ = 0.5f * (zI + mI), yi_ = -0.5f * (zR - mR);     // F(pAB_A)
                aRe[kx] += sAB * (xr_ * yr_ + xi_ * yi_);
                aIm[kx] += sAB * (xi_ * yr_ - xr_ * yi_);
                if constexpr (SHARED) {
                    aAA[kx] += sA * (yr_ * yr_ + yi_ * yi_);
                    aBB[kx] += sB * (xr_ * xr_ + xi_ * xi_);
                }
            }
        }
        if (!SHARED && g.triple) {
            // z2 = pAA + i pBB from a second read of the window
            float z2r[N], z2i[N];
#pragma unroll
            for (int y = 0; y < N; ++y) {
                const float a = valid ? pa(y) : 0.0f, b = valid ? pb(y) : 0.0f;
                z2r[y] = (a - mAA) * g.wAA[y * N + t];
                z2i[y] = (b - mBB) * g.wBB[y * N + t];
            }
            __syncthreads();
            fwd2d<N>(z2r, z2i, Sr, Si, t, use);
            if (use) {
#pragma unroll
                for (int kx = 0; kx < HW; ++kx) {
                    const int mx = (N - kx) % N;
                    const float zR = Sr[ky * PX + kx], zI = Si[ky * PX + kx];
                    const float mR = Sr[my * PX + mx], mI = Si[my * PX + mx];
                    const float ar = 0.5f * (zR + mR), ai = 0.5f * (zI - mI);
                    const float br = 0.5f * (zI + mI), bi = -0.5f * (zR - mR);
                    aAA[kx] += sA * (ar * ar + ai * ai);
                    aBB[kx] += sB * (br * br + bi * bi);
                }
            }
        }
        __syncthreads();
        // the call's last images go straight to the readout below
        if ((n + 1) % kEnsFlush == 0 && n + 1 < g.nImages) flush();
    }
    // This call's totals: the registers, plus the FP64 sums when an earlier
    // flush spilled into them (the same value the flush-and-read-back gave).
    // They are staged row by row in the window's LDS tile, so the column
    // readout never goes through HBM; calls of up to kEnsFlush images touch
    // no FP64 sums at all.
    if (!first && valid) {
        const double *row = acc + (size_t)ky * HW;
        asm volatile("" : "+v"(row));
#pragma unroll
        for (int k = 0; k < HW; ++k) {
            aRe[k] = (float)(row[k] + (double)aRe[k]);
            aIm[k] = (float)(row[N * HW + k] + (double)aIm[k]);
            aAA[k] = (float)(row[2 * N * HW + k] + (double)aAA[k]);
            aBB[k] = (float)(row[3 * N * HW + k] + (double)aBB[k]);
            asm volatile("" ::: "memory");
        }
    }

    // ---- readout: one c2r per accumulated spectrum, shift, crop, += -------
    const int oh = g.fitY, ow = g.fitX;
    const int sy = (N - oh) / 2, sx = (N - ow) / 2;
    for (int which = 0; which < (g.triple ? 3 : 1); ++which) {
        if (valid) {
#pragma unroll
            for (int k = 0; k < HW; ++k) {
                if (which == 0) {
                    Sr[ky * PX + k] = aRe[k];
                    Si[ky * PX + k] = aIm[k];
                } else {
                    Sr[ky * PX + k] = which == 1 ? aAA[k] : aBB[k];
                }
            }
        }
        __syncthreads();
        if (valid && t < HW) {
            // each lane reads and then rewrites only its own column t
            float cr[N], ci[N];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                cr[k] = Sr[k * PX + t];
                ci[k] = which == 0 ? Si[k * PX + t] : 0.0f;
            }
            fft_reg<N, true>(cr, ci);
#pragma unroll
            for (int y = 0; y < N; ++y) {
                Sr[y * PX + t] = cr[y];
                Si[y * PX + t] = ci[y];
            }
        }
        __syncthreads();
        if (valid && t < N / 2) {
            const int ya = t, yb = t + N / 2;
            float xr[N], xi[N];
#pragma unroll
            for (int k = 0; k <= N / 2; ++k) {
                float g1r = Sr[ya * PX + k], g1i = Si[ya * PX + k];
                float g2r = Sr[yb * PX + k], g2i = Si[yb * PX + k];
                if (k == 0 || k == N / 2) { g1i = 0.0f; g2i = 0.0f; }
                xr[k] = g1r - g2i;
                xi[k] = g1i + g2r;
                if (k != 0 && k != N / 2) {
                    xr[N - k] = g1r + g2i;
                    xi[N - k] = -g1i + g2r;
                }
            }
            fft_reg<N, true>(xr, xi);
            const float mul = 1.0f / (float)numel;
            // the shifted rows ya and yb go back into this lane's own rows of
            // the tile, so the plane leaves in row order below
#pragma unroll
            for (int x = 0; x < N; ++x) {
                const int xs = (x + N / 2) % N;
                Sr[yb * PX + xs] = xr[x] * mul;
                Sr[ya * PX + xs] = xi[x] * mul;
            }
        }
        __syncthreads();
        if (valid) {
            // the cropped plane added to the caller's output, consecutive
            // lanes on consecutive addresses (row-per-lane stores left most
            // of every written line untouched)
            float *out = (which == 0 ? g.outAB : (which == 1 ? g.outAA : g.outBB)) + (size_t)w * oh * ow;
            for (int e = t; e < oh * ow; e += T) {
                const int y = e / ow, x = e - y * ow;
                out[e] += Sr[(y + sy) * PX + x + sx];
            }
        }
        __syncthreads();
    }
}

template <int N>
static int launch_ens(EnsArgs g, hipStream_t st)
{
    using C = EnsCfg<N>;
    const int nWin = g.nWy * g.nWx;
    const int nb = (nWin + C::WPB - 1) / C::WPB;
    if (nb <= 0) return 0;
    g.nblocks = nb;
    const int grid = (nb + 7) / 8 * 8;
    if constexpr (N <= 32) {
        // the register form reads each frame as one buffer resource
        if ((size_t)g.H * g.W * sizeof(float) > (size_t)0x7fffffff) return PIVAMD_ERR_ARG;
        if (g.sharedW)
            hipLaunchKernelGGL((ens_kernel_reg<N, true>), dim3(grid), dim3(C::BLOCK), 0, st, g);
        else
            hipLaunchKernelGGL((ens_kernel_reg<N, false>), dim3(grid), dim3(C::BLOCK), 0, st, g);
    } else
        hipLaunchKernelGGL((ens_kernel<N>), dim3(grid), dim3(C::BLOCK), 0, st, g);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

}  // namespace pivamd

using namespace pivamd;

extern "C" size_t pivamd_ensemble_workspace(int nWy, int nWx, int winY, int winX)
{
    return (size_t)nWy * nWx * 4 * winY * (winX / 2 + 1) * sizeof(double);
}

// triple = 1: AB, AA, BB (accumulate_triple); triple = 0: AB only with
// weights (wA, wB) and no mean removal (accumulate).  Sums are added to the
// caller's output planes; sumW* are the weight sums of the mean removal.
extern "C" int pivamd_accumulate(const float *d_A, const float *d_B, const float *d_mask, int H, int W,
                                 int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                                 const float *d_wAab, const float *d_wBab, const float *d_wAA,
                                 const float *d_wBB, int winY, int winX, int fitY, int fitX, int meanSub,
                                 int perPairNorm, int triple, float sumWBab, float sumWAA, float sumWBB,
                                 float *d_outAB, float *d_outAA, float *d_outBB, void *d_work,
                                 size_t work_bytes, hipStream_t st)
{
    if (winY != winX) return PIVAMD_ERR_SIZE;
    if (n_images < 0 || nWy < 0 || nWx < 0 || H <= 0 || W <= 0) return PIVAMD_ERR_ARG;
    if (fitY <= 0 || fitX <= 0 || fitY > winY || fitX > winX) return PIVAMD_ERR_ARG;
    if (work_bytes < pivamd_ensemble_workspace(nWy, nWx, winY, winX) || !d_work) return PIVAMD_ERR_ARG;
    if (!d_outAB || (triple && (!d_outAA || !d_outBB))) return PIVAMD_ERR_ARG;
    EnsArgs g;
    g.imgA = d_A; g.imgB = d_B; g.mask = d_mask; g.ctrX = d_ctrX; g.ctrY = d_ctrY;
    g.wAab = d_wAab; g.wBab = d_wBab; g.wAA = triple ? d_wAA : d_wAab; g.wBB = triple ? d_wBB : d_wBab;
    g.outAB = d_outAB; g.outAA = d_outAA; g.outBB = d_outBB; g.acc = (double *)d_work;
    g.H = H; g.W = W; g.nImages = n_images; g.nWy = nWy; g.nWx = nWx; g.fitY = fitY; g.fitX = fitX;
    g.meanSub = triple ? meanSub : 0; g.perPairNorm = triple ? perPairNorm : 0; g.triple = triple;
    g.sumWBab = sumWBab; g.sumWAA = sumWAA; g.sumWBB = sumWBB; g.nblocks = 0;
    g.sharedW = triple && d_wAab == d_wAA && d_wBab == d_wBB && sumWBab == sumWBB;
    switch (winY) {
    case 8: return launch_ens<8>(g, st);
    case 12: return launch_ens<12>(g, st);
    case 16: return launch_ens<16>(g, st);
    case 24: return launch_ens<24>(g, st);
    case 32: return launch_ens<32>(g, st);
    case 48: return launch_ens<48>(g, st);
    case 64: return launch_ens<64>(g, st);
    case 96: return launch_ens<96>(g, st);
    case 128: return launch_ens<128>(g, st);
    default: return PIVAMD_ERR_SIZE;
    }
}

// ---------------------------------------------------------------------------
// Ensemble finalize (pivtools_cli/piv/piv_backend/single_pass_accumulator.py
// :555-795, 1067-1290), on the device.  The arithmetic is float32 in the
// reference's order (sums / N, minus the background correlation of the mean
// images, optional plane-mean removal, optional envelope divide, division
// by sqrt(max(AA(0) BB(0), 1e-12))); the planes go out in FP64 for the
// k-space fit.
// ---------------------------------------------------------------------------
namespace pivamd {

__global__ void sum_images_kernel(const float *img, int n, size_t hw, double *sum)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < hw; i += (size_t)gridDim.x * blockDim.x) {
        double s = sum[i];
        for (int k = 0; k < n; ++k) s += (double)img[(size_t)k * hw + i];
        sum[i] = s;
    }
}

__global__ void mean_image_kernel(const double *sum, double n, size_t hw, float *out)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < hw; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)(sum[i] / n);
}

struct FinArgs {
    const float *sAA, *sBB, *sAB, *bAA, *bBB, *bAB, *envAuto, *envAB;
    double *oAA, *oBB, *oAB;
    float nImages, abScale;
    int P, centre, dcZero;
};

__device__ __forceinline__ float fin_block_sum(float v, float *red)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float s = 0.0f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    return s;
}

// One workgroup per window.
__global__ void __launch_bounds__(256) ens_finalize_kernel(FinArgs a)
{
    __shared__ float red[4];
    const size_t off = (size_t)blockIdx.x * a.P;
    const int t = threadIdx.x, P = a.P;
    const float *s3[3] = {a.sAA + off, a.sBB + off, a.sAB + off};
    const float *b3[3] = {a.bAA ? a.bAA + off : nullptr, a.bBB ? a.bBB + off : nullptr,
                          a.bAB ? a.bAB + off : nullptr};
    float mean[3] = {0.0f, 0.0f, 0.0f};
    if (a.dcZero) {
        for (int p = 0; p < 3; ++p) {
            float s = 0.0f;
            for (int i = t; i < P; i += blockDim.x) s += s3[p][i] / a.nImages - (b3[p] ? b3[p][i] : 0.0f);
            mean[p] = fin_block_sum(s, red) / (float)P;
        }
    }
    // centre values of AA and BB after the same steps
    float cv[2];
    for (int p = 0; p < 2; ++p) {
        const int c = a.centre;
        float r = s3[p][c] / a.nImages;
        if (b3[p]) r = r - b3[p][c];
        if (a.dcZero) r = r - mean[p];
        if (a.envAuto) r = r / a.envAuto[c];
        cv[p] = r;
    }
    float prod = cv[0] * cv[1];
    prod = prod > 1e-12f ? prod : 1e-12f;
    const float nf = sqrtf(prod);
    double *o3[3] = {a.oAA + off, a.oBB + off, a.oAB + off};
    for (int p = 0; p < 3; ++p) {
        const float *env = p < 2 ? a.envAuto : a.envAB;
        for (int i = t; i < P; i += blockDim.x) {
            float r = s3[p][i] / a.nImages;
            if (b3[p]) r = r - b3[p][i];
            if (a.dcZero) r = r - mean[p];
            if (env) r = r / env[i];
            if (p == 2) r = r * a.abScale;
            o3[p][i] = (double)(r / nf);
        }
    }
}

// Zero-bordered float32 copy of n frames (16-bit counts or float32) into
// (Hp, Wp) planes at (top, left): the single run type's image padding
// (window_utils.py:447-520, constant 0), fused with the count conversion.
template <typename T>
__global__ void pad_frames_kernel(const T *src, size_t n, int H, int W, int top, int left, int Hp, int Wp,
                                  float *dst)
{
    const size_t plane = (size_t)Hp * Wp, total = n * plane;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t f = i / plane;
        const int rem = (int)(i - f * plane);
        const int y = rem / Wp - top, x = rem % Wp - left;
        float v = 0.0f;
        if (y >= 0 && y < H && x >= 0 && x < W) v = (float)src[(f * H + y) * W + x];
        dst[i] = v;
    }
}

// Displacements and stresses from the k-space fit rows
// (single_pass_accumulator.py:1067-1290 without the optional outlier
// methods): status != 0 or the 3/4-window rule gives NaN; masked windows 0.
__global__ void ens_extract_kernel(const double *gauss, const int *status, const unsigned char *mask, int nWin,
                                   float cx1, float cy1, float maxdx, float maxdy, double peakScale, float *ux,
                                   float *uy, float *uu, float *vv, float *uv, float *peak, int *reason)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nWin) return;
    const double *g = gauss + (size_t)w * 16;
    auto safe = [](double v, float lim) {
        float f = (float)(v < -lim ? -lim : (v > lim ? lim : v));
        return isfinite(f) ? f : 0.0f;
    };
    float x = (float)g[14] - cx1, y = (float)g[15] - cy1;
    int st = status[w];
    if (!isfinite(x) || !isfinite(y) || fabsf(x) > maxdx || fabsf(y) > maxdy) {
        x = NAN;
        y = NAN;
        st = 6;
    }
    const float aA = safe(g[0], 1e10f), aB = safe(g[1], 1e10f), aAB = safe(g[2], 1e10f);
    float gm = aA * aB;
    gm = sqrtf(gm > 1e-12f ? gm : 1e-12f);
    float ph = aAB / gm;
    // single run type: sqrt(sum area / A area) applied in FP64 to the float32
    // ratio, as single_pass_accumulator.py:1199-1205 (exact for a scale of 1)
    ph = (float)((double)ph * peakScale);
    ph = ph < 0.0f ? 0.0f : (ph > 1.0f ? 1.0f : ph);
    const float sAx = safe(g[6], 1e6f), sAy = safe(g[7], 1e6f), sAxy = safe(g[8], 1e6f);
    float U = safe(g[9], 1e6f) - sAx, V = safe(g[10], 1e6f) - sAy, C = safe(g[11], 1e6f) - sAxy;
    U = U > 0.0f ? U : 0.0f;
    V = V > 0.0f ? V : 0.0f;
    const bool m = mask && mask[w];
    if (m) {
        x = y = U = V = C = ph = 0.0f;
        st = -1;
    } else if (st != 0) {
        x = y = U = V = C = ph = NAN;
    }
    ux[w] = x;
    uy[w] = y;
    uu[w] = U;
    vv[w] = V;
    uv[w] = C;
    peak[w] = ph;
    reason[w] = st;
}

}  // namespace pivamd

extern "C" int pivamd_sum_images(const float *d_imgs, int n, unsigned long hw, double *d_sum, hipStream_t st)
{
    if (n <= 0 || hw == 0) return PIVAMD_OK;
    const unsigned long blocks = (hw + 255) / 256;
    hipLaunchKernelGGL(sum_images_kernel, dim3(blocks < 4096 ? blocks : 4096), dim3(256), 0, st, d_imgs, n,
                       (size_t)hw, d_sum);
    return hipGetLastError() == hipSuccess ? PIVAMD_OK : PIVAMD_ERR_LAUNCH;
}

extern "C" int pivamd_mean_image(const double *d_sum, double n, unsigned long hw, float *d_out, hipStream_t st)
{
    if (hw == 0) return PIVAMD_OK;
    const unsigned long blocks = (hw + 255) / 256;
    hipLaunchKernelGGL(mean_image_kernel, dim3(blocks < 4096 ? blocks : 4096), dim3(256), 0, st, d_sum, n,
                       (size_t)hw, d_out);
    return hipGetLastError() == hipSuccess ? PIVAMD_OK : PIVAMD_ERR_LAUNCH;
}

extern "C" int pivamd_ensemble_finalize(const float *d_sAA, const float *d_sBB, const float *d_sAB,
                                        const float *d_bAA, const float *d_bBB, const float *d_bAB,
                                        const float *d_envAuto, const float *d_envAB, int n_windows,
                                        int corr_h, int corr_w, float n_images, float ab_scale, int dc_zero,
                                        double *d_oAA, double *d_oBB, double *d_oAB, hipStream_t st)
{
    if (n_windows <= 0) return PIVAMD_OK;
    if (corr_h <= 0 || corr_w <= 0 || !(n_images > 0.0f)) return PIVAMD_ERR_ARG;
    FinArgs a{d_sAA, d_sBB, d_sAB, d_bAA, d_bBB, d_bAB, d_envAuto, d_envAB, d_oAA, d_oBB, d_oAB,
              n_images, ab_scale, corr_h * corr_w, (corr_h / 2) * corr_w + corr_w / 2, dc_zero};
    hipLaunchKernelGGL(ens_finalize_kernel, dim3(n_windows), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? PIVAMD_OK : PIVAMD_ERR_LAUNCH;
}

extern "C" int pivamd_ensemble_extract(const double *d_gauss, const int *d_status, const unsigned char *d_mask,
                                       int n_windows, int corr_h, int corr_w, double peak_scale, float *d_ux,
                                       float *d_uy, float *d_uu, float *d_vv, float *d_uv, float *d_peak,
                                       int *d_reason, hipStream_t st)
{
    if (n_windows <= 0) return PIVAMD_OK;
    if (!(peak_scale > 0.0)) return PIVAMD_ERR_ARG;
    const float cx1 = corr_w / 2.0f + 1.0f, cy1 = corr_h / 2.0f + 1.0f;
    hipLaunchKernelGGL(ens_extract_kernel, dim3((n_windows + 255) / 256), dim3(256), 0, st, d_gauss, d_status,
                       d_mask, n_windows, cx1, cy1, 0.75f * corr_w, 0.75f * corr_h, peak_scale, d_ux, d_uy, d_uu,
                       d_vv, d_uv, d_peak, d_reason);
    return hipGetLastError() == hipSuccess ? PIVAMD_OK : PIVAMD_ERR_LAUNCH;
}

extern "C" int pivamd_pad_frames(const void *d_src, int src_u16, int n, int H, int W, int pad_top, int pad_left,
                                 int Hp, int Wp, float *d_dst, hipStream_t st)
{
    if (n < 0 || H <= 0 || W <= 0 || pad_top < 0 || pad_left < 0 || Hp < H + pad_top || Wp < W + pad_left)
        return PIVAMD_ERR_ARG;
    if ((size_t)Hp * Wp >= ((size_t)1 << 31)) return PIVAMD_ERR_SIZE;   // in-frame index is 32-bit
    if (n == 0) return PIVAMD_OK;
    if (!d_src || !d_dst) return PIVAMD_ERR_ARG;
    const size_t total = (size_t)n * Hp * Wp;
    const size_t blocks = (total + 255) / 256;
    const dim3 grid((unsigned)(blocks < 8192 ? blocks : 8192));
    if (src_u16)
        hipLaunchKernelGGL(pad_frames_kernel<unsigned short>, grid, dim3(256), 0, st,
                           (const unsigned short *)d_src, (size_t)n, H, W, pad_top, pad_left, Hp, Wp, d_dst);
    else
        hipLaunchKernelGGL(pad_frames_kernel<float>, grid, dim3(256), 0, st, (const float *)d_src, (size_t)n, H,
                           W, pad_top, pad_left, Hp, Wp, d_dst);
    return hipGetLastError() == hipSuccess ? PIVAMD_OK : PIVAMD_ERR_LAUNCH;
}
