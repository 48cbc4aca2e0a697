// kspace.hip -- k-space transfer-function LM fit of ensemble correlation
// planes (gfx950).
//
// Replaces kspace_lm_fit_batch (pivtools_cli/lib/kspace_lm_fit.c:669-886).
// Per window, with the three averaged planes R_AA, R_BB, R_AB:
//   F = fftshift(fft2(ifftshift(R))) taken in natural k order
//   F_ref = sqrt(|F_AA| |F_BB|),  T = F_AB / F_ref,  weights F_ref / max F_ref
//   (DC and bins with F_ref <= 1e-12 excluded)
//   model g exp(-2 pi^2 k^T S k [- b4x kx^4 - b4y ky^4]) (1 - N0 D) exp(-2 pi i k.mu)
//   D = P / F_ref (coloured floor, P given centred) or 1 / F_ref (flat)
//   damped Gauss-Newton (lambda 1e-3, /3 on accept, x10 on reject, 250
//   iterations, xtol = ftol = 1e-8), bounds S_xx, S_yy, N0 >= 0, g in [1e-3, 1e3]
// then the status gates and the 16-column output row of the reference.
//
// Mapping to CDNA4: one workgroup per window.  The centred transform is a
// direct FP64 DFT per axis through LDS (twiddles and k grids in LDS).  The
// per-bin observation T, weight and regressor D stay in LDS (or, above 64x64
// planes, in a per-workgroup global slice) for the whole fit.  Each LM
// evaluation sweeps the bins with the block's lanes in FP64, reduces
// J^T J, J^T r and the cost with wave shuffles and one LDS stage, and every
// lane then solves the damped normal equations itself (packed Cholesky in
// registers) so that the parameter vector is uniform across the block
// without a broadcast.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pivamd_dev.h"

namespace pivamd {
namespace {

constexpr int kMaxIter = 250;
constexpr double kXtol = 1e-8;
constexpr double kFtol = 1e-8;
constexpr int kMinValid = 10;
constexpr double kMaxDispFrac = 0.75;
constexpr double kCostAccept = 1.0;
constexpr double kExpClip = 300.0;
constexpr double kLam0 = 1e-3;
constexpr double kLamMin = 1e-12;
constexpr double kLamStuck = 1e12;
constexpr double kGainLo = 1e-3;
constexpr double kGainHi = 1e3;
constexpr double kSeedSxx = 1.0;
constexpr double kSeedSyy = 0.2;
constexpr double kN0Hi = 10.0;
constexpr double kTailKr = 0.35;
constexpr double kN0ClipLo = 1e-3;
constexpr double kN0ClipHi = 10.0;
constexpr double k2Pi = 6.283185307179586476925286766559;
constexpr double k2Pi2 = 2.0 * 9.869604401089358618834490999876;
constexpr int kMaxAxis = 128;

enum { kMasked = -1, kOk = 0, kNoConverge = 1, kLowSnr = 2, kBigDisp = 3 };

struct KsArgs {
    const double *aa, *bb, *ab;
    const unsigned char *mask;
    const double *pwin;
    size_t pwinStride;   // H * W (one shape per window) or 0 (one shared shape)
    int nWin, H, W;
    const int *nproc;
    double *gauss, *init;
    int *status;
    double *gain, *n0, *b4x, *b4y, *cpp;
    int *nvalid, *iter;
    unsigned char *conv;
    double *scratch;
    double *dbg_dv, *dbg_n0;   // prep-only mode: D (natural order) and the N0 seed
};

// Shared tables and reduction buffer (the bin arrays come on top: dynamic
// LDS, or a global slice).
template <int TPB, int NS>
struct KsShared {
    double twr[2][kMaxAxis], twi[2][kMaxAxis];   // [0] = H axis, [1] = W axis
    double kx[kMaxAxis], ky[kMaxAxis];
    double red[TPB / 64][NS + 2];
    double ring[8];
    int hist[16];
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
    return v;
}

// Block-wide max / min / argmax; every lane returns the same value.  The
// leading barrier protects the buffer from the previous reduction's readers.
template <int TPB, int NS>
__device__ __forceinline__ double block_max(double v, double (*red)[NS + 2])
{
    constexpr int NW = TPB / 64;
    v = wave_max(v);
    if constexpr (NW > 1) {
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][0] = v;
        __syncthreads();
        v = red[0][0];
#pragma unroll
        for (int w = 1; w < NW; ++w) v = fmax(v, red[w][0]);
    }
    return v;
}

template <int TPB, int NS>
__device__ __forceinline__ double block_min(double v, double (*red)[NS + 2])
{
    constexpr int NW = TPB / 64;
    v = wave_min(v);
    if constexpr (NW > 1) {
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][0] = v;
        __syncthreads();
        v = red[0][0];
#pragma unroll
        for (int w = 1; w < NW; ++w) v = fmin(v, red[w][0]);
    }
    return v;
}

// First-occurrence argmax (strict '>' scan from index 0): larger value
// wins, equal values go to the smaller index, NaN never wins.
template <int TPB, int NS>
__device__ __forceinline__ int block_argmax(double v, int i, double (*red)[NS + 2])
{
    constexpr int NW = TPB / 64;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const double ov = __shfl_xor(v, m, 64);
        const int oi = __shfl_xor(i, m, 64);
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
    if constexpr (NW > 1) {
        __syncthreads();
        if ((threadIdx.x & 63) == 0) {
            red[threadIdx.x >> 6][0] = v;
            red[threadIdx.x >> 6][1] = (double)i;
        }
        __syncthreads();
        v = red[0][0];
        i = (int)red[0][1];
#pragma unroll
        for (int w = 1; w < NW; ++w) {
            const double ov = red[w][0];
            const int oi = (int)red[w][1];
            if (ov > v || (ov == v && oi < i)) {
                v = ov;
                i = oi;
            }
        }
    }
    return i;
}

__device__ __forceinline__ double log3pt(double lo, double ce, double hi)
{
    if (lo <= 0.0 || ce <= 0.0 || hi <= 0.0) return 0.0;
    const double a = log(lo), b = log(ce), c = log(hi);
    const double den = 2.0 * (a - 2.0 * b + c);
    if (fabs(den) <= 1e-12) return 0.0;
    return (a - c) / den;
}

__device__ __forceinline__ double clampd(double v, double lo, double hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

// Parameter bounds of kspace_lm_fit.c:689-696: S_xx, S_yy, N0 >= 0,
// g in [1e-3, 1e3], N0 <= 10 on the coloured floor; the others are free
// (no clamp is applied to them at all).
template <int I, int K>
__device__ __forceinline__ void apply_bounds(double (&x)[K], bool coloured)
{
    if constexpr (I < K) {
        double v = x[I];
        if constexpr (I == 2 || I == 3) {
            if (v < 0.0) v = 0.0;
        } else if constexpr (I == 5) {
            if (v < kGainLo) v = kGainLo;
            if (v > kGainHi) v = kGainHi;
        } else if constexpr (I == 6) {
            if (v < 0.0) v = 0.0;
            if (coloured && v > kN0Hi) v = kN0Hi;
        }
        x[I] = v;
        apply_bounds<I + 1, K>(x, coloured);
    }
}

// Bin arrays of one window (LDS, or a per-workgroup global slice):
//   fr  F_ref, later the floor regressor D
//   tr  centred real plane during a transform, later Re T
//   ti  Im T
//   yr  row-transform real part, later the fit weights
//   yi  row-transform imaginary part, later the floor-seed candidates
struct Bins {
    double *fr, *tr, *ti, *yr, *yi;
};

// Row then column DFT of one centred plane (input loaded with the ifftshift
// index map).  emit(b, re, im) receives bin b of the natural-order spectrum;
// it runs after the row pass, so it may write fr, tr or ti but not yr/yi.
template <int TPB, typename SH, typename EMIT>
__device__ __forceinline__ void dft2_centred(const double *plane, int H, int W, const Bins &s, SH &sh,
                                             EMIT emit)
{
    const int P = H * W, t = threadIdx.x;
    __syncthreads();
    for (int b = t; b < P; b += TPB) {
        const int r = b / W, c = b - r * W;
        s.tr[b] = plane[((r + H / 2) % H) * W + (c + W / 2) % W];
    }
    __syncthreads();
    for (int b = t; b < P; b += TPB) {
        const int r = b / W, kc = b - r * W;
        const double *row = s.tr + r * W;
        double sr = 0.0, si = 0.0;
        int m = 0;
        for (int c = 0; c < W; ++c) {
            const double x = row[c];
            sr += x * sh.twr[1][m];
            si += x * sh.twi[1][m];
            m += kc;
            if (m >= W) m -= W;
        }
        s.yr[b] = sr;
        s.yi[b] = si;
    }
    __syncthreads();
    for (int b = t; b < P; b += TPB) {
        const int kr = b / W, kc = b - kr * W;
        double sr = 0.0, si = 0.0;
        int m = 0;
        for (int r = 0; r < H; ++r) {
            const double ar = s.yr[r * W + kc], ai = s.yi[r * W + kc];
            const double cr = sh.twr[0][m], ci = sh.twi[0][m];
            sr += ar * cr - ai * ci;
            si += ar * ci + ai * cr;
            m += kr;
            if (m >= H) m -= H;
        }
        emit(b, sr, si);
    }
}

// k-th smallest (0-based) of the values v[b] >= 0 (negative entries are
// not candidates), by a 4-bit radix descent over the IEEE bit patterns.
template <int TPB, typename SH>
__device__ __forceinline__ double block_select(const double *v, int P, int k, SH &sh)
{
    uint64_t prefix = 0;
    for (int shift = 60; shift >= 0; shift -= 4) {
        const uint64_t hi_mask = shift == 60 ? 0ull : (~0ull << (shift + 4));
        __syncthreads();
        if (threadIdx.x < 16) sh.hist[threadIdx.x] = 0;
        __syncthreads();
        for (int b = threadIdx.x; b < P; b += TPB) {
            const double x = v[b];
            if (!(x >= 0.0)) continue;
            const uint64_t u = (uint64_t)__double_as_longlong(x);
            if ((u & hi_mask) == prefix) atomicAdd(&sh.hist[(u >> shift) & 15u], 1);
        }
        __syncthreads();
        int d = 0;
        for (; d < 15; ++d) {
            const int c = sh.hist[d];
            if (k < c) break;
            k -= c;
        }
        prefix |= (uint64_t)d << shift;
    }
    return __longlong_as_double((long long)prefix);
}

// Residual sweep over this lane's bins.  Returns the lane's share of the
// cost; with JTJ also accumulates the packed upper triangle of J^T J and
// then J^T r into acc.
template <int TPB, int K, bool KX4, bool KY4, bool JTJ>
__device__ __forceinline__ double sweep(const double (&x)[K], const Bins &s, const double *kx,
                                        const double *ky, int H, int W, double *acc)
{
    constexpr int NT = K * (K + 1) / 2;
    if constexpr (JTJ) {
#pragma unroll
        for (int i = 0; i < NT + K; ++i) acc[i] = 0.0;
    }
    const double mux = x[0], muy = x[1], sxx = x[2], syy = x[3], sxy = x[4], g = x[5], n0 = x[6];
    double b4x = 0.0, b4y = 0.0;
    if constexpr (KX4) b4x = x[7];
    if constexpr (KY4) b4y = x[7 + (KX4 ? 1 : 0)];
    const int P = H * W;
    double cost = 0.0;
#pragma unroll 1
    for (int b = threadIdx.x; b < P; b += TPB) {
        const double w = s.yr[b];
        if (w == 0.0) continue;
        const int r = b / W, c = b - r * W;
        const double kyv = ky[r], ky2 = kyv * kyv;
        const double kxv = kx[c], kx2 = kxv * kxv, kxy = kxv * kyv;
        double e = -k2Pi2 * (sxx * kx2 + 2.0 * sxy * kxy + syy * ky2);
        if constexpr (KX4) e -= b4x * kx2 * kx2;
        if constexpr (KY4) e -= b4y * ky2 * ky2;
        if (e > kExpClip) e = kExpClip;
        const double dec = exp(e);
        const double d = s.fr[b];
        const double att = 1.0 - n0 * d;
        const double ph = -k2Pi * (kxv * mux + kyv * muy);
        double sp, cp;
        sincos(ph, &sp, &cp);
        const double gd = g * dec * att;
        const double rr = w * (s.tr[b] - gd * cp);
        const double ri = w * (s.ti[b] - gd * sp);
        cost += rr * rr + ri * ri;
        if constexpr (JTJ) {
            double jr[K], ji[K];
            const double px = -k2Pi * kxv, py = -k2Pi * kyv;
            jr[0] = -w * gd * (-sp) * px;
            ji[0] = -w * gd * cp * px;
            jr[1] = -w * gd * (-sp) * py;
            ji[1] = -w * gd * cp * py;
            const double d2 = gd * (-k2Pi2 * kx2), d3 = gd * (-k2Pi2 * ky2), d4 = gd * (-k2Pi2 * 2.0 * kxy);
            jr[2] = -w * d2 * cp;
            ji[2] = -w * d2 * sp;
            jr[3] = -w * d3 * cp;
            ji[3] = -w * d3 * sp;
            jr[4] = -w * d4 * cp;
            ji[4] = -w * d4 * sp;
            jr[5] = -w * dec * att * cp;
            ji[5] = -w * dec * att * sp;
            const double dn = g * dec * (-d);
            jr[6] = -w * dn * cp;
            ji[6] = -w * dn * sp;
            if constexpr (KX4) {
                const double db = gd * (-(kx2 * kx2));
                jr[7] = -w * db * cp;
                ji[7] = -w * db * sp;
            }
            if constexpr (KY4) {
                constexpr int q = 7 + (KX4 ? 1 : 0);
                const double db = gd * (-(ky2 * ky2));
                jr[q] = -w * db * cp;
                ji[q] = -w * db * sp;
            }
            int p = 0;
#pragma unroll
            for (int a = 0; a < K; ++a) {
#pragma unroll
                for (int bb = a; bb < K; ++bb) acc[p++] += jr[a] * jr[bb] + ji[a] * ji[bb];
            }
#pragma unroll
            for (int a = 0; a < K; ++a) acc[NT + a] += jr[a] * rr + ji[a] * ri;
        }
    }
    return cost;
}

// Block sum of M values into the LDS array dst (every lane may read dst
// after the trailing barrier).
template <int TPB, int M, typename SH>
__device__ __forceinline__ void block_sum_to(double *v, SH &sh, double *dst)
{
    constexpr int NW = TPB / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < M; ++i) v[i] = wave_sum(v[i]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < M; ++i) sh.red[wv][i] = v[i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < M; i += TPB) {
        double s = sh.red[0][i];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += sh.red[w][i];
        dst[i] = s;
    }
    __syncthreads();
}

// Solve (A + lam diag(max(A_ii, 1e-30))) d = -g by Cholesky; A is the
// packed upper triangle followed by g, read from LDS.  False when the
// damped matrix is not positive definite.
template <int K>
__device__ __forceinline__ bool damped_solve(const double *A, double lam, double (&d)[K])
{
    constexpr int NT = K * (K + 1) / 2;
    auto up = [](int i, int j) { return i * K - i * (i - 1) / 2 + (j - i); };
    double L[NT];   // lower factor: L(i, k), k <= i, stored at L[up(k, i)]
#pragma unroll
    for (int i = 0; i < K; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double t = A[up(j, i)];
            if (i == j) {
                const double dg = t < 1e-30 ? 1e-30 : t;
                t += lam * dg;
            }
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[up(k, i)] * L[up(k, j)];
            if (i == j) {
                if (!(t > 0.0) || !isfinite(t)) return false;
                L[up(i, i)] = sqrt(t);
            } else {
                L[up(j, i)] = t / L[up(j, j)];
            }
        }
    }
    double y[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        double t = -A[NT + i];
#pragma unroll
        for (int k = 0; k < i; ++k) t -= L[up(k, i)] * y[k];
        y[i] = t / L[up(i, i)];
    }
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {
        double t = y[i];
#pragma unroll
        for (int k = i + 1; k < K; ++k) t -= L[up(i, k)] * d[k];
        d[i] = t / L[up(i, i)];
    }
    return true;
}

template <int TPB, int NS>
struct KsSharedFit : KsShared<TPB, NS> {
    double A[NS];
    double scal[4];
};

// Two waves per SIMD: unbounded, the per-lane FP64 normal equations and
// Cholesky factor took about 310 registers and one workgroup per CU; at 256
// the compiler spills about 20 values (scratch, outside the bin sweeps).
template <int TPB, bool GLOBAL, bool KX4, bool KY4>
__global__ __launch_bounds__(TPB, 2) void kspace_kernel(KsArgs a)
{
    constexpr int K = 7 + (KX4 ? 1 : 0) + (KY4 ? 1 : 0);
    constexpr int NT = K * (K + 1) / 2;
    constexpr int NS = NT + K;
    __shared__ KsSharedFit<TPB, NS> sh;
    extern __shared__ double dyn[];

    const int H = a.H, W = a.W, P = H * W, t = threadIdx.x;
    const int cy = H / 2, cx = W / 2;
    const double cxp = W / 2.0 + 1.0, cyp = H / 2.0 + 1.0;
    const bool coloured = a.pwin != nullptr;
    const bool all_masked = *a.nproc == 0;

    double *base = GLOBAL ? a.scratch + (size_t)blockIdx.x * 5 * P : dyn;
    const Bins s{base, base + P, base + 2 * P, base + 3 * P, base + 4 * P};

    for (int i = t; i < H; i += TPB) {
        sh.twr[0][i] = cos(k2Pi * i / H);
        sh.twi[0][i] = -sin(k2Pi * i / H);
        sh.ky[i] = (i < (H + 1) / 2 ? (double)i : (double)(i - H)) / (double)H;
    }
    for (int i = t; i < W; i += TPB) {
        sh.twr[1][i] = cos(k2Pi * i / W);
        sh.twi[1][i] = -sin(k2Pi * i / W);
        sh.kx[i] = (i < (W + 1) / 2 ? (double)i : (double)(i - W)) / (double)W;
    }
    __syncthreads();

    for (int w = blockIdx.x; w < a.nWin; w += gridDim.x) {
        double *gf = a.gauss + (size_t)w * 16, *ig = a.init + (size_t)w * 16;
        const bool masked = a.mask[w] != 0;
        if (t < 16) {
            double v = 0.0;
            if (t >= 6 && t <= 8) v = NAN;
            if (t == 12 || t == 14) v = cxp;
            if (t == 13 || t == 15) v = cyp;
            gf[t] = v;
            ig[t] = (all_masked || !masked) ? v : 0.0;
        }
        if (t == 0) {
            a.status[w] = kMasked;
            a.gain[w] = a.n0[w] = a.b4x[w] = a.b4y[w] = a.cpp[w] = NAN;
            a.nvalid[w] = 0;
            a.iter[w] = 0;
            a.conv[w] = 0;
        }
        if (masked) continue;

        const size_t off = (size_t)w * P;
        const double *paa = a.aa + off, *pbb = a.bb + off, *pab = a.ab + off;

        // ---- amplitudes and the mu seed from the raw R_AB plane ----
        double bv = -INFINITY;
        int bi = 0x7fffffff;
        for (int b = t; b < P; b += TPB) {
            const double v = pab[b];
            if (v > bv) {
                bv = v;
                bi = b;
            }
        }
        int pk = block_argmax<TPB, NS>(bv, bi, sh.red);
        if (isnan(pab[0]) || pk == 0x7fffffff) pk = 0;
        const double ampA = paa[cy * W + cx], ampB = pbb[cy * W + cx], ampAB = pab[pk];
        const int py = pk / W, px = pk - py * W;
        double mux0 = 0.0, muy0 = 0.0;
        if (px > 0 && px < W - 1) mux0 = log3pt(pab[pk - 1], pab[pk], pab[pk + 1]);
        if (py > 0 && py < H - 1) muy0 = log3pt(pab[pk - W], pab[pk], pab[pk + W]);
        mux0 = (double)px + mux0 - (double)cx;
        muy0 = (double)py + muy0 - (double)cy;

        // ---- centred transforms, F_ref and T ----
        dft2_centred<TPB>(paa, H, W, s, sh, [&](int b, double re, double im) { s.fr[b] = hypot(re, im); });
        dft2_centred<TPB>(pbb, H, W, s, sh,
                          [&](int b, double re, double im) { s.fr[b] = sqrt(s.fr[b] * hypot(re, im)); });
        double fmax = 0.0;
        dft2_centred<TPB>(pab, H, W, s, sh, [&](int b, double re, double im) {
            const double f = s.fr[b];
            const double inv = 1.0 / (f > 1e-30 ? f : 1e-30);
            s.tr[b] = re * inv;
            s.ti[b] = im * inv;
            if (f > fmax) fmax = f;
        });
        __syncthreads();   // every lane done reading the transform scratch
        fmax = block_max<TPB, NS>(fmax, sh.red);
        if (fmax < 1e-30) fmax = 1e-30;
        double nv = 0.0;
        for (int b = t; b < P; b += TPB) {
            const double f = s.fr[b];
            const bool valid = f > 1e-12 && b != 0;
            s.yr[b] = valid ? f / fmax : 0.0;
            nv += valid ? 1.0 : 0.0;
        }
        block_sum_to<TPB, 1>(&nv, sh, sh.scal);
        const int n_valid = (int)sh.scal[0];
        if (t == 0) {
            a.nvalid[w] = n_valid;
            gf[0] = ig[0] = ampA;
            gf[1] = ig[1] = ampB;
            gf[2] = ig[2] = ampAB;
            ig[9] = kSeedSxx;
            ig[10] = kSeedSyy;
            ig[14] = cxp + mux0;
            ig[15] = cyp + muy0;
        }
        const bool prep_only = a.dbg_dv != nullptr;
        if (!prep_only && !(ampA >= 1e-12 && ampB >= 1e-12)) {
            if (t == 0) a.status[w] = kLowSnr;
            continue;
        }
        if (!prep_only && n_valid < kMinValid) {
            if (t == 0) a.status[w] = kNoConverge;
            continue;
        }

        // ---- gain seed: median |T| on the ring around DC ----
        if (t < 8) {
            const int dy = (t < 3) ? -1 : (t < 5 ? 0 : 1);
            const int dx = (t < 3) ? t - 1 : (t < 5 ? (t == 3 ? -1 : 1) : t - 6);
            const int b = ((dy + H) % H) * W + (dx + W) % W;
            sh.ring[t] = hypot(s.tr[b], s.ti[b]);
        }
        __syncthreads();
        double ring[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) ring[i] = sh.ring[i];
#pragma unroll
        for (int i = 1; i < 8; ++i)
#pragma unroll
            for (int j = i; j > 0; --j)
                if (ring[j] < ring[j - 1]) {
                    const double tmp = ring[j];
                    ring[j] = ring[j - 1];
                    ring[j - 1] = tmp;
                }
        const double gain0 = clampd(0.5 * (ring[3] + ring[4]), 1e-2, 10.0);

        // ---- floor regressor D (over F_ref), coloured N0 seed ----
        double ntail = 0.0;
        for (int b = t; b < P; b += TPB) {
            const int r = b / W, c = b - r * W;
            const double f = s.fr[b];
            const double fc = f > 1e-30 ? f : 1e-30;
            if (coloured) {
                const double p = a.pwin[(size_t)w * a.pwinStride + (size_t)((r + H / 2) % H) * W + (c + W / 2) % W];
                s.fr[b] = p / fc;
                const double kr2 = sh.ky[r] * sh.ky[r] + sh.kx[c] * sh.kx[c];
                const bool tail = kr2 >= kTailKr * kTailKr;
                s.yi[b] = tail ? f / (p > 1e-30 ? p : 1e-30) : -1.0;
                ntail += tail ? 1.0 : 0.0;
            } else {
                s.fr[b] = 1.0 / fc;
            }
        }
        double n0seed = 0.0;
        if (coloured) {
            block_sum_to<TPB, 1>(&ntail, sh, sh.scal);
            const int n = (int)sh.scal[0];
            const int k1 = (n & 1) ? n / 2 : n / 2 - 1;
            const double m1 = block_select<TPB>(s.yi, P, k1, sh);
            double med = m1;
            if (!(n & 1)) {
                double le = 0.0, above = INFINITY;
                for (int b = t; b < P; b += TPB) {
                    const double v = s.yi[b];
                    if (!(v >= 0.0)) continue;
                    if (v <= m1) le += 1.0;
                    else above = fmin(above, v);
                }
                block_sum_to<TPB, 1>(&le, sh, sh.scal + 1);
                const double m2 = sh.scal[1] > (double)(n / 2) ? m1 : block_min<TPB, NS>(above, sh.red);
                med = 0.5 * (m1 + m2);
            }
            n0seed = clampd(med, kN0ClipLo, kN0ClipHi);
        }
        __syncthreads();   // D complete before the sweeps
        if (prep_only) {
            for (int b = t; b < P; b += TPB) a.dbg_dv[b] = s.fr[b];
            if (t == 0) *a.dbg_n0 = n0seed;
            continue;
        }

        // ---- Levenberg-Marquardt ----
        double x[K];
#pragma unroll
        for (int i = 0; i < K; ++i) x[i] = 0.0;
        x[0] = mux0;
        x[1] = muy0;
        x[2] = kSeedSxx;
        x[3] = kSeedSyy;
        x[5] = gain0;
        x[6] = n0seed;
        apply_bounds<0, K>(x, coloured);
        double cost = sweep<TPB, K, KX4, KY4, false>(x, s, sh.kx, sh.ky, H, W, nullptr);
        block_sum_to<TPB, 1>(&cost, sh, sh.scal);
        cost = sh.scal[0];
        int it = 0;
        bool conv = false;
        if (isfinite(cost)) {
            double lam = kLam0;
            bool fresh = false;
            while (it < kMaxIter) {
                if (!fresh) {
                    double acc[NS];
                    sweep<TPB, K, KX4, KY4, true>(x, s, sh.kx, sh.ky, H, W, acc);
                    block_sum_to<TPB, NS>(acc, sh, sh.A);
                    fresh = true;
                }
                double d[K], xn[K];
                bool take = false;
                double cn = INFINITY;
                if (damped_solve<K>(sh.A, lam, d)) {
                    bool fin = true;
#pragma unroll
                    for (int i = 0; i < K; ++i) xn[i] = x[i] + d[i];
                    apply_bounds<0, K>(xn, coloured);
#pragma unroll
                    for (int i = 0; i < K; ++i) fin = fin && isfinite(xn[i]);
                    if (fin) {
                        double c1 = sweep<TPB, K, KX4, KY4, false>(xn, s, sh.kx, sh.ky, H, W, nullptr);
                        block_sum_to<TPB, 1>(&c1, sh, sh.scal);
                        cn = sh.scal[0];
                        take = isfinite(cn) && cn <= cost;
                    }
                }
                ++it;
                if (take) {
                    bool small = true;
#pragma unroll
                    for (int i = 0; i < K; ++i)
                        if (fabs(xn[i] - x[i]) > kXtol * (kXtol + fabs(xn[i]))) small = false;
                    const bool flat = (cost - cn) <= kFtol * (cost > 1e-300 ? cost : 1e-300);
#pragma unroll
                    for (int i = 0; i < K; ++i) x[i] = xn[i];
                    cost = cn;
                    lam = lam / 3.0 < kLamMin ? kLamMin : lam / 3.0;
                    fresh = false;
                    if (small || flat) {
                        conv = true;
                        break;
                    }
                } else {
                    lam *= 10.0;
                    if (lam > kLamStuck) break;
                }
            }
        }
        if (t == 0) {
            a.iter[w] = it;
            a.conv[w] = conv ? 1 : 0;
            const double cpp = cost / (double)n_valid;
            a.cpp[w] = cpp;
            if (!(conv || (it >= kMaxIter && cpp < kCostAccept))) {
                a.status[w] = kNoConverge;
            } else if (fabs(x[0]) > kMaxDispFrac * W || fabs(x[1]) > kMaxDispFrac * H) {
                a.status[w] = kBigDisp;
            } else {
                a.status[w] = kOk;
                gf[9] = x[2] > 0.0 ? x[2] : 0.0;
                gf[10] = x[3] > 0.0 ? x[3] : 0.0;
                gf[11] = x[4];
                gf[14] = cxp + x[0];
                gf[15] = cyp + x[1];
                a.gain[w] = x[5];
                a.n0[w] = x[6];
                if constexpr (KX4) a.b4x[w] = x[7];
                if constexpr (KY4) a.b4y[w] = x[7 + (KX4 ? 1 : 0)];
            }
        }
    }
}

// One centred transform, natural-order spectrum out (gate entry point).
template <int TPB>
__global__ __launch_bounds__(TPB) void kspace_fft_kernel(const double *plane, int H, int W, double *ore,
                                                         double *oim, double *scratch)
{
    __shared__ KsShared<TPB, 1> sh;
    const int P = H * W, t = threadIdx.x;
    for (int i = t; i < H; i += TPB) {
        sh.twr[0][i] = cos(k2Pi * i / H);
        sh.twi[0][i] = -sin(k2Pi * i / H);
    }
    for (int i = t; i < W; i += TPB) {
        sh.twr[1][i] = cos(k2Pi * i / W);
        sh.twi[1][i] = -sin(k2Pi * i / W);
    }
    const Bins s{scratch, scratch + P, scratch + 2 * P, scratch + 3 * P, scratch + 4 * P};
    dft2_centred<TPB>(plane, H, W, s, sh, [&](int b, double re, double im) {
        ore[b] = re;
        oim[b] = im;
    });
}

__global__ void count_unmasked_kernel(const unsigned char *mask, int n, int *out)
{
    int c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c += mask[i] == 0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

__global__ void zero_int_kernel(int *p) { *p = 0; }

bool axis_ok(int n)
{
    switch (n) {
    case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96: case 128: return true;
    default: return false;
    }
}

constexpr int kGlobalBlocks = 1024;
constexpr int kLdsBins = 1024;   // planes up to 1024 bins keep their bin arrays in LDS

template <int TPB, bool GLOBAL, bool KX4, bool KY4>
hipError_t launch_ks(const KsArgs &a, hipStream_t st)
{
    const int P = a.H * a.W;
    const size_t dyn = GLOBAL ? 0 : (size_t)5 * P * sizeof(double);
    const int grid = GLOBAL ? (a.nWin < kGlobalBlocks ? a.nWin : kGlobalBlocks) : a.nWin;
    if (grid <= 0) return hipSuccess;
    auto k = kspace_kernel<TPB, GLOBAL, KX4, KY4>;
    if (!GLOBAL && dyn > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(TPB), dyn, st, a);
    return hipGetLastError();
}

template <int TPB, bool GLOBAL>
hipError_t launch_shape(const KsArgs &a, int kx4, int ky4, hipStream_t st)
{
    if (kx4 && ky4) return launch_ks<TPB, GLOBAL, true, true>(a, st);
    if (kx4) return launch_ks<TPB, GLOBAL, true, false>(a, st);
    if (ky4) return launch_ks<TPB, GLOBAL, false, true>(a, st);
    return launch_ks<TPB, GLOBAL, false, false>(a, st);
}

}  // namespace
}  // namespace pivamd

using namespace pivamd;

extern "C" size_t pivamd_kspace_workspace(int n_windows, int corr_h, int corr_w)
{
    const size_t P = (size_t)corr_h * corr_w;
    size_t bytes = 256;   // unmasked counter
    if (P > (size_t)kLdsBins) {
        const size_t g = n_windows < kGlobalBlocks ? (size_t)n_windows : (size_t)kGlobalBlocks;
        bytes += g * 5 * P * sizeof(double);
    }
    return bytes;
}

extern "C" int pivamd_kspace_fft2_centered(const double *d_plane, int corr_h, int corr_w, double *d_re,
                                           double *d_im, void *d_work, size_t work_bytes, hipStream_t st)
{
    if (!axis_ok(corr_h) || !axis_ok(corr_w)) return PIVAMD_ERR_SIZE;
    if (!d_work || work_bytes < (size_t)5 * corr_h * corr_w * sizeof(double)) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL((kspace_fft_kernel<256>), dim3(1), dim3(256), 0, st, d_plane, corr_h, corr_w, d_re,
                       d_im, (double *)d_work);
    return hipGetLastError() == hipSuccess ? PIVAMD_OK : PIVAMD_ERR_LAUNCH;
}

static int kspace_run(const double *d_aa, const double *d_bb, const double *d_ab,
                      const unsigned char *d_mask, const double *d_pwin, int n_windows, int corr_h,
                      int corr_w, int use_kx4, int use_ky4, double *d_gauss, int *d_status, double *d_init,
                      double *d_gain, double *d_n0, double *d_b4x, double *d_b4y, double *d_cpp,
                      int *d_nvalid, int *d_iter, unsigned char *d_conv, double *dbg_dv, double *dbg_n0,
                      int pwin_shared, void *d_work, size_t work_bytes, hipStream_t st);

extern "C" int pivamd_kspace_fit(const double *d_aa, const double *d_bb, const double *d_ab,
                                 const unsigned char *d_mask, const double *d_pwin, int n_windows,
                                 int corr_h, int corr_w, int use_kx4, int use_ky4, double *d_gauss,
                                 int *d_status, double *d_init, double *d_gain, double *d_n0,
                                 double *d_b4x, double *d_b4y, double *d_cpp, int *d_nvalid,
                                 int *d_iter, unsigned char *d_conv, int pwin_shared, void *d_work,
                                 size_t work_bytes, hipStream_t st)
{
    return kspace_run(d_aa, d_bb, d_ab, d_mask, d_pwin, n_windows, corr_h, corr_w, use_kx4, use_ky4, d_gauss,
                      d_status, d_init, d_gain, d_n0, d_b4x, d_b4y, d_cpp, d_nvalid, d_iter, d_conv, nullptr,
                      nullptr, pwin_shared, d_work, work_bytes, st);
}

// Prep-only pass over one window: D in natural order and the coloured N0
// seed (0 for the flat floor); the per-window outputs go to d_out16 scratch.
extern "C" int pivamd_kspace_debug_prep(const double *d_aa, const double *d_bb, const double *d_ab,
                                        const double *d_pcentred, int corr_h, int corr_w, double *d_dv,
                                        double *d_n0seed, void *d_work, size_t work_bytes, hipStream_t st)
{
    // per-window output rows live after the workspace proper
    const size_t need = pivamd_kspace_workspace(1, corr_h, corr_w);
    if (!d_work || work_bytes < need + 512) return PIVAMD_ERR_ARG;
    char *o = (char *)d_work + need;
    double *g = (double *)o, *ig = g + 16, *dg = ig + 16;   // 32 doubles of rows, 5 of diagnostics
    int *ist = (int *)(dg + 5);
    unsigned char *msk = (unsigned char *)(ist + 4);
    if (hipMemsetAsync(msk, 0, 1, st) != hipSuccess) return PIVAMD_ERR_LAUNCH;
    return kspace_run(d_aa, d_bb, d_ab, msk, d_pcentred, 1, corr_h, corr_w, 0, 0, g, ist, ig, dg, dg + 1,
                      dg + 2, dg + 3, dg + 4, ist + 1, ist + 2, msk + 1, d_dv, d_n0seed, 0, d_work, need, st);
}

static int kspace_run(const double *d_aa, const double *d_bb, const double *d_ab,
                      const unsigned char *d_mask, const double *d_pwin, int n_windows, int corr_h,
                      int corr_w, int use_kx4, int use_ky4, double *d_gauss, int *d_status, double *d_init,
                      double *d_gain, double *d_n0, double *d_b4x, double *d_b4y, double *d_cpp,
                      int *d_nvalid, int *d_iter, unsigned char *d_conv, double *dbg_dv, double *dbg_n0,
                      int pwin_shared, void *d_work, size_t work_bytes, hipStream_t st)
{
    if (!axis_ok(corr_h) || !axis_ok(corr_w)) return PIVAMD_ERR_SIZE;
    if (n_windows < 0 || (n_windows > 0 && (!d_aa || !d_bb || !d_ab || !d_mask))) return PIVAMD_ERR_ARG;
    if (n_windows == 0) return PIVAMD_OK;
    if (!d_work || work_bytes < pivamd_kspace_workspace(n_windows, corr_h, corr_w)) return PIVAMD_ERR_ARG;
    KsArgs a{};
    a.aa = d_aa;
    a.bb = d_bb;
    a.ab = d_ab;
    a.mask = d_mask;
    a.pwin = d_pwin;
    a.pwinStride = pwin_shared ? 0 : (size_t)corr_h * corr_w;
    a.nWin = n_windows;
    a.H = corr_h;
    a.W = corr_w;
    int *cnt = (int *)d_work;
    a.nproc = cnt;
    a.gauss = d_gauss;
    a.init = d_init;
    a.status = d_status;
    a.gain = d_gain;
    a.n0 = d_n0;
    a.b4x = d_b4x;
    a.b4y = d_b4y;
    a.cpp = d_cpp;
    a.nvalid = d_nvalid;
    a.iter = d_iter;
    a.conv = d_conv;
    a.scratch = (double *)((char *)d_work + 256);
    a.dbg_dv = dbg_dv;
    a.dbg_n0 = dbg_n0;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, st, cnt);
    const int cb = (n_windows + 255) / 256;
    hipLaunchKernelGGL(count_unmasked_kernel, dim3(cb < 1024 ? cb : 1024), dim3(256), 0, st, d_mask, n_windows, cnt);
    const int P = corr_h * corr_w;
    hipError_t e;
    if (P <= 256) e = launch_shape<64, false>(a, use_kx4, use_ky4, st);
    else if (P <= kLdsBins) e = launch_shape<256, false>(a, use_kx4, use_ky4, st);
    else e = launch_shape<256, true>(a, use_kx4, use_ky4, st);
    return e == hipSuccess ? PIVAMD_OK : PIVAMD_ERR_LAUNCH;
}
