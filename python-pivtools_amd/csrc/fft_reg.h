// fft_reg.h -- register-resident complex FFTs for CDNA4 (gfx950).
//
// One lane owns one whole 1D transform of length N (8..128, 2^k or 3*2^k)
// held in registers: xr[N], xi[N].  The transform is a self-sorting
// Stockham decimation-in-frequency with radix 4, then 2, then 3 stages,
// fully unrolled at compile time.  Every index and every twiddle is a
// compile-time constant, so the twiddles become instruction literals
// (no table traffic at all) and quarter-turn twiddles cost no multiply.
//
// Conventions match FFTW's unnormalised transforms: forward uses
// exp(-2*pi*i*k*x/N), inverse exp(+2*pi*i*k*x/N), no scaling.
#pragma once

#include <hip/hip_runtime.h>
#include <utility>

namespace pivamd {

// ---------------------------------------------------------------------------
// compile-time trigonometry (double precision Taylor series on [0, pi/2))
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin_poly(double x)
{
    double term = x, sum = x;
    for (int k = 1; k < 24; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}

constexpr double ct_cos_poly(double x)
{
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 24; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}

// cos(2*pi*t/n) and sin(2*pi*t/n), with the quadrant split done in integers
constexpr double ct_cos2pi(long t, long n)
{
    t %= n;
    if (t < 0) t += n;
    const long q = (4 * t) / n, r = 4 * t - q * n;
    const double x = (kPi / 2.0) * (double)r / (double)n;
    const double c = ct_cos_poly(x), s = ct_sin_poly(x);
    return q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}

constexpr double ct_sin2pi(long t, long n)
{
    t %= n;
    if (t < 0) t += n;
    const long q = (4 * t) / n, r = 4 * t - q * n;
    const double x = (kPi / 2.0) * (double)r / (double)n;
    const double c = ct_cos_poly(x), s = ct_sin_poly(x);
    return q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
}

template <int I, int END, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < END) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, END>(f);
    }
}

constexpr int stockham_radix(int n) { return (n % 4 == 0) ? 4 : ((n % 2 == 0) ? 2 : 3); }

constexpr int stockham_stages(int n)
{
    int s = 0;
    while (n > 1) {
        n /= stockham_radix(n);
        ++s;
    }
    return s;
}

// ---------------------------------------------------------------------------
// one Stockham stage: length-n sub-blocks, stride s, x -> y, then recurse
// ---------------------------------------------------------------------------
template <int N, bool INV, int n, int s>
__device__ __forceinline__ void stockham_stage(float (&xr)[N], float (&xi)[N], float (&yr)[N],
                                               float (&yi)[N])
{
    if constexpr (n > 1) {
        constexpr int r = stockham_radix(n);
        constexpr int m = n / r;
        constexpr int scale = N / n;
        static_for<0, m>([&](auto P) {
            constexpr int p = decltype(P)::value;
            static_for<0, s>([&](auto Q) {
                constexpr int q = decltype(Q)::value;
                float br[4], bi[4];
                if constexpr (r == 2) {
                    const float a0r = xr[q + s * p], a0i = xi[q + s * p];
                    const float a1r = xr[q + s * (p + m)], a1i = xi[q + s * (p + m)];
                    br[0] = a0r + a1r; bi[0] = a0i + a1i;
                    br[1] = a0r - a1r; bi[1] = a0i - a1i;
                } else if constexpr (r == 4) {
                    const float a0r = xr[q + s * p], a0i = xi[q + s * p];
                    const float a1r = xr[q + s * (p + m)], a1i = xi[q + s * (p + m)];
                    const float a2r = xr[q + s * (p + 2 * m)], a2i = xi[q + s * (p + 2 * m)];
                    const float a3r = xr[q + s * (p + 3 * m)], a3i = xi[q + s * (p + 3 * m)];
                    const float t0r = a0r + a2r, t0i = a0i + a2i;
                    const float t1r = a0r - a2r, t1i = a0i - a2i;
                    const float t2r = a1r + a3r, t2i = a1i + a3i;
                    const float t3r = a1r - a3r, t3i = a1i - a3i;
                    // t3 * (-i) forward, t3 * (+i) inverse
                    const float u3r = INV ? -t3i : t3i;
                    const float u3i = INV ? t3r : -t3r;
                    br[0] = t0r + t2r; bi[0] = t0i + t2i;
                    br[2] = t0r - t2r; bi[2] = t0i - t2i;
                    br[1] = t1r + u3r; bi[1] = t1i + u3i;
                    br[3] = t1r - u3r; bi[3] = t1i - u3i;
                } else {
                    constexpr float c3 = -0.5f;
                    constexpr float s3 = INV ? 0.86602540378443864676f : -0.86602540378443864676f;
                    const float a0r = xr[q + s * p], a0i = xi[q + s * p];
                    const float a1r = xr[q + s * (p + m)], a1i = xi[q + s * (p + m)];
                    const float a2r = xr[q + s * (p + 2 * m)], a2i = xi[q + s * (p + 2 * m)];
                    const float sr = a1r + a2r, si = a1i + a2i;
                    const float dr = a1r - a2r, di = a1i - a2i;
                    const float mr = a0r + c3 * sr, mi = a0i + c3 * si;
                    const float er = -s3 * di, ei = s3 * dr;
                    br[0] = a0r + sr; bi[0] = a0i + si;
                    br[1] = mr + er;  bi[1] = mi + ei;
                    br[2] = mr - er;  bi[2] = mi - ei;
                }
                static_for<0, r>([&](auto J) {
                    constexpr int j = decltype(J)::value;
                    constexpr int t = (p * j * scale) % N;
                    constexpr int o = q + s * (r * p + j);
                    if constexpr (t == 0) {
                        yr[o] = br[j]; yi[o] = bi[j];
                    } else if constexpr (4 * t == N) {          // w = -i (fwd), +i (inv)
                        yr[o] = INV ? -bi[j] : bi[j];
                        yi[o] = INV ? br[j] : -br[j];
                    } else if constexpr (2 * t == N) {          // w = -1
                        yr[o] = -br[j]; yi[o] = -bi[j];
                    } else if constexpr (4 * t == 3 * N) {      // w = +i (fwd), -i (inv)
                        yr[o] = INV ? bi[j] : -bi[j];
                        yi[o] = INV ? -br[j] : br[j];
                    } else {
                        constexpr float wr = (float)ct_cos2pi(t, N);
                        constexpr float wi = INV ? (float)ct_sin2pi(t, N) : (float)-ct_sin2pi(t, N);
                        yr[o] = br[j] * wr - bi[j] * wi;
                        yi[o] = br[j] * wi + bi[j] * wr;
                    }
                });
            });
        });
        stockham_stage<N, INV, m, s * r>(yr, yi, xr, xi);
    }
}

// In-place (in registers) complex FFT of length N.
template <int N, bool INV>
__device__ __forceinline__ void fft_reg(float (&xr)[N], float (&xi)[N])
{
    float yr[N], yi[N];
    stockham_stage<N, INV, N, 1>(xr, xi, yr, yi);
    if constexpr (stockham_stages(N) % 2 == 1) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            xr[k] = yr[k];
            xi[k] = yi[k];
        }
    }
}

}  // namespace pivamd

namespace pivamd {

// ---------------------------------------------------------------------------
// One transform of length N shared by R lanes (R = 1, 2 or 4: the R lanes of
// a line are consecutive and quad-aligned; R = 8: lanes {0-3, 8-11} + 16m
// or {4-7, 12-15} + 16m, see the R = 8 exchange below).
//   in : lane q holds x[n1*R + q], n1 = 0..N/R-1
//   out: lane q holds X[k1 + (N/R) * out_block(q)], k1 = 0..N/R-1
// X[k1 + N1 k2] = sum_n2 W_R^(n2 k2) W_N^(n2 k1) FFT_N1(x[. R + n2])[k1]:
// each lane does the N1-point sub-transform in registers, applies its
// twiddle W_N^(q k1), then the radix-R combination runs across the R lanes
// as R = 2: one DPP exchange; R = 4: two exchange stages (DIF, so the
// blocks come out bit-reversed).
//
// The twiddles come from a table (SplitTw: entry q*N1 + k1 = W_N^(q k1),
// forward sign) that the kernel copies to LDS once; lane q reads its row
// with immediate offsets.  A per-lane select among R compile-time
// constants costs 2(R-1) instructions per value, the table one LDS read per
// two twiddles.  Each exchange step is own*sign + partner: the product is
// exact (sign = +-1), and keeping it a separate multiply lets the DPP move
// fold into the add (v_add_f32_dpp) instead of feeding an FMA.
// ---------------------------------------------------------------------------
template <int N, int R>
struct SplitTw {
    float v[2 * N];
    constexpr SplitTw() : v{}
    {
        for (int q = 0; q < R; ++q)
            for (int k = 0; k < N / R; ++k) {
                v[2 * (q * (N / R) + k)] = (float)ct_cos2pi((long)q * k, N);
                v[2 * (q * (N / R) + k) + 1] = (float)-ct_sin2pi((long)q * k, N);
            }
    }
};

template <int N, int R>
__constant__ SplitTw<N, R> kSplitTw{};

__device__ __forceinline__ float dpp_swap1(float v)
{
    // quad_perm [1,0,3,2]
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}

__device__ __forceinline__ float dpp_swap2(float v)
{
    // quad_perm [2,3,0,1]
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
}

__device__ __forceinline__ float dpp_ror8(float v)
{
    // row_ror:8 (lane i of a 16-lane row reads lane i ^ 8)
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x128, 0xF, 0xF, true));
}

// partner + sign * own, rounded once (the product is exact)
__device__ __forceinline__ float xadd(float partner, float sign, float own)
{
#pragma clang fp contract(off)
    const float t = sign * own;
    return partner + t;
}

template <int R>
__device__ __forceinline__ int out_block(int q)
{
    if constexpr (R == 8) return ((q & 1) << 2) | (q & 2) | ((q >> 2) & 1);
    else if constexpr (R == 4) return ((q & 1) << 1) | (q >> 1);
    else return q;
}

// tw: this lane's table row, kSplitTw<N, R> entries q*N1 .. q*N1 + N1-1 (in
// LDS); unused for R = 1.
template <int N, int R, bool INV>
__device__ __forceinline__ void fft_split(float (&xr)[N / R], float (&xi)[N / R], int q, const float2 *tw)
{
    constexpr int N1 = N / R;
    fft_reg<N1, INV>(xr, xi);
    if constexpr (R > 1) {
        static_for<1, N1>([&](auto K) {
            constexpr int k1 = decltype(K)::value;
            const float2 w = tw[k1];
            const float c = w.x, s = INV ? -w.y : w.y;
            const float r = xr[k1], i = xi[k1];
            xr[k1] = r * c - i * s;
            xi[k1] = r * s + i * c;
        });
        if constexpr (R == 2) {
            const float sg = (q & 1) ? -1.0f : 1.0f;
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                xr[k] = xadd(dpp_swap1(xr[k]), sg, xr[k]);
                xi[k] = xadd(dpp_swap1(xi[k]), sg, xi[k]);
            }
        } else if constexpr (R == 8) {
            // radix-2 DIF over q in three exchange stages.  The caller maps
            // q bit 2 to lane bit 3 (partner 8 lanes away: row_ror:8) and
            // q bits 0-1 to the quad (quad_perm).  Stage A: lanes q >= 4 take
            // (partner - own) * W_8^(q & 3); stage B: (partner - own) * W_4
            // on q & 3 == 3; stage C: partner -+ own.  Lane q ends with the
            // block bitrev3(q) = out_block<8>(q).
            constexpr float h = 0.70710678118654752440f;
            const int j = q & 3;
            const bool hi = (q & 4) != 0;
            float wr = 1.0f, wi = 0.0f;
            if (hi) {
                wr = j == 0 ? 1.0f : (j == 1 ? h : (j == 2 ? 0.0f : -h));
                wi = j == 0 ? 0.0f : (j == 1 ? -h : (j == 2 ? -1.0f : -h));
                if (INV) wi = -wi;
            }
            const float sa = hi ? -1.0f : 1.0f;
            const float sb = (q & 2) ? -1.0f : 1.0f;
            const float sc = (q & 1) ? -1.0f : 1.0f;
            const bool rot = j == 3;
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                const float ur = xadd(dpp_ror8(xr[k]), sa, xr[k]);
                const float ui = xadd(dpp_ror8(xi[k]), sa, xi[k]);
                float tr = ur * wr - ui * wi;
                float ti = ur * wi + ui * wr;
                float vr = xadd(dpp_swap2(tr), sb, tr);
                float vi = xadd(dpp_swap2(ti), sb, ti);
                // W_4 = -i (forward) or +i (inverse) on q & 3 == 3
                const float rr = INV ? -vi : vi, ri = INV ? vr : -vr;
                vr = rot ? rr : vr;
                vi = rot ? ri : vi;
                xr[k] = xadd(dpp_swap1(vr), sc, vr);
                xi[k] = xadd(dpp_swap1(vi), sc, vi);
            }
        } else {
            const float sa = (q & 2) ? -1.0f : 1.0f;
            const float sb = (q & 1) ? -1.0f : 1.0f;
            const bool rot = (q == 3);
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                float tr = xadd(dpp_swap2(xr[k]), sa, xr[k]);
                float ti = xadd(dpp_swap2(xi[k]), sa, xi[k]);
                // lane 3 multiplies by W_4 = -i (forward) or +i (inverse)
                const float rr = INV ? -ti : ti, ri = INV ? tr : -tr;
                tr = rot ? rr : tr;
                ti = rot ? ri : ti;
                xr[k] = xadd(dpp_swap1(tr), sb, tr);
                xi[k] = xadd(dpp_swap1(ti), sb, ti);
            }
        }
    }
}

}  // namespace pivamd
