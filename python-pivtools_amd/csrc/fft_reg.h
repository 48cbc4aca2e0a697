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
