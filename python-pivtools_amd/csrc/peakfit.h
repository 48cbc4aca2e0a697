// peakfit.h -- per-window sub-pixel peak fit on the device.
//
// Follows the scalar fitter of pivtools_cli/lib/peak_locate_lm.c:
//   threept_estimate          :35-66   (3-point log-Gaussian, also the LM seed)
//   4/5/6-DOF models          :69-103
//   residual / Jacobian       :108-333 (float model, FP64 sums, cached model)
//   solve_lm_step             :342-393 (FP64 Cholesky)
//   lm_gauss{4,5,6}_fit       :412-712 (lambda 0.01, x0.5 / x2, max_iter 50,
//                                       tol 1e-6, trust rule, width guard)
// The arithmetic is kept unfused (no FMA contraction) and exp/log are
// evaluated in FP64 and rounded, which reproduces the host libm float
// results, so the device fit agrees with the host fit on the same patch.
// The batch flavour (selector value 1, pivtools_cli/lib/peak_locate_lm_batch.c)
// swaps in the Cephes degree-5 exponential of peak_simd.h:159-185 and the
// batch trust-rule association (1e-3*25)*A*A (peak_locate_lm_batch.c:426).
#pragma once

#include <hip/hip_runtime.h>

namespace pivamd {

#pragma clang fp contract(off)

constexpr int PK = 5;
constexpr float kFltEps = 1.19209290e-07f;

__device__ __forceinline__ float exp_libm(float x) { return (float)exp((double)x); }

// Cephes-style expf of the lockstep batch fitter (peak_simd.h:159-185).
__device__ __forceinline__ float exp_cephes(float x)
{
    const float xc = fmaxf(x, -87.3365478515625f);
    const float rn = (1.44269504088896341f * xc + 12582912.0f) - 12582912.0f;
    const int n = (int)rn;
    const float r = xc - rn * 0.693359375f - rn * -2.12194440e-4f;
    const float r2 = r * r;
    float p = 1.9875691500e-4f;
    p = p * r + 1.3981999507e-3f;
    p = p * r + 8.3334519073e-3f;
    p = p * r + 4.1665795894e-2f;
    p = p * r + 1.6666665459e-1f;
    p = p * r + 5.0000001201e-1f;
    const float er = p * r2 + r + 1.0f;
    return er * __int_as_float((n + 127) << 23);
}

template <bool CEPHES>
__device__ __forceinline__ float pexp(float x)
{
    if constexpr (CEPHES) return exp_cephes(x);
    else return exp_libm(x);
}

__device__ __forceinline__ void threept(const float *xc, float *pk, float *A, float *sx, float *sy)
{
    float xf[3], yf[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        xf[i] = xc[(i - 1 + 2) * PK + 2];
        yf[i] = xc[2 * PK + (i - 1 + 2)];
        xf[i] = (float)log((double)((xf[i] < kFltEps) ? kFltEps : xf[i]));
        yf[i] = (float)log((double)((yf[i] < kFltEps) ? kFltEps : yf[i]));
    }
    const float dx = 2 * xf[0] - 4 * xf[1] + 2 * xf[2];
    const float dy = 2 * yf[0] - 4 * yf[1] + 2 * yf[2];
    pk[0] = (fabs((double)dx) > (double)kFltEps) ? (xf[0] - xf[2]) / dx : 0.0f;
    pk[1] = (fabs((double)dy) > (double)kFltEps) ? (yf[0] - yf[2]) / dy : 0.0f;
    *A = xc[2 * PK + 2];
    *sx = (float)sqrt((double)(-4.0f / (dx + kFltEps)));
    *sy = (float)sqrt((double)(-4.0f / (dy + kFltEps)));
}

template <int NP, bool CEPHES>
__device__ __forceinline__ float model(float i, float j, const float *p)
{
    if constexpr (NP == 4) {
        const float di = (i - p[1]) / p[3], dj = (j - p[2]) / p[3];
        return p[0] * pexp<CEPHES>(-(di * di + dj * dj));
    } else if constexpr (NP == 5) {
        const float di = (i - p[1]) / p[3], dj = (j - p[2]) / p[4];
        return p[0] * pexp<CEPHES>(-(di * di + dj * dj));
    } else {
        const float di = i - p[1], dj = j - p[2];
        return p[0] * pexp<CEPHES>(-0.5f * (di * di / p[3] + dj * dj / p[4] + 2.0f * di * dj * p[5]));
    }
}

template <int NP, bool CEPHES>
__device__ double resid_only(const float *xc, const float *p, float *pred)
{
    double rs = 0.0;
    for (int ii = 0; ii < PK; ++ii) {
        const float i = (float)(ii - 2);
        for (int jj = 0; jj < PK; ++jj) {
            const float j = (float)(jj - 2);
            const float v = model<NP, CEPHES>(i, j, p);
            pred[ii * PK + jj] = v;
            const float r = v - xc[ii * PK + jj];
            rs += (double)r * (double)r;
        }
    }
    return rs;
}

template <int NP>
__device__ void build_normal(const float *xc, const float *p, const float *pred, double *JtJ, double *Jtr)
{
    for (int a = 0; a < NP * NP; ++a) JtJ[a] = 0.0;
    for (int a = 0; a < NP; ++a) Jtr[a] = 0.0;
    for (int ii = 0; ii < PK; ++ii) {
        const float i = (float)(ii - 2);
        for (int jj = 0; jj < PK; ++jj) {
            const float j = (float)(jj - 2);
            const int idx = ii * PK + jj;
            const float v = pred[idx];
            const float r = v - xc[idx];
            float J[NP];
            if constexpr (NP == 4) {
                const float s = p[3];
                const float di = (i - p[1]) / s, dj = (j - p[2]) / s;
                const float r2 = di * di + dj * dj;
                J[0] = v / p[0];
                J[1] = 2.0f * v * di / s;
                J[2] = 2.0f * v * dj / s;
                J[3] = 2.0f * v * r2 / s;
            } else if constexpr (NP == 5) {
                const float di = (i - p[1]) / p[3], dj = (j - p[2]) / p[4];
                J[0] = v / p[0];
                J[1] = 2.0f * v * di / p[3];
                J[2] = 2.0f * v * dj / p[4];
                J[3] = 2.0f * v * di * di / p[3];
                J[4] = 2.0f * v * dj * dj / p[4];
            } else {
                const float di = i - p[1], dj = j - p[2];
                J[0] = v / p[0];
                J[1] = v * (di / p[3] + dj * p[5]);
                J[2] = v * (dj / p[4] + di * p[5]);
                J[3] = 0.5f * v * di * di / (p[3] * p[3]);
                J[4] = 0.5f * v * dj * dj / (p[4] * p[4]);
                J[5] = -v * di * dj;
            }
#pragma unroll
            for (int a = 0; a < NP; ++a) {
                Jtr[a] += (double)J[a] * (double)r;
#pragma unroll
                for (int b = 0; b <= a; ++b) JtJ[a * NP + b] += (double)J[a] * (double)J[b];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < NP; ++a)
#pragma unroll
        for (int b = a + 1; b < NP; ++b) JtJ[a * NP + b] = JtJ[b * NP + a];
}

template <int NP>
__device__ int solve_step(const double *JtJ, const double *Jtr, float lambda, float *delta)
{
    double A[NP * NP], L[NP * NP], y[NP], d[NP];
#pragma unroll
    for (int a = 0; a < NP * NP; ++a) { A[a] = JtJ[a]; L[a] = 0.0; }
#pragma unroll
    for (int i = 0; i < NP; ++i) A[i * NP + i] *= (1.0 + (double)lambda);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double s = A[i * NP + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[i * NP + k] * L[j * NP + k];
            if (i == j) {
                if (s <= 0.0) return -1;
                L[i * NP + j] = sqrt(s);
            } else {
                L[i * NP + j] = s / L[j * NP + j];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        double s = -Jtr[i];
#pragma unroll
        for (int j = 0; j < i; ++j) s -= L[i * NP + j] * y[j];
        y[i] = s / L[i * NP + i];
    }
#pragma unroll
    for (int i = NP - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int j = i + 1; j < NP; ++j) s -= L[j * NP + i] * d[j];
        d[i] = s / L[i * NP + i];
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) delta[i] = (float)d[i];
    return 0;
}

// LM fit of an NP-parameter model to the 5x5 patch xc; returns 0 if
// trustworthy.  pk[0..1] = sub-pixel offsets, sig[0..2] = widths,
// pout[0..5] = final model parameters (for the multi-peak subtraction).
template <int NP, bool CEPHES>
__device__ int lm_fit(const float *xc, float *pk, float *sig, float *pout)
{
    float p[6] = {0, 0, 0, 0, 0, 0}, pn[6], delta[6];
    double JtJ[NP * NP], Jtr[NP];
    float pred[PK * PK];
    float lambda = 0.01f;
    int fit_ok = 0;
    float A, sx, sy;
    threept(xc, pk, &A, &sx, &sy);
    p[0] = A;
    p[1] = fminf(fmaxf(pk[0], -2.0f), 2.0f);
    p[2] = fminf(fmaxf(pk[1], -2.0f), 2.0f);
    if constexpr (NP == 4) {
        const float s = sqrtf(sx * sx + sy * sy);
        p[3] = fminf(fmaxf(s, 0.5f), 3.0f);
    } else if constexpr (NP == 5) {
        p[3] = fminf(fmaxf(sx, 0.5f), 3.0f);
        p[4] = fminf(fmaxf(sy, 0.5f), 3.0f);
    } else {
        p[5] = 0.0f;
        p[3] = fminf(fmaxf(sx * sx, 0.25f), 9.0f);
        p[4] = fminf(fmaxf(sy * sy, 0.25f), 9.0f);
    }
    double residual = resid_only<NP, CEPHES>(xc, p, pred);
    build_normal<NP>(xc, p, pred, JtJ, Jtr);
    for (int iter = 0; iter < 50; ++iter) {
        if (solve_step<NP>(JtJ, Jtr, lambda, delta) != 0) break;
#pragma unroll
        for (int k = 0; k < NP; ++k) pn[k] = p[k] + delta[k];
        pn[0] = fmaxf(pn[0], p[0] * 0.5f);
        pn[1] = fminf(fmaxf(pn[1], -2.5f), 2.5f);
        pn[2] = fminf(fmaxf(pn[2], -2.5f), 2.5f);
        if constexpr (NP == 4) {
            pn[3] = fminf(fmaxf(pn[3], 0.25f), 4.0f);
        } else if constexpr (NP == 5) {
            pn[3] = fminf(fmaxf(pn[3], 0.25f), 4.0f);
            pn[4] = fminf(fmaxf(pn[4], 0.25f), 4.0f);
        } else {
            pn[3] = fminf(fmaxf(pn[3], 0.1f), 16.0f);
            pn[4] = fminf(fmaxf(pn[4], 0.1f), 16.0f);
            const float smax = 0.95f / sqrtf(pn[3] * pn[4]);
            pn[5] = fminf(fmaxf(pn[5], -smax), smax);
        }
        const double new_residual = resid_only<NP, CEPHES>(xc, pn, pred);
        if (new_residual < residual) {
#pragma unroll
            for (int k = 0; k < NP; ++k) p[k] = pn[k];
            const double improvement = (residual - new_residual) / (residual + (double)kFltEps);
            residual = new_residual;
            lambda *= 0.5f;
            build_normal<NP>(xc, p, pred, JtJ, Jtr);
            if (improvement < 1e-6) { fit_ok = 1; break; }
        } else {
            lambda *= 2.0f;
            if (lambda > 1e6f) break;
        }
    }
    if (!fit_ok) {
        if constexpr (CEPHES) fit_ok = (residual <= (double)((1e-3f * 25.0f) * p[0] * p[0]));
        else fit_ok = (residual <= (double)(1e-3f * p[0] * p[0] * 25.0f));
    }
    if (fit_ok) {
        if constexpr (NP == 4) fit_ok = (p[3] > 0.25f && p[3] < 4.0f);
        else if constexpr (NP == 5) fit_ok = (p[3] > 0.25f && p[3] < 4.0f && p[4] > 0.25f && p[4] < 4.0f);
        else fit_ok = (p[3] > 0.1f && p[3] < 16.0f && p[4] > 0.1f && p[4] < 16.0f);
    }
    pk[0] = p[1];
    pk[1] = p[2];
    if constexpr (NP == 4) { sig[0] = p[3]; sig[1] = p[3]; sig[2] = 0.0f; }
    else if constexpr (NP == 5) { sig[0] = p[3]; sig[1] = p[4]; sig[2] = 0.0f; }
    else { sig[0] = p[3]; sig[1] = p[4]; sig[2] = p[5]; }
#pragma unroll
    for (int k = 0; k < 6; ++k) pout[k] = p[k];
    return fit_ok ? 0 : -1;
}

// Model value on the patch for the multi-peak subtraction (5-DOF form for
// the 3-point fit, the fitted model for LM fits).
template <int NP, bool CEPHES>
__device__ float model_at(int ii, int jj, const float *p)
{
    return model<NP, CEPHES>((float)(ii - 2), (float)(jj - 2), p);
}

#pragma clang fp contract(on)

}  // namespace pivamd
