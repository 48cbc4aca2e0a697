/*
 * piv_oracle.c -- CPU restatement of the PIVtools correlation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it to produce
 * the expected answer.  Nothing in the shipped library links or calls it.
 *
 * Every function restates the algorithm of the reference C path and names
 * the file:line it follows (paths relative to the reference repository):
 *
 *   orc_convolve                  pivtools_cli/lib/xcorr.c:22-89
 *   orc_bulkxcorr2d               pivtools_cli/lib/PIV_2d_cross_correlate.c:203-397
 *   orc_bulkxcorr2d_accumulate    pivtools_cli/lib/PIV_2d_cross_correlate.c:502-607
 *   orc_bulkxcorr2d_accumulate_triple
 *                                 pivtools_cli/lib/PIV_2d_cross_correlate.c:432-484,627-833
 *   separable r2c / c2r / shift   pivtools_cli/lib/codelet_fft.c:155-203,290-366
 *   orc_lsqpeaklocate_lm          pivtools_cli/lib/peak_locate_lm.c:35-66,108-393,405-712,767-918
 *   orc_peak_detectability        pivtools_cli/lib/peak_locate_lm.c:727-744
 *   orc_fused_symmetric_warp(_batch)
 *                                 pivtools_cli/lib/fused_warp.c:52-582 (scalar path, impl 0)
 *
 * The 1D transforms are a float32 Stockham mixed-radix (2/3/4) FFT written
 * here; the reference's generated codelets are not reproduced (they are
 * generated code and are not built).  Where the reference is pure scalar C
 * (the peak fitter and the scalar warp) the arithmetic order is kept so the
 * compiled reference sources agree bit for bit.
 *
 * Parity pins (what checks this file, see tests/):
 *   FFT plane      brute-force FP64 circular correlation + numpy FP64 FFT
 *   plane weight   closed form (N-|m|)(N-|n|) for the square window
 *   peak fitter    the reference's own peak_locate_lm.c built in oracle/_ref
 *                  (bit-exact), closed-form Gaussians, known-answer planes
 *   warp           the reference's own fused_warp.c built in oracle/_ref
 *                  (bit-exact, scalar impl), integer-shift round trip
 *   accumulate     FP64 numpy fftshift(ifft2(fft2(B)*conj(fft2(A)))),
 *                  offset/gain invariants
 */
#include <float.h>
#include <math.h>
#include <stdint.h>

/* ORC_CR_LIBM (liboracle_cr.so): expf and logf correctly rounded, i.e.
 * evaluated in double and rounded once -- the definition the device fitter
 * uses.  Without it (liboracle.so) the host libm's expf/logf are called, as
 * the reference's C does; glibc's are not correctly rounded for every input,
 * so the two flavours can differ in the last bit of an exponential, and an
 * LM trust decision sitting on its threshold can flip between them. */
#ifdef ORC_CR_LIBM
static inline float orc_expf_cr(float x) { return (float)exp((double)x); }
static inline float orc_logf_cr(float x) { return (float)log((double)x); }
#define expf orc_expf_cr
#define logf orc_logf_cr
#endif
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PK 5

/* ------------------------------------------------------------------------ */
/* float32 Stockham FFT (own code)                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n;
    float *twr, *twi;   /* exp(-2*pi*i*t/n), t in [0,n), from double */
} orc_fft_plan;

static int orc_size_ok(int n)
{
    switch (n) {
    case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96:
    case 128: case 192: case 256:
        return 1;
    default:
        return 0;
    }
}

static void orc_plan_init(orc_fft_plan *p, int n)
{
    p->n = n;
    p->twr = (float *)malloc(sizeof(float) * n);
    p->twi = (float *)malloc(sizeof(float) * n);
    for (int t = 0; t < n; ++t) {
        double a = -2.0 * M_PI * (double)t / (double)n;
        p->twr[t] = (float)cos(a);
        p->twi[t] = (float)sin(a);
    }
}

static void orc_plan_free(orc_fft_plan *p)
{
    free(p->twr);
    free(p->twi);
}

/* One Stockham (self-sorting DIF) stage of radix r on a length-n block with
 * stride s; input x, output y.  Twiddles come from the top-level table
 * (index scaled by N/n).  sign = -1 forward, +1 inverse (unnormalised). */
static void orc_stockham(const orc_fft_plan *P, int n, int s, int sign,
                         float *xr, float *xi, float *yr, float *yi, int eo)
{
    if (n == 1) {
        if (eo)
            for (int q = 0; q < s; ++q) { yr[q] = xr[q]; yi[q] = xi[q]; }
        return;
    }
    const int r = (n % 4 == 0) ? 4 : ((n % 2 == 0) ? 2 : 3);
    const int m = n / r;
    const int scale = P->n / n;
    const float c3 = -0.5f, s3 = (float)(0.86602540378443864676 * (double)sign);
    for (int p = 0; p < m; ++p) {
        for (int q = 0; q < s; ++q) {
            float ar[4], ai[4], br[4], bi[4];
            for (int k = 0; k < r; ++k) {
                ar[k] = xr[q + s * (p + k * m)];
                ai[k] = xi[q + s * (p + k * m)];
            }
            if (r == 2) {
                br[0] = ar[0] + ar[1]; bi[0] = ai[0] + ai[1];
                br[1] = ar[0] - ar[1]; bi[1] = ai[0] - ai[1];
            } else if (r == 4) {
                float t0r = ar[0] + ar[2], t0i = ai[0] + ai[2];
                float t1r = ar[0] - ar[2], t1i = ai[0] - ai[2];
                float t2r = ar[1] + ar[3], t2i = ai[1] + ai[3];
                float t3r = ar[1] - ar[3], t3i = ai[1] - ai[3];
                /* multiply t3 by (sign * i): forward -i, inverse +i */
                float u3r = (sign < 0) ? t3i : -t3i;
                float u3i = (sign < 0) ? -t3r : t3r;
                br[0] = t0r + t2r; bi[0] = t0i + t2i;
                br[2] = t0r - t2r; bi[2] = t0i - t2i;
                br[1] = t1r + u3r; bi[1] = t1i + u3i;
                br[3] = t1r - u3r; bi[3] = t1i - u3i;
            } else {
                float sr = ar[1] + ar[2], si = ai[1] + ai[2];
                float dr = ar[1] - ar[2], di = ai[1] - ai[2];
                float mr = ar[0] + c3 * sr, mi = ai[0] + c3 * si;
                /* (s3 * i) * d */
                float er = -s3 * di, ei = s3 * dr;
                br[0] = ar[0] + sr; bi[0] = ai[0] + si;
                br[1] = mr + er;    bi[1] = mi + ei;
                br[2] = mr - er;    bi[2] = mi - ei;
            }
            for (int j = 0; j < r; ++j) {
                int t = (p * j * scale) % P->n;
                float wr = P->twr[t];
                float wi = (sign < 0) ? P->twi[t] : -P->twi[t];
                float vr = br[j] * wr - bi[j] * wi;
                float vi = br[j] * wi + bi[j] * wr;
                if (t == 0) { vr = br[j]; vi = bi[j]; }
                yr[q + s * (r * p + j)] = vr;
                yi[q + s * (r * p + j)] = vi;
            }
        }
    }
    orc_stockham(P, m, s * r, sign, yr, yi, xr, xi, !eo);
}

/* In-place complex FFT of length P->n (scratch of 2n floats). */
static void orc_cfft(const orc_fft_plan *P, float *re, float *im, float *scr, int sign)
{
    orc_stockham(P, P->n, 1, sign, re, im, scr, scr + P->n, 0);
}

/* ------------------------------------------------------------------------ */
/* Separable 2D transforms (codelet_fft.c:155-203, 290-366)                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    int H, W, hw;
    orc_fft_plan row, col;
    float *bufr, *bufi, *scr;
} orc_plan2d;

static void orc_plan2d_init(orc_plan2d *p, int H, int W)
{
    p->H = H; p->W = W; p->hw = W / 2 + 1;
    orc_plan_init(&p->row, W);
    orc_plan_init(&p->col, H);
    int mx = H > W ? H : W;
    p->bufr = (float *)malloc(sizeof(float) * mx);
    p->bufi = (float *)malloc(sizeof(float) * mx);
    p->scr = (float *)malloc(sizeof(float) * 2 * mx);
}

static void orc_plan2d_free(orc_plan2d *p)
{
    orc_plan_free(&p->row);
    orc_plan_free(&p->col);
    free(p->bufr); free(p->bufi); free(p->scr);
}

/* r2c: rfft along the W-length rows, then cfft along the H-length columns.
 * Output split re/im, [H][W/2+1]. */
static void orc_r2c_2d(orc_plan2d *p, const float *in, float *outr, float *outi)
{
    const int H = p->H, W = p->W, hw = p->hw;
    for (int r = 0; r < H; ++r) {
        for (int x = 0; x < W; ++x) { p->bufr[x] = in[r * W + x]; p->bufi[x] = 0.0f; }
        orc_cfft(&p->row, p->bufr, p->bufi, p->scr, -1);
        for (int k = 0; k < hw; ++k) { outr[r * hw + k] = p->bufr[k]; outi[r * hw + k] = p->bufi[k]; }
    }
    for (int k = 0; k < hw; ++k) {
        for (int r = 0; r < H; ++r) { p->bufr[r] = outr[r * hw + k]; p->bufi[r] = outi[r * hw + k]; }
        orc_cfft(&p->col, p->bufr, p->bufi, p->scr, -1);
        for (int r = 0; r < H; ++r) { outr[r * hw + k] = p->bufr[r]; outi[r * hw + k] = p->bufi[r]; }
    }
}

/* c2r: icfft along columns, then the Hermitian irfft along rows with the DC
 * and Nyquist imaginary parts forced to zero (codelet_gen/gen_codelet.py:198-200).
 * Unnormalised.  Destroys inr/ini. */
static void orc_c2r_2d(orc_plan2d *p, float *inr, float *ini, float *out)
{
    const int H = p->H, W = p->W, hw = p->hw;
    for (int k = 0; k < hw; ++k) {
        for (int r = 0; r < H; ++r) { p->bufr[r] = inr[r * hw + k]; p->bufi[r] = ini[r * hw + k]; }
        orc_cfft(&p->col, p->bufr, p->bufi, p->scr, +1);
        for (int r = 0; r < H; ++r) { inr[r * hw + k] = p->bufr[r]; ini[r * hw + k] = p->bufi[r]; }
    }
    for (int r = 0; r < H; ++r) {
        for (int k = 0; k < hw; ++k) { p->bufr[k] = inr[r * hw + k]; p->bufi[k] = ini[r * hw + k]; }
        p->bufi[0] = 0.0f;
        p->bufi[W / 2] = 0.0f;
        for (int k = hw; k < W; ++k) { p->bufr[k] = p->bufr[W - k]; p->bufi[k] = -p->bufi[W - k]; }
        orc_cfft(&p->row, p->bufr, p->bufi, p->scr, +1);
        for (int x = 0; x < W; ++x) out[r * W + x] = p->bufr[x];
    }
}

/* x 1/numel then fftshift (zero lag at [H/2][W/2]); codelet_fft.c:155-167 */
static void orc_finish(int H, int W, float *raw, float *out)
{
    const int numel = H * W;
    const float mul = 1.0f / (float)numel;
    for (int j = 0; j < numel; ++j) raw[j] *= mul;
    for (int row = 0; row < H; ++row) {
        int rs = (row + H / 2) % H;
        for (int j = 0; j < W / 2; ++j) {
            out[row * W + W / 2 + j] = raw[rs * W + j];
            out[row * W + j] = raw[rs * W + W / 2 + j];
        }
    }
}

/* plane = shift( IFFT( FFT(s0) .* conj(FFT(s1)) ) ) / numel  (codelet_fft.c:335-352) */
typedef struct {
    orc_plan2d p;
    float *f0r, *f0i, *f1r, *f1i, *cr, *ci, *raw;
} orc_xc;

static void orc_xc_init(orc_xc *x, int H, int W)
{
    orc_plan2d_init(&x->p, H, W);
    size_t nf = (size_t)H * (W / 2 + 1);
    x->f0r = (float *)malloc(sizeof(float) * nf); x->f0i = (float *)malloc(sizeof(float) * nf);
    x->f1r = (float *)malloc(sizeof(float) * nf); x->f1i = (float *)malloc(sizeof(float) * nf);
    x->cr = (float *)malloc(sizeof(float) * nf);  x->ci = (float *)malloc(sizeof(float) * nf);
    x->raw = (float *)malloc(sizeof(float) * H * W);
}

static void orc_xc_free(orc_xc *x)
{
    orc_plan2d_free(&x->p);
    free(x->f0r); free(x->f0i); free(x->f1r); free(x->f1i); free(x->cr); free(x->ci); free(x->raw);
}

static void orc_xc_forward(orc_xc *x, const float *in, int slot)
{
    if (slot == 0) orc_r2c_2d(&x->p, in, x->f0r, x->f0i);
    else           orc_r2c_2d(&x->p, in, x->f1r, x->f1i);
}

static void orc_xc_emit_xcorr(orc_xc *x, float *out)
{
    const int nf = x->p.H * x->p.hw;
    for (int i = 0; i < nf; ++i) {
        const float Ar = x->f0r[i], Ai = x->f0i[i];
        const float Br = x->f1r[i], Bi = x->f1i[i];
        x->cr[i] = Ar * Br + Ai * Bi;
        x->ci[i] = Ai * Br - Ar * Bi;
    }
    orc_c2r_2d(&x->p, x->cr, x->ci, x->raw);
    orc_finish(x->p.H, x->p.W, x->raw, out);
}

static void orc_xc_emit_power(orc_xc *x, int slot, float *out)
{
    const int nf = x->p.H * x->p.hw;
    const float *Fr = slot == 0 ? x->f0r : x->f1r;
    const float *Fi = slot == 0 ? x->f0i : x->f1i;
    for (int i = 0; i < nf; ++i) {
        x->cr[i] = Fr[i] * Fr[i] + Fi[i] * Fi[i];
        x->ci[i] = 0.0f;
    }
    orc_c2r_2d(&x->p, x->cr, x->ci, x->raw);
    orc_finish(x->p.H, x->p.W, x->raw, out);
}

/* ------------------------------------------------------------------------ */
/* Correlation-plane weight (xcorr.c:22-89; PIV_2d_cross_correlate.c:219-222) */
/* ------------------------------------------------------------------------ */

/* conv[i][j] = sum_{y,x} a[y+i-H/2][x+j-W/2] * b[y][x] (zero outside), i.e.
 * the centre H x W crop of the 2N zero-padded correlation.  Summed in FP64. */
unsigned orc_convolve(const float *a, const float *b, float *conv, const int *N)
{
    const int H = N[0], W = N[1];
    if (!orc_size_ok(2 * H) || !orc_size_ok(2 * W)) return 4;
    for (int i = 0; i < H; ++i) {
        for (int j = 0; j < W; ++j) {
            const int my = i - H / 2, mx = j - W / 2;
            double s = 0.0;
            for (int y = 0; y < H; ++y) {
                int ya = y + my;
                if (ya < 0 || ya >= H) continue;
                for (int x = 0; x < W; ++x) {
                    int xa = x + mx;
                    if (xa < 0 || xa >= W) continue;
                    s += (double)a[ya * W + xa] * (double)b[y * W + x];
                }
            }
            conv[i * W + j] = (float)s;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Peak search and sub-pixel fits (peak_locate_lm.c)                        */
/* ------------------------------------------------------------------------ */

/* threept_estimate, peak_locate_lm.c:35-66 */
static void orc_threept(const float *xc, const int *N, float *pk, float *A_out, float *sx_out, float *sy_out)
{
    float xf[3], yf[3];
    for (int i = 0; i < 3; ++i) {
        xf[i] = xc[(i - 1 + (N[0] - 1) / 2) * N[1] + (N[1] - 1) / 2];
        yf[i] = xc[(N[0] - 1) / 2 * N[1] + (i - 1 + (N[1] - 1) / 2)];
        xf[i] = (float)log((xf[i] < FLT_EPSILON) ? FLT_EPSILON : xf[i]);
        yf[i] = (float)log((yf[i] < FLT_EPSILON) ? FLT_EPSILON : yf[i]);
    }
    float dx = 2 * xf[0] - 4 * xf[1] + 2 * xf[2];
    float dy = 2 * yf[0] - 4 * yf[1] + 2 * yf[2];
    pk[0] = (fabs(dx) > FLT_EPSILON) ? (xf[0] - xf[2]) / dx : 0.0f;
    pk[1] = (fabs(dy) > FLT_EPSILON) ? (yf[0] - yf[2]) / dy : 0.0f;
    *A_out = xc[(N[0] - 1) / 2 * N[1] + (N[1] - 1) / 2];
    *sx_out = (float)sqrt(-4.0f / (dx + FLT_EPSILON));
    *sy_out = (float)sqrt(-4.0f / (dy + FLT_EPSILON));
}

/* The three Gaussian models (peak_locate_lm.c:69-103). */
static inline float orc_g4(float i, float j, float A, float i0, float j0, float s)
{
    float di = (i - i0) / s, dj = (j - j0) / s;
    return A * expf(-(di * di + dj * dj));
}
static inline float orc_g5(float i, float j, float A, float i0, float j0, float sx, float sy)
{
    float di = (i - i0) / sx, dj = (j - j0) / sy;
    return A * expf(-(di * di + dj * dj));
}
static inline float orc_g6(float i, float j, float A, float i0, float j0, float sx, float sy, float sxy)
{
    float di = i - i0, dj = j - j0;
    return A * expf(-0.5f * (di * di / sx + dj * dj / sy + 2.0f * di * dj * sxy));
}

/* Residual / Jacobian of the 4/5/6-DOF models with the cached model values
 * (peak_locate_lm.c:108-333).  np = number of parameters, p = parameters
 * (A, i0, j0, then widths).  compute_jac = 0: residual only, fills pred. */
static double orc_resid_jac(int np, const float *xc, const float *p, double *JtJ, double *Jtr,
                            int compute_jac, float *pred)
{
    const int n0 = ORC_PK, n1 = ORC_PK;
    double rs = 0.0;
    if (!compute_jac) {
        for (int ii = 0; ii < n0; ++ii) {
            float i = (float)(ii - (n0 - 1) / 2);
            int row = ii * n1;
            for (int jj = 0; jj < n1; ++jj) {
                float j = (float)(jj - (n1 - 1) / 2);
                float v;
                if (np == 4) v = orc_g4(i, j, p[0], p[1], p[2], p[3]);
                else if (np == 5) v = orc_g5(i, j, p[0], p[1], p[2], p[3], p[4]);
                else v = orc_g6(i, j, p[0], p[1], p[2], p[3], p[4], p[5]);
                pred[row + jj] = v;
                float r = v - xc[row + jj];
                rs += (double)r * (double)r;
            }
        }
        return rs;
    }
    memset(JtJ, 0, sizeof(double) * np * np);
    memset(Jtr, 0, sizeof(double) * np);
    for (int ii = 0; ii < n0; ++ii) {
        float i = (float)(ii - (n0 - 1) / 2);
        for (int jj = 0; jj < n1; ++jj) {
            float j = (float)(jj - (n1 - 1) / 2);
            int idx = ii * n1 + jj;
            float v = pred[idx];
            float r = v - xc[idx];
            rs += (double)r * (double)r;
            float J[6];
            if (np == 4) {
                float s = p[3];
                float di = (i - p[1]) / s, dj = (j - p[2]) / s;
                float r2 = di * di + dj * dj;
                J[0] = v / p[0];
                J[1] = 2.0f * v * di / s;
                J[2] = 2.0f * v * dj / s;
                J[3] = 2.0f * v * r2 / s;
            } else if (np == 5) {
                float sx = p[3], sy = p[4];
                float di = (i - p[1]) / sx, dj = (j - p[2]) / sy;
                J[0] = v / p[0];
                J[1] = 2.0f * v * di / sx;
                J[2] = 2.0f * v * dj / sy;
                J[3] = 2.0f * v * di * di / sx;
                J[4] = 2.0f * v * dj * dj / sy;
            } else {
                float sx = p[3], sy = p[4], sxy = p[5];
                float di = i - p[1], dj = j - p[2];
                J[0] = v / p[0];
                J[1] = v * (di / sx + dj * sxy);
                J[2] = v * (dj / sy + di * sxy);
                J[3] = 0.5f * v * di * di / (sx * sx);
                J[4] = 0.5f * v * dj * dj / (sy * sy);
                J[5] = -v * di * dj;
            }
            for (int a = 0; a < np; ++a) {
                Jtr[a] += (double)J[a] * (double)r;
                for (int b = 0; b <= a; ++b) JtJ[a * np + b] += (double)J[a] * (double)J[b];
            }
        }
    }
    for (int a = 0; a < np; ++a)
        for (int b = a + 1; b < np; ++b) JtJ[a * np + b] = JtJ[b * np + a];
    return rs;
}

/* (JtJ + lambda diag) delta = -Jtr by FP64 Cholesky (peak_locate_lm.c:342-393) */
static int orc_solve(const double *JtJ, const double *Jtr, float lambda, float *delta, int n)
{
    double A[36], L[36], y[6], d[6];
    memcpy(A, JtJ, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) A[i * n + i] *= (1.0 + (double)lambda);
    memset(L, 0, sizeof(L));
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = A[i * n + j];
            for (int k = 0; k < j; ++k) s -= L[i * n + k] * L[j * n + k];
            if (i == j) {
                if (s <= 0.0) return -1;
                L[i * n + j] = sqrt(s);
            } else {
                L[i * n + j] = s / L[j * n + j];
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        double s = -Jtr[i];
        for (int j = 0; j < i; ++j) s -= L[i * n + j] * y[j];
        y[i] = s / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < n; ++j) s -= L[j * n + i] * d[j];
        d[i] = s / L[i * n + i];
    }
    for (int i = 0; i < n; ++i) delta[i] = (float)d[i];
    return 0;
}

/* LM fit of the 4/5/6-DOF model to a 5x5 patch (peak_locate_lm.c:412-712).
 * Returns 0 trustworthy, -1 failed.  Writes pk[2] offsets and sig[3];
 * fitval (nullable) receives the model on the patch. */
static int orc_lm_fit(int np, const float *xc, float *pk, float *fitval, float *sig)
{
    const int N5[2] = {ORC_PK, ORC_PK};
    float p[6], pn[6], delta[6];
    double JtJ[36], Jtr[6];
    float pred[ORC_PK * ORC_PK];
    float lambda = 0.01f;
    double residual, new_residual;
    const int max_iter = 50;
    const double tol = 1e-6;
    int fit_ok = 0;
    float A, sx, sy;
    orc_threept(xc, N5, pk, &A, &sx, &sy);
    p[0] = A;
    p[1] = fminf(fmaxf(pk[0], -2.0f), 2.0f);
    p[2] = fminf(fmaxf(pk[1], -2.0f), 2.0f);
    if (np == 4) {
        float s = sqrtf(sx * sx + sy * sy);
        p[3] = fminf(fmaxf(s, 0.5f), 3.0f);
    } else if (np == 5) {
        p[3] = fminf(fmaxf(sx, 0.5f), 3.0f);
        p[4] = fminf(fmaxf(sy, 0.5f), 3.0f);
    } else {
        p[5] = 0.0f;
        p[3] = fminf(fmaxf(sx * sx, 0.25f), 9.0f);
        p[4] = fminf(fmaxf(sy * sy, 0.25f), 9.0f);
    }
    residual = orc_resid_jac(np, xc, p, NULL, NULL, 0, pred);
    orc_resid_jac(np, xc, p, JtJ, Jtr, 1, pred);
    for (int iter = 0; iter < max_iter; ++iter) {
        if (orc_solve(JtJ, Jtr, lambda, delta, np) != 0) break;
        for (int k = 0; k < np; ++k) pn[k] = p[k] + delta[k];
        pn[0] = fmaxf(pn[0], p[0] * 0.5f);
        pn[1] = fminf(fmaxf(pn[1], -2.5f), 2.5f);
        pn[2] = fminf(fmaxf(pn[2], -2.5f), 2.5f);
        if (np == 4) {
            pn[3] = fminf(fmaxf(pn[3], 0.25f), 4.0f);
        } else if (np == 5) {
            pn[3] = fminf(fmaxf(pn[3], 0.25f), 4.0f);
            pn[4] = fminf(fmaxf(pn[4], 0.25f), 4.0f);
        } else {
            pn[3] = fminf(fmaxf(pn[3], 0.1f), 16.0f);
            pn[4] = fminf(fmaxf(pn[4], 0.1f), 16.0f);
            float smax = 0.95f / sqrtf(pn[3] * pn[4]);
            pn[5] = fminf(fmaxf(pn[5], -smax), smax);
        }
        new_residual = orc_resid_jac(np, xc, pn, NULL, NULL, 0, pred);
        if (new_residual < residual) {
            for (int k = 0; k < np; ++k) p[k] = pn[k];
            double improvement = (residual - new_residual) / (residual + FLT_EPSILON);
            residual = new_residual;
            lambda *= 0.5f;
            orc_resid_jac(np, xc, p, JtJ, Jtr, 1, pred);
            if (improvement < tol) { fit_ok = 1; break; }
        } else {
            lambda *= 2.0f;
            if (lambda > 1e6f) break;
        }
    }
    if (!fit_ok)
        fit_ok = (residual <= 1e-3f * p[0] * p[0] * (float)(N5[0] * N5[1]));
    if (fit_ok) {
        if (np == 4) fit_ok = (p[3] > 0.25f && p[3] < 4.0f);
        else if (np == 5) fit_ok = (p[3] > 0.25f && p[3] < 4.0f && p[4] > 0.25f && p[4] < 4.0f);
        else fit_ok = (p[3] > 0.1f && p[3] < 16.0f && p[4] > 0.1f && p[4] < 16.0f);
    }
    pk[0] = p[1];
    pk[1] = p[2];
    if (np == 4) { sig[0] = p[3]; sig[1] = p[3]; sig[2] = 0.0f; }
    else if (np == 5) { sig[0] = p[3]; sig[1] = p[4]; sig[2] = 0.0f; }
    else { sig[0] = p[3]; sig[1] = p[4]; sig[2] = p[5]; }
    if (fitval) {
        for (int ii = 0; ii < ORC_PK; ++ii) {
            float i = (float)(ii - (ORC_PK - 1) / 2);
            for (int jj = 0; jj < ORC_PK; ++jj) {
                float j = (float)(jj - (ORC_PK - 1) / 2);
                float v;
                if (np == 4) v = orc_g4(i, j, p[0], p[1], p[2], p[3]);
                else if (np == 5) v = orc_g5(i, j, p[0], p[1], p[2], p[3], p[4]);
                else v = orc_g6(i, j, p[0], p[1], p[2], p[3], p[4], p[5]);
                fitval[ii * ORC_PK + jj] = v;
            }
        }
    }
    return fit_ok ? 0 : -1;
}

/* peak_locate_lm.c:727-744 */
float orc_peak_detectability(const float *xc, const int *N)
{
    float p1 = 0.0f, p2 = 0.0f;
    for (int i = N[0] / 4; i < N[0] * 3 / 4; ++i) {
        for (int j = N[1] / 4; j < N[1] * 3 / 4; ++j) {
            float v = xc[i * N[1] + j];
            if (v <= p2 || v <= 0.0f) continue;
            int is_max = 1;
            for (int di = -1; di <= 1 && is_max; ++di)
                for (int dj = -1; dj <= 1; ++dj)
                    if (xc[(i + di) * N[1] + (j + dj)] > v) { is_max = 0; break; }
            if (!is_max) continue;
            if (v > p1) { p2 = p1; p1 = v; }
            else p2 = v;
        }
    }
    return (p2 > 0.0f) ? p1 / p2 : 0.0f;
}

/* peak_locate_lm.c:767-918 */
void orc_lsqpeaklocate_lm(const float *xcorr, const int *N, float *peak_loc, int nPeaks,
                          int iFitType, float *std_dev, const float *w)
{
    float *copy = NULL;
    const float *src = xcorr;
    float sub[ORC_PK * ORC_PK], fitval[ORC_PK * ORC_PK];
    const int N5[2] = {ORC_PK, ORC_PK};
    if (nPeaks > 1) {
        copy = (float *)malloc(sizeof(float) * N[0] * N[1]);
        memcpy(copy, xcorr, sizeof(float) * N[0] * N[1]);
        src = copy;
    }
    for (int ip = 0; ip < nPeaks; ++ip) {
        int i0 = 0, j0 = 0;
        float h = 0;
        for (int i = N[0] / 4; i < N[0] * 3 / 4; ++i) {
            for (int j = N[1] / 4; j < N[1] * 3 / 4; ++j) {
                float v = src[i * N[1] + j];
                if (v <= h) continue;
                int is_max = 1;
                for (int di = -1; di <= 1 && is_max; ++di)
                    for (int dj = -1; dj <= 1; ++dj)
                        if (src[(i + di) * N[1] + (j + dj)] > v) { is_max = 0; break; }
                if (is_max) { h = v; i0 = i; j0 = j; }
            }
        }
        if (h <= 0 || i0 < (ORC_PK - 1) / 2 || i0 >= N[0] - (ORC_PK - 1) / 2 ||
            j0 < (ORC_PK - 1) / 2 || j0 >= N[1] - (ORC_PK - 1) / 2) {
            peak_loc[0 * nPeaks + ip] = NAN;
            peak_loc[1 * nPeaks + ip] = NAN;
            peak_loc[2 * nPeaks + ip] = 0;
            std_dev[0 * nPeaks + ip] = 0;
            std_dev[1 * nPeaks + ip] = 0;
            std_dev[2 * nPeaks + ip] = 0;
            continue;
        }
        for (int i = 0; i < ORC_PK; ++i)
            for (int j = 0; j < ORC_PK; ++j) {
                int idx = (i0 + i - (ORC_PK - 1) / 2) * N[1] + (j0 + j - (ORC_PK - 1) / 2);
                sub[i * ORC_PK + j] = w ? src[idx] * w[idx] : src[idx];
            }
        float pk[2], sig[3];
        int st = 0;
        if (ip == 0 && iFitType >= 4 && iFitType <= 6) {
            st = orc_lm_fit(iFitType, sub, pk, fitval, sig);
        } else {
            float A, sx, sy;
            orc_threept(sub, N5, pk, &A, &sx, &sy);
            sig[0] = sx; sig[1] = sy; sig[2] = 0.0f;
            for (int i = 0; i < ORC_PK; ++i) {
                float fi = (float)(i - (ORC_PK - 1) / 2);
                for (int j = 0; j < ORC_PK; ++j) {
                    float fj = (float)(j - (ORC_PK - 1) / 2);
                    fitval[i * ORC_PK + j] = orc_g5(fi, fj, A, pk[0], pk[1], sx, sy);
                }
            }
        }
        if (st != 0) {
            peak_loc[0 * nPeaks + ip] = NAN;
            peak_loc[1 * nPeaks + ip] = NAN;
            peak_loc[2 * nPeaks + ip] = 0;
            std_dev[0 * nPeaks + ip] = 0;
            std_dev[1 * nPeaks + ip] = 0;
            std_dev[2 * nPeaks + ip] = 0;
        } else {
            peak_loc[0 * nPeaks + ip] = pk[0] + i0;
            peak_loc[1 * nPeaks + ip] = pk[1] + j0;
            peak_loc[2 * nPeaks + ip] = h;
            std_dev[0 * nPeaks + ip] = sig[0];
            std_dev[1 * nPeaks + ip] = sig[1];
            std_dev[2 * nPeaks + ip] = sig[2];
        }
        if (nPeaks > 1) {
            for (int i = 0; i < ORC_PK; ++i)
                for (int j = 0; j < ORC_PK; ++j) {
                    int idx = (i0 + i - (ORC_PK - 1) / 2) * N[1] + (j0 + j - (ORC_PK - 1) / 2);
                    float f = fitval[i * ORC_PK + j];
                    if (w) f /= w[idx];
                    float d = copy[idx] - f;
                    /* MAX(0, d) of common.h: a NaN d stays NaN */
                    copy[idx] = (0 > d) ? 0 : d;
                }
        }
    }
    free(copy);
}

/* ------------------------------------------------------------------------ */
/* bulkxcorr2d (PIV_2d_cross_correlate.c:59-112, 203-397)                   */
/* ------------------------------------------------------------------------ */

unsigned char orc_bulkxcorr2d(const float *fImageA, const float *fImageB, const float *fMask,
                              const int *nImageSize, int N_images, const float *fWinCtrsX,
                              const float *fWinCtrsY, const int *nWindows, float *wA, int bEnsemble,
                              const float *wB, const int *nWindowSize, int nPeaks, int iPeakFinder,
                              float *pkX, float *pkY, float *pkH, float *fSx, float *fSy, float *fSxy,
                              float *pkRatio, float *planeOut)
{
    const int nWin = nWindows[0] * nWindows[1];
    const int H = nWindowSize[0], W = nWindowSize[1];
    const int numel = H * W;
    if (!orc_size_ok(H) || !orc_size_ok(W) || H > 128 || W > 128) return 4;
    float *cw = (float *)malloc(sizeof(float) * numel);
    unsigned e = orc_convolve(wB, wB, cw, nWindowSize);
    if (e) { free(cw); return (unsigned char)e; }
    for (int i = 0; i < numel; ++i) cw[i] = numel / cw[i];
    const long total = (long)N_images * nWin;
#pragma omp parallel
    {
        orc_xc xc;
        orc_xc_init(&xc, H, W);
        float *pA = (float *)malloc(sizeof(float) * numel);
        float *pB = (float *)malloc(sizeof(float) * numel);
        float *plane = (float *)malloc(sizeof(float) * numel);
        float *ploc = (float *)malloc(sizeof(float) * 3 * nPeaks);
        float *pstd = (float *)malloc(sizeof(float) * 3 * nPeaks);
#pragma omp for schedule(dynamic, 16)
        for (long idx = 0; idx < total; ++idx) {
            int n = (int)(idx / nWin);
            int w = (int)(idx % nWin);
            int ii = w % nWindows[1];
            int jj = w / nWindows[1];
            if (fMask[jj * nWindows[1] + ii] == 1) continue;
            int row_min = (int)floor(fWinCtrsY[jj] - ((float)H - 1.0) / 2.0 + 0.5);
            int col_min = (int)floor(fWinCtrsX[ii] - ((float)W - 1.0) / 2.0 + 0.5);
            if (row_min < 0 || col_min < 0 || row_min + H > nImageSize[0] || col_min + W > nImageSize[1])
                continue;
            const float *IA = fImageA + (size_t)n * nImageSize[0] * nImageSize[1];
            const float *IB = fImageB + (size_t)n * nImageSize[0] * nImageSize[1];
            for (int i = 0; i < H; ++i)
                for (int j = 0; j < W; ++j) {
                    size_t g = (size_t)(row_min + i) * nImageSize[1] + (col_min + j);
                    pA[i * W + j] = IA[g];
                    pB[i * W + j] = IB[g];
                }
            float mA = 0.0f, mB = 0.0f;
            for (int i = 0; i < numel; ++i) {
                pA[i] *= wA[i];
                pB[i] *= wB[i];
                mA += pA[i];
                mB += pB[i];
            }
            mA /= numel;
            mB /= numel;
            float eA = 0.0f, eB = 0.0f;
            if (!bEnsemble) {
                for (int i = 0; i < numel; ++i) {
                    pA[i] -= mA;
                    pB[i] -= mB;
                    eA += pA[i] * pA[i];
                    eB += pB[i] * pB[i];
                }
            } else {
                for (int i = 0; i < numel; ++i) {
                    eA += pA[i] * pA[i];
                    eB += pB[i] * pB[i];
                }
            }
            float enorm = 1.0f / sqrtf(eA * eB);
            orc_xc_forward(&xc, pB, 0);
            orc_xc_forward(&xc, pA, 1);
            orc_xc_emit_xcorr(&xc, plane);
            if (planeOut) memcpy(planeOut + (size_t)idx * numel, plane, sizeof(float) * numel);
            if (!bEnsemble) {
                orc_lsqpeaklocate_lm(plane, nWindowSize, ploc, nPeaks, iPeakFinder, pstd, cw);
                pkRatio[(size_t)n * nWin + w] = orc_peak_detectability(plane, nWindowSize);
                for (int p = 0; p < nPeaks; ++p) {
                    size_t o = (size_t)n * nPeaks * nWin + (size_t)p * nWin + w;
                    pkX[o] = ploc[1 * nPeaks + p] - W / 2.0f;
                    pkY[o] = ploc[0 * nPeaks + p] - H / 2.0f;
                    fSx[o] = pstd[0 * nPeaks + p];
                    fSy[o] = pstd[1 * nPeaks + p];
                    fSxy[o] = pstd[2 * nPeaks + p];
                    pkH[o] = ploc[2 * nPeaks + p] * enorm;
                }
            }
        }
        orc_xc_free(&xc);
        free(pA); free(pB); free(plane); free(ploc); free(pstd);
    }
    free(cw);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Ensemble accumulation (PIV_2d_cross_correlate.c:432-484, 502-833)        */
/* ------------------------------------------------------------------------ */

unsigned char orc_bulkxcorr2d_accumulate(const float *IAs, const float *IBs, const float *fMask,
                                         const int *nImageSize, int N_images, const float *ctrX,
                                         const float *ctrY, const int *nWindows, const float *wA,
                                         const float *wB, const int *nWindowSize,
                                         const int *nFit, float *out)
{
    const int nWin = nWindows[0] * nWindows[1];
    const int H = nWindowSize[0], W = nWindowSize[1], numel = H * W;
    if (!orc_size_ok(H) || !orc_size_ok(W) || H > 128 || W > 128) return 4;
    const int oh = nFit ? nFit[0] : H, ow = nFit ? nFit[1] : W;
    const int sy = (H - oh) / 2, sx = (W - ow) / 2;
    const size_t npix = (size_t)nImageSize[0] * nImageSize[1];
#pragma omp parallel
    {
        orc_xc xc;
        orc_xc_init(&xc, H, W);
        float *pA = (float *)malloc(sizeof(float) * numel);
        float *pB = (float *)malloc(sizeof(float) * numel);
        float *plane = (float *)malloc(sizeof(float) * numel);
#pragma omp for schedule(dynamic, 4)
        for (int w = 0; w < nWin; ++w) {
            int ii = w % nWindows[1], jj = w / nWindows[1];
            if (fMask[w] == 1) continue;
            int row_min = (int)floor(ctrY[jj] - ((float)H - 1.0f) / 2.0f + 0.5f);
            int col_min = (int)floor(ctrX[ii] - ((float)W - 1.0f) / 2.0f + 0.5f);
            if (row_min < 0 || col_min < 0 || row_min + H > nImageSize[0] || col_min + W > nImageSize[1])
                continue;
            float *o = out + (size_t)w * oh * ow;
            for (int n = 0; n < N_images; ++n) {
                const float *IA = IAs + (size_t)n * npix, *IB = IBs + (size_t)n * npix;
                for (int i = 0; i < H; ++i)
                    for (int j = 0; j < W; ++j) {
                        size_t g = (size_t)(row_min + i) * nImageSize[1] + (col_min + j);
                        pA[i * W + j] = IA[g] * wA[i * W + j];
                        pB[i * W + j] = IB[g] * wB[i * W + j];
                    }
                orc_xc_forward(&xc, pB, 0);
                orc_xc_forward(&xc, pA, 1);
                orc_xc_emit_xcorr(&xc, plane);
                for (int i = 0; i < oh; ++i)
                    for (int j = 0; j < ow; ++j) o[i * ow + j] += plane[(sy + i) * W + (sx + j)];
            }
        }
        orc_xc_free(&xc);
        free(pA); free(pB); free(plane);
    }
    return 0;
}

unsigned char orc_bulkxcorr2d_accumulate_triple(
    const float *IAs, const float *IBs, const float *fMask, const int *nImageSize, int N_images,
    const float *ctrX, const float *ctrY, const int *nWindows, const float *wA_AB, const float *wB_AB,
    const float *wAutoA, const float *wAutoB, const int *nWindowSize, const int *nFit,
    int bMeanSubtract, int bPerPairNorm, float *outAB, float *outAA, float *outBB)
{
    const int nWin = nWindows[0] * nWindows[1];
    const int H = nWindowSize[0], W = nWindowSize[1], numel = H * W;
    if (!orc_size_ok(H) || !orc_size_ok(W) || H > 128 || W > 128) return 4;
    const int oh = nFit ? nFit[0] : H, ow = nFit ? nFit[1] : W;
    const int sy = (H - oh) / 2, sx = (W - ow) / 2;
    const size_t npix = (size_t)nImageSize[0] * nImageSize[1];
    float sWB = 0.0f, sWAA = 0.0f, sWBB = 0.0f;
    if (bMeanSubtract)
        for (int i = 0; i < numel; ++i) { sWB += wB_AB[i]; sWAA += wAutoA[i]; sWBB += wAutoB[i]; }
#pragma omp parallel
    {
        orc_xc xc;
        orc_xc_init(&xc, H, W);
        float *rA = (float *)malloc(sizeof(float) * numel), *rB = (float *)malloc(sizeof(float) * numel);
        float *ab_b = (float *)malloc(sizeof(float) * numel), *ab_a = (float *)malloc(sizeof(float) * numel);
        float *aa = (float *)malloc(sizeof(float) * numel), *bb = (float *)malloc(sizeof(float) * numel);
        float *plane = (float *)malloc(sizeof(float) * numel);
#pragma omp for schedule(dynamic, 4)
        for (int w = 0; w < nWin; ++w) {
            int ii = w % nWindows[1], jj = w / nWindows[1];
            if (fMask[w] == 1) continue;
            int row_min = (int)floor(ctrY[jj] - ((float)H - 1.0f) / 2.0f + 0.5f);
            int col_min = (int)floor(ctrX[ii] - ((float)W - 1.0f) / 2.0f + 0.5f);
            if (row_min < 0 || col_min < 0 || row_min + H > nImageSize[0] || col_min + W > nImageSize[1])
                continue;
            float *oAB = outAB + (size_t)w * oh * ow;
            float *oAA = outAA + (size_t)w * oh * ow;
            float *oBB = outBB + (size_t)w * oh * ow;
            for (int n = 0; n < N_images; ++n) {
                const float *IA = IAs + (size_t)n * npix, *IB = IBs + (size_t)n * npix;
                for (int i = 0; i < H; ++i)
                    for (int j = 0; j < W; ++j) {
                        size_t g = (size_t)(row_min + i) * nImageSize[1] + (col_min + j);
                        rA[i * W + j] = IA[g];
                        rB[i * W + j] = IB[g];
                    }
                float mB = 0.0f, mAA = 0.0f, mBB = 0.0f;
                if (bMeanSubtract) {
                    float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
                    for (int i = 0; i < numel; ++i) {
                        s1 += rB[i] * wB_AB[i];
                        s2 += rA[i] * wAutoA[i];
                        s3 += rB[i] * wAutoB[i];
                    }
                    mB = (sWB > 0.0f) ? s1 / sWB : 0.0f;
                    mAA = (sWAA > 0.0f) ? s2 / sWAA : 0.0f;
                    mBB = (sWBB > 0.0f) ? s3 / sWBB : 0.0f;
                }
                float eAA = 0.0f, eBB = 0.0f;
                for (int i = 0; i < numel; ++i) {
                    ab_b[i] = (rB[i] - mB) * wB_AB[i];
                    ab_a[i] = (rA[i] - mAA) * wA_AB[i];
                    aa[i] = (rA[i] - mAA) * wAutoA[i];
                    bb[i] = (rB[i] - mBB) * wAutoB[i];
                    eAA += aa[i] * aa[i];
                    eBB += bb[i] * bb[i];
                }
                int skip = bPerPairNorm && (eAA <= 0.0f || eBB <= 0.0f);
                if (skip) continue;
                float sAB = bPerPairNorm ? 1.0f / sqrtf(eAA * eBB) : 1.0f;
                float sAA = bPerPairNorm ? 1.0f / eAA : 1.0f;
                float sBB = bPerPairNorm ? 1.0f / eBB : 1.0f;
                orc_xc_forward(&xc, ab_b, 0);
                orc_xc_forward(&xc, ab_a, 1);
                orc_xc_emit_xcorr(&xc, plane);
                for (int i = 0; i < oh; ++i)
                    for (int j = 0; j < ow; ++j) oAB[i * ow + j] += sAB * plane[(sy + i) * W + (sx + j)];
                orc_xc_forward(&xc, aa, 0);
                orc_xc_emit_power(&xc, 0, plane);
                for (int i = 0; i < oh; ++i)
                    for (int j = 0; j < ow; ++j) oAA[i * ow + j] += sAA * plane[(sy + i) * W + (sx + j)];
                orc_xc_forward(&xc, bb, 0);
                orc_xc_emit_power(&xc, 0, plane);
                for (int i = 0; i < oh; ++i)
                    for (int j = 0; j < ow; ++j) oBB[i * ow + j] += sBB * plane[(sy + i) * W + (sx + j)];
            }
        }
        orc_xc_free(&xc);
        free(rA); free(rB); free(ab_b); free(ab_a); free(aa); free(bb); free(plane);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Fused symmetric warp, scalar path (fused_warp.c:52-582, impl 0)          */
/* ------------------------------------------------------------------------ */

static inline int orc_clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static inline void orc_keys4(float d, float w[4])
{
    float s0 = 1.0f + d, q2 = 1.0f - d, s3 = 2.0f - d;
    w[1] = (1.25f * d - 2.25f) * d * d + 1.0f;
    w[2] = (1.25f * q2 - 2.25f) * q2 * q2 + 1.0f;
    w[0] = ((-0.75f * s0 + 3.75f) * s0 - 6.0f) * s0 + 3.0f;
    w[3] = ((-0.75f * s3 + 3.75f) * s3 - 6.0f) * s3 + 3.0f;
}

static inline float orc_bicubic(const float *img, float fy, float fx, int H, int W)
{
    float fyf = floorf(fy), fxf = floorf(fx);
    int iy = (int)fyf - 1, ix = (int)fxf - 1;
    float wy[4], wx[4], val = 0.0f;
    orc_keys4(fy - fyf, wy);
    orc_keys4(fx - fxf, wx);
    for (int m = 0; m < 4; m++) {
        int row = iy + m;
        if (row < 0 || row >= H) continue;
        for (int n = 0; n < 4; n++) {
            int col = ix + n;
            if (col < 0 || col >= W) continue;
            val += wy[m] * wx[n] * img[row * W + col];
        }
    }
    return val;
}

static inline float orc_lanczos_w(float t)
{
    float at = fabsf(t);
    if (at < 1e-6f) return 1.0f;
    if (at >= 3.0f) return 0.0f;
    float pt = (float)M_PI * at;
    float pta = pt / 3.0f;
    return (sinf(pt) / pt) * (sinf(pta) / pta);
}

#define ORC_LUT 4096

static inline float orc_lanczos(const float *img, float fy, float fx, int H, int W, const float (*lut)[6])
{
    float fyf = floorf(fy), fxf = floorf(fx);
    int iy = (int)fyf - 2, ix = (int)fxf - 2;
    float dys = fy - fyf, dxs = fx - fxf;
    float wy[6], wx[6], val = 0.0f;
    for (int pass = 0; pass < 2; ++pass) {
        float d = pass ? dxs : dys;
        float *wv = pass ? wx : wy;
        float pos = d * (float)ORC_LUT;
        int idx = (int)pos;
        float fr = pos - (float)idx;
        if (idx >= ORC_LUT) { idx = ORC_LUT - 1; fr = 0.0f; }
        for (int k = 0; k < 6; k++) wv[k] = lut[idx][k] + fr * (lut[idx + 1][k] - lut[idx][k]);
    }
    for (int m = 0; m < 6; m++) {
        int row = iy + m;
        if (row < 0 || row >= H) continue;
        for (int n = 0; n < 6; n++) {
            int col = ix + n;
            if (col < 0 || col >= W) continue;
            val += wy[m] * wx[n] * img[row * W + col];
        }
    }
    return val;
}

/* np.interp(p, ctrs, arange(nC)) with clamp (fused_warp.c:259-307) */
static void orc_pred_lut(float *lut, int N, const float *ctrs, int nC)
{
    if (nC == 1) { for (int p = 0; p < N; p++) lut[p] = 0.0f; return; }
    int seg = 0;
    for (int p = 0; p < N; p++) {
        float c = (float)p;
        if (c <= ctrs[0]) { lut[p] = 0.0f; continue; }
        if (c >= ctrs[nC - 1]) { lut[p] = (float)(nC - 1); continue; }
        while (seg < nC - 2 && ctrs[seg + 1] < c) seg++;
        float den = ctrs[seg + 1] - ctrs[seg];
        if (den < 1e-12f) { lut[p] = (float)seg; continue; }
        lut[p] = (float)seg + (c - ctrs[seg]) / den;
    }
}

int orc_fused_symmetric_warp_batch(const float *ia, const float *ib, float *oa, float *ob,
                                   const float *pdy, const float *pdx, int N, int H, int W,
                                   int nPY, int nPX, const float *cy, const float *cx,
                                   int interp_mode, int shared_predictor, int round_shifts)
{
    if (N <= 0 || H <= 0 || W <= 0 || nPY <= 0 || nPX <= 0) return 1;
    if (!ia || !ib || !oa || !ob || !pdy || !pdx || !cy || !cx) return 1;
    float *ly = (float *)malloc(sizeof(float) * H), *lx = (float *)malloc(sizeof(float) * W);
    orc_pred_lut(ly, H, cy, nPY);
    orc_pred_lut(lx, W, cx, nPX);
    float (*lut)[6] = NULL;
    if (interp_mode != 0) {
        lut = (float (*)[6])malloc(sizeof(float) * 6 * (ORC_LUT + 1));
        for (int i = 0; i <= ORC_LUT; i++) {
            float d = (float)i / (float)ORC_LUT;
            for (int k = 0; k < 6; k++) lut[i][k] = orc_lanczos_w(d - (float)(k - 2));
        }
    }
    const int stride = shared_predictor ? 0 : nPY * nPX;
    const long rows = (long)N * H;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < rows; t++) {
        int ni = (int)(t / H), i = (int)(t % H);
        const float *dy = pdy + (size_t)ni * stride, *dx = pdx + (size_t)ni * stride;
        const float *a = ia + (size_t)ni * H * W, *b = ib + (size_t)ni * H * W;
        float *oA = oa + (size_t)ni * H * W, *oB = ob + (size_t)ni * H * W;
        float fiy = ly[i], fiyf = floorf(fiy);
        int iyb = (int)fiyf - 1;
        float wy[4];
        orc_keys4(fiy - fiyf, wy);
        for (int j = 0; j < W; j++) {
            float fx = lx[j], fxf = floorf(fx);
            int ix = (int)fxf - 1;
            float wx[4], va = 0.0f, vb = 0.0f;
            orc_keys4(fx - fxf, wx);
            for (int m = 0; m < 4; m++) {
                int row = orc_clampi(iyb + m, 0, nPY - 1);
                for (int n = 0; n < 4; n++) {
                    int col = orc_clampi(ix + n, 0, nPX - 1);
                    float ww = wy[m] * wx[n];
                    int k = row * nPX + col;
                    va += ww * dy[k];
                    vb += ww * dx[k];
                }
            }
            float hy = 0.5f * va, hx = 0.5f * vb;
            if (round_shifts) { hy = roundf(hy); hx = roundf(hx); }
            size_t o = (size_t)i * W + j;
            if (interp_mode == 0) {
                oA[o] = orc_bicubic(a, (float)i - hy, (float)j - hx, H, W);
                oB[o] = orc_bicubic(b, (float)i + hy, (float)j + hx, H, W);
            } else {
                oA[o] = orc_lanczos(a, (float)i - hy, (float)j - hx, H, W, (const float (*)[6])lut);
                oB[o] = orc_lanczos(b, (float)i + hy, (float)j + hx, H, W, (const float (*)[6])lut);
            }
        }
    }
    free(ly); free(lx); free(lut);
    return 0;
}

int orc_fused_symmetric_warp(const float *ia, const float *ib, float *oa, float *ob,
                             const float *pdy, const float *pdx, int H, int W, int nPY, int nPX,
                             const float *cy, const float *cx, int interp_mode, int round_shifts)
{
    return orc_fused_symmetric_warp_batch(ia, ib, oa, ob, pdy, pdx, 1, H, W, nPY, nPX, cy, cx,
                                          interp_mode, 1, round_shifts);
}

/* Thread count used by the OpenMP loops (for reporting). */
int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
