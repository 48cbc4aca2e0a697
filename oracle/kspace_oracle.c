/*
 * kspace_oracle.c -- CPU restatement of the k-space transfer-function LM fit.
 *
 * TEST INFRASTRUCTURE ONLY.  The parity tests and the CPU-baseline leg of the
 * benchmark call it as the checker; nothing in the shipped library links it.
 *
 * Follows pivtools_cli/lib/kspace_lm_fit.c (paths relative to the reference
 * repository):
 *   constants                         :90-123
 *   natural-order k grids, tail set   :213-227
 *   centred 2D FFT (ifftshift load)   :257-289
 *   mu seed (3-point log-Gaussian)    :291-321
 *   gain seed (ring-1 median)         :330-346
 *   floor regressor, coloured N0 seed :366-409
 *   cost / cost + J^T J               :411-532
 *   Cholesky step                     :535-562
 *   per-window LM                     :574-666
 *   batch entry, gates, output rows   :669-886
 *   debug prep / FFT entries          :903-947
 *
 * The transform is a direct FP64 DFT per axis written here (rows, then
 * columns, as the reference orders them); the reference's generated double
 * codelets are not reproduced.  Everything else keeps the reference's
 * arithmetic in FP64 with libm.
 *
 * Parity pins (tests/test_kspace_oracle.py):
 *   centred FFT       numpy fftshift(fft2(ifftshift(R)))
 *   model + LM        planes built in closed form from a known transfer
 *                     function (mu, Sigma, gain) are recovered to 1e-9
 *   seeds             3-point log-Gaussian exact on sampled Gaussians, ring
 *                     median and tail median by numpy.median
 *   status contract   masked / low-SNR / too-few-bins / big-displacement cases
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KS_MAX_ITER 250
#define KS_XTOL 1e-8
#define KS_FTOL 1e-8
#define KS_MIN_VALID 10
#define KS_MAX_DISP_FRAC 0.75
#define KS_COST_ACCEPT 1.0
#define KS_EXP_CLIP 300.0
#define KS_LAM0 1e-3
#define KS_LAM_MIN 1e-12
#define KS_LAM_STUCK 1e12
#define KS_GAIN_LO 1e-3
#define KS_GAIN_HI 1e3
#define KS_SEED_SXX 1.0
#define KS_SEED_SYY 0.2
#define KS_N0_HI 10.0
#define KS_TAIL_KR 0.35
#define KS_N0_CLIP_LO 1e-3
#define KS_N0_CLIP_HI 10.0
#define KS_NPAR_MAX 9

enum { KS_MASKED = -1, KS_OK = 0, KS_NO_CONVERGE = 1, KS_LOW_SNR = 2, KS_BIG_DISP = 3 };

static const double KS_2PI = 6.283185307179586476925286766559;
static const double KS_2PI2 = 2.0 * 9.869604401089358618834490999876;

static int ks_axis_ok(int n)
{
    switch (n) {
    case 8: case 12: case 16: case 24: case 32: case 48: case 64: case 96: case 128: return 1;
    default: return 0;
    }
}

/* ------------------------------------------------------------------------ */
/* per-window working set                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int H, W, P;
    double *twH_r, *twH_i, *twW_r, *twW_i;  /* exp(-2 pi i m / n) */
    double *re[3], *im[3];                  /* spectra, natural order */
    double *lr, *li, *outr, *outi;          /* one line in/out */
    double *fr, *tr, *ti, *wt, *dv, *pn, *tail;
    double *kx, *ky;
    int *tail_idx, n_tail;
} ks_ws;

static void ks_ws_free(ks_ws *s)
{
    free(s->twH_r); free(s->twH_i); free(s->twW_r); free(s->twW_i);
    for (int i = 0; i < 3; ++i) { free(s->re[i]); free(s->im[i]); }
    free(s->lr); free(s->li); free(s->outr); free(s->outi);
    free(s->fr); free(s->tr); free(s->ti); free(s->wt); free(s->dv); free(s->pn);
    free(s->tail); free(s->kx); free(s->ky); free(s->tail_idx);
}

static int ks_ws_init(ks_ws *s, int H, int W)
{
    memset(s, 0, sizeof(*s));
    s->H = H; s->W = W; s->P = H * W;
    const size_t P = (size_t)s->P;
    const int L = H > W ? H : W;
    s->twH_r = malloc(sizeof(double) * H); s->twH_i = malloc(sizeof(double) * H);
    s->twW_r = malloc(sizeof(double) * W); s->twW_i = malloc(sizeof(double) * W);
    for (int i = 0; i < 3; ++i) { s->re[i] = malloc(sizeof(double) * P); s->im[i] = malloc(sizeof(double) * P); }
    s->lr = malloc(sizeof(double) * L); s->li = malloc(sizeof(double) * L);
    s->outr = malloc(sizeof(double) * L); s->outi = malloc(sizeof(double) * L);
    s->fr = malloc(sizeof(double) * P); s->tr = malloc(sizeof(double) * P); s->ti = malloc(sizeof(double) * P);
    s->wt = malloc(sizeof(double) * P); s->dv = malloc(sizeof(double) * P); s->pn = malloc(sizeof(double) * P);
    s->tail = malloc(sizeof(double) * P); s->tail_idx = malloc(sizeof(int) * P);
    s->kx = malloc(sizeof(double) * W); s->ky = malloc(sizeof(double) * H);
    if (!s->twH_r || !s->twH_i || !s->twW_r || !s->twW_i || !s->re[2] || !s->im[2] || !s->lr || !s->li ||
        !s->outr || !s->outi || !s->fr || !s->tr || !s->ti || !s->wt || !s->dv || !s->pn || !s->tail ||
        !s->tail_idx || !s->kx || !s->ky || !s->re[0] || !s->re[1] || !s->im[0] || !s->im[1])
        return -1;
    for (int m = 0; m < H; ++m) {
        s->twH_r[m] = cos(KS_2PI * m / H);
        s->twH_i[m] = -sin(KS_2PI * m / H);
    }
    for (int m = 0; m < W; ++m) {
        s->twW_r[m] = cos(KS_2PI * m / W);
        s->twW_i[m] = -sin(KS_2PI * m / W);
    }
    /* numpy fftfreq in natural order */
    for (int k = 0; k < W; ++k) s->kx[k] = (k < (W + 1) / 2 ? (double)k : (double)(k - W)) / (double)W;
    for (int k = 0; k < H; ++k) s->ky[k] = (k < (H + 1) / 2 ? (double)k : (double)(k - H)) / (double)H;
    s->n_tail = 0;
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c)
            if (s->ky[r] * s->ky[r] + s->kx[c] * s->kx[c] >= KS_TAIL_KR * KS_TAIL_KR)
                s->tail_idx[s->n_tail++] = r * W + c;
    return 0;
}

/* direct DFT of one line of length n with the table exp(-2 pi i m/n) */
static void ks_dft(int n, const double *twr, const double *twi, const double *xr, const double *xi,
                   double *yr, double *yi)
{
    for (int k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < n; ++j) {
            const int m = (int)(((long)k * j) % n);
            sr += xr[j] * twr[m] - xi[j] * twi[m];
            si += xr[j] * twi[m] + xi[j] * twr[m];
        }
        yr[k] = sr;
        yi[k] = si;
    }
}

/* fftshift(fft2(ifftshift(R))) in natural order: load with the ifftshift
 * index map, rows then columns (kspace_lm_fit.c:257-289). */
static void ks_fft2_centred(ks_ws *s, const double *plane, int slot)
{
    const int H = s->H, W = s->W;
    double *re = s->re[slot], *im = s->im[slot];
    for (int r = 0; r < H; ++r) {
        const double *src = plane + (size_t)((r + H / 2) % H) * W;
        for (int c = 0; c < W; ++c) {
            s->lr[c] = src[(c + W / 2) % W];
            s->li[c] = 0.0;
        }
        ks_dft(W, s->twW_r, s->twW_i, s->lr, s->li, re + (size_t)r * W, im + (size_t)r * W);
    }
    for (int c = 0; c < W; ++c) {
        for (int r = 0; r < H; ++r) {
            s->lr[r] = re[(size_t)r * W + c];
            s->li[r] = im[(size_t)r * W + c];
        }
        ks_dft(H, s->twH_r, s->twH_i, s->lr, s->li, s->outr, s->outi);
        for (int r = 0; r < H; ++r) {
            re[(size_t)r * W + c] = s->outr[r];
            im[(size_t)r * W + c] = s->outi[r];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* seeds                                                                    */
/* ------------------------------------------------------------------------ */
static double ks_log3(double lo, double ce, double hi)
{
    if (lo <= 0.0 || ce <= 0.0 || hi <= 0.0) return 0.0;
    const double a = log(lo), b = log(ce), c = log(hi);
    const double den = 2.0 * (a - 2.0 * b + c);
    if (fabs(den) <= 1e-12) return 0.0;
    return (a - c) / den;
}

/* first-occurrence argmax of the raw R_AB plane (strict '>' from index 0),
 * 3-point log-Gaussian refinement on each axis, offsets from (H/2, W/2) */
static void ks_mu_seed(const double *ab, int H, int W, double *mx, double *my, double *amp)
{
    int best = 0;
    double v = ab[0];
    for (int i = 1; i < H * W; ++i)
        if (ab[i] > v) { v = ab[i]; best = i; }
    const int py = best / W, px = best % W;
    double fx = 0.0, fy = 0.0;
    if (px > 0 && px < W - 1) fx = ks_log3(ab[py * W + px - 1], ab[py * W + px], ab[py * W + px + 1]);
    if (py > 0 && py < H - 1) fy = ks_log3(ab[(py - 1) * W + px], ab[py * W + px], ab[(py + 1) * W + px]);
    *mx = (double)px + fx - (double)(W / 2);
    *my = (double)py + fy - (double)(H / 2);
    *amp = v;
}

static int ks_cmp(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static double ks_median_sorted(const double *v, int n)
{
    return (n & 1) ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

static double ks_gain_seed(const ks_ws *s)
{
    double v[8];
    int n = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            if (!dy && !dx) continue;
            const size_t i = (size_t)((dy + s->H) % s->H) * s->W + (size_t)((dx + s->W) % s->W);
            v[n++] = hypot(s->tr[i], s->ti[i]);
        }
    qsort(v, 8, sizeof(double), ks_cmp);
    double g = ks_median_sorted(v, 8);
    return g < 1e-2 ? 1e-2 : (g > 10.0 ? 10.0 : g);
}

/* D = P / max(F_ref, 1e-30) (coloured, P given centred) or 1 / max(F_ref, 1e-30) */
static void ks_regressor(ks_ws *s, const double *pc)
{
    const int H = s->H, W = s->W;
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            const size_t i = (size_t)r * W + c;
            const double f = s->fr[i] > 1e-30 ? s->fr[i] : 1e-30;
            if (pc) {
                const double p = pc[(size_t)((r + H / 2) % H) * W + (c + W / 2) % W];
                s->pn[i] = p;
                s->dv[i] = p / f;
            } else {
                s->dv[i] = 1.0 / f;
            }
        }
}

static double ks_n0_seed(ks_ws *s)
{
    for (int t = 0; t < s->n_tail; ++t) {
        const int i = s->tail_idx[t];
        s->tail[t] = s->fr[i] / (s->pn[i] > 1e-30 ? s->pn[i] : 1e-30);
    }
    qsort(s->tail, (size_t)s->n_tail, sizeof(double), ks_cmp);
    double m = ks_median_sorted(s->tail, s->n_tail);
    return m < KS_N0_CLIP_LO ? KS_N0_CLIP_LO : (m > KS_N0_CLIP_HI ? KS_N0_CLIP_HI : m);
}

/* ------------------------------------------------------------------------ */
/* model: T(k) = g exp(-2 pi^2 k^T S k - b4x kx^4 - b4y ky^4) (1 - N0 D)    */
/*              exp(-2 pi i k.mu)                                            */
/* ------------------------------------------------------------------------ */
typedef struct { int K, kx4, ky4; } ks_shape;

/* residual (and, when A != NULL, J^T J upper+lower and J^T r) over all
 * weighted bins in row-major order */
static double ks_eval(const ks_ws *s, ks_shape sh, const double *x, double *A, double *gv)
{
    const double mux = x[0], muy = x[1], sxx = x[2], syy = x[3], sxy = x[4], g = x[5], n0 = x[6];
    const double b4x = sh.kx4 ? x[7] : 0.0;
    const double b4y = sh.ky4 ? x[7 + sh.kx4] : 0.0;
    const int K = sh.K;
    if (A) {
        memset(A, 0, sizeof(double) * K * K);
        memset(gv, 0, sizeof(double) * K);
    }
    double cost = 0.0;
    for (int r = 0; r < s->H; ++r) {
        const double ky = s->ky[r], ky2 = ky * ky;
        for (int c = 0; c < s->W; ++c) {
            const size_t i = (size_t)r * s->W + c;
            const double w = s->wt[i];
            if (w == 0.0) continue;
            const double kx = s->kx[c], kx2 = kx * kx, kxy = kx * ky;
            double e = -KS_2PI2 * (sxx * kx2 + 2.0 * sxy * kxy + syy * ky2);
            if (sh.kx4) e -= b4x * kx2 * kx2;
            if (sh.ky4) e -= b4y * ky2 * ky2;
            if (e > KS_EXP_CLIP) e = KS_EXP_CLIP;
            const double dec = exp(e);
            const double d = s->dv[i];
            const double att = 1.0 - n0 * d;
            const double ph = -KS_2PI * (kx * mux + ky * muy);
            const double cp = cos(ph), sp = sin(ph);
            const double gd = g * dec * att;
            const double rr = w * (s->tr[i] - gd * cp);
            const double ri = w * (s->ti[i] - gd * sp);
            cost += rr * rr + ri * ri;
            if (!A) continue;
            double jr[KS_NPAR_MAX], ji[KS_NPAR_MAX];
            const double px = -KS_2PI * kx, py = -KS_2PI * ky;
            jr[0] = -w * gd * (-sp) * px;  ji[0] = -w * gd * cp * px;
            jr[1] = -w * gd * (-sp) * py;  ji[1] = -w * gd * cp * py;
            const double d2 = gd * (-KS_2PI2 * kx2), d3 = gd * (-KS_2PI2 * ky2), d4 = gd * (-KS_2PI2 * 2.0 * kxy);
            jr[2] = -w * d2 * cp;  ji[2] = -w * d2 * sp;
            jr[3] = -w * d3 * cp;  ji[3] = -w * d3 * sp;
            jr[4] = -w * d4 * cp;  ji[4] = -w * d4 * sp;
            jr[5] = -w * dec * att * cp;  ji[5] = -w * dec * att * sp;
            const double dn = g * dec * (-d);
            jr[6] = -w * dn * cp;  ji[6] = -w * dn * sp;
            int q = 7;
            if (sh.kx4) {
                const double db = gd * (-(kx2 * kx2));
                jr[q] = -w * db * cp; ji[q] = -w * db * sp; ++q;
            }
            if (sh.ky4) {
                const double db = gd * (-(ky2 * ky2));
                jr[q] = -w * db * cp; ji[q] = -w * db * sp;
            }
            for (int a = 0; a < K; ++a) {
                gv[a] += jr[a] * rr + ji[a] * ri;
                for (int b = a; b < K; ++b) A[a * K + b] += jr[a] * jr[b] + ji[a] * ji[b];
            }
        }
    }
    if (A)
        for (int a = 1; a < K; ++a)
            for (int b = 0; b < a; ++b) A[a * K + b] = A[b * K + a];
    return cost;
}

/* Cholesky solve M d = b; -1 when M is not positive definite */
static int ks_chol(int K, const double *M, const double *b, double *d)
{
    double L[KS_NPAR_MAX * KS_NPAR_MAX], y[KS_NPAR_MAX];
    for (int i = 0; i < K; ++i)
        for (int j = 0; j <= i; ++j) {
            double t = M[i * K + j];
            for (int k = 0; k < j; ++k) t -= L[i * K + k] * L[j * K + k];
            if (i == j) {
                if (t <= 0.0 || !isfinite(t)) return -1;
                L[i * K + i] = sqrt(t);
            } else {
                L[i * K + j] = t / L[j * K + j];
            }
        }
    for (int i = 0; i < K; ++i) {
        double t = b[i];
        for (int k = 0; k < i; ++k) t -= L[i * K + k] * y[k];
        y[i] = t / L[i * K + i];
    }
    for (int i = K - 1; i >= 0; --i) {
        double t = y[i];
        for (int k = i + 1; k < K; ++k) t -= L[k * K + i] * d[k];
        d[i] = t / L[i * K + i];
    }
    return 0;
}

static double ks_clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* damped Gauss-Newton with lambda /3 on accept, x10 on reject; returns the
 * number of iterations, sets *conv and the final cost */
static int ks_lm(const ks_ws *s, ks_shape sh, const double *seed, const double *lo, const double *hi,
                 double *x, double *cost_out, int *conv)
{
    const int K = sh.K;
    for (int i = 0; i < K; ++i) x[i] = ks_clampd(seed[i], lo[i], hi[i]);
    double cost = ks_eval(s, sh, x, NULL, NULL);
    int it = 0;
    *conv = 0;
    if (!isfinite(cost)) {
        *cost_out = cost;
        return 0;
    }
    double lam = KS_LAM0, A[KS_NPAR_MAX * KS_NPAR_MAX], gv[KS_NPAR_MAX];
    int fresh = 0;
    while (it < KS_MAX_ITER) {
        if (!fresh) {
            ks_eval(s, sh, x, A, gv);
            fresh = 1;
        }
        double M[KS_NPAR_MAX * KS_NPAR_MAX], rhs[KS_NPAR_MAX], d[KS_NPAR_MAX], xn[KS_NPAR_MAX];
        memcpy(M, A, sizeof(double) * K * K);
        for (int i = 0; i < K; ++i) {
            const double dg = A[i * K + i] < 1e-30 ? 1e-30 : A[i * K + i];
            M[i * K + i] += lam * dg;
            rhs[i] = -gv[i];
        }
        int take = 0;
        double cn = INFINITY;
        if (ks_chol(K, M, rhs, d) == 0) {
            int fin = 1;
            for (int i = 0; i < K; ++i) {
                xn[i] = ks_clampd(x[i] + d[i], lo[i], hi[i]);
                if (!isfinite(xn[i])) fin = 0;
            }
            if (fin) {
                cn = ks_eval(s, sh, xn, NULL, NULL);
                take = isfinite(cn) && cn <= cost;
            }
        }
        ++it;
        if (take) {
            int small = 1;
            for (int i = 0; i < K; ++i)
                if (fabs(xn[i] - x[i]) > KS_XTOL * (KS_XTOL + fabs(xn[i]))) { small = 0; break; }
            const int flat = (cost - cn) <= KS_FTOL * (cost > 1e-300 ? cost : 1e-300);
            memcpy(x, xn, sizeof(double) * K);
            cost = cn;
            lam = lam / 3.0 < KS_LAM_MIN ? KS_LAM_MIN : lam / 3.0;
            fresh = 0;
            if (small || flat) {
                *conv = 1;
                break;
            }
        } else {
            lam *= 10.0;
            if (lam > KS_LAM_STUCK) break;
        }
    }
    *cost_out = cost;
    return it;
}

/* transfer ratio, weights and valid count from the three spectra */
static int ks_transfer(ks_ws *s)
{
    double fmax = 0.0;
    for (int i = 0; i < s->P; ++i) {
        const double f = sqrt(hypot(s->re[0][i], s->im[0][i]) * hypot(s->re[1][i], s->im[1][i]));
        s->fr[i] = f;
        const double inv = 1.0 / (f > 1e-30 ? f : 1e-30);
        s->tr[i] = s->re[2][i] * inv;
        s->ti[i] = s->im[2][i] * inv;
        if (f > fmax) fmax = f;
    }
    if (fmax < 1e-30) fmax = 1e-30;
    int nv = 0;
    for (int i = 0; i < s->P; ++i) {
        const int ok = s->fr[i] > 1e-12 && i != 0;
        s->wt[i] = ok ? s->fr[i] / fmax : 0.0;
        nv += ok;
    }
    return nv;
}

int orc_kspace_lm_fit_batch(const double *R_AA, const double *R_BB, const double *R_AB,
                            const unsigned char *mask, const double *P_win, int nwin, int H, int W,
                            int kx4, int ky4, int n_threads, double *gauss, int *status, double *init,
                            double *dgain, double *dn0, double *db4x, double *db4y, double *dcpp,
                            int *dnvalid, int *diter, unsigned char *dconv)
{
    if (!ks_axis_ok(H) || !ks_axis_ok(W)) return -1;
    const int P = H * W;
    const double cxp = W / 2.0 + 1.0, cyp = H / 2.0 + 1.0;
    const ks_shape sh = {7 + (kx4 != 0) + (ky4 != 0), kx4 != 0, ky4 != 0};
    double lo[KS_NPAR_MAX], hi[KS_NPAR_MAX];
    for (int i = 0; i < KS_NPAR_MAX; ++i) { lo[i] = -INFINITY; hi[i] = INFINITY; }
    lo[2] = 0.0; lo[3] = 0.0; lo[5] = KS_GAIN_LO; hi[5] = KS_GAIN_HI; lo[6] = 0.0;
    if (P_win) hi[6] = KS_N0_HI;
    if (P_win) {
        ks_ws probe;
        const int bad = ks_ws_init(&probe, H, W);
        const int nt = probe.n_tail;
        ks_ws_free(&probe);
        if (bad) return -2;
        if (nt == 0) return -3;
    }
    int nproc = 0;
    for (int w = 0; w < nwin; ++w) nproc += !mask[w];
    for (int w = 0; w < nwin; ++w) {
        double *gf = gauss + (size_t)w * 16, *ig = init + (size_t)w * 16;
        for (int c = 0; c < 16; ++c) gf[c] = ig[c] = 0.0;
        gf[6] = gf[7] = gf[8] = NAN;
        gf[12] = gf[14] = cxp;
        gf[13] = gf[15] = cyp;
        if (nproc == 0 || !mask[w]) {
            ig[6] = ig[7] = ig[8] = NAN;
            ig[12] = ig[14] = cxp;
            ig[13] = ig[15] = cyp;
        }
        status[w] = KS_MASKED;
        dgain[w] = dn0[w] = db4x[w] = db4y[w] = dcpp[w] = NAN;
        dnvalid[w] = diter[w] = 0;
        dconv[w] = 0;
    }
    int failed = 0;
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : omp_get_max_threads())
    {
        ks_ws s;
        const int ok = ks_ws_init(&s, H, W) == 0;
        if (!ok) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(dynamic, 16)
        for (int w = 0; w < nwin; ++w) {
            if (!ok || mask[w]) continue;
            const double *aa = R_AA + (size_t)w * P, *bb = R_BB + (size_t)w * P, *ab = R_AB + (size_t)w * P;
            double *gf = gauss + (size_t)w * 16, *ig = init + (size_t)w * 16;
            const double ampA = aa[(H / 2) * W + W / 2], ampB = bb[(H / 2) * W + W / 2];
            double seed[KS_NPAR_MAX] = {0}, ampAB;
            ks_mu_seed(ab, H, W, &seed[0], &seed[1], &ampAB);
            gf[0] = ig[0] = ampA;
            gf[1] = ig[1] = ampB;
            gf[2] = ig[2] = ampAB;
            ks_fft2_centred(&s, aa, 0);
            ks_fft2_centred(&s, bb, 1);
            ks_fft2_centred(&s, ab, 2);
            const int nv = ks_transfer(&s);
            dnvalid[w] = nv;
            seed[2] = KS_SEED_SXX;
            seed[3] = KS_SEED_SYY;
            ig[9] = seed[2];
            ig[10] = seed[3];
            ig[14] = cxp + seed[0];
            ig[15] = cyp + seed[1];
            if (!(ampA >= 1e-12 && ampB >= 1e-12)) { status[w] = KS_LOW_SNR; continue; }
            if (nv < KS_MIN_VALID) { status[w] = KS_NO_CONVERGE; continue; }
            seed[5] = ks_gain_seed(&s);
            ks_regressor(&s, P_win ? P_win + (size_t)w * P : NULL);
            if (P_win) seed[6] = ks_n0_seed(&s);
            double x[KS_NPAR_MAX], cost;
            int conv;
            const int it = ks_lm(&s, sh, seed, lo, hi, x, &cost, &conv);
            diter[w] = it;
            dconv[w] = (unsigned char)conv;
            const double cpp = cost / (double)nv;
            dcpp[w] = cpp;
            if (!(conv || (it >= KS_MAX_ITER && cpp < KS_COST_ACCEPT))) { status[w] = KS_NO_CONVERGE; continue; }
            if (fabs(x[0]) > KS_MAX_DISP_FRAC * W || fabs(x[1]) > KS_MAX_DISP_FRAC * H) {
                status[w] = KS_BIG_DISP;
                continue;
            }
            status[w] = KS_OK;
            gf[9] = x[2] > 0.0 ? x[2] : 0.0;
            gf[10] = x[3] > 0.0 ? x[3] : 0.0;
            gf[11] = x[4];
            gf[14] = cxp + x[0];
            gf[15] = cyp + x[1];
            dgain[w] = x[5];
            dn0[w] = x[6];
            if (kx4) db4x[w] = x[7];
            if (ky4) db4y[w] = x[7 + (kx4 != 0)];
        }
        ks_ws_free(&s);
    }
    return failed ? -2 : 0;
}

int orc_kspace_fft2_centered(const double *plane, int H, int W, double *out_re, double *out_im)
{
    if (!ks_axis_ok(H) || !ks_axis_ok(W)) return -1;
    ks_ws s;
    if (ks_ws_init(&s, H, W)) {
        ks_ws_free(&s);
        return -2;
    }
    ks_fft2_centred(&s, plane, 0);
    memcpy(out_re, s.re[0], sizeof(double) * s.P);
    memcpy(out_im, s.im[0], sizeof(double) * s.P);
    ks_ws_free(&s);
    return 0;
}

int orc_kspace_debug_prep(const double *R_AA, const double *R_BB, const double *R_AB, const double *Pc,
                          int H, int W, double *out_dv, double *out_n0)
{
    if (!ks_axis_ok(H) || !ks_axis_ok(W)) return -1;
    ks_ws s;
    if (ks_ws_init(&s, H, W)) {
        ks_ws_free(&s);
        return -2;
    }
    if (Pc && s.n_tail == 0) {
        ks_ws_free(&s);
        return -3;
    }
    ks_fft2_centred(&s, R_AA, 0);
    ks_fft2_centred(&s, R_BB, 1);
    ks_fft2_centred(&s, R_AB, 2);
    ks_transfer(&s);
    ks_regressor(&s, Pc);
    *out_n0 = Pc ? ks_n0_seed(&s) : 0.0;
    memcpy(out_dv, s.dv, sizeof(double) * s.P);
    ks_ws_free(&s);
    return 0;
}
