// xcorr.hip -- batched window cross-correlation with peak fit (gfx950).
//
// One kernel replaces one call of the reference's bulkxcorr2d
// (pivtools_cli/lib/PIV_2d_cross_correlate.c:203-397): for every
// (image, window) it extracts the window, applies the weight, removes the
// weighted mean, forms the correlation plane
//     plane = fftshift( IFFT( FFT(B) .* conj(FFT(A)) ) ) / numel
// (codelet_fft.c:290-366, slot0 = B, slot1 = A), finds the peak on the raw
// plane, fits it on the 5x5 patch times the plane weight, and writes the
// seven per-window outputs (PIV_2d_cross_correlate.c:59-112).
//
// Layout and mapping.  A workgroup holds WPB windows; window w owns T =
// max(NY, NX) consecutive lanes and a padded complex NY x (NX+1) LDS tile.
// The 2D transform is separable and every 1D transform lives in one lane's
// registers (fft_reg.h):
//   1. lane x loads column x of A and B straight from HBM (coalesced across
//      lanes), packs z = a + i*b, and does the forward FFT down the column;
//   2. lane y does the forward FFT along row y of the tile;
//   3. lane kx <= NX/2 separates FA/FB from Z(k) and Z(-k), forms
//      C = FB * conj(FA) for its column, and does the inverse column FFT;
//   4. lane y < NY/2 rebuilds the Hermitian rows y and y+NY/2 into one
//      complex row, does one inverse FFT for both, scales by 1/numel and
//      stores the two real rows fftshifted: the plane never leaves LDS;
//   5. all lanes scan the quarter-region candidates (3x3 local maxima) and
//      the window's leader lane runs the gate and the sub-pixel fit.
// Per window this is 1.5 complex 2D FFTs of work, the same as two r2c and
// one c2r.  Blocks are renumbered so that each XCD (L2 domain) works on a
// contiguous run of windows, which keeps overlapping windows in one L2.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_reg.h"
#include "peakfit.h"
#include "pivamd_dev.h"

namespace pivamd {

struct XcorrArgs {
    const float *imgA, *imgB, *mask, *ctrX, *ctrY, *wA, *wB, *pw;
    float *pkX, *pkY, *pkH, *sx, *sy, *sxy, *ratio, *planes;
    float *rec, *work_planes;   // workspace: per-window peak records, multi-peak planes
    long long total;        // N_images * nWin
    int H, W, nWy, nWx, nPeaks, finder, impl, ensemble, nblocks;
};

constexpr int kMaxLds = 150 * 1024;

// Launch geometry of one square window size N: R lanes share each line
// (row or column) transform, T = N*R lanes per window, WPB windows per
// workgroup as LDS allows (one complex N x (N+1) tile per window).
template <int N>
struct XcCfg {
    static constexpr int R = N >= 96 ? 4 : (N >= 48 ? 2 : 1);
    static constexpr int N1 = N / R;
    static constexpr int T = N * R;
    static constexpr int PX = N + 1;
    static constexpr int tile = N * PX;
    static constexpr int lds_win = 2 * tile * 4 + 5 * T * 4;
    static constexpr int wpb_thr = (256 / T) > 0 ? (256 / T) : 1;
    static constexpr int wpb_lds = kMaxLds / lds_win;
    static constexpr int WPB = wpb_thr < wpb_lds ? wpb_thr : wpb_lds;
    static constexpr int BLOCK = WPB * T;
};

// Sum of (a, b) over the T lanes of one window; every lane gets the same
// bits.  T | 64: xor butterfly inside the wave.  64 | T: butterfly per wave,
// then the per-wave partials through LDS in wave order.  Otherwise every
// lane sums the T partials from LDS in lane order.
template <int T>
__device__ __forceinline__ void win_sum2(float &a, float &b, float *ra, float *rb, int t)
{
    if constexpr (64 % T == 0) {
#pragma unroll
        for (int o = 1; o < T; o <<= 1) {
            a += __shfl_xor(a, o, T);
            b += __shfl_xor(b, o, T);
        }
    } else if constexpr (T % 64 == 0) {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
        }
        if ((t & 63) == 0) {
            ra[t >> 6] = a;
            rb[t >> 6] = b;
        }
        __syncthreads();
        float sa = 0.0f, sb = 0.0f;
#pragma unroll
        for (int k = 0; k < T / 64; ++k) {
            sa += ra[k];
            sb += rb[k];
        }
        __syncthreads();
        a = sa;
        b = sb;
    } else {
        ra[t] = a;
        rb[t] = b;
        __syncthreads();
        float sa = 0.0f, sb = 0.0f;
        for (int k = 0; k < T; ++k) {
            sa += ra[k];
            sb += rb[k];
        }
        __syncthreads();
        a = sa;
        b = sb;
    }
}

// Candidate summary of a set of plane pixels: the largest 3x3 local maximum
// (first in row-major candidate order on ties), the two largest candidate
// values (detectability) and whether a NaN was seen.  merge() is
// commutative and associative and only selects, so any reduction order
// gives the same bits.
struct Cand {
    float v;
    int c;
    float p1, p2;
    int nan;
    __device__ void merge(const Cand &o)
    {
        if (o.v > v || (o.v == v && o.v > 0.0f && o.c < c)) { v = o.v; c = o.c; }
        const float n1 = fmaxf(p1, o.p1);
        const float n2 = fmaxf(fminf(p1, o.p1), fmaxf(p2, o.p2));
        p1 = n1;
        p2 = n2;
        nan |= o.nan;
    }
    __device__ Cand shfl(int m, int w) const
    {
        Cand o;
        o.v = __shfl_xor(v, m, w);
        o.c = __shfl_xor(c, m, w);
        o.p1 = __shfl_xor(p1, m, w);
        o.p2 = __shfl_xor(p2, m, w);
        o.nan = __shfl_xor(nan, m, w);
        return o;
    }
};

__device__ __forceinline__ int remap_block(int b, int nb)
{
    // gridDim is a multiple of 8; consecutive workgroups go to consecutive
    // XCDs, so give XCD x the contiguous range [x*per, (x+1)*per).
    const int per = (int)gridDim.x >> 3;
    const int bb = (b & 7) * per + (b >> 3);
    return bb < nb ? bb : -1;
}

// Gate-passed peak at plane (i0, j0): 5x5 patch times the plane weight
// (pw, row stride nx; nullable), then the selected fit.  P has row stride px.
template <bool CEPHES>
__device__ int leader_fit(const float *P, int px, int nx, const float *pw, int i0, int j0, int ip,
                          int finder, float *loc, float *sig, float *fitv)
{
    float sub[PK * PK];
#pragma unroll
    for (int i = 0; i < PK; ++i)
#pragma unroll
        for (int j = 0; j < PK; ++j) {
            const int r = i0 + i - 2, c = j0 + j - 2;
            const float v = P[r * px + c];
            sub[i * PK + j] = pw ? v * pw[r * nx + c] : v;
        }
    float pk[2], sg[3], par[6];
    int st = 0;
    if (ip == 0 && finder == 6) {
        st = lm_fit<6, CEPHES>(sub, pk, sg, par);
        if (fitv) for (int k = 0; k < PK * PK; ++k) fitv[k] = model_at<6, CEPHES>(k / PK, k % PK, par);
    } else if (ip == 0 && finder == 5) {
        st = lm_fit<5, CEPHES>(sub, pk, sg, par);
        if (fitv) for (int k = 0; k < PK * PK; ++k) fitv[k] = model_at<5, CEPHES>(k / PK, k % PK, par);
    } else if (ip == 0 && finder == 4) {
        st = lm_fit<4, CEPHES>(sub, pk, sg, par);
        if (fitv) for (int k = 0; k < PK * PK; ++k) fitv[k] = model_at<4, CEPHES>(k / PK, k % PK, par);
    } else {
        float A, sx, sy;
        threept(sub, pk, &A, &sx, &sy);
        sg[0] = sx; sg[1] = sy; sg[2] = 0.0f;
        if (fitv) {
            const float q[6] = {A, pk[0], pk[1], sx, sy, 0.0f};
            for (int k = 0; k < PK * PK; ++k) fitv[k] = model_at<5, false>(k / PK, k % PK, q);
        }
    }
    if (st != 0) {
        loc[0] = __builtin_nanf(""); loc[1] = __builtin_nanf("");
        sig[0] = 0.0f; sig[1] = 0.0f; sig[2] = 0.0f;
    } else {
        loc[0] = pk[0] + (float)i0;
        loc[1] = pk[1] + (float)j0;
        sig[0] = sg[0]; sig[1] = sg[1]; sig[2] = sg[2];
    }
    return st;
}

// Sequential peak search and detectability, used only when the candidate
// region holds a NaN (the reference's comparison chain is order-dependent
// there; peak_locate_lm.c:801-813 and :727-744).
__device__ void seq_search(const float *P, int NY, int NX, int PX, float *h_out, int *i_out, int *j_out)
{
    int i0 = 0, j0 = 0;
    float h = 0.0f;
    for (int i = NY / 4; i < NY * 3 / 4; ++i)
        for (int j = NX / 4; j < NX * 3 / 4; ++j) {
            const float v = P[i * PX + j];
            if (v <= h) continue;
            int is_max = 1;
            for (int di = -1; di <= 1 && is_max; ++di)
                for (int dj = -1; dj <= 1; ++dj)
                    if (P[(i + di) * PX + (j + dj)] > v) { is_max = 0; break; }
            if (is_max) { h = v; i0 = i; j0 = j; }
        }
    *h_out = h; *i_out = i0; *j_out = j0;
}

__device__ float seq_detect(const float *P, int NY, int NX, int PX)
{
    float p1 = 0.0f, p2 = 0.0f;
    for (int i = NY / 4; i < NY * 3 / 4; ++i)
        for (int j = NX / 4; j < NX * 3 / 4; ++j) {
            const float v = P[i * PX + j];
            if (v <= p2 || v <= 0.0f) continue;
            int is_max = 1;
            for (int di = -1; di <= 1 && is_max; ++di)
                for (int dj = -1; dj <= 1; ++dj)
                    if (P[(i + di) * PX + (j + dj)] > v) { is_max = 0; break; }
            if (!is_max) continue;
            if (v > p1) { p2 = p1; p1 = v; }
            else p2 = v;
        }
    return (p2 > 0.0f) ? p1 / p2 : 0.0f;
}

// Output of bulkxcorr2d for one window (PIV_2d_cross_correlate.c:95-110).
struct XcorrOut {
    const XcorrArgs *g;
    size_t o_base, r_idx;
    int nWin;
    float enorm, hy, hx;
    __device__ void ratio(float v) const { g->ratio[r_idx] = v; }
    __device__ void peak(int ip, float row, float col, float h, const float *sg) const
    {
        const size_t o = o_base + (size_t)ip * nWin;
        g->pkX[o] = col - hx;
        g->pkY[o] = row - hy;
        g->sx[o] = sg[0];
        g->sy[o] = sg[1];
        g->sxy[o] = sg[2];
        g->pkH[o] = h * enorm;
    }
};

// Output of lsqpeaklocate_lm for one plane (peak_locate_lm.c:894-910).
struct PlaneOut {
    float *loc, *std, *rat;
    int n, nPeaks;
    __device__ void ratio(float v) const { if (rat) rat[0] = v; }
    __device__ void peak(int ip, float row, float col, float h, const float *sg) const
    {
        loc[(0 * nPeaks + ip) * n] = row;
        loc[(1 * nPeaks + ip) * n] = col;
        loc[(2 * nPeaks + ip) * n] = h;
        std[(0 * nPeaks + ip) * n] = sg[0];
        std[(1 * nPeaks + ip) * n] = sg[1];
        std[(2 * nPeaks + ip) * n] = sg[2];
    }
};

// Peak search, gate, fit and detectability for one plane P (NY x NX, row
// stride PX, writable: later peaks search the fit-subtracted plane).  Called
// by all lanes of the workgroup (it holds barriers); the T lanes of this
// window scan the quarter-region candidates c = t, t+T, ... and the lane
// t == 0 reduces them in lane order and fits (peak_locate_lm.c:767-918).
// Scratch r0..r4 hold T entries each.
template <typename Out>
__device__ void search_and_fit(float *P, int NY, int NX, int PX, int pwx, const float *pw, int nPeaks,
                               int finder, int impl, bool valid, int t, int T, float *r0, float *r1,
                               float *r2, float *r3, float *r4, const Out &out)
{
    const int CW = NX / 2, NC = (NY / 2) * (NX / 2);
    for (int ip = 0; ip < nPeaks; ++ip) {
        float bv = 0.0f, p1 = 0.0f, p2 = 0.0f;
        int bc = 0x7fffffff;
        float nan_seen = 0.0f;
        if (valid) {
            for (int c = t; c < NC; c += T) {
                const int i = NY / 4 + c / CW, j = NX / 4 + c % CW;
                const float v = P[i * PX + j];
                if (v != v) { nan_seen = 1.0f; continue; }
                if (!(v > 0.0f)) continue;
                bool is_max = true;
#pragma unroll
                for (int di = -1; di <= 1; ++di)
#pragma unroll
                    for (int dj = -1; dj <= 1; ++dj)
                        if (P[(i + di) * PX + (j + dj)] > v) is_max = false;
                if (!is_max) continue;
                if (v > bv) { bv = v; bc = c; }
                if (v > p1) { p2 = p1; p1 = v; }
                else if (v > p2) p2 = v;
            }
        }
        r0[t] = bv;
        r1[t] = __int_as_float(bc);
        r2[t] = p1;
        r3[t] = p2;
        r4[t] = nan_seen;
        __syncthreads();
        if (valid && t == 0) {
            float h = 0.0f, q1 = 0.0f, q2 = 0.0f;
            int hc = 0x7fffffff;
            bool anynan = false;
            for (int k = 0; k < T; ++k) {
                const float v = r0[k];
                const int c = __float_as_int(r1[k]);
                if (v > h || (v == h && v > 0.0f && c < hc)) { h = v; hc = c; }
                const float a1 = r2[k], a2 = r3[k];
                // top-2 of {q1, q2, a1, a2} (a1 >= a2)
                if (a1 > q1) { q2 = fmaxf(q1, a2); q1 = a1; }
                else if (a1 > q2) q2 = a1;
                anynan = anynan || (r4[k] != 0.0f);
            }
            int i0 = 0, j0 = 0;
            if (hc != 0x7fffffff) { i0 = NY / 4 + hc / CW; j0 = NX / 4 + hc % CW; }
            float det = (q2 > 0.0f) ? q1 / q2 : 0.0f;
            if (anynan) {
                seq_search(P, NY, NX, PX, &h, &i0, &j0);
                if (ip == 0) det = seq_detect(P, NY, NX, PX);
            }
            if (ip == 0) out.ratio(det);
            float loc[2], sg[3], fitv[PK * PK];
            float hout = h;
            // reference gate: h <= 0 (false for NaN) or the 5x5 patch off the plane
            const bool gate = (h <= 0.0f) || i0 < 2 || i0 >= NY - 2 || j0 < 2 || j0 >= NX - 2;
            if (gate) {
                loc[0] = __builtin_nanf(""); loc[1] = __builtin_nanf("");
                sg[0] = sg[1] = sg[2] = 0.0f;
                hout = 0.0f;
            } else {
                const bool want_fit = nPeaks > 1;
                int st;
                if (impl == 1 && nPeaks == 1 && finder >= 4)
                    st = leader_fit<true>(P, PX, pwx, pw, i0, j0, ip, finder, loc, sg, nullptr);
                else
                    st = leader_fit<false>(P, PX, pwx, pw, i0, j0, ip, finder, loc, sg,
                                           want_fit ? fitv : nullptr);
                if (st != 0) hout = 0.0f;
                if (want_fit && ip + 1 < nPeaks) {
                    // subtract the fitted model (un-compensated); MAX(0, d) keeps a NaN d
                    for (int i = 0; i < PK; ++i)
                        for (int j = 0; j < PK; ++j) {
                            const int r = i0 + i - 2, c = j0 + j - 2;
                            float f = fitv[i * PK + j];
                            if (pw) f /= pw[r * pwx + c];
                            const float d = P[r * PX + c] - f;
                            P[r * PX + c] = (0.0f > d) ? 0.0f : d;
                        }
                }
            }
            out.peak(ip, loc[0], loc[1], hout, sg);
        }
        __syncthreads();
    }
}

// Per-window peak record written by xcorr_kernel and consumed by fit_kernel.
constexpr int kRec = 32;
constexpr int kRecH = 25, kRecI0 = 26, kRecJ0 = 27, kRecEnorm = 28, kRecFlag = 29;

// Lane-order reduction of the per-lane candidate summaries (largest 3x3
// local maximum, first in row-major order on ties; top-2 for the
// detectability); falls back to the sequential scan when a NaN was seen.
__device__ void reduce_candidates(const float *r0, const float *r1, const float *r2, const float *r3,
                                  const float *r4, int T, int NY, int NX, const float *P, int PX,
                                  float *h_out, int *i_out, int *j_out, float *det_out)
{
    const int CW = NX / 2;
    float h = 0.0f, q1 = 0.0f, q2 = 0.0f;
    int hc = 0x7fffffff;
    bool anynan = false;
    for (int k = 0; k < T; ++k) {
        const float v = r0[k];
        const int c = __float_as_int(r1[k]);
        if (v > h || (v == h && v > 0.0f && c < hc)) { h = v; hc = c; }
        const float a1 = r2[k], a2 = r3[k];
        if (a1 > q1) { q2 = fmaxf(q1, a2); q1 = a1; }
        else if (a1 > q2) q2 = a1;
        anynan = anynan || (r4[k] != 0.0f);
    }
    int i0 = 0, j0 = 0;
    if (hc != 0x7fffffff) { i0 = NY / 4 + hc / CW; j0 = NX / 4 + hc % CW; }
    float det = (q2 > 0.0f) ? q1 / q2 : 0.0f;
    if (anynan) {
        seq_search(P, NY, NX, PX, &h, &i0, &j0);
        det = seq_detect(P, NY, NX, PX);
    }
    *h_out = h; *i_out = i0; *j_out = j0; *det_out = det;
}

template <int N>
__global__ void __launch_bounds__((XcCfg<N>::BLOCK))
xcorr_kernel(XcorrArgs g)
{
    using C = XcCfg<N>;
    constexpr int R = C::R, N1 = C::N1, T = C::T, PX = C::PX, WPB = C::WPB;
    constexpr int numel = N * N;
    __shared__ float s_re[WPB * C::tile];
    __shared__ float s_im[WPB * C::tile];
    __shared__ float s_red[5][WPB * T];

    const int b = remap_block((int)blockIdx.x, g.nblocks);
    if (b < 0) return;    // whole block exits together

    const int wloc = (int)threadIdx.x / T;
    const int t = (int)threadIdx.x - wloc * T;
    const int l = t / R;              // line (row or column) of this lane
    const int q = t - l * R;          // lane within the line
    const int ob = out_block<R>(q);   // output block of fft_split
    float *Sr = s_re + wloc * C::tile;
    float *Si = s_im + wloc * C::tile;
    float *r0 = &s_red[0][wloc * T], *r1 = &s_red[1][wloc * T];
    float *r2 = &s_red[2][wloc * T], *r3 = &s_red[3][wloc * T];
    float *r4 = &s_red[4][wloc * T];

    const int nWin = g.nWy * g.nWx;
    const long long gw = (long long)b * WPB + wloc;
    int n = 0, w = 0, row0 = 0, col0 = 0;
    bool valid = gw < g.total;
    if (valid) {
        n = (int)(gw / nWin);
        w = (int)(gw - (long long)n * nWin);
        const int ii = w % g.nWx, jj = w / g.nWx;
        if (g.mask && g.mask[jj * g.nWx + ii] == 1.0f) valid = false;
        row0 = (int)floor((double)g.ctrY[jj] - ((double)(float)N - 1.0) / 2.0 + 0.5);
        col0 = (int)floor((double)g.ctrX[ii] - ((double)(float)N - 1.0) / 2.0 + 0.5);
        if (row0 < 0 || col0 < 0 || row0 + N > g.H || col0 + N > g.W) valid = false;
    }

    // ---- 1. column x = l: rows n1*R + q, weight, mean, energy, FFT --------
    float zr[N1], zi[N1];
    float sa = 0.0f, sb = 0.0f;
    if (valid) {
        const size_t img = (size_t)n * g.H * g.W;
        const float *pa = g.imgA + img + (size_t)(row0 + q) * g.W + col0 + l;
        const float *pb = g.imgB + img + (size_t)(row0 + q) * g.W + col0 + l;
        const float *wa = g.wA + q * N + l, *wb = g.wB + q * N + l;
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const float a = pa[(size_t)k * R * g.W] * wa[k * R * N];
            const float bb = pb[(size_t)k * R * g.W] * wb[k * R * N];
            zr[k] = a;
            zi[k] = bb;
            sa += a;
            sb += bb;
        }
    } else {
#pragma unroll
        for (int k = 0; k < N1; ++k) { zr[k] = 0.0f; zi[k] = 0.0f; }
    }
    win_sum2<T>(sa, sb, r0, r1, t);
    const float meanA = sa / (float)numel;
    const float meanB = sb / (float)numel;
    float ea = 0.0f, eb = 0.0f;
    if (valid) {
        if (!g.ensemble) {
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                zr[k] -= meanA;
                zi[k] -= meanB;
                ea += zr[k] * zr[k];
                eb += zi[k] * zi[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                ea += zr[k] * zr[k];
                eb += zi[k] * zi[k];
            }
        }
    }
    win_sum2<T>(ea, eb, r0, r1, t);
    const float enorm = 1.0f / sqrtf(ea * eb);

    if (valid) {
        fft_split<N, R, false>(zr, zi, q);
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            Sr[(k + N1 * ob) * PX + l] = zr[k];
            Si[(k + N1 * ob) * PX + l] = zi[k];
        }
    }
    __syncthreads();

    // ---- 2. row ky = l --------------------------------------------------------
    if (valid) {
        float xr[N1], xi[N1];
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            xr[k] = Sr[l * PX + k * R + q];
            xi[k] = Si[l * PX + k * R + q];
        }
        fft_split<N, R, false>(xr, xi, q);
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            Sr[l * PX + k + N1 * ob] = xr[k];
            Si[l * PX + k + N1 * ob] = xi[k];
        }
    }
    __syncthreads();

    // ---- 3. column kx = l <= N/2: C = FB conj(FA), inverse column FFT ------
    if (valid && l <= N / 2) {
        const int mx = (N - l) % N;
        float cr[N1], ci[N1];
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int ky = k * R + q;
            const int my = (N - ky) % N;
            const float zR = Sr[ky * PX + l], zI = Si[ky * PX + l];
            const float mR = Sr[my * PX + mx], mI = Si[my * PX + mx];
            // FA = (Z + conj(Zm)) / 2 ; FB = (Z - conj(Zm)) / (2i)
            const float ar = 0.5f * (zR + mR), ai = 0.5f * (zI - mI);
            const float br = 0.5f * (zI + mI), bi = -0.5f * (zR - mR);
            cr[k] = br * ar + bi * ai;
            ci[k] = bi * ar - br * ai;
        }
        fft_split<N, R, true>(cr, ci, q);
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            Sr[(k + N1 * ob) * PX + l] = cr[k];
            Si[(k + N1 * ob) * PX + l] = ci[k];
        }
    }
    __syncthreads();

    // ---- 4. rows j and j + N/2 (j = l < N/2): Hermitian rows -> real rows -----
    if (valid && l < N / 2) {
        const int ya = l, yb = l + N / 2;
        float xr[N1], xi[N1];
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int kx = k * R + q;
            const bool up = kx > N / 2;
            const int kk = up ? N - kx : kx;
            float g1r = Sr[ya * PX + kk], g1i = Si[ya * PX + kk];
            float g2r = Sr[yb * PX + kk], g2i = Si[yb * PX + kk];
            if (kk == 0 || kk == N / 2) { g1i = 0.0f; g2i = 0.0f; }
            if (up) { g1i = -g1i; g2i = -g2i; }     // G(N-k) = conj(G(k)) per real row
            xr[k] = g1r - g2i;
            xi[k] = g1i + g2r;
        }
        fft_split<N, R, true>(xr, xi, q);
        const float mul = 1.0f / (float)numel;
        // plane[(y + N/2) % N][(x + N/2) % N] = c[y][x] / numel
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int xs = (k + N1 * ob + N / 2) % N;
            Sr[yb * PX + xs] = xr[k] * mul;
            Sr[ya * PX + xs] = xi[k] * mul;
        }
    }
    __syncthreads();

    const float *P = Sr;
    if (g.planes && valid) {
        float *dst = g.planes + (size_t)gw * numel;
        for (int e = t; e < numel; e += T) dst[e] = P[(e / N) * PX + (e % N)];
    }
    if (g.ensemble) return;    // uniform: the flag is a kernel argument

    // ---- 5. peak search; the fit runs in fit_kernel -----------------------
    if (g.nPeaks > 1) {
        // multi-peak: the plane goes to the workspace for multipeak_kernel
        if (valid) {
            float *dst = g.work_planes + (size_t)gw * numel;
            for (int e = t; e < numel; e += T) dst[e] = P[(e / N) * PX + (e % N)];
        }
        if (t == 0 && gw < g.total) {
            float *rec = g.rec + (size_t)gw * kRec;
            rec[kRecFlag] = valid ? 1.0f : 0.0f;
            rec[kRecEnorm] = enorm;
        }
        return;
    }
    constexpr int CW = N / 2, NC = (N / 2) * (N / 2);
    Cand cd = {0.0f, 0x7fffffff, 0.0f, 0.0f, 0};
    if (valid) {
        for (int c = t; c < NC; c += T) {
            const int i = N / 4 + c / CW, j = N / 4 + c % CW;
            const float v = P[i * PX + j];
            if (v != v) { cd.nan = 1; continue; }
            if (!(v > 0.0f)) continue;
            bool is_max = true;
#pragma unroll
            for (int di = -1; di <= 1; ++di)
#pragma unroll
                for (int dj = -1; dj <= 1; ++dj)
                    if (P[(i + di) * PX + (j + dj)] > v) is_max = false;
            if (!is_max) continue;
            if (v > cd.v) { cd.v = v; cd.c = c; }
            if (v > cd.p1) { cd.p2 = cd.p1; cd.p1 = v; }
            else if (v > cd.p2) cd.p2 = v;
        }
    }
    if constexpr (64 % T == 0) {
#pragma unroll
        for (int o = 1; o < T; o <<= 1) cd.merge(cd.shfl(o, T));
    } else if constexpr (T % 64 == 0) {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) cd.merge(cd.shfl(o, 64));
        if ((t & 63) == 0) {
            r0[t >> 6] = cd.v;
            r1[t >> 6] = __int_as_float(cd.c);
            r2[t >> 6] = cd.p1;
            r3[t >> 6] = cd.p2;
            r4[t >> 6] = (float)cd.nan;
        }
        __syncthreads();
        if (t == 0) {
            for (int k = 1; k < T / 64; ++k) {
                Cand o = {r0[k], __float_as_int(r1[k]), r2[k], r3[k], r4[k] != 0.0f};
                cd.merge(o);
            }
        }
    } else {
        r0[t] = cd.v;
        r1[t] = __int_as_float(cd.c);
        r2[t] = cd.p1;
        r3[t] = cd.p2;
        r4[t] = (float)cd.nan;
        __syncthreads();
        if (t == 0) {
            for (int k = 1; k < T; ++k) {
                Cand o = {r0[k], __float_as_int(r1[k]), r2[k], r3[k], r4[k] != 0.0f};
                cd.merge(o);
            }
        }
    }
    if (t == 0 && gw < g.total) {
        float *rec = g.rec + (size_t)gw * kRec;
        if (!valid) {
            rec[kRecFlag] = 0.0f;
        } else {
            float h = cd.v, det = (cd.p2 > 0.0f) ? cd.p1 / cd.p2 : 0.0f;
            int i0 = 0, j0 = 0;
            if (cd.c != 0x7fffffff) { i0 = N / 4 + cd.c / CW; j0 = N / 4 + cd.c % CW; }
            if (cd.nan) {
                seq_search(P, N, N, PX, &h, &i0, &j0);
                det = seq_detect(P, N, N, PX);
            }
            g.ratio[(size_t)n * nWin + w] = det;
            const bool gate = (h <= 0.0f) || i0 < 2 || i0 >= N - 2 || j0 < 2 || j0 >= N - 2;
            if (gate) {
                const size_t o = (size_t)n * nWin + w;
                g.pkX[o] = __builtin_nanf("") - N / 2.0f;
                g.pkY[o] = __builtin_nanf("") - N / 2.0f;
                g.sx[o] = 0.0f; g.sy[o] = 0.0f; g.sxy[o] = 0.0f;
                g.pkH[o] = 0.0f * enorm;
                rec[kRecFlag] = 0.0f;
            } else {
#pragma unroll
                for (int i = 0; i < PK; ++i)
#pragma unroll
                    for (int j = 0; j < PK; ++j) {
                        const int r = i0 + i - 2, c = j0 + j - 2;
                        const float v = P[r * PX + c];
                        rec[i * PK + j] = g.pw ? v * g.pw[r * N + c] : v;
                    }
                rec[kRecH] = h;
                rec[kRecI0] = (float)i0;
                rec[kRecJ0] = (float)j0;
                rec[kRecEnorm] = enorm;
                rec[kRecFlag] = 1.0f;
            }
        }
    }
}

#ifdef PIVAMD_XCORR_COMMON
// One plane per workgroup (lsqpeaklocate_lm / lsqpeaklocate_lm_batch /
// peak_detectability over planes already in device memory).
__global__ void __launch_bounds__(256) peak_kernel(const float *planes, int NY, int NX, int nPeaks,
                                                   int finder, int impl, const float *pw, float *loc,
                                                   float *std, float *rat, int n)
{
    extern __shared__ float s_dyn[];
    __shared__ float s_r[5][256];
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    const int numel = NY * NX;
    float *P = s_dyn;
    for (int e = t; e < numel; e += 256) P[e] = planes[(size_t)p * numel + e];
    __syncthreads();
    PlaneOut out;
    out.loc = loc + p;
    out.std = std + p;
    out.rat = rat ? rat + p : nullptr;
    out.n = n;
    out.nPeaks = nPeaks;
    search_and_fit(P, NY, NX, NX, NX, pw, nPeaks, finder, impl, true, t, 256, s_r[0], s_r[1], s_r[2],
                   s_r[3], s_r[4], out);
}

// One lane per window: the sub-pixel fit of every gate-passed window
// (PIV_2d_cross_correlate.c:171-190 output assembly).
__global__ void __launch_bounds__(256) fit_kernel(const float *rec_all, long long total, int nWin,
                                                  int winY, int winX, int finder, int impl, float *pkX,
                                                  float *pkY, float *pkH, float *sx, float *sy, float *sxy)
{
    const long long gw = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gw >= total) return;
    const float *rec = rec_all + (size_t)gw * kRec;
    if (rec[kRecFlag] != 1.0f) return;
    float sub[PK * PK];
#pragma unroll
    for (int k = 0; k < PK * PK; ++k) sub[k] = rec[k];
    const float h = rec[kRecH];
    const int i0 = (int)rec[kRecI0], j0 = (int)rec[kRecJ0];
    const float enorm = rec[kRecEnorm];
    float pk[2], sg[3], par[6];
    int st = 0;
    if (finder >= 4 && finder <= 6) {
        if (impl == 1) {
            if (finder == 6) st = lm_fit<6, true>(sub, pk, sg, par);
            else if (finder == 5) st = lm_fit<5, true>(sub, pk, sg, par);
            else st = lm_fit<4, true>(sub, pk, sg, par);
        } else {
            if (finder == 6) st = lm_fit<6, false>(sub, pk, sg, par);
            else if (finder == 5) st = lm_fit<5, false>(sub, pk, sg, par);
            else st = lm_fit<4, false>(sub, pk, sg, par);
        }
    } else {
        float A, s0, s1;
        threept(sub, pk, &A, &s0, &s1);
        sg[0] = s0; sg[1] = s1; sg[2] = 0.0f;
    }
    const size_t o = (size_t)gw;   // nPeaks == 1: n*nWin + w
    if (st != 0) {
        pkX[o] = __builtin_nanf("") - winX / 2.0f;
        pkY[o] = __builtin_nanf("") - winY / 2.0f;
        sx[o] = 0.0f; sy[o] = 0.0f; sxy[o] = 0.0f;
        pkH[o] = 0.0f * enorm;
    } else {
        pkX[o] = (pk[1] + (float)j0) - winX / 2.0f;
        pkY[o] = (pk[0] + (float)i0) - winY / 2.0f;
        sx[o] = sg[0]; sy[o] = sg[1]; sxy[o] = sg[2];
        pkH[o] = h * enorm;
    }
}

// nPeaks > 1: one workgroup per window over the plane left in the workspace.
__global__ void __launch_bounds__(256) multipeak_kernel(const float *planes, const float *rec_all,
                                                        long long total, int nWin, int NY, int NX,
                                                        int nPeaks, int finder, const float *pw,
                                                        XcorrArgs g)
{
    extern __shared__ float s_dyn[];
    __shared__ float s_r[5][256];
    const long long gw = blockIdx.x;
    if (gw >= total) return;
    const float *rec = rec_all + (size_t)gw * kRec;
    if (rec[kRecFlag] != 1.0f) return;    // uniform across the block
    const int t = threadIdx.x, numel = NY * NX;
    float *P = s_dyn;
    for (int e = t; e < numel; e += 256) P[e] = planes[(size_t)gw * numel + e];
    __syncthreads();
    const int n = (int)(gw / nWin), w = (int)(gw - (long long)n * nWin);
    XcorrOut out;
    out.g = &g;
    out.o_base = (size_t)n * nPeaks * nWin + w;
    out.r_idx = (size_t)n * nWin + w;
    out.nWin = nWin;
    out.enorm = rec[kRecEnorm];
    out.hy = NY / 2.0f;
    out.hx = NX / 2.0f;
    search_and_fit(P, NY, NX, NX, NX, pw, nPeaks, finder, 0, true, t, 256, s_r[0], s_r[1], s_r[2],
                   s_r[3], s_r[4], out);
}

// numel / (wB (*) wB) on the centre crop of the 2N zero-padded correlation
// (xcorr.c:22-89, PIV_2d_cross_correlate.c:219-222), summed in FP64.
__global__ void plane_weight_kernel(const float *wB, int NY, int NX, float *out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NY * NX) return;
    const int i = e / NX, j = e % NX;
    const int my = i - NY / 2, mx = j - NX / 2;
    double s = 0.0;
    for (int y = 0; y < NY; ++y) {
        const int ya = y + my;
        if (ya < 0 || ya >= NY) continue;
        for (int x = 0; x < NX; ++x) {
            const int xa = x + mx;
            if (xa < 0 || xa >= NX) continue;
            s += (double)wB[ya * NX + xa] * (double)wB[y * NX + x];
        }
    }
    out[e] = (float)(NY * NX) / (float)s;
}

// conv[i][j] = sum a[y+i-NY/2][x+j-NX/2] * b[y][x]: the centre crop of the
// 2N zero-padded correlation computed by convolve (xcorr.c:22-89), in FP64.
__global__ void conv_kernel(const float *a, const float *b, int NY, int NX, float *out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NY * NX) return;
    const int my = e / NX - NY / 2, mx = e % NX - NX / 2;
    double s = 0.0;
    for (int y = 0; y < NY; ++y) {
        const int ya = y + my;
        if (ya < 0 || ya >= NY) continue;
        for (int x = 0; x < NX; ++x) {
            const int xa = x + mx;
            if (xa < 0 || xa >= NX) continue;
            s += (double)a[ya * NX + xa] * (double)b[y * NX + x];
        }
    }
    out[e] = (float)s;
}

#endif  // PIVAMD_XCORR_COMMON

// Launch of the correlation kernel for one square size; defined in the
// translation unit compiled with PIVAMD_XCORR_SIZE.
template <int N>
int launch_main(XcorrArgs g, hipStream_t st);

#ifdef PIVAMD_XCORR_SIZE
template <>
int launch_main<PIVAMD_XCORR_SIZE>(XcorrArgs g, hipStream_t st)
{
    constexpr int N = PIVAMD_XCORR_SIZE;
    using C = XcCfg<N>;
    const long long nb = (g.total + C::WPB - 1) / C::WPB;
    if (nb <= 0) return 0;
    if (nb > (1LL << 30)) return PIVAMD_ERR_ARG;
    g.nblocks = (int)nb;
    const int grid = (int)((nb + 7) / 8 * 8);
    hipLaunchKernelGGL((xcorr_kernel<N>), dim3(grid), dim3(C::BLOCK), 0, st, g);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}
#endif

#ifdef PIVAMD_XCORR_COMMON
template <> int launch_main<8>(XcorrArgs, hipStream_t);
template <> int launch_main<12>(XcorrArgs, hipStream_t);
template <> int launch_main<16>(XcorrArgs, hipStream_t);
template <> int launch_main<24>(XcorrArgs, hipStream_t);
template <> int launch_main<32>(XcorrArgs, hipStream_t);
template <> int launch_main<48>(XcorrArgs, hipStream_t);
template <> int launch_main<64>(XcorrArgs, hipStream_t);
template <> int launch_main<96>(XcorrArgs, hipStream_t);
template <> int launch_main<128>(XcorrArgs, hipStream_t);

// Correlation kernel, then the fit (nPeaks == 1) or multi-peak kernel.
static int launch_xcorr(int N, XcorrArgs g, hipStream_t st)
{
    int rc;
    switch (N) {
    case 8: rc = launch_main<8>(g, st); break;
    case 12: rc = launch_main<12>(g, st); break;
    case 16: rc = launch_main<16>(g, st); break;
    case 24: rc = launch_main<24>(g, st); break;
    case 32: rc = launch_main<32>(g, st); break;
    case 48: rc = launch_main<48>(g, st); break;
    case 64: rc = launch_main<64>(g, st); break;
    case 96: rc = launch_main<96>(g, st); break;
    case 128: rc = launch_main<128>(g, st); break;
    default: return PIVAMD_ERR_SIZE;
    }
    if (rc != 0 || g.ensemble || g.total <= 0) return rc;
    const int nWin = g.nWy * g.nWx;
    if (g.nPeaks == 1) {
        const long long fb = (g.total + 255) / 256;
        hipLaunchKernelGGL(fit_kernel, dim3((unsigned)fb), dim3(256), 0, st, g.rec, g.total, nWin, N, N,
                           g.finder, g.impl, g.pkX, g.pkY, g.pkH, g.sx, g.sy, g.sxy);
    } else {
        hipLaunchKernelGGL(multipeak_kernel, dim3((unsigned)g.total), dim3(256),
                           (size_t)N * N * sizeof(float), st, g.work_planes, g.rec, g.total, nWin, N, N,
                           g.nPeaks, g.finder, g.pw, g);
    }
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}
#endif

}  // namespace pivamd

#ifdef PIVAMD_XCORR_COMMON

using namespace pivamd;

extern "C" int pivamd_size_supported(int ny, int nx)
{
#define PIVAMD_SQ(N) if (ny == N && nx == N) return 1;
    PIVAMD_SQ(8) PIVAMD_SQ(12) PIVAMD_SQ(16) PIVAMD_SQ(24) PIVAMD_SQ(32)
    PIVAMD_SQ(48) PIVAMD_SQ(64) PIVAMD_SQ(96) PIVAMD_SQ(128)
#undef PIVAMD_SQ
    return 0;
}

extern "C" int pivamd_convolve(const float *d_a, const float *d_b, int ny, int nx, float *d_out,
                               hipStream_t st)
{
    const int n = ny * nx;
    if (n <= 0) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL(conv_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_a, d_b, ny, nx, d_out);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

extern "C" int pivamd_plane_weight(const float *d_wB, int ny, int nx, float *d_out, hipStream_t st)
{
    const int n = ny * nx;
    hipLaunchKernelGGL(plane_weight_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_wB, ny, nx, d_out);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

extern "C" size_t pivamd_xcorr_workspace(int n_images, int nWy, int nWx, int winY, int winX, int nPeaks,
                                         int ensemble)
{
    if (ensemble) return 0;
    const size_t total = (size_t)n_images * nWy * nWx;
    size_t bytes = total * kRec * sizeof(float);
    if (nPeaks > 1) bytes += total * (size_t)winY * winX * sizeof(float);
    return bytes;
}

extern "C" int pivamd_xcorr(const float *d_imgA, const float *d_imgB, const float *d_mask, int H, int W,
                            int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                            const float *d_wA, const float *d_wB, const float *d_planeWeight, int ensemble,
                            int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX,
                            float *d_pkY, float *d_pkH, float *d_sx, float *d_sy, float *d_sxy,
                            float *d_ratio, float *d_planes, void *d_work, size_t work_bytes, hipStream_t st)
{
    if (n_images < 0 || nWy < 0 || nWx < 0 || nPeaks < 1 || H <= 0 || W <= 0) return PIVAMD_ERR_ARG;
    if (work_bytes < pivamd_xcorr_workspace(n_images, nWy, nWx, winY, winX, nPeaks, ensemble))
        return PIVAMD_ERR_ARG;
    XcorrArgs g;
    g.imgA = d_imgA; g.imgB = d_imgB; g.mask = d_mask; g.ctrX = d_ctrX; g.ctrY = d_ctrY;
    g.wA = d_wA; g.wB = d_wB; g.pw = d_planeWeight;
    g.pkX = d_pkX; g.pkY = d_pkY; g.pkH = d_pkH; g.sx = d_sx; g.sy = d_sy; g.sxy = d_sxy;
    g.ratio = d_ratio; g.planes = d_planes;
    g.rec = (float *)d_work;
    g.work_planes = nPeaks > 1 && d_work ? (float *)d_work + (size_t)n_images * nWy * nWx * kRec : nullptr;
    g.total = (long long)n_images * nWy * nWx;
    g.H = H; g.W = W; g.nWy = nWy; g.nWx = nWx; g.nPeaks = nPeaks; g.finder = finder;
    g.impl = impl; g.ensemble = ensemble; g.nblocks = 0;
    if (!ensemble && (!d_pkX || !d_pkY || !d_pkH || !d_sx || !d_sy || !d_sxy || !d_ratio || !d_work))
        return PIVAMD_ERR_ARG;
if (winY != winX || !pivamd_size_supported(winY, winX)) return PIVAMD_ERR_SIZE;
    return launch_xcorr(winY, g, st);
}

extern "C" int pivamd_peakfit(const float *d_planes, int n, int ny, int nx, int nPeaks, int finder,
                              int impl, const float *d_planeWeight, float *d_loc, float *d_std,
                              float *d_ratio, hipStream_t st)
{
    if (n < 0 || ny < 4 || nx < 4 || nPeaks < 1 || !d_planes || !d_loc || !d_std) return PIVAMD_ERR_ARG;
    if ((long long)ny * nx > 32768) return PIVAMD_ERR_SIZE;
    if (n == 0) return 0;
    const size_t shm = (size_t)ny * nx * sizeof(float);
    hipLaunchKernelGGL(peak_kernel, dim3(n), dim3(256), shm, st, d_planes, ny, nx, nPeaks, finder, impl,
                       d_planeWeight, d_loc, d_std, d_ratio, n);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

#endif  // PIVAMD_XCORR_COMMON
