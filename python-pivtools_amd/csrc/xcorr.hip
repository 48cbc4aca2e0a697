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
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <type_traits>

#include "fft_reg.h"
#include "peakfit.h"
#include "pivamd_dev.h"

namespace pivamd {

struct XcorrArgs {
    const float *imgA, *imgB, *mask, *ctrX, *ctrY, *wA, *wB, *pw;
    const unsigned short *imgA16, *imgB16;   // 16-bit counts instead of imgA/imgB when set
    float *pkX, *pkY, *pkH, *sx, *sy, *sxy, *ratio, *planes;
    float *rec, *work_planes;   // workspace: per-window peak records, multi-peak planes
    long long total;        // N_images * nWin
    int H, W, nWy, nWx, nPeaks, finder, impl, ensemble, nblocks;
    int dft;                // 1: a split-precision MFMA DFT form, 0: FP32 FFT
    int variant;            // FFT lanes per line R (0: default), or a kVariant* form
};

constexpr int kMaxLds = 158 * 1024;

// One ragged correlation launch: the correlation calls (jobs) of one square
// window size over different batches -- other image sizes, window grids,
// buffers -- share one grid of workgroups.  Job j owns workgroups
// [start[j], start[j + 1]); the table travels by value in the kernel
// arguments, so a captured HIP graph replays it as recorded.
constexpr int kMaxRaggedJobs = 8;
struct RaggedArgs {
    XcorrArgs job[kMaxRaggedJobs];
    int start[kMaxRaggedJobs + 1];
    int n_jobs;
    int nblocks;
};

// Job of workgroup b (uniform over the workgroup).
__device__ __forceinline__ int ragged_job(const RaggedArgs &ra, int b)
{
    int j = 0;
#pragma unroll
    for (int k = 1; k < kMaxRaggedJobs; ++k) j += (k < ra.n_jobs && b >= ra.start[k]) ? 1 : 0;
    return j;
}



// Lanes per line transform (R) for window size N: the default, and the
// values built as selectable variants (fft_reg holds N/R points per lane).
constexpr int xc_radix(int N) { return N >= 48 ? 4 : 1; }

constexpr bool xc_radix_built(int N, int R)
{
    return R == 1 ? N <= 32
         : R == 2 ? (N >= 16 && N <= 64)
         : R == 4 ? (N >= 32 && N <= 128 && N != 24)
         : R == 8 ? N == 128
         : false;
}

// Variant codes of pivamd_xcorr_set_variant: 0 = default, R = FP32 FFT with
// R lanes per line, kVariantMfma = split-precision MFMA DFT, kVariantPersist
// + 4 = the persistent prefetching form of the 4-lane FFT (N >= 64, where a
// workgroup holds one window), kVariantHalf + R = the R-lane FFT through
// one real LDS plane per window (R = 4: N = 64, 96, 128; R = 1: N <= 32):
// more windows per CU.
constexpr int kVariantMfma = 16;
constexpr int kVariantPersist = 32;
constexpr int kVariantHalf = 64;     // + R: xcorr_kernel_half (one real LDS plane per window)
constexpr int kVariantGemm = 96;     // + 0/1/2: N = 16 register-chained MFMA DFT, fp16 / bf16 / bf16 x3

// Tile layout of the complex stages.  R = 1: element (r, c) at r*(N+1) + c.
// R > 1: at r*TPX + SKEW*(r/N1) + c + PADC*(c/N1), i.e. rows padded to TPX,
// PADC words after every block of N1 columns and the row blocks skewed by
// SKEW words.  The R lanes of a line touch elements N1 apart (their output
// blocks) or adjacent (their interleaved inputs); with the plain N+1 pitch
// the former fall on one bank (2-4 way conflicts on every stage).  The
// constants minimise the modelled conflicts of all six tile access patterns
// (64 banks per wave and 32 banks per half wave); the block indices r/N1 and
// c/N1 are compile-time or per-lane constants at every access, so the
// addresses stay a lane base plus an immediate.  The final real plane is
// written with the plain N+1 pitch (after a barrier) for the epilogue.
struct XcLayout {
    int tpx, padc, skew;
};

constexpr XcLayout xc_layout(int N, int R)
{
    return R == 1 ? XcLayout{N + 1, 0, 0}
         : (N == 128 && R == 4) ? XcLayout{140, 3, 8}
         : (N == 128 && R == 8) ? XcLayout{136, 1, 4}
         : (N == 96 && R == 4) ? XcLayout{100, 1, 8}
         : (N == 32 && R == 4) ? XcLayout{44, 3, 40}
         : (N == 64 && R == 4) ? XcLayout{76, 3, 8}
         : (N == 64 && R == 2) ? XcLayout{66, 1, 16}
         : (N == 48 && R == 2) ? XcLayout{50, 1, 0}
         : (N == 48 && R == 4) ? XcLayout{52, 1, 0}
         : (N == 24 && R == 2) ? XcLayout{30, 1, 0}
         : (N == 32 && R == 2) ? XcLayout{34, 1, 0}
         : (N == 16 && R == 2) ? XcLayout{17, 0, 8}
         : XcLayout{N + 1, 0, 0};
}

// Launch geometry of one square window size N: R lanes share each line
// (row or column) transform, T = N*R lanes per window, WPB windows per
// workgroup as LDS allows (one complex tile and RED reduction slots per
// window).
template <int N, int R_ = xc_radix(N)>
struct XcCfg {
    static constexpr int R = R_;
    static constexpr int N1 = N / R;
    static constexpr int T = N * R;
    static constexpr int PX = N + 1;      // plane pitch
    static constexpr int TPX = xc_layout(N, R).tpx;
    static constexpr int PADC = xc_layout(N, R).padc;
    static constexpr int SKEW = xc_layout(N, R).skew;
    static constexpr int tile_words = (N - 1) * TPX + SKEW * (R - 1) + (N - 1) + PADC * (R - 1) + 1;
    static constexpr int tile = tile_words > N * PX ? tile_words : N * PX;
    // reduction slots: shuffles only (T | 64), one per wave (64 | T), one per lane
    static constexpr int RED = 64 % T == 0 ? 1 : (T % 64 == 0 ? T / 64 : T);
    static constexpr int TWP = N1 + 2;    // twiddle row pitch (float2): rows q start 4 banks apart
    static constexpr int lds_win = 2 * tile * 4 + 5 * RED * 4;
    static constexpr int lds_blk = R > 1 ? R * TWP * 8 : 0;
    static constexpr int wpb_thr = (256 / T) > 0 ? (256 / T) : 1;
    static constexpr int wpb_lds = (kMaxLds - lds_blk) / lds_win;
    static constexpr int WPB = wpb_thr < wpb_lds ? wpb_thr : wpb_lds;
    static constexpr int BLOCK = WPB * T;
    // Window tiles in the workgroup's LDS.  LDS serves a wave in two halves
    // of 32 lanes over 32 banks.  With one lane per line and H = 32 / T
    // windows per half wave (N = 8, 16, 32), the tile size N (N + 1) puts
    // those H windows N banks apart, which keeps the column and row stages
    // free of conflicts.  The packed inverse stages hold 2H windows per half
    // wave; every second group of H windows is therefore moved by another
    // 16 / H banks (and each pair of groups by 32 words, so tiles never
    // overlap), which puts the 2H windows 16 / H banks apart.
    static constexpr int H = (R == 1 && T <= 32 && 32 % T == 0) ? 32 / T : 0;
    static __host__ __device__ constexpr int wbase(int wl)
    {
        return wl * tile + (H ? 32 * (wl / (2 * H)) + (16 / H) * ((wl / H) % 2) : 0);
    }
    static constexpr int tiles = wbase(WPB - 1) + tile;
    // lane t of a window -> line l and lane-in-line q.  R = 8 puts q bit 2 on
    // lane bit 3, so every exchange partner is one DPP away (xor 1, 2: quad
    // permutes; xor 8: row_ror:8).
    static __device__ __forceinline__ void lane_map(int t, int &l, int &q)
    {
        if constexpr (R == 8) {
            q = (t & 3) | (((t >> 3) & 1) << 2);
            l = ((t >> 2) & 1) | ((t >> 4) << 1);
        } else {
            l = t / R;
            q = t - l * R;
        }
    }
};

// Tile word of element (r, c); rb = r / N1 and cb = c / N1 are passed in
// by the caller, which knows them as compile-time or per-lane constants.
template <int N, int R>
__device__ __forceinline__ int tix(int r, int rb, int c, int cb)
{
    using C = XcCfg<N, R>;
    return r * C::TPX + C::SKEW * rb + c + C::PADC * cb;
}

// Block (of N1) of index jR - q, 0 <= q < R, j >= 1: jR/N1 unless jR starts
// a block and q > 0.
template <int N1, int R>
__device__ __forceinline__ int mirror_block(int j, int q)
{
    return ((j * R) % N1 == 0 && q > 0) ? (j * R) / N1 - 1 : (j * R) / N1;
}

// Sum of (a, b) over the T lanes of one window; every lane gets the same
// bits.  T | 64: xor butterfly inside the wave.  64 | T: butterfly per wave,
// then the per-wave partials through LDS in wave order.  Otherwise every
// lane sums the T partials from LDS in lane order.  There is no barrier
// after the LDS reads: a caller must not write ra/rb again before its next
// workgroup barrier (consecutive calls use different slots).
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
    // the partner given by a DPP lane pattern within a row of 16 lanes
    template <int CTRL>
    __device__ Cand dpp() const
    {
        Cand o;
        o.v = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
        o.c = __builtin_amdgcn_mov_dpp(c, CTRL, 0xF, 0xF, true);
        o.p1 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(p1), CTRL, 0xF, 0xF, true));
        o.p2 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(p2), CTRL, 0xF, 0xF, true));
        o.nan = __builtin_amdgcn_mov_dpp(nan, CTRL, 0xF, 0xF, true);
        return o;
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
// The 3-point fits read only the cross through the peak, so with one peak and
// a 3-point finder the record is 8 words at an 8-word stride: the weighted
// plane at (-1, 0), (0, -1), (0, 0), (0, 1), (1, 0) from the peak, the height,
// the energy norm, and the peak position as the integer i0 * 256 + j0 (-1: no
// fit).  A quarter of the bytes through HBM, written and read.
constexpr int kRecC = 8;
constexpr int kRecCH = 5, kRecCEnorm = 6, kRecCPos = 7;
__host__ __device__ constexpr bool rec_compact(int finder, int nPeaks)
{
    return nPeaks == 1 && !(finder >= 4 && finder <= 6);
}

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

// Everything after the plane is in LDS (P, row stride PX, fftshifted, the
// reference's raw plane): optional plane dump, then either the multi-peak
// hand-off or the candidate search and the per-window peak record for
// fit_kernel.  Called by all T lanes of the window.
template <int N, int T>
__device__ __forceinline__ void plane_epilogue(const XcorrArgs &g, const float *P, int PX, bool valid, long long gw,
                                               int n, int w, float enorm, int t, float *r0, float *r1, float *r2,
                                               float *r3, float *r4)
{
    constexpr int numel = N * N;
    const int nWin = g.nWy * g.nWx;
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
    constexpr int RPL = NC / T;
    if constexpr (T % CW == 0 && NC % T == 0 && RPL >= 1) {
        if (valid) {
            // Lane t scans RPL consecutive rows of one column.  The 3x3 maxima
            // come from row-wise 3-maxima of RPL + 2 rows (3 (RPL + 2) reads in
            // place of 9 RPL); including the centre changes nothing, since the
            // test is "no neighbour above v" and fmaxf skips NaN as before.  The
            // visiting order stays increasing c within the lane.
            const int jc = t % CW, ib = (t / CW) * RPL;
            const float *col = P + (N / 4 + ib - 1) * PX + N / 4 + jc;
            float hm[RPL + 2], cen[RPL];
#pragma unroll
            for (int r = 0; r < RPL + 2; ++r) {
                const float *p = col + r * PX;
                const float b = p[0];
                hm[r] = fmaxf(fmaxf(p[-1], b), p[1]);
                if (r >= 1 && r <= RPL) cen[r - 1] = b;
            }
#pragma unroll
            for (int k = 0; k < RPL; ++k) {
                const float v = cen[k];
                const float m = fmaxf(fmaxf(hm[k], hm[k + 1]), hm[k + 2]);
                const int c = (ib + k) * CW + jc;
                const bool vnan = v != v;
                cd.nan = cd.nan || vnan;
                const bool cand = !vnan && v > 0.0f && !(m > v);
                const bool best = cand && v > cd.v;
                cd.v = best ? v : cd.v;
                cd.c = best ? c : cd.c;
                const bool g1 = cand && v > cd.p1;
                const bool g2 = cand && !g1 && v > cd.p2;
                cd.p2 = g1 ? cd.p1 : (g2 ? v : cd.p2);
                cd.p1 = g1 ? v : cd.p1;
            }
        }
    } else if (valid) {
        // Branch-free and unrolled, so the nine LDS reads of every candidate
        // issue together.  A neighbour counts only if it is > v; fmaxf skips
        // NaN neighbours, which a > test never counts either.  Candidates are
        // visited in increasing c, with the same strict updates as the
        // reference's loop (first maximum wins, top-2 for the ratio).
#pragma unroll
        for (int it = 0; it < (NC + T - 1) / T; ++it) {
            const int c = t + it * T;
            if constexpr (NC % T != 0) {
                if (c >= NC) break;
            }
            const int i = N / 4 + c / CW, j = N / 4 + c % CW;
            const float *pc = P + i * PX + j;
            const float v = pc[0];
            const float m = fmaxf(fmaxf(fmaxf(pc[-PX - 1], pc[-PX]), fmaxf(pc[-PX + 1], pc[-1])),
                                  fmaxf(fmaxf(pc[1], pc[PX - 1]), fmaxf(pc[PX], pc[PX + 1])));
            const bool vnan = v != v;
            cd.nan = cd.nan || vnan;
            const bool cand = !vnan && v > 0.0f && !(m > v);
            const bool best = cand && v > cd.v;
            cd.v = best ? v : cd.v;
            cd.c = best ? c : cd.c;
            const bool g1 = cand && v > cd.p1;
            const bool g2 = cand && !g1 && v > cd.p2;
            cd.p2 = g1 ? cd.p1 : (g2 ? v : cd.p2);
            cd.p1 = g1 ? v : cd.p1;
        }
    }
    if constexpr (64 % T == 0) {
        // Within a row of 16 lanes the partners come by DPP (quad xor 1 and 2,
        // then the half-row and row mirrors, each pairing two halves already
        // reduced); only wider windows take LDS-routed shuffles.  The merge is
        // order-free (larger value, then smaller index; top two of the multiset),
        // so any pairing that covers the window gives the same summary.
#pragma unroll
        for (int o = 1; o < T; o <<= 1) {
            if (o == 1) cd.merge(cd.template dpp<0xB1>());
            else if (o == 2) cd.merge(cd.template dpp<0x4E>());
            else if (o == 4) cd.merge(cd.template dpp<0x141>());
            else if (o == 8) cd.merge(cd.template dpp<0x140>());
            else cd.merge(cd.shfl(o, T));
        }
    } else if constexpr (T % 64 == 0) {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            if (o == 1) cd.merge(cd.template dpp<0xB1>());
            else if (o == 2) cd.merge(cd.template dpp<0x4E>());
            else if (o == 4) cd.merge(cd.template dpp<0x141>());
            else if (o == 8) cd.merge(cd.template dpp<0x140>());
            else cd.merge(cd.shfl(o, 64));
        }
        if ((t & 63) == 0) {
            r0[t >> 6] = cd.v;
            r1[t >> 6] = __int_as_float(cd.c);
            r2[t >> 6] = cd.p1;
            r3[t >> 6] = cd.p2;
            r4[t >> 6] = (float)cd.nan;
        }
        __syncthreads();
        // every lane merges the per-wave summaries in wave order (the same
        // bits in all lanes, no second barrier)
        cd = Cand{r0[0], __float_as_int(r1[0]), r2[0], r3[0], r4[0] != 0.0f};
        for (int k = 1; k < T / 64; ++k) {
            Cand o = {r0[k], __float_as_int(r1[k]), r2[k], r3[k], r4[k] != 0.0f};
            cd.merge(o);
        }
    } else {
        r0[t] = cd.v;
        r1[t] = __int_as_float(cd.c);
        r2[t] = cd.p1;
        r3[t] = cd.p2;
        r4[t] = (float)cd.nan;
        __syncthreads();
        cd = Cand{r0[0], __float_as_int(r1[0]), r2[0], r3[0], r4[0] != 0.0f};
        for (int k = 1; k < T; ++k) {
            Cand o = {r0[k], __float_as_int(r1[k]), r2[k], r3[k], r4[k] != 0.0f};
            cd.merge(o);
        }
    }
    if (gw >= g.total) return;
    const bool compact = rec_compact(g.finder, g.nPeaks);    // uniform: kernel arguments
    float *rec = g.rec + (size_t)gw * (compact ? kRecC : kRec);
    if (!valid) {
        if (t == 0) {
            if (compact) rec[kRecCPos] = __int_as_float(-1);
            else rec[kRecFlag] = 0.0f;
        }
        return;
    }
    float h = cd.v, det = (cd.p2 > 0.0f) ? cd.p1 / cd.p2 : 0.0f;
    int i0 = 0, j0 = 0;
    if (cd.c != 0x7fffffff) { i0 = N / 4 + cd.c / CW; j0 = N / 4 + cd.c % CW; }
    if (cd.nan) {
        // a NaN in the plane: the sequential scan decides (rare, one lane)
        if (t != 0) return;
        seq_search(P, N, N, PX, &h, &i0, &j0);
        det = seq_detect(P, N, N, PX);
    }
    const bool gate = (h <= 0.0f) || i0 < 2 || i0 >= N - 2 || j0 < 2 || j0 >= N - 2;
    const size_t o = (size_t)n * nWin + w;
    if (gate) {
        if (t == 0) {
            g.ratio[o] = det;
            g.pkX[o] = __builtin_nanf("") - N / 2.0f;
            g.pkY[o] = __builtin_nanf("") - N / 2.0f;
            g.sx[o] = 0.0f; g.sy[o] = 0.0f; g.sxy[o] = 0.0f;
            g.pkH[o] = 0.0f * enorm;
            if (compact) rec[kRecCPos] = __int_as_float(-1);
            else rec[kRecFlag] = 0.0f;
        }
        return;
    }
    const int nl = cd.nan ? 1 : T;
    if (compact) {
        // the cross through the peak times the plane weight (5x5 patch
        // entries 7, 11, 12, 13, 17), spread over the window's lanes
        for (int e = t; e < 5; e += nl) {
            const int r = i0 + (e == 0 ? -1 : (e == 4 ? 1 : 0)), c = j0 + (e >= 1 && e <= 3 ? e - 2 : 0);
            const float v = P[r * PX + c];
            rec[e] = g.pw ? v * g.pw[r * N + c] : v;
        }
        if (t == 0) {
            g.ratio[o] = det;
            rec[kRecCH] = h;
            rec[kRecCEnorm] = enorm;
            rec[kRecCPos] = __int_as_float(i0 * 256 + j0);
        }
        return;
    }
    // the 5x5 patch times the plane weight, spread over the window's lanes
    for (int e = t; e < PK * PK; e += nl) {
        const int r = i0 + e / PK - 2, c = j0 + e % PK - 2;
        const float v = P[r * PX + c];
        rec[e] = g.pw ? v * g.pw[r * N + c] : v;
    }
    if (t == 0) {
        g.ratio[o] = det;
        rec[kRecH] = h;
        rec[kRecI0] = (float)i0;
        rec[kRecJ0] = (float)j0;
        rec[kRecEnorm] = enorm;
        rec[kRecFlag] = 1.0f;
    }
}

// Everything of one window after its weighted values are in registers:
// mean removal, energy, the four transform stages and the epilogue.  zr/zi
// hold the weighted A and B column samples (rows n1*R + q of column l), sa/sb
// this lane's partial sums of them.  Called by all lanes of the workgroup
// (it has barriers).
// With one lane per line (R = 1) and several windows per workgroup, the
// inverse stages run packed: stage 3 has N/2 + 1 columns and stage 4 N/2 row
// pairs per window, so the workgroup's lanes take them window after window
// (sv: per-window valid flags in LDS; Sbr/Sbi: the workgroup's tiles) and the
// waves left without work skip those stages.
template <int N, int R>
__device__ __forceinline__ void xc_window(const XcorrArgs &g, float (&zr)[N / R], float (&zi)[N / R], float sa,
                                          float sb, bool valid, long long gw, int n, int w, int t, int l, int q,
                                          int ob, int lb, const float2 *tw, float *Sr, float *Si, float *r0,
                                          float *r1, float *r2, float *r3, float *r4, float *Sbr = nullptr,
                                          float *Sbi = nullptr, int *sv = nullptr, int wloc = 0)
{
    using C = XcCfg<N, R>;
    constexpr int N1 = C::N1, T = C::T, PX = C::PX;
    constexpr int numel = N * N;
    constexpr bool kPack = R == 1 && C::WPB > 1;
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
    win_sum2<T>(ea, eb, r2, r3, t);     // other slots: the first sum's reads need no barrier
    const float enorm = 1.0f / sqrtf(ea * eb);

    if (valid) {
        fft_split<N, R, false>(zr, zi, q, tw);
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int e = tix<N, R>(k + N1 * ob, ob, l, lb);
            Sr[e] = zr[k];
            Si[e] = zi[k];
        }
    }
    if constexpr (kPack) {
        if (t == 0) sv[wloc] = valid ? 1 : 0;
    }
    __syncthreads();

    // ---- 2. row ky = l --------------------------------------------------------
    if (valid) {
        float xr[N1], xi[N1];
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int e = tix<N, R>(l, lb, k * R + q, (k * R) / N1);
            xr[k] = Sr[e];
            xi[k] = Si[e];
        }
        fft_split<N, R, false>(xr, xi, q, tw);
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int e = tix<N, R>(l, lb, k + N1 * ob, ob);
            Sr[e] = xr[k];
            Si[e] = xi[k];
        }
    }
    __syncthreads();

    // ---- 3. column kx = l <= N/2: C = FB conj(FA), inverse column FFT ------
    auto stage3 = [&](int l, float *Sr, float *Si) {
        const int mx = (N - l) % N, mxb = mx / N1;
        float cr[N1], ci[N1];
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int ky = k * R + q;
            // (N - ky) mod N = (N1 - k) R - q; affine in k except at ky = 0
            const int my = k > 0 ? N - ky : (q == 0 ? 0 : N - q);
            const int myb = k > 0 ? mirror_block<N1, R>(N1 - k, q) : (q == 0 ? 0 : R - 1);
            const int ez = tix<N, R>(ky, (k * R) / N1, l, lb), em = tix<N, R>(my, myb, mx, mxb);
            const float zR = Sr[ez], zI = Si[ez];
            const float mR = Sr[em], mI = Si[em];
            // FA = (Z + conj(Zm)) / 2 ; FB = (Z - conj(Zm)) / (2i)
            const float ar = 0.5f * (zR + mR), ai = 0.5f * (zI - mI);
            const float br = 0.5f * (zI + mI), bi = -0.5f * (zR - mR);
            cr[k] = br * ar + bi * ai;
            ci[k] = bi * ar - br * ai;
        }
        fft_split<N, R, true>(cr, ci, q, tw);
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int e = tix<N, R>(k + N1 * ob, ob, l, lb);
            Sr[e] = cr[k];
            Si[e] = ci[k];
        }
    };
    if constexpr (kPack) {
        // columns 0 .. N/2-1 of every window, window after window (2H whole
        // windows per half wave, at distinct bank offsets), then the N/2
        // column of every window
        constexpr int C3 = N / 2;
        const int v = (int)threadIdx.x;
        int wl = -1, col = 0;
        if (v < C::WPB * C3) {
            wl = v / C3;
            col = v - wl * C3;
        } else if (v < C::WPB * C3 + C::WPB) {
            wl = v - C::WPB * C3;
            col = N / 2;
        }
        if (wl >= 0 && sv[wl]) stage3(col, Sbr + C::wbase(wl), Sbi + C::wbase(wl));
    } else {
        if (valid && l <= N / 2) stage3(l, Sr, Si);
    }
    __syncthreads();

    // ---- 4. rows j and j + N/2 (j = l < N/2): Hermitian rows -> real rows -----
    auto load4 = [&](int l, const float *Sr, const float *Si, float (&xr)[N1], float (&xi)[N1]) {
        const int ya = l, yb = l + N / 2;
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            // kx = k R + q; with 0 <= q < R the tests below fold to compile
            // time except at k = 0 and k = N1/2
            const int kx = k * R + q;
            const bool up = 2 * k > N1 || (2 * k == N1 && q > 0);
            const int kk = 2 * k < N1 ? kx : N - kx;        // N - kx = N/2 at kx = N/2
            const int kb = 2 * k < N1 ? (k * R) / N1 : mirror_block<N1, R>(N1 - k, q);
            const int ea = tix<N, R>(ya, lb, kk, kb), eb = tix<N, R>(yb, lb + R / 2, kk, kb);
            float g1r = Sr[ea], g1i = Si[ea];
            float g2r = Sr[eb], g2i = Si[eb];
            if ((k == 0 || 2 * k == N1) && q == 0) { g1i = 0.0f; g2i = 0.0f; }   // kx = 0, N/2
            if (up) { g1i = -g1i; g2i = -g2i; }     // G(N-k) = conj(G(k)) per real row
            xr[k] = g1r - g2i;
            xi[k] = g1i + g2r;
        }
    };
    auto store4 = [&](int l, float *Sr, float (&xr)[N1], float (&xi)[N1]) {
        const int ya = l, yb = l + N / 2;
        fft_split<N, R, true>(xr, xi, q, tw);
        const float mul = 1.0f / (float)numel;
        // plane[(y + N/2) % N][(x + N/2) % N] = c[y][x] / numel
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int xs = (k + N1 * ob + N / 2) % N;
            Sr[yb * PX + xs] = xr[k] * mul;
            Sr[ya * PX + xs] = xi[k] * mul;
        }
    };
    if constexpr (kPack) {
        // R = 1: each row pair is read and written by the same lane (the
        // plane overlays only that lane's two rows), so no barrier between
        constexpr int C4 = N / 2;
        const int v = (int)threadIdx.x;
        if (v < C::WPB * C4) {
            const int wl = v / C4;
            if (sv[wl]) {
                float xr[N1], xi[N1];
                load4(v - wl * C4, Sbr + C::wbase(wl), Sbi + C::wbase(wl), xr, xi);
                store4(v - wl * C4, Sbr + C::wbase(wl), xr, xi);
            }
        }
    } else {
        const bool act4 = valid && l < N / 2;
        float xr[N1], xi[N1];
        if (act4) load4(l, Sr, Si, xr, xi);
        // R > 1: the plane (pitch PX) overlays tile words other lines still read
        if constexpr (R > 1) __syncthreads();
        if (act4) store4(l, Sr, xr, xi);
    }
    __syncthreads();

    plane_epilogue<N, T>(g, Sr, PX, valid, gw, n, w, enorm, t, r0, r1, r2, r3, r4);
}
// xc_window with one real LDS plane of C::tile words per window instead of
// the complex pair (72 KB instead of 144 KB at N = 128).  Every exchange
// between stages moves the real parts through the plane, then the imaginary
// parts, with a barrier between, so two 512-lane workgroups (16 waves) share
// a CU where the complex tile allowed one.  The arithmetic is xc_window's, in
// the same order; split-line layout only (R > 1, one window per workgroup).
template <int N, int R>
__device__ __forceinline__ void xc_window_half(const XcorrArgs &g, float (&zr)[N / R], float (&zi)[N / R],
                                               float sa, float sb, bool valid, long long gw, int n, int w, int t,
                                               int l, int q, int ob, int lb, const float2 *tw, float *S, float *r0,
                                               float *r1, float *r2, float *r3, float *r4)
{
    using C = XcCfg<N, R>;
    constexpr int N1 = C::N1, T = C::T, PX = C::PX;
    constexpr int numel = N * N;
    static_assert(R > 1 && C::WPB == 1, "split-line layout, one window per workgroup");
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
    win_sum2<T>(ea, eb, r2, r3, t);
    const float enorm = 1.0f / sqrtf(ea * eb);

    // ---- 1 -> 2: column spectra to rows, real parts then imaginary parts ----
    if (valid) {
        fft_split<N, R, false>(zr, zi, q, tw);
#pragma unroll
        for (int k = 0; k < N1; ++k) S[tix<N, R>(k + N1 * ob, ob, l, lb)] = zr[k];
    }
    __syncthreads();
    float xr[N1], xi[N1];
    if (valid) {
#pragma unroll
        for (int k = 0; k < N1; ++k) xr[k] = S[tix<N, R>(l, lb, k * R + q, (k * R) / N1)];
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int k = 0; k < N1; ++k) S[tix<N, R>(k + N1 * ob, ob, l, lb)] = zi[k];
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int k = 0; k < N1; ++k) xi[k] = S[tix<N, R>(l, lb, k * R + q, (k * R) / N1)];
        // ---- 2. row ky = l ----
        fft_split<N, R, false>(xr, xi, q, tw);
    }
    __syncthreads();

    // ---- 2 -> 3: column kx = l <= N/2 and its mirror, C = FB conj(FA) -----
    if (valid) {
#pragma unroll
        for (int k = 0; k < N1; ++k) S[tix<N, R>(l, lb, k + N1 * ob, ob)] = xr[k];
    }
    __syncthreads();
    const bool act3 = valid && l <= N / 2;
    const int mx = (N - l) % N, mxb = mx / N1;
    auto ez_of = [&](int k) { return tix<N, R>(k * R + q, (k * R) / N1, l, lb); };
    auto em_of = [&](int k) {
        const int ky = k * R + q;
        const int my = k > 0 ? N - ky : (q == 0 ? 0 : N - q);
        const int myb = k > 0 ? mirror_block<N1, R>(N1 - k, q) : (q == 0 ? 0 : R - 1);
        return tix<N, R>(my, myb, mx, mxb);
    };
    float cr[N1], ci[N1];     // first zR + mR and zR - mR, then the cross spectrum
    if (act3) {
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const float zR = S[ez_of(k)], mR = S[em_of(k)];
            cr[k] = zR + mR;
            ci[k] = zR - mR;
        }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int k = 0; k < N1; ++k) S[tix<N, R>(l, lb, k + N1 * ob, ob)] = xi[k];
    }
    __syncthreads();
    if (act3) {
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const float zI = S[ez_of(k)], mI = S[em_of(k)];
            // FA = (Z + conj(Zm)) / 2 ; FB = (Z - conj(Zm)) / (2i)
            const float ar = 0.5f * cr[k], ai = 0.5f * (zI - mI);
            const float br = 0.5f * (zI + mI), bi = -0.5f * ci[k];
            cr[k] = br * ar + bi * ai;
            ci[k] = bi * ar - br * ai;
        }
        // ---- 3. inverse column FFT ----
        fft_split<N, R, true>(cr, ci, q, tw);
    }
    __syncthreads();

    // ---- 3 -> 4: rows j and j + N/2 (j = l < N/2): Hermitian rows -> real rows
    if (act3) {
#pragma unroll
        for (int k = 0; k < N1; ++k) S[tix<N, R>(k + N1 * ob, ob, l, lb)] = cr[k];
    }
    __syncthreads();
    const bool act4 = valid && l < N / 2;
    const int ya = l, yb = l + N / 2;
    auto kk_of = [&](int k) {
        const int kx = k * R + q;
        return 2 * k < N1 ? kx : N - kx;
    };
    auto kb_of = [&](int k) { return 2 * k < N1 ? (k * R) / N1 : mirror_block<N1, R>(N1 - k, q); };
    float g1[N1], g2[N1];     // real parts of rows ya and yb, then the packed row
    if (act4) {
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            g1[k] = S[tix<N, R>(ya, lb, kk_of(k), kb_of(k))];
            g2[k] = S[tix<N, R>(yb, lb + R / 2, kk_of(k), kb_of(k))];
        }
    }
    __syncthreads();
    if (act3) {
#pragma unroll
        for (int k = 0; k < N1; ++k) S[tix<N, R>(k + N1 * ob, ob, l, lb)] = ci[k];
    }
    __syncthreads();
    if (act4) {
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const bool up = 2 * k > N1 || (2 * k == N1 && q > 0);
            float g1i = S[tix<N, R>(ya, lb, kk_of(k), kb_of(k))];
            float g2i = S[tix<N, R>(yb, lb + R / 2, kk_of(k), kb_of(k))];
            if ((k == 0 || 2 * k == N1) && q == 0) { g1i = 0.0f; g2i = 0.0f; }   // kx = 0, N/2
            if (up) { g1i = -g1i; g2i = -g2i; }     // G(N-k) = conj(G(k)) per real row
            const float g1r = g1[k], g2r = g2[k];
            g1[k] = g1r - g2i;
            g2[k] = g1i + g2r;
        }
        // ---- 4. inverse row FFT ----
        fft_split<N, R, true>(g1, g2, q, tw);
    }
    __syncthreads();    // the plane (pitch PX) overlays tile words other lines still read
    if (act4) {
        const float mul = 1.0f / (float)numel;
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            const int xs = (k + N1 * ob + N / 2) % N;
            S[yb * PX + xs] = g1[k] * mul;
            S[ya * PX + xs] = g2[k] * mul;
        }
    }
    __syncthreads();

    plane_epilogue<N, T>(g, S, PX, valid, gw, n, w, enorm, t, r0, r1, r2, r3, r4);
}

// xc_window_half for one lane per line (R = 1) with several windows per
// workgroup: one real LDS plane per window, real then imaginary parts at
// every stage exchange, and the inverse stages packed over the workgroup's
// windows as in xc_window (sv: per-window valid flags; Sb: the workgroup's
// planes).  Half the LDS of the complex tiles, so twice the workgroups fit
// a CU by LDS.  Same arithmetic in the same order as xc_window.
template <int N>
__device__ __forceinline__ void xc_window_half1(const XcorrArgs &g, float (&zr)[N], float (&zi)[N], float sa,
                                                float sb, bool valid, long long gw, int n, int w, int t, int l,
                                                float *S, float *Sb, int *sv, int wloc, float *r0, float *r1,
                                                float *r2, float *r3, float *r4)
{
    using C = XcCfg<N, 1>;
    constexpr int T = C::T, PX = C::PX, TPX = C::TPX, WPB = C::WPB;
    constexpr int numel = N * N;
    static_assert(C::R == 1 && WPB > 1 && TPX == PX, "one lane per line, packed inverse stages");
    win_sum2<T>(sa, sb, r0, r1, t);
    const float meanA = sa / (float)numel;
    const float meanB = sb / (float)numel;
    float ea = 0.0f, eb = 0.0f;
    if (valid) {
        if (!g.ensemble) {
#pragma unroll
            for (int k = 0; k < N; ++k) {
                zr[k] -= meanA;
                zi[k] -= meanB;
                ea += zr[k] * zr[k];
                eb += zi[k] * zi[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < N; ++k) {
                ea += zr[k] * zr[k];
                eb += zi[k] * zi[k];
            }
        }
    }
    win_sum2<T>(ea, eb, r2, r3, t);
    const float enorm = 1.0f / sqrtf(ea * eb);
    if (t == 0) sv[wloc] = valid ? 1 : 0;

    // ---- 1 -> 2: column spectra to rows ------------------------------------
    if (valid) {
        fft_split<N, 1, false>(zr, zi, 0, nullptr);
#pragma unroll
        for (int k = 0; k < N; ++k) S[k * TPX + l] = zr[k];
    }
    __syncthreads();
    float xr[N], xi[N];
    if (valid) {
#pragma unroll
        for (int k = 0; k < N; ++k) xr[k] = S[l * TPX + k];
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int k = 0; k < N; ++k) S[k * TPX + l] = zi[k];
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int k = 0; k < N; ++k) xi[k] = S[l * TPX + k];
        fft_split<N, 1, false>(xr, xi, 0, nullptr);      // ---- 2. row ky = l
    }
    __syncthreads();

    // ---- 2 -> 3: packed columns kx <= N/2 of every window -------------------
    if (valid) {
#pragma unroll
        for (int k = 0; k < N; ++k) S[l * TPX + k] = xr[k];
    }
    constexpr int C3 = N / 2;
    const int v = (int)threadIdx.x;
    int wl = -1, col = 0;
    if (v < WPB * C3) {
        wl = v / C3;
        col = v - wl * C3;
    } else if (v < WPB * C3 + WPB) {
        wl = v - WPB * C3;
        col = N / 2;
    }
    __syncthreads();
    const bool act3 = wl >= 0 && sv[wl];
    float *S3 = Sb + C::wbase(wl < 0 ? 0 : wl);
    const int mx = (N - col) % N;
    float cr[N], ci[N];
    if (act3) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int my = (N - k) % N;
            const float zR = S3[k * TPX + col], mR = S3[my * TPX + mx];
            cr[k] = zR + mR;
            ci[k] = zR - mR;
        }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int k = 0; k < N; ++k) S[l * TPX + k] = xi[k];
    }
    __syncthreads();
    if (act3) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int my = (N - k) % N;
            const float zI = S3[k * TPX + col], mI = S3[my * TPX + mx];
            // FA = (Z + conj(Zm)) / 2 ; FB = (Z - conj(Zm)) / (2i)
            const float ar = 0.5f * cr[k], ai = 0.5f * (zI - mI);
            const float br = 0.5f * (zI + mI), bi = -0.5f * ci[k];
            cr[k] = br * ar + bi * ai;
            ci[k] = bi * ar - br * ai;
        }
        fft_split<N, 1, true>(cr, ci, 0, nullptr);       // ---- 3. inverse column FFT
    }
    __syncthreads();

    // ---- 3 -> 4: packed row pairs j, j + N/2 of every window ----------------
    if (act3) {
#pragma unroll
        for (int k = 0; k < N; ++k) S3[k * TPX + col] = cr[k];
    }
    constexpr int C4 = N / 2;
    const int w4 = v < WPB * C4 ? v / C4 : -1;
    const int j4 = v - (w4 < 0 ? 0 : w4) * C4;
    __syncthreads();
    const bool act4 = w4 >= 0 && sv[w4];
    float *S4 = Sb + C::wbase(w4 < 0 ? 0 : w4);
    const int ya = j4, yb = j4 + N / 2;
    float g1[N], g2[N];
    if (act4) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int kk = 2 * k < N ? k : N - k;
            g1[k] = S4[ya * TPX + kk];
            g2[k] = S4[yb * TPX + kk];
        }
    }
    __syncthreads();
    if (act3) {
#pragma unroll
        for (int k = 0; k < N; ++k) S3[k * TPX + col] = ci[k];
    }
    __syncthreads();
    if (act4) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int kk = 2 * k < N ? k : N - k;
            float g1i = S4[ya * TPX + kk], g2i = S4[yb * TPX + kk];
            if (k == 0 || 2 * k == N) { g1i = 0.0f; g2i = 0.0f; }     // kx = 0, N/2
            if (2 * k > N) { g1i = -g1i; g2i = -g2i; }                // G(N-k) = conj(G(k))
            const float g1r = g1[k], g2r = g2[k];
            g1[k] = g1r - g2i;
            g2[k] = g1i + g2r;
        }
        fft_split<N, 1, true>(g1, g2, 0, nullptr);       // ---- 4. inverse row FFT
        // the plane rows ya, yb overlay exactly the tile rows this lane read
        const float mul = 1.0f / (float)numel;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int xs = (k + N / 2) % N;
            S4[yb * PX + xs] = g1[k] * mul;
            S4[ya * PX + xs] = g2[k] * mul;
        }
    }
    __syncthreads();

    plane_epilogue<N, T>(g, S, PX, valid, gw, n, w, enorm, t, r0, r1, r2, r3, r4);
}

// The K weighted samples of one lane's column of a window: elements off,
// off + step, ... of the image stack, times the weights w[0], w[wstep], ...
// As raw buffer loads (one resource over the whole stack, uniform for the
// workgroup; the lane's byte offset in a VGPR; the row step a uniform
// scalar offset) no load needs its own 64-bit address arithmetic.  Stacks of
// 2^32 bytes or more take plain loads.  same_w (uniform: the caller passed
// one weight array as A and B, as the instantaneous path does) reads each
// weight once.
template <int K, typename TI>
__device__ __forceinline__ void load_window(const TI *sa_, const TI *sb_, size_t elems, size_t off, size_t step,
                                            const float *wa, const float *wb, int wstep, bool same_w,
                                            float (&zr)[K], float (&zi)[K], float &sa, float &sb)
{
    // each sample is weighted as soon as it arrives (few live registers)
    auto run = [&](auto same, auto load) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            float a, b;
            load(k, a, b);
            const float w1 = wa[k * wstep];
            const float w2 = decltype(same)::value ? w1 : wb[k * wstep];
            a = a * w1;
            b = b * w2;
            zr[k] = a;
            zi[k] = b;
            sa += a;
            sb += b;
        }
    };
    auto go = [&](auto load) {
        if (same_w) run(std::true_type{}, load);
        else run(std::false_type{}, load);
    };
    const size_t bytes = elems * sizeof(TI);
    if (bytes < ((size_t)1 << 32)) {
        const __amdgpu_buffer_rsrc_t ra =
            __builtin_amdgcn_make_buffer_rsrc((void *)sa_, (short)0, (int)(uint32_t)bytes, 0x00020000);
        const __amdgpu_buffer_rsrc_t rb =
            __builtin_amdgcn_make_buffer_rsrc((void *)sb_, (short)0, (int)(uint32_t)bytes, 0x00020000);
        const uint32_t vo = (uint32_t)(off * sizeof(TI)), so = (uint32_t)(step * sizeof(TI));
        go([&](int k, float &a, float &b) {
            if constexpr (sizeof(TI) == 2) {
                a = (float)__builtin_amdgcn_raw_buffer_load_b16(ra, vo, (uint32_t)k * so, 0);
                b = (float)__builtin_amdgcn_raw_buffer_load_b16(rb, vo, (uint32_t)k * so, 0);
            } else {
                a = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ra, vo, (uint32_t)k * so, 0));
                b = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rb, vo, (uint32_t)k * so, 0));
            }
        });
    } else {
        const TI *pa = sa_ + off, *pb = sb_ + off;
        go([&](int k, float &a, float &b) {
            a = (float)pa[(size_t)k * step];
            b = (float)pb[(size_t)k * step];
        });
    }
}

// Window geometry of global window gw (image n, window w, top-left corner);
// false when masked, out of the image or past the end.
template <int N>
__device__ __forceinline__ bool xc_geom(const XcorrArgs &g, long long gw, int &n, int &w, int &row0, int &col0)
{
    n = 0; w = 0; row0 = 0; col0 = 0;
    if (gw >= g.total) return false;
    const int nWin = g.nWy * g.nWx;
    n = (int)(gw / nWin);
    w = (int)(gw - (long long)n * nWin);
    const int ii = w % g.nWx, jj = w / g.nWx;
    bool valid = true;
    if (g.mask && g.mask[jj * g.nWx + ii] == 1.0f) valid = false;
    row0 = (int)floor((double)g.ctrY[jj] - ((double)(float)N - 1.0) / 2.0 + 0.5);
    col0 = (int)floor((double)g.ctrX[ii] - ((double)(float)N - 1.0) / 2.0 + 0.5);
    if (row0 < 0 || col0 < 0 || row0 + N > g.H || col0 + N > g.W) valid = false;
    return valid;
}

template <int N, int R_>
__device__ __forceinline__ void xcorr_body(const XcorrArgs &g, const int b)
{
    using C = XcCfg<N, R_>;
    constexpr int R = C::R, N1 = C::N1, T = C::T, WPB = C::WPB;
    __shared__ float s_re[C::tiles];
    __shared__ float s_im[C::tiles];
    __shared__ float s_red[5][WPB * C::RED];
    __shared__ __attribute__((aligned(16))) float2 s_tw[R > 1 ? R * C::TWP : 1];   // split twiddles

    if constexpr (R > 1) {
        for (int e = (int)threadIdx.x; e < N; e += C::BLOCK)
            s_tw[(e / N1) * C::TWP + e % N1] = make_float2(kSplitTw<N, R>.v[2 * e], kSplitTw<N, R>.v[2 * e + 1]);
        __syncthreads();
    }

    const int wloc = (int)threadIdx.x / T;
    const int t = (int)threadIdx.x - wloc * T;
    int l, q;                         // line (row or column), lane within the line
    C::lane_map(t, l, q);
    const int ob = out_block<R>(q);   // output block of fft_split
    const int lb = l / N1;            // block of this lane's line index
    const float2 *tw = s_tw + (R > 1 ? q * C::TWP : 0);
    float *Sr = s_re + C::wbase(wloc);
    float *Si = s_im + C::wbase(wloc);
    float *r0 = &s_red[0][wloc * C::RED], *r1 = &s_red[1][wloc * C::RED];
    float *r2 = &s_red[2][wloc * C::RED], *r3 = &s_red[3][wloc * C::RED];
    float *r4 = &s_red[4][wloc * C::RED];

    __shared__ int s_valid[WPB];
    const long long gw = (long long)b * WPB + wloc;
    int n, w, row0, col0;
    const bool valid = xc_geom<N>(g, gw, n, w, row0, col0);

    // ---- 1. column x = l: rows n1*R + q, weight -----------------------------
    float zr[N1], zi[N1];
    float sa = 0.0f, sb = 0.0f;
    if (valid) {
        const size_t plane = (size_t)g.H * g.W, elems = (size_t)(g.total / ((long long)g.nWy * g.nWx)) * plane;
        const size_t off = (size_t)n * plane + (size_t)(row0 + q) * g.W + col0 + l, step = (size_t)R * g.W;
        const float *wa = g.wA + q * N + l, *wb = g.wB + q * N + l;
        // 16-bit counts convert exactly, so both inputs give the same bits
        if (g.imgA16) load_window<N1>(g.imgA16, g.imgB16, elems, off, step, wa, wb, R * N, g.wA == g.wB, zr, zi, sa, sb);
        else load_window<N1>(g.imgA, g.imgB, elems, off, step, wa, wb, R * N, g.wA == g.wB, zr, zi, sa, sb);
    } else {
#pragma unroll
        for (int k = 0; k < N1; ++k) { zr[k] = 0.0f; zi[k] = 0.0f; }
    }
    xc_window<N, R>(g, zr, zi, sa, sb, valid, gw, n, w, t, l, q, ob, lb, tw, Sr, Si, r0, r1, r2, r3, r4, s_re,
                    s_im, s_valid, wloc);
}

template <int N, int R_>
__global__ void __launch_bounds__((XcCfg<N, R_>::BLOCK))
xcorr_kernel(XcorrArgs g)
{
    const int b = remap_block((int)blockIdx.x, g.nblocks);
    if (b < 0) return;    // whole block exits together
    xcorr_body<N, R_>(g, b);
}

// The correlation kernel with xc_window_half: one window per 512-lane
// workgroup in a single real plane, bounded to 128 VGPRs (four waves per
// SIMD) so that two workgroups share a CU.
template <int N, int R_>
__device__ __forceinline__ void xcorr_half_body(const XcorrArgs &g, const int b)
{
    using C = XcCfg<N, R_>;
    constexpr int R = C::R, N1 = C::N1;
    static_assert(C::WPB == 1 && R > 1, "one window per workgroup, split lines");
    __shared__ float s_pl[C::tile];
    __shared__ float s_red[5][C::RED];
    __shared__ __attribute__((aligned(16))) float2 s_tw[R * C::TWP];

    for (int e = (int)threadIdx.x; e < N; e += C::BLOCK)
        s_tw[(e / N1) * C::TWP + e % N1] = make_float2(kSplitTw<N, R>.v[2 * e], kSplitTw<N, R>.v[2 * e + 1]);
    __syncthreads();

    const int t = (int)threadIdx.x;
    int l, q;
    C::lane_map(t, l, q);
    const int ob = out_block<R>(q);
    const int lb = l / N1;
    const float2 *tw = s_tw + q * C::TWP;
    const long long gw = (long long)b;
    int n, w, row0, col0;
    const bool valid = xc_geom<N>(g, gw, n, w, row0, col0);

    float zr[N1], zi[N1];
    float sa = 0.0f, sb = 0.0f;
    if (valid) {
        const size_t plane = (size_t)g.H * g.W, elems = (size_t)(g.total / ((long long)g.nWy * g.nWx)) * plane;
        const size_t off = (size_t)n * plane + (size_t)(row0 + q) * g.W + col0 + l, step = (size_t)R * g.W;
        const float *wa = g.wA + q * N + l, *wb = g.wB + q * N + l;
        // 16-bit counts convert exactly, so both inputs give the same bits
        if (g.imgA16) load_window<N1>(g.imgA16, g.imgB16, elems, off, step, wa, wb, R * N, g.wA == g.wB, zr, zi, sa, sb);
        else load_window<N1>(g.imgA, g.imgB, elems, off, step, wa, wb, R * N, g.wA == g.wB, zr, zi, sa, sb);
    } else {
#pragma unroll
        for (int k = 0; k < N1; ++k) { zr[k] = 0.0f; zi[k] = 0.0f; }
    }
    xc_window_half<N, R>(g, zr, zi, sa, sb, valid, gw, n, w, t, l, q, ob, lb, tw, s_pl, s_red[0], s_red[1],
                         s_red[2], s_red[3], s_red[4]);
}

template <int N, int R_>
__global__ void __launch_bounds__((XcCfg<N, R_>::BLOCK), (2 * XcCfg<N, R_>::BLOCK / 256))
xcorr_kernel_half(XcorrArgs g)
{
    const int b = remap_block((int)blockIdx.x, g.nblocks);
    if (b < 0) return;    // whole block exits together
    xcorr_half_body<N, R_>(g, b);
}

// The correlation kernel with xc_window_half1 (one lane per line, several
// windows per workgroup, one real plane per window).
template <int N>
__device__ __forceinline__ void xcorr_half1_body(const XcorrArgs &g, const int b)
{
    using C = XcCfg<N, 1>;
    constexpr int WPB = C::WPB;
    __shared__ float s_pl[C::tiles];
    __shared__ float s_red[5][WPB * C::RED];
    __shared__ int s_valid[WPB];

    const int wloc = (int)threadIdx.x / N;
    const int t = (int)threadIdx.x - wloc * N;
    const int l = t;
    float *S = s_pl + C::wbase(wloc);
    float *r0 = &s_red[0][wloc * C::RED], *r1 = &s_red[1][wloc * C::RED];
    float *r2 = &s_red[2][wloc * C::RED], *r3 = &s_red[3][wloc * C::RED];
    float *r4 = &s_red[4][wloc * C::RED];
    const long long gw = (long long)b * WPB + wloc;
    int n, w, row0, col0;
    const bool valid = xc_geom<N>(g, gw, n, w, row0, col0);

    float zr[N], zi[N];
    float sa = 0.0f, sb = 0.0f;
    if (valid) {
        const size_t plane = (size_t)g.H * g.W, elems = (size_t)(g.total / ((long long)g.nWy * g.nWx)) * plane;
        const size_t off = (size_t)n * plane + (size_t)row0 * g.W + col0 + l, step = (size_t)g.W;
        const float *wa = g.wA + l, *wb = g.wB + l;
        if (g.imgA16) load_window<N>(g.imgA16, g.imgB16, elems, off, step, wa, wb, N, g.wA == g.wB, zr, zi, sa, sb);
        else load_window<N>(g.imgA, g.imgB, elems, off, step, wa, wb, N, g.wA == g.wB, zr, zi, sa, sb);
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) { zr[k] = 0.0f; zi[k] = 0.0f; }
    }
    xc_window_half1<N>(g, zr, zi, sa, sb, valid, gw, n, w, t, l, S, s_pl, s_valid, wloc, r0, r1, r2, r3, r4);
}

template <int N>
__global__ void __launch_bounds__((XcCfg<N, 1>::BLOCK), 3)
xcorr_kernel_half1(XcorrArgs g)
{
    const int b = remap_block((int)blockIdx.x, g.nblocks);
    if (b < 0) return;    // whole block exits together
    xcorr_half1_body<N>(g, b);
}

// Ragged forms of the three kernels: workgroup b of the flattened grid
// (renumbered so each XCD takes a contiguous run) runs tile b - start[j]
// of its job j.
template <int N, int R_>
__global__ void __launch_bounds__((XcCfg<N, R_>::BLOCK))
xcorr_ragged(RaggedArgs ra)
{
    const int b = remap_block((int)blockIdx.x, ra.nblocks);
    if (b < 0) return;
    const int j = ragged_job(ra, b);
    xcorr_body<N, R_>(ra.job[j], b - ra.start[j]);
}

template <int N, int R_>
__global__ void __launch_bounds__((XcCfg<N, R_>::BLOCK), (2 * XcCfg<N, R_>::BLOCK / 256))
xcorr_ragged_half(RaggedArgs ra)
{
    const int b = remap_block((int)blockIdx.x, ra.nblocks);
    if (b < 0) return;
    const int j = ragged_job(ra, b);
    xcorr_half_body<N, R_>(ra.job[j], b - ra.start[j]);
}

template <int N>
__global__ void __launch_bounds__((XcCfg<N, 1>::BLOCK), 3)
xcorr_ragged_half1(RaggedArgs ra)
{
    const int b = remap_block((int)blockIdx.x, ra.nblocks);
    if (b < 0) return;
    const int j = ragged_job(ra, b);
    xcorr_half1_body<N>(ra.job[j], b - ra.start[j]);
}

// Persistent form for one window per workgroup (WPB = 1): each workgroup
// walks a contiguous run of windows (runs of one XCD are adjacent) and
// loads the next window's raw samples into registers while the current one
// goes through its transform stages, so the global-load latency of stage 1
// is hidden behind LDS and VALU work.  TI: float or 16-bit frames (the raw
// samples stay in their load registers until the next iteration converts
// them).  The arithmetic per window is that of xcorr_kernel.
template <int N, int R_, typename TI>
__global__ void __launch_bounds__((XcCfg<N, R_>::BLOCK))
xcorr_kernel_pf(XcorrArgs g)
{
    using C = XcCfg<N, R_>;
    constexpr int R = C::R, N1 = C::N1;
    static_assert(C::WPB == 1, "one window per workgroup");
    __shared__ float s_re[C::tile];
    __shared__ float s_im[C::tile];
    __shared__ float s_red[5][C::RED];
    __shared__ __attribute__((aligned(16))) float2 s_tw[R > 1 ? R * C::TWP : 1];

    if constexpr (R > 1) {
        for (int e = (int)threadIdx.x; e < N; e += C::BLOCK)
            s_tw[(e / N1) * C::TWP + e % N1] = make_float2(kSplitTw<N, R>.v[2 * e], kSplitTw<N, R>.v[2 * e + 1]);
        __syncthreads();
    }
    const int t = (int)threadIdx.x;
    int l, q;
    C::lane_map(t, l, q);
    const int ob = out_block<R>(q);
    const int lb = l / N1;
    const float2 *tw = s_tw + (R > 1 ? q * C::TWP : 0);

    // contiguous run of windows of this workgroup; runs of one XCD adjacent
    const int P = (int)gridDim.x;
    const int pidx = ((int)blockIdx.x & 7) * (P >> 3) + ((int)blockIdx.x >> 3);
    const long long per = (g.total + P - 1) / P;
    long long gw = (long long)pidx * per;
    const long long gend = gw + per < g.total ? gw + per : g.total;
    if (gw >= gend) return;    // whole block

    const TI *srcA = nullptr, *srcB = nullptr;
    if constexpr (sizeof(TI) == 2) { srcA = (const TI *)g.imgA16; srcB = (const TI *)g.imgB16; }
    else { srcA = (const TI *)g.imgA; srcB = (const TI *)g.imgB; }
    const float *wa = g.wA + q * N + l, *wb = g.wB + q * N + l;

    TI ra[N1], rb[N1];     // raw samples of the window in flight
    int n, w, row0, col0;
    bool valid = xc_geom<N>(g, gw, n, w, row0, col0);
    auto fetch = [&](bool v, int n_, int r0_, int c0_) {
        if (v) {
            const size_t img = (size_t)n_ * g.H * g.W + (size_t)(r0_ + q) * g.W + c0_ + l;
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                ra[k] = srcA[img + (size_t)k * R * g.W];
                rb[k] = srcB[img + (size_t)k * R * g.W];
            }
        }
    };
    fetch(valid, n, row0, col0);
    for (; gw < gend; ++gw) {
        float zr[N1], zi[N1];
        float sa = 0.0f, sb = 0.0f;
        // An opaque zero added to the lane indices and weight pointers keeps
        // their address arithmetic inside the loop: hoisted out of it, the
        // per-element addresses would stay live across the loop and spill.
        int z0;
        asm volatile("v_mov_b32 %0, 0" : "=v"(z0));
        if (valid) {
            const float *wai = wa + z0, *wbi = wb + z0;
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                const float a = (float)ra[k] * wai[k * R * N];
                const float bb = (float)rb[k] * wbi[k * R * N];
                zr[k] = a;
                zi[k] = bb;
                sa += a;
                sb += bb;
            }
        } else {
#pragma unroll
            for (int k = 0; k < N1; ++k) { zr[k] = 0.0f; zi[k] = 0.0f; }
        }
        const bool cur_valid = valid;
        const int cur_n = n, cur_w = w;
        if (gw + 1 < gend) {
            valid = xc_geom<N>(g, gw + 1, n, w, row0, col0);
            fetch(valid, n, row0, col0);
        }
        xc_window<N, R>(g, zr, zi, sa, sb, cur_valid, gw, cur_n, cur_w, t + z0, l + z0, q + z0, ob + z0,
                        lb + z0, tw + z0, s_re + z0, s_im + z0, s_red[0], s_red[1], s_red[2], s_red[3], s_red[4]);
        __syncthreads();    // the next window overwrites the plane and the reduction slots
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
    if (rec_compact(finder, 1)) {
        const float *rc = rec_all + (size_t)gw * kRecC;
        const int pos = __float_as_int(rc[kRecCPos]);
        if (pos < 0) return;
        float sub[PK * PK];    // threept reads entries 7, 11, 12, 13 and 17 only
        sub[7] = rc[0]; sub[11] = rc[1]; sub[12] = rc[2]; sub[13] = rc[3]; sub[17] = rc[4];
        float pk[2], A, s0, s1;
        threept(sub, pk, &A, &s0, &s1);
        const size_t o = (size_t)gw;
        pkX[o] = (pk[1] + (float)(pos & 255)) - winX / 2.0f;
        pkY[o] = (pk[0] + (float)(pos >> 8)) - winY / 2.0f;
        sx[o] = s0; sy[o] = s1; sxy[o] = 0.0f;
        pkH[o] = rc[kRecCH] * rc[kRecCEnorm];
        return;
    }
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

// ---- rectangular windows (NY != NX, each a built size) ---------------------
// The codelet engine binds row length W and column length H independently
// (pivtools_cli/lib/codelet_fft.c:52-70).  Rectangular windows take this
// general kernel: one 256-lane workgroup per window, the complex tile
// z = a + ib in LDS, and each 1D transform stage as direct DFT sums over the
// line (twiddles from a per-length table computed in FP64 and rounded), every
// lane accumulating its share of the stage's outputs in registers before the
// tile is overwritten.  It does O(N) work per output instead of the FFT's
// O(log N), so it trades speed for generality; the plane it writes (raw,
// fftshifted, / numel) goes to the multi-peak search and fit.
constexpr int kRectT = 256;
constexpr int kRectMaxE = (128 * 96 + kRectT - 1) / kRectT;   // outputs per lane at most

__device__ __forceinline__ void rect_sum2(float &a, float &b, float *red, int t)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((t & 63) == 0) {
        red[t >> 6] = a;
        red[4 + (t >> 6)] = b;
    }
    __syncthreads();
    a = ((red[0] + red[1]) + red[2]) + red[3];
    b = ((red[4] + red[5]) + red[6]) + red[7];
    __syncthreads();
}

// One in-place stage of 1D DFTs over the tile: along rows (length NX) or
// columns (length NY); inverse selects the + sign.  real_only: keep only the
// real part of each output (the last inverse stage).
__device__ __forceinline__ void rect_stage(float *Sr, float *Si, int NY, int NX, bool rows, bool inverse,
                                           const float2 *twr, const float2 *twc, int t)
{
    const int L = rows ? NX : NY;
    const float2 *tw = rows ? twr : twc;
    const int total = NY * NX;
    const float sg = inverse ? 1.0f : -1.0f;
    float outr[kRectMaxE], outi[kRectMaxE];
#pragma unroll
    for (int m = 0; m < kRectMaxE; ++m) {
        const int e = t + m * kRectT;
        float xr = 0.0f, xi = 0.0f;
        if (e < total) {
            const int y = e / NX, x = e - y * NX;
            const int k = rows ? x : y;                 // output frequency on the line
            const int base = rows ? y * NX : x, step = rows ? 1 : NX;
            int idx = 0;
            for (int j = 0; j < L; ++j) {
                const float ar = Sr[base + j * step], ai = Si[base + j * step];
                const float c = tw[idx].x, sn = sg * tw[idx].y;
                xr += ar * c - ai * sn;
                xi += ar * sn + ai * c;
                idx += k;
                if (idx >= L) idx -= L;
            }
        }
        outr[m] = xr;
        outi[m] = xi;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kRectMaxE; ++m) {
        const int e = t + m * kRectT;
        if (e < total) {
            Sr[e] = outr[m];
            Si[e] = outi[m];
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kRectT) xcorr_rect_kernel(XcorrArgs g, int NY, int NX)
{
    __shared__ float s_rect[2 * 128 * 96];
    __shared__ float s_red[8];
    __shared__ float2 s_twr[128], s_twc[128];
    const int t = (int)threadIdx.x;
    const long long gw = blockIdx.x;
    if (gw >= g.total) return;
    const int total = NY * NX;
    float *Sr = s_rect, *Si = s_rect + total;
    for (int j = t; j < NX; j += kRectT) {
        double sn, cs;
        sincospi(2.0 * (double)j / (double)NX, &sn, &cs);
        s_twr[j] = make_float2((float)cs, (float)sn);
    }
    for (int j = t; j < NY; j += kRectT) {
        double sn, cs;
        sincospi(2.0 * (double)j / (double)NY, &sn, &cs);
        s_twc[j] = make_float2((float)cs, (float)sn);
    }
    const int nWin = g.nWy * g.nWx;
    const int n = (int)(gw / nWin), w = (int)(gw - (long long)n * nWin);
    const int ii = w % g.nWx, jj = w / g.nWx;
    bool valid = !(g.mask && g.mask[jj * g.nWx + ii] == 1.0f);
    const int row0 = (int)floor((double)g.ctrY[jj] - ((double)(float)NY - 1.0) / 2.0 + 0.5);
    const int col0 = (int)floor((double)g.ctrX[ii] - ((double)(float)NX - 1.0) / 2.0 + 0.5);
    if (row0 < 0 || col0 < 0 || row0 + NY > g.H || col0 + NX > g.W) valid = false;
    if (!valid) {                       // uniform over the block
        if (!g.ensemble && t == 0) g.rec[(size_t)gw * kRec + kRecFlag] = 0.0f;
        return;
    }
    // window samples times the weights, A real and B imaginary
    float sa = 0.0f, sb = 0.0f;
    const size_t img = (size_t)n * g.H * g.W;
    for (int e = t; e < total; e += kRectT) {
        const int y = e / NX, x = e - y * NX;
        const size_t o = img + (size_t)(row0 + y) * g.W + col0 + x;
        const float a = (g.imgA16 ? (float)g.imgA16[o] : g.imgA[o]) * g.wA[e];
        const float b = (g.imgA16 ? (float)g.imgB16[o] : g.imgB[o]) * g.wB[e];
        Sr[e] = a;
        Si[e] = b;
        sa += a;
        sb += b;
    }
    rect_sum2(sa, sb, s_red, t);
    const float meanA = sa / (float)total, meanB = sb / (float)total;
    float ea = 0.0f, eb = 0.0f;
    for (int e = t; e < total; e += kRectT) {
        float a = Sr[e], b = Si[e];
        if (!g.ensemble) {
            a -= meanA;
            b -= meanB;
            Sr[e] = a;
            Si[e] = b;
        }
        ea += a * a;
        eb += b * b;
    }
    rect_sum2(ea, eb, s_red, t);
    const float enorm = 1.0f / sqrtf(ea * eb);

    rect_stage(Sr, Si, NY, NX, true, false, s_twr, s_twc, t);     // rows, forward
    rect_stage(Sr, Si, NY, NX, false, false, s_twr, s_twc, t);    // columns, forward
    // C = FB conj(FA), FA = (Z + conj(Zm)) / 2, FB = (Z - conj(Zm)) / (2i)
    {
        float cr[kRectMaxE], ci[kRectMaxE];
#pragma unroll
        for (int m = 0; m < kRectMaxE; ++m) {
            const int e = t + m * kRectT;
            float vr = 0.0f, vi = 0.0f;
            if (e < total) {
                const int ky = e / NX, kx = e - ky * NX;
                const int em = ((NY - ky) % NY) * NX + (NX - kx) % NX;
                const float zR = Sr[e], zI = Si[e], mR = Sr[em], mI = Si[em];
                const float ar = 0.5f * (zR + mR), ai = 0.5f * (zI - mI);
                const float br = 0.5f * (zI + mI), bi = -0.5f * (zR - mR);
                vr = br * ar + bi * ai;
                vi = bi * ar - br * ai;
            }
            cr[m] = vr;
            ci[m] = vi;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kRectMaxE; ++m) {
            const int e = t + m * kRectT;
            if (e < total) {
                Sr[e] = cr[m];
                Si[e] = ci[m];
            }
        }
        __syncthreads();
    }
    rect_stage(Sr, Si, NY, NX, false, true, s_twr, s_twc, t);     // columns, inverse
    // rows, inverse, real part only: the fftshifted plane / numel
    float *dst = g.ensemble ? g.planes + (size_t)gw * total : g.work_planes + (size_t)gw * total;
    const float mul = 1.0f / (float)total;
    for (int e = t; e < total; e += kRectT) {
        const int y = e / NX, x = e - y * NX;
        float v = 0.0f;
        int idx = 0;
        for (int j = 0; j < NX; ++j) {
            v += Sr[y * NX + j] * s_twr[idx].x - Si[y * NX + j] * s_twr[idx].y;
            idx += x;
            if (idx >= NX) idx -= NX;
        }
        const int o = ((y + NY / 2) % NY) * NX + (x + NX / 2) % NX;
        dst[o] = v * mul;
        if (!g.ensemble && g.planes) g.planes[(size_t)gw * total + o] = v * mul;
    }
    if (!g.ensemble && t == 0) {
        float *rec = g.rec + (size_t)gw * kRec;
        rec[kRecEnorm] = enorm;
        rec[kRecFlag] = 1.0f;
    }
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


#if defined(PIVAMD_XCORR_SIZE) && (PIVAMD_XCORR_SIZE % 16 == 0) && (PIVAMD_XCORR_SIZE <= 64)
// ---------------------------------------------------------------------------
// Split-precision MFMA form of the same correlation (N = 16, 32, 48, 64).
//
// The 2D transforms are dense products with the DFT matrix F (F[m][k] =
// exp(-2 pi i m k / N), symmetric), computed on the matrix cores in fp16
// with a hi/lo split of every operand (x = hi + lo, both fp16) and FP32
// accumulation of hi*hi + hi*lo + lo*hi, which keeps about 22 significant
// bits per product:
//   Z = A + i B                  (weighted, mean-removed windows)
//   P   = F Z                    (columns)          4 real GEMMs
//   Q^T = F P^T                  (rows)             4
//   FA, FB from Q(k), Q(-k);  C = FB conj(FA)
//   W   = conj(F) C                                  4
//   plane^T = Re[conj(F) W^T] / N^2                  2
// Every stage multiplies a constant by data, so F is always the A operand
// (each wave keeps its 16-row tile of Fr, Fi and -Fi, hi and lo, in
// registers for the whole kernel) and the data always enters as the B
// operand from LDS, stored split and k-contiguous by the stage before.
// Power-of-two scales keep every fp16 operand within 32768 (window values
// <= 512, cross spectrum <= 512) and are removed exactly at the end.
// One wave per 16-row output tile, N/16 waves per window; the plane lands
// in the same fftshifted LDS tile as the FFT form and shares its epilogue.
// ---------------------------------------------------------------------------
typedef _Float16 mf_h4 __attribute__((ext_vector_type(4)));
typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

template <int N>
struct MfCfg {
    static constexpr int KV = (N % 32 == 0) ? 8 : 4;        // halfs per lane per operand
    static constexpr int KM = 4 * KV;                       // K of one MFMA (32 or 16)
    static constexpr int KCH = N / KM;                      // k-chunks per product
    static constexpr int TL = N / 16;                       // 16x16 tiles per dimension
    static constexpr int WPW = TL;                          // waves per window
    static constexpr int T = 64 * WPW;                      // lanes per window
    static constexpr int PX = N + 1;
    static constexpr int NN = N * N;
    // LDS per window (bytes): data operands (2 complex split matrices of
    // fp16: 8 NN each) + the FP32 Q / plane region (8 NN)
    static constexpr int lds_win = 8 * NN + 8 * NN + 8 * NN;
    static constexpr int wpb_lds = (64 * 1024) / lds_win;
    static constexpr int wpb_thr = 512 / T;
    static constexpr int WPB0 = wpb_lds < wpb_thr ? wpb_lds : wpb_thr;
    static constexpr int WPB = WPB0 > 0 ? WPB0 : 1;
    static constexpr int BLOCK = WPB * T;
};

template <int KV>
using mf_hv = typename std::conditional<KV == 8, mf_h8, mf_h4>::type;

template <int KV>
__device__ __forceinline__ mf_f4 mf_mma(const mf_hv<KV> &a, const mf_hv<KV> &b, mf_f4 c)
{
    if constexpr (KV == 8) return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// hi*hi + hi*lo + lo*hi
template <int KV>
__device__ __forceinline__ mf_f4 mf_mma3(const mf_hv<KV> &ah, const mf_hv<KV> &al, const mf_hv<KV> &bh,
                                         const mf_hv<KV> &bl, mf_f4 c)
{
    c = mf_mma<KV>(ah, bh, c);
    c = mf_mma<KV>(ah, bl, c);
    c = mf_mma<KV>(al, bh, c);
    return c;
}

__device__ __forceinline__ void mf_split(float x, _Float16 &hi, _Float16 &lo)
{
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

// One split complex matrix in LDS: four N x N fp16 planes (re hi, re lo,
// im hi, im lo), each stored [j][k] (k contiguous) for the B operand.
template <int N>
struct MfMat {
    _Float16 *rh, *rl, *ih, *il;
    __device__ MfMat(_Float16 *base) : rh(base), rl(base + N * N), ih(base + 2 * N * N), il(base + 3 * N * N) {}
};

// Constant fragments of this wave's row tile: F (forward), hi/lo parts.
template <int N>
struct MfConst {
    using C = MfCfg<N>;
    mf_hv<C::KV> frh[C::KCH], frl[C::KCH], fih[C::KCH], fil[C::KCH], nih[C::KCH], nil[C::KCH];
    __device__ void init(int rowtile, int lane)
    {
        const int m = 16 * rowtile + (lane & 15);
#pragma unroll
        for (int c = 0; c < C::KCH; ++c) {
#pragma unroll
            for (int v = 0; v < C::KV; ++v) {
                const int k = c * C::KM + C::KV * (lane >> 4) + v;
                const int t = (m * k) % N;
                double sn, cs;
                sincospi(2.0 * (double)t / (double)N, &sn, &cs);
                const double re = cs, im = -sn;
                const _Float16 rh = (_Float16)re, ih = (_Float16)im;
                frh[c][v] = rh;
                frl[c][v] = (_Float16)(re - (double)rh);
                fih[c][v] = ih;
                fil[c][v] = (_Float16)(im - (double)ih);
                nih[c][v] = -fih[c][v];
                nil[c][v] = -fil[c][v];
            }
        }
    }
};

// Out = K1 X1 + K2 X2 for one output tile (row tile of this wave, column
// tile J): X1, X2 are B operands from split planes (x1h/x1l, x2h/x2l), K1, K2
// constant fragments (hi/lo).
template <int N>
__device__ __forceinline__ mf_f4 mf_pair(const mf_hv<MfCfg<N>::KV> *k1h, const mf_hv<MfCfg<N>::KV> *k1l,
                                         const _Float16 *x1h, const _Float16 *x1l,
                                         const mf_hv<MfCfg<N>::KV> *k2h, const mf_hv<MfCfg<N>::KV> *k2l,
                                         const _Float16 *x2h, const _Float16 *x2l, int J, int lane)
{
    using C = MfCfg<N>;
    using hv = mf_hv<C::KV>;
    mf_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int j = 16 * J + (lane & 15);
#pragma unroll
    for (int c = 0; c < C::KCH; ++c) {
        const int k0 = c * C::KM + C::KV * (lane >> 4);
        const hv b1h = *(const hv *)(x1h + j * N + k0);
        const hv b1l = *(const hv *)(x1l + j * N + k0);
        const hv b2h = *(const hv *)(x2h + j * N + k0);
        const hv b2l = *(const hv *)(x2l + j * N + k0);
        acc = mf_mma3<C::KV>(k1h[c], k1l[c], b1h, b1l, acc);
        acc = mf_mma3<C::KV>(k2h[c], k2l[c], b2h, b2l, acc);
    }
    return acc;
}

template <int N>
__global__ void __launch_bounds__((MfCfg<N>::BLOCK)) xcorr_mfma_kernel(XcorrArgs g)
{
    using C = MfCfg<N>;
    constexpr int T = C::T, WPB = C::WPB, NN = C::NN, PX = C::PX;
    constexpr int lds_half = 4 * NN;                  // halfs in one split complex matrix
    __shared__ __attribute__((aligned(16))) _Float16 s_mat[WPB][2][lds_half];   // Z, then C | P, then W
    __shared__ float s_q[WPB][2 * NN];                // Q (re, im), then the plane tile
    __shared__ float s_red[5][WPB * T];

    const int b = remap_block((int)blockIdx.x, g.nblocks);
    if (b < 0) return;
    const int wloc = (int)threadIdx.x / T;
    const int t = (int)threadIdx.x - wloc * T;
    const int lane = t & 63, wave = t >> 6;           // wave = this wave's row tile
    float *r0 = &s_red[0][wloc * T], *r1 = &s_red[1][wloc * T], *r2 = &s_red[2][wloc * T];
    float *r3 = &s_red[3][wloc * T], *r4 = &s_red[4][wloc * T];
    MfMat<N> M0(s_mat[wloc][0]), M1(s_mat[wloc][1]);
    float *Qr = s_q[wloc], *Qi = s_q[wloc] + NN;

    MfConst<N> K;
    K.init(wave, lane);

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

    // ---- window: weight, mean, energy (element e = t + T m, row-major) -----
    constexpr int EPL = NN / T;                       // elements per lane
    float za[EPL], zb[EPL];
    float sa = 0.0f, sb = 0.0f;
    if (valid) {
        const size_t img = (size_t)n * g.H * g.W;
        auto load = [&](const auto *A, const auto *B) {
#pragma unroll
            for (int m = 0; m < EPL; ++m) {
                const int e = t + T * m, y = e / N, x = e - (e / N) * N;
                const size_t o = img + (size_t)(row0 + y) * g.W + col0 + x;
                za[m] = (float)A[o] * g.wA[e];
                zb[m] = (float)B[o] * g.wB[e];
                sa += za[m];
                sb += zb[m];
            }
        };
        if (g.imgA16) load(g.imgA16, g.imgB16);
        else load(g.imgA, g.imgB);
    } else {
#pragma unroll
        for (int m = 0; m < EPL; ++m) { za[m] = 0.0f; zb[m] = 0.0f; }
    }
    win_sum2<T>(sa, sb, r0, r1, t);
    const float meanA = sa / (float)NN, meanB = sb / (float)NN;
    float ea = 0.0f, eb = 0.0f;
#pragma unroll
    for (int m = 0; m < EPL; ++m) {
        if (valid && !g.ensemble) {
            za[m] -= meanA;
            zb[m] -= meanB;
        }
        ea += za[m] * za[m];
        eb += zb[m] * zb[m];
    }
    win_sum2<T>(ea, eb, r2, r3, t);     // other slots: the first sum's reads need no barrier
    const float enorm = 1.0f / sqrtf(ea * eb);
    // scales: sqrt(max energy) -> <= 512 for the window, N^2 sqrt(eA eB) -> <= 512 for C
    // log2 bounds: sqrt(e) < 2^(ilogb + 1); |C| <= N^2 sqrt(eA eB) after the
    // window scale 2^ez on both frames
    const float emax = fmaxf(ea, eb);
    const bool scalable = emax > 0.0f && isfinite(emax);
    const int ez = scalable ? 8 - ilogbf(sqrtf(emax)) : 0;
    const bool cross = scalable && ea > 0.0f && eb > 0.0f;
    const int lc = cross ? (ilogbf(sqrtf(ea)) + 1) + (ilogbf(sqrtf(eb)) + 1) + 2 * ez + 2 * (ilogbf((float)N) + 1)
                         : 0;
    const int ec = cross ? 9 - lc : 0;
#pragma unroll
    for (int m = 0; m < EPL; ++m) {
        const int e = t + T * m, y = e / N, x = e - (e / N) * N;
        // Z stored [j = column x][k = row y]
        _Float16 h, l;
        mf_split(ldexpf(za[m], ez), h, l);
        M0.rh[x * N + y] = h;
        M0.rl[x * N + y] = l;
        mf_split(ldexpf(zb[m], ez), h, l);
        M0.ih[x * N + y] = h;
        M0.il[x * N + y] = l;
    }
    __syncthreads();

    // ---- P = F Z ; stored row-major P[r][c] into M1 ---------------------------
    const int rb = 16 * wave + 4 * (lane >> 4), cl = lane & 15;
#pragma unroll
    for (int J = 0; J < C::TL; ++J) {
        const mf_f4 pr = mf_pair<N>(K.frh, K.frl, M0.rh, M0.rl, K.nih, K.nil, M0.ih, M0.il, J, lane);
        const mf_f4 pi = mf_pair<N>(K.fih, K.fil, M0.rh, M0.rl, K.frh, K.frl, M0.ih, M0.il, J, lane);
        const int c = 16 * J + cl;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = rb + v;
            mf_split(pr[v], M1.rh[r * N + c], M1.rl[r * N + c]);
            mf_split(pi[v], M1.ih[r * N + c], M1.il[r * N + c]);
        }
    }
    __syncthreads();

    // ---- Q^T = F P^T ; Q[ky][kx] = Qt[kx][ky] in FP32 --------------------------
#pragma unroll
    for (int J = 0; J < C::TL; ++J) {
        const mf_f4 qr = mf_pair<N>(K.frh, K.frl, M1.rh, M1.rl, K.nih, K.nil, M1.ih, M1.il, J, lane);
        const mf_f4 qi = mf_pair<N>(K.fih, K.fil, M1.rh, M1.rl, K.frh, K.frl, M1.ih, M1.il, J, lane);
        const int ky = 16 * J + cl;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Qr[ky * N + rb + v] = qr[v];
            Qi[ky * N + rb + v] = qi[v];
        }
    }
    __syncthreads();

    // ---- C = FB conj(FA), scaled by 2^ec; stored [kx][ky] into M0 ------------
#pragma unroll
    for (int m = 0; m < EPL; ++m) {
        const int e = t + T * m, ky = e / N, kx = e - (e / N) * N;
        const int my = (N - ky) % N, mx = (N - kx) % N;
        const float zR = Qr[ky * N + kx], zI = Qi[ky * N + kx];
        const float mR = Qr[my * N + mx], mI = Qi[my * N + mx];
        const float ar = 0.5f * (zR + mR), ai = 0.5f * (zI - mI);
        const float br = 0.5f * (zI + mI), bi = -0.5f * (zR - mR);
        const float cr = ldexpf(br * ar + bi * ai, ec), ci = ldexpf(bi * ar - br * ai, ec);
        mf_split(cr, M0.rh[kx * N + ky], M0.rl[kx * N + ky]);
        mf_split(ci, M0.ih[kx * N + ky], M0.il[kx * N + ky]);
    }
    __syncthreads();

    // ---- W = conj(F) C = (Fr Cr + Fi Ci) + i (Fr Ci - Fi Cr); row-major into M1 --
#pragma unroll
    for (int J = 0; J < C::TL; ++J) {
        const mf_f4 wr = mf_pair<N>(K.frh, K.frl, M0.rh, M0.rl, K.fih, K.fil, M0.ih, M0.il, J, lane);
        const mf_f4 wi = mf_pair<N>(K.nih, K.nil, M0.rh, M0.rl, K.frh, K.frl, M0.ih, M0.il, J, lane);
        const int c = 16 * J + cl;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = rb + v;
            mf_split(wr[v], M1.rh[r * N + c], M1.rl[r * N + c]);
            mf_split(wi[v], M1.ih[r * N + c], M1.il[r * N + c]);
        }
    }
    __syncthreads();

    // ---- plane^T = Fr W^T_r + Fi W^T_i ; plane[y][x] = planeT[x][y] -----------
    float *P = s_q[wloc];                              // plane tile, row stride PX
    const int eback = -(2 * ez + ec);               // then 1 / numel, as the FFT form scales
    const float mul = 1.0f / (float)NN;
#pragma unroll
    for (int J = 0; J < C::TL; ++J) {
        const mf_f4 pt = mf_pair<N>(K.frh, K.frl, M1.rh, M1.rl, K.fih, K.fil, M1.ih, M1.il, J, lane);
        const int y = 16 * J + cl;                    // planeT column = plane row
        const int ys = (y + N / 2) % N;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int x = rb + v;
            P[ys * PX + (x + N / 2) % N] = ldexpf(pt[v], eback) * mul;
        }
    }
    __syncthreads();
    plane_epilogue<N, T>(g, P, PX, valid, gw, n, w, enorm, t, r0, r1, r2, r3, r4);
}

static int launch_mfma(XcorrArgs g, hipStream_t st)
{
    constexpr int N = PIVAMD_XCORR_SIZE;
    using C = MfCfg<N>;
    const long long nb = (g.total + C::WPB - 1) / C::WPB;
    if (nb <= 0) return 0;
    if (nb > (1LL << 30)) return PIVAMD_ERR_ARG;
    g.nblocks = (int)nb;
    const int grid = (int)((nb + 7) / 8 * 8);
    hipLaunchKernelGGL((xcorr_mfma_kernel<N>), dim3(grid), dim3(C::BLOCK), 0, st, g);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

#if PIVAMD_XCORR_SIZE == 16
// ---------------------------------------------------------------------------
// The 16-point correlation as register-chained 16x16x32 MFMA products.  One
// wave per window; the DFT matrix fragments stay in registers, the window
// batch is the grid, and no workgroup barrier or LDS staging sits between
// the products:
//   T   = Z F^T                  A = window (z = a + i b), B = constant
//   Q   = F T                    A = constant,  B = T (the C registers)
//   C   = FB conj(FA)            FA, FB from Q(k), Q(-k) by lane permutes
//   U   = conj(F) C              A = constant,  B = C (the C registers)
//   P^T = Re[conj(F) U^T]        B = U^T through a wave-private LDS transpose
// The output tile of v_mfma_f32_16x16x32 puts rows 4g..4g+3 of column c in
// lane 16g + c, and its B operand takes k-slots 8g..8g+7 in that lane.  Giving
// the k-slots of lane group g the order (real rows 4g..4g+3, imaginary rows
// 4g..4g+3), with the constant operand's slots in the same order, turns one
// complex product into a single K = 32 MFMA per output part and split term,
// fed straight from the previous product's accumulators.
//
// Precision forms (operands x = x0 + x1 [+ x2], FP32 accumulation, smaller
// terms first):
//   kGemmF16:    fp16 hi/lo, x0 y0 + x0 y1 + x1 y0 (about 22 bits; the
//                power-of-two scales of the staged MFMA form keep fp16 in range)
//   kGemmBf16:   bf16 hi/lo, the same three products (about 16 bits)
//   kGemmBf16x3: bf16 hi/mid/lo, the six products of order <= 2 (about 24 bits)
// ---------------------------------------------------------------------------
constexpr int kGemmF16 = 0, kGemmBf16 = 1, kGemmBf16x3 = 2;
constexpr int kGemmWindowsPerWave = 4;    // windows per wave: the constants are built once, and the
                                          // epilogue runs 16 lanes per window
constexpr int kGemmWaves = 4;             // waves per workgroup
constexpr int kGemmTP = 20;               // transpose pitch: conflict-free writes and b128 reads

typedef __bf16 gm_b8 __attribute__((ext_vector_type(8)));

template <int P>
struct GmT {
    using E = _Float16;
    using V = mf_h8;
    static constexpr int NT = 2;
};
template <>
struct GmT<kGemmBf16> {
    using E = __bf16;
    using V = gm_b8;
    static constexpr int NT = 2;
};
template <>
struct GmT<kGemmBf16x3> {
    using E = __bf16;
    using V = gm_b8;
    static constexpr int NT = 3;
};

template <int P>
__device__ __forceinline__ mf_f4 gm_mma(const typename GmT<P>::V &a, const typename GmT<P>::V &b, mf_f4 c)
{
    if constexpr (P == kGemmF16) return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Split operand: NT parts of one 8-slot fragment.
template <int P>
struct GmOp {
    typename GmT<P>::V p[GmT<P>::NT];
};

// slot v of the split of x (FP32 value)
template <int P>
__device__ __forceinline__ void gm_split(float x, GmOp<P> &o, int v)
{
    using E = typename GmT<P>::E;
    const E h = (E)x;
    o.p[0][v] = h;
    const float r = x - (float)h;
    const E m = (E)r;
    o.p[1][v] = m;
    if constexpr (GmT<P>::NT == 3) o.p[2][v] = (E)(r - (float)m);
}

// A B over the split parts, products of order <= NT - 1, smaller terms first.
template <int P>
__device__ __forceinline__ mf_f4 gm_prod(const GmOp<P> &a, const GmOp<P> &b)
{
    mf_f4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (GmT<P>::NT == 3) {
        c = gm_mma<P>(a.p[0], b.p[2], c);
        c = gm_mma<P>(a.p[1], b.p[1], c);
        c = gm_mma<P>(a.p[2], b.p[0], c);
    }
    c = gm_mma<P>(a.p[0], b.p[1], c);
    c = gm_mma<P>(a.p[1], b.p[0], c);
    c = gm_mma<P>(a.p[0], b.p[0], c);
    return c;
}

// cos(2 pi t / 16), t = 0..15, rounded to float: the split parts of every
// constant (at most 24 significant bits in all) come from these values
__constant__ float gm_cos16[16] = {
    1.0f, 0.92387953f, 0.70710678f, 0.38268343f, 0.0f, -0.38268343f, -0.70710678f, -0.92387953f,
    -1.0f, -0.92387953f, -0.70710678f, -0.38268343f, 0.0f, 0.38268343f, 0.70710678f, 0.92387953f};

// Constant fragment of lane (g, c) for the blocks [X | Y] (slots j < 4: X at
// column 4g + j, j >= 4: Y at column 4g + j - 4), F[m][k] = exp(-2 pi i mk/16).
// sel: 0 = [Fr | -Fi], 1 = [Fi | Fr], 2 = [Fr | Fi], 3 = [-Fi | Fr].  F is
// symmetric, so the same fragment serves as the A operand (row c) and as the B
// operand (column c).  Every entry is a cosine at a shifted angle: Fr(t) =
// cos(t), -Fi(t) = sin(t) = cos(t - 4), Fi(t) = cos(t + 4) in units of 2 pi / 16.
template <int P>
__device__ __forceinline__ void gm_const(GmOp<P> &o, int sel, int g, int c)
{
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const int t = c * (4 * g + (v & 3));
        int sh;
        if (v < 4) sh = (sel == 0 || sel == 2) ? 0 : sel == 1 ? 4 : 12;
        else sh = (sel == 1 || sel == 3) ? 0 : sel == 0 ? 12 : 4;
        gm_split<P>(gm_cos16[(t + sh) & 15], o, v);
    }
}

template <int P>
__global__ void __launch_bounds__(64 * kGemmWaves) xcorr_gemm16_kernel(XcorrArgs g)
{
    constexpr int N = 16, NN = 256, PX = N + 1, TP = kGemmTP, WV = kGemmWindowsPerWave;
    constexpr int PT = N * PX;                            // one plane tile
    // per wave: the U^T transpose (2 x 16 x TP) and the planes of its WV windows
    __shared__ __attribute__((aligned(16))) float s_t[kGemmWaves][2 * N * TP + WV * PT];

    const int b = remap_block((int)blockIdx.x, g.nblocks);
    if (b < 0) return;
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int gq = lane >> 4, cl = lane & 15;
    float *S = s_t[wv];
    float *Pls = s_t[wv] + 2 * N * TP;

    GmOp<P> K1, K2, K3, K4;
    gm_const<P>(K1, 0, gq, cl);
    gm_const<P>(K2, 1, gq, cl);
    gm_const<P>(K3, 2, gq, cl);
    gm_const<P>(K4, 3, gq, cl);
    // lanes holding Q(-k): rows (16 - 4g - j) & 15 of column (16 - c) & 15
    const int mc = (16 - cl) & 15;
    const int src0 = 16 * ((4 - gq) & 3) + mc, src1 = 16 * (3 - gq) + mc;

    const int nWin = g.nWy * g.nWx;
    const long long gw0 = ((long long)b * kGemmWaves + wv) * WV;
    // the window that lane group gq finishes in the epilogue
    bool my_valid = false;
    int my_n = 0, my_w = 0;
    float my_enorm = 0.0f;
#pragma unroll
    for (int it = 0; it < WV; ++it) {
        const long long gw = gw0 + it;
        if (gw >= g.total) break;           // wave-uniform; no barrier follows
        const int n = (int)(gw / nWin);
        const int w = (int)(gw - (long long)n * nWin);
        bool valid = true;
        int row0, col0;
        {
            const int ii = w % g.nWx, jj = w / g.nWx;
            if (g.mask && g.mask[jj * g.nWx + ii] == 1.0f) valid = false;
            row0 = (int)floor((double)g.ctrY[jj] - ((double)(float)N - 1.0) / 2.0 + 0.5);
            col0 = (int)floor((double)g.ctrX[ii] - ((double)(float)N - 1.0) / 2.0 + 0.5);
            if (row0 < 0 || col0 < 0 || row0 + N > g.H || col0 + N > g.W) valid = false;
        }

        // ---- window: lane (g, c) holds row c, columns 4g..4g+3 of A and B ----
        float za[4] = {0.0f, 0.0f, 0.0f, 0.0f}, zb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const int e0 = cl * N + 4 * gq;
        if (valid) {
            const size_t o = (size_t)n * g.H * g.W + (size_t)(row0 + cl) * g.W + col0 + 4 * gq;
            if (g.imgA16) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    za[j] = (float)g.imgA16[o + j];
                    zb[j] = (float)g.imgB16[o + j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    za[j] = g.imgA[o + j];
                    zb[j] = g.imgB[o + j];
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                za[j] *= g.wA[e0 + j];
                zb[j] *= g.wB[e0 + j];
            }
        }
        float sa = (za[0] + za[1]) + (za[2] + za[3]), sb = (zb[0] + zb[1]) + (zb[2] + zb[3]);
        win_sum2<64>(sa, sb, nullptr, nullptr, lane);
        const float meanA = sa / (float)NN, meanB = sb / (float)NN;
        float ea = 0.0f, eb = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (valid && !g.ensemble) {
                za[j] -= meanA;
                zb[j] -= meanB;
            }
            ea += za[j] * za[j];
            eb += zb[j] * zb[j];
        }
        win_sum2<64>(ea, eb, nullptr, nullptr, lane);
        const float enorm = 1.0f / sqrtf(ea * eb);
        // the staged MFMA form's scales: window values <= 512, |C| <= 512
        const float emax = fmaxf(ea, eb);
        const bool scalable = emax > 0.0f && isfinite(emax);
        const int ez = scalable ? 8 - ilogbf(sqrtf(emax)) : 0;
        const bool cross = scalable && ea > 0.0f && eb > 0.0f;
        const int lc = cross ? (ilogbf(sqrtf(ea)) + 1) + (ilogbf(sqrtf(eb)) + 1) + 2 * ez + 2 * (ilogbf((float)N) + 1)
                             : 0;
        const int ec = cross ? 9 - lc : 0;

        GmOp<P> X;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            gm_split<P>(ldexpf(za[j], ez), X, j);
            gm_split<P>(ldexpf(zb[j], ez), X, 4 + j);
        }
        // ---- T = Z F^T; Q = F T ---------------------------------------------
        const mf_f4 tr = gm_prod<P>(X, K1), ti = gm_prod<P>(X, K2);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            gm_split<P>(tr[j], X, j);
            gm_split<P>(ti[j], X, 4 + j);
        }
        const mf_f4 qr = gm_prod<P>(K1, X), qi = gm_prod<P>(K2, X);

        // ---- C = FB conj(FA), scaled by 2^ec --------------------------------
        // register j of Q(-k) comes from register (4 - j) & 3 of the partner lane
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int sj = (4 - j) & 3, src = j == 0 ? src0 : src1;
            const float mR = __shfl(qr[sj], src, 64), mI = __shfl(qi[sj], src, 64);
            const float zR = qr[j], zI = qi[j];
            const float ar = 0.5f * (zR + mR), ai = 0.5f * (zI - mI);
            const float br = 0.5f * (zI + mI), bi = -0.5f * (zR - mR);
            gm_split<P>(ldexpf(br * ar + bi * ai, ec), X, j);
            gm_split<P>(ldexpf(bi * ar - br * ai, ec), X, 4 + j);
        }
        // ---- U = conj(F) C; U^T through LDS ------------------------------------
        const mf_f4 ur = gm_prod<P>(K3, X), ui = gm_prod<P>(K4, X);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            S[(4 * gq + j) * TP + cl] = ur[j];
            S[N * TP + (4 * gq + j) * TP + cl] = ui[j];
        }
        // same wave: LDS operations complete in order, no barrier
        const float4 vr = *(const float4 *)&S[cl * TP + 4 * gq];
        const float4 vi = *(const float4 *)&S[N * TP + cl * TP + 4 * gq];
        gm_split<P>(vr.x, X, 0); gm_split<P>(vr.y, X, 1); gm_split<P>(vr.z, X, 2); gm_split<P>(vr.w, X, 3);
        gm_split<P>(vi.x, X, 4); gm_split<P>(vi.y, X, 5); gm_split<P>(vi.z, X, 6); gm_split<P>(vi.w, X, 7);
        // ---- P^T = Re[conj(F) U^T]: lane (g, c) holds plane row c, columns 4g + j
        const mf_f4 pt = gm_prod<P>(K3, X);
        const int eback = -(2 * ez + ec);
        const float mul = 1.0f / (float)NN;
        float *Pl = Pls + it * PT;           // fftshifted plane tile, row stride PX
        const int ys = (cl + N / 2) & 15;
#pragma unroll
        for (int j = 0; j < 4; ++j) Pl[ys * PX + ((4 * gq + j + N / 2) & 15)] = ldexpf(pt[j], eback) * mul;
        if (gq == it) {
            my_valid = valid;
            my_n = n;
            my_w = w;
            my_enorm = enorm;
        }
    }
    // ---- epilogue: 16 lanes per window, the wave's windows side by side -----
    const long long my_gw = gw0 + gq;
    if (my_gw < g.total)
        plane_epilogue<N, 16>(g, Pls + gq * PT, PX, my_valid, my_gw, my_n, my_w, my_enorm, cl, nullptr, nullptr,
                              nullptr, nullptr, nullptr);
}

template <int P>
static int launch_gemm16(XcorrArgs g, hipStream_t st)
{
    constexpr int per = kGemmWaves * kGemmWindowsPerWave;
    const long long nb = (g.total + per - 1) / per;
    if (nb <= 0) return 0;
    if (nb > (1LL << 30)) return PIVAMD_ERR_ARG;
    g.nblocks = (int)nb;
    const int grid = (int)((nb + 7) / 8 * 8);
    hipLaunchKernelGGL((xcorr_gemm16_kernel<P>), dim3(grid), dim3(64 * kGemmWaves), 0, st, g);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}
#endif
#endif

// Launch of the correlation kernel for one square size; defined in the
// translation unit compiled with PIVAMD_XCORR_SIZE.
template <int N>
int launch_main(XcorrArgs g, hipStream_t st);

// The ragged launch of one square size (kernel chosen by variant as in
// launch_main; the FFT forms only); same translation unit.
template <int N>
int launch_ragged_main(RaggedArgs &ra, int variant, hipStream_t st);

#ifdef PIVAMD_XCORR_SIZE
template <int N, int R>
static int launch_fft(XcorrArgs g, hipStream_t st)
{
    if constexpr (!xc_radix_built(N, R)) {
        return PIVAMD_ERR_ARG;
    } else {
        using C = XcCfg<N, R>;
        const long long nb = (g.total + C::WPB - 1) / C::WPB;
        if (nb <= 0) return 0;
        if (nb > (1LL << 30)) return PIVAMD_ERR_ARG;
        g.nblocks = (int)nb;
        const int grid = (int)((nb + 7) / 8 * 8);
        hipLaunchKernelGGL((xcorr_kernel<N, R>), dim3(grid), dim3(C::BLOCK), 0, st, g);
        return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
    }
}

template <int N, int R>
static int launch_half(XcorrArgs g, hipStream_t st)
{
    if constexpr (!xc_radix_built(N, R) || XcCfg<N, R>::WPB != 1 || R == 1) {
        return PIVAMD_ERR_ARG;
    } else {
        using C = XcCfg<N, R>;
        const long long nb = g.total;
        if (nb <= 0) return 0;
        if (nb > (1LL << 30)) return PIVAMD_ERR_ARG;
        g.nblocks = (int)nb;
        const int grid = (int)((nb + 7) / 8 * 8);
        hipLaunchKernelGGL((xcorr_kernel_half<N, R>), dim3(grid), dim3(C::BLOCK), 0, st, g);
        return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
    }
}

template <int N>
static int launch_half1(XcorrArgs g, hipStream_t st)
{
    if constexpr (!xc_radix_built(N, 1) || XcCfg<N, 1>::WPB < 2) {
        return PIVAMD_ERR_ARG;
    } else {
        using C = XcCfg<N, 1>;
        const long long nb = (g.total + C::WPB - 1) / C::WPB;
        if (nb <= 0) return 0;
        if (nb > (1LL << 30)) return PIVAMD_ERR_ARG;
        g.nblocks = (int)nb;
        const int grid = (int)((nb + 7) / 8 * 8);
        hipLaunchKernelGGL((xcorr_kernel_half1<N>), dim3(grid), dim3(C::BLOCK), 0, st, g);
        return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
    }
}

template <int N, int R>
static int launch_pf(XcorrArgs g, hipStream_t st)
{
    if constexpr (!xc_radix_built(N, R) || XcCfg<N, R>::WPB != 1) {
        return PIVAMD_ERR_ARG;
    } else {
        using C = XcCfg<N, R>;
        if (g.total <= 0) return 0;
        static int occ[2] = {0, 0}, cus = 0;
        auto kf = xcorr_kernel_pf<N, R, float>;
        auto ku = xcorr_kernel_pf<N, R, unsigned short>;
        const int which = g.imgA16 ? 1 : 0;
        if (!cus) {
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return PIVAMD_ERR_RUNTIME;
        }
        if (!occ[which]) {
            int o = 0;
            const hipError_t e = which ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, ku, C::BLOCK, 0)
                                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, kf, C::BLOCK, 0);
            if (e != hipSuccess || o < 1) return PIVAMD_ERR_RUNTIME;
            occ[which] = o;
        }
        long long P = (long long)cus * occ[which];
        if (P > g.total) P = g.total;
        P = (P + 7) / 8 * 8;
        g.nblocks = (int)P;
        if (which) hipLaunchKernelGGL(ku, dim3((unsigned)P), dim3(C::BLOCK), 0, st, g);
        else hipLaunchKernelGGL(kf, dim3((unsigned)P), dim3(C::BLOCK), 0, st, g);
        return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
    }
}

// Tiles of every job at WPB windows per workgroup, then one launch.
template <int WPB, typename K>
static int launch_ragged_kernel(RaggedArgs &ra, K kernel, int block, hipStream_t st)
{
    long long acc = 0;
    for (int j = 0; j < ra.n_jobs; ++j) {
        ra.start[j] = (int)acc;
        acc += (ra.job[j].total + WPB - 1) / WPB;
        if (acc > (1LL << 30)) return PIVAMD_ERR_ARG;
    }
    ra.start[ra.n_jobs] = (int)acc;
    ra.nblocks = (int)acc;
    if (acc == 0) return 0;
    const int grid = (int)((acc + 7) / 8 * 8);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, ra);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

template <int N, int R>
static int ragged_fft(RaggedArgs &ra, hipStream_t st)
{
    if constexpr (!xc_radix_built(N, R)) return PIVAMD_ERR_ARG;
    else return launch_ragged_kernel<XcCfg<N, R>::WPB>(ra, xcorr_ragged<N, R>, XcCfg<N, R>::BLOCK, st);
}

template <int N>
static int ragged_half(RaggedArgs &ra, hipStream_t st)
{
    if constexpr (!xc_radix_built(N, 4) || XcCfg<N, 4>::WPB != 1) return PIVAMD_ERR_ARG;
    else return launch_ragged_kernel<1>(ra, xcorr_ragged_half<N, 4>, XcCfg<N, 4>::BLOCK, st);
}

template <int N>
static int ragged_half1(RaggedArgs &ra, hipStream_t st)
{
    if constexpr (!xc_radix_built(N, 1) || XcCfg<N, 1>::WPB < 2) return PIVAMD_ERR_ARG;
    else return launch_ragged_kernel<XcCfg<N, 1>::WPB>(ra, xcorr_ragged_half1<N>, XcCfg<N, 1>::BLOCK, st);
}

template <>
int launch_ragged_main<PIVAMD_XCORR_SIZE>(RaggedArgs &ra, int variant, hipStream_t st)
{
    constexpr int N = PIVAMD_XCORR_SIZE;
    switch (variant) {
    case kVariantHalf + 4: return ragged_half<N>(ra, st);
    case kVariantHalf + 1: return ragged_half1<N>(ra, st);
    case 1: return ragged_fft<N, 1>(ra, st);
    case 2: return ragged_fft<N, 2>(ra, st);
    case 4: return ragged_fft<N, 4>(ra, st);
    case 8: return ragged_fft<N, 8>(ra, st);
    case 0:
        if constexpr (N == 48 || N == 64 || N == 128) return ragged_half<N>(ra, st);
        else if constexpr (N == 16 || N == 32) return ragged_half1<N>(ra, st);
        else return ragged_fft<N, xc_radix(N)>(ra, st);
    default: return PIVAMD_ERR_ARG;    // the MFMA and persistent forms have no ragged launch
    }
}

template <>
int launch_main<PIVAMD_XCORR_SIZE>(XcorrArgs g, hipStream_t st)
{
    constexpr int N = PIVAMD_XCORR_SIZE;
    switch (g.variant) {
    case kVariantPersist + 4: return launch_pf<N, 4>(g, st);
    case kVariantHalf + 4: return launch_half<N, 4>(g, st);
    case kVariantHalf + 1: return launch_half1<N>(g, st);
#if (PIVAMD_XCORR_SIZE % 16 == 0) && (PIVAMD_XCORR_SIZE <= 64)
    case kVariantMfma: return launch_mfma(g, st);
#endif
#if PIVAMD_XCORR_SIZE == 16
    case kVariantGemm + kGemmF16: return launch_gemm16<kGemmF16>(g, st);
    case kVariantGemm + kGemmBf16: return launch_gemm16<kGemmBf16>(g, st);
    case kVariantGemm + kGemmBf16x3: return launch_gemm16<kGemmBf16x3>(g, st);
#endif
    case 1: return launch_fft<N, 1>(g, st);
    case 2: return launch_fft<N, 2>(g, st);
    case 4: return launch_fft<N, 4>(g, st);
    case 8: return launch_fft<N, 8>(g, st);
    case 0:
        if constexpr (N == 48 || N == 64 || N == 128) return launch_half<N, 4>(g, st);   // more windows per CU
        else if constexpr (N == 16 || N == 32) return launch_half1<N>(g, st);
        else return launch_fft<N, xc_radix(N)>(g, st);
    default: return PIVAMD_ERR_ARG;
    }
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

template <> int launch_ragged_main<8>(RaggedArgs &, int, hipStream_t);
template <> int launch_ragged_main<12>(RaggedArgs &, int, hipStream_t);
template <> int launch_ragged_main<16>(RaggedArgs &, int, hipStream_t);
template <> int launch_ragged_main<24>(RaggedArgs &, int, hipStream_t);
template <> int launch_ragged_main<32>(RaggedArgs &, int, hipStream_t);
template <> int launch_ragged_main<48>(RaggedArgs &, int, hipStream_t);
template <> int launch_ragged_main<64>(RaggedArgs &, int, hipStream_t);
template <> int launch_ragged_main<96>(RaggedArgs &, int, hipStream_t);
template <> int launch_ragged_main<128>(RaggedArgs &, int, hipStream_t);

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

static bool built_axis(int n)
{
    return n == 8 || n == 12 || n == 16 || n == 24 || n == 32 || n == 48 || n == 64 || n == 96 || n == 128;
}

// Square windows of a built size run the FFT kernels; rectangular ones whose
// row and column lengths are both built sizes run xcorr_rect_kernel (its LDS
// tile holds up to 128 x 96).
extern "C" int pivamd_size_supported(int ny, int nx)
{
    if (!built_axis(ny) || !built_axis(nx)) return 0;
    return (ny == nx || ny * nx <= 128 * 96) ? 1 : 0;
}

// Correlation of rectangular windows: xcorr_rect_kernel, then the search
// and fit of every plane (multipeak_kernel, one peak or several).
static int launch_rect(int NY, int NX, XcorrArgs g, hipStream_t st)
{
    if (g.total <= 0) return 0;
    if (g.total > (1LL << 31) - 1) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL(xcorr_rect_kernel, dim3((unsigned)g.total), dim3(kRectT), 0, st, g, NY, NX);
    if (hipGetLastError() != hipSuccess) return PIVAMD_ERR_LAUNCH;
    if (g.ensemble) return 0;
    hipLaunchKernelGGL(multipeak_kernel, dim3((unsigned)g.total), dim3(256), (size_t)NY * NX * sizeof(float), st,
                       g.work_planes, g.rec, g.total, g.nWy * g.nWx, NY, NX, g.nPeaks, g.finder, g.pw, g);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
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
    if (nPeaks > 1 || winY != winX) bytes += total * (size_t)winY * winX * sizeof(float);
    return bytes;
}

// Transform used by pivamd_xcorr: 0 = FP32 FFT (default, the reference's
// arithmetic class), 1 = split-fp16 MFMA DFT for 16/32/48/64-point windows
// (16: the register-chained form, 32-64: the LDS-staged form), 2 = split-bf16
// (hi/lo) and 3 = split-bf16 (hi/mid/lo) register-chained DFT for 16-point
// windows; other sizes keep the FFT.  Read when a launch is issued.
static int g_dft_mode = 0;

// Selected variant per window size (index N); 0 = default.
static int g_variant[129];

static bool variant_ok(int N, int v)
{
    if (!pivamd_size_supported(N, N)) return false;
    if (v == 0) return true;
    if (v == kVariantMfma) return N % 16 == 0 && N <= 64;
    if (v >= kVariantGemm && v <= kVariantGemm + 2) return N == 16;
    if (v == kVariantPersist + 4) return xc_radix_built(N, 4) && N * 4 >= 256;   // WPB = 1
    if (v == kVariantHalf + 4) return N == 48 || N == 64 || N == 96 || N == 128;
    if (v == kVariantHalf + 1) return xc_radix_built(N, 1);
    return (v == 1 || v == 2 || v == 4 || v == 8) && xc_radix_built(N, v);
}

extern "C" int pivamd_xcorr_variants(int N, int *out, int max_out)
{
    static const int cand[] = {1, 2, 4, 8, kVariantMfma, kVariantPersist + 4, kVariantHalf + 1, kVariantHalf + 4,
                               kVariantGemm, kVariantGemm + 1, kVariantGemm + 2};
    int n = 0;
    for (int v : cand)
        if (variant_ok(N, v)) {
            if (out && n < max_out) out[n] = v;
            ++n;
        }
    return n;
}

extern "C" int pivamd_xcorr_default_variant(int N)
{
    if (!pivamd_size_supported(N, N)) return -1;
    return (N == 48 || N == 64 || N == 128) ? kVariantHalf + 4 : (N == 16 || N == 32) ? kVariantHalf + 1
                                                                                        : xc_radix(N);
}

extern "C" int pivamd_xcorr_set_variant(int N, int v)
{
    if (!variant_ok(N, v)) return -1;
    const int prev = g_variant[N];
    g_variant[N] = v;
    return prev;
}

extern "C" int pivamd_xcorr_get_variant(int N)
{
    return pivamd_size_supported(N, N) ? g_variant[N] : -1;
}

extern "C" int pivamd_set_dft_mode(int mode)
{
    const int prev = g_dft_mode;
    if (mode >= 0 && mode <= 3) g_dft_mode = mode;
    return prev;
}

static int xcorr_entry(const float *d_imgA, const float *d_imgB, const unsigned short *d_imgA16,
                       const unsigned short *d_imgB16, const float *d_mask, int H, int W,
                       int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                       const float *d_wA, const float *d_wB, const float *d_planeWeight, int ensemble,
                       int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX,
                       float *d_pkY, float *d_pkH, float *d_sx, float *d_sy, float *d_sxy,
                       float *d_ratio, float *d_planes, void *d_work, size_t work_bytes, hipStream_t st)
{
    if (n_images < 0 || nWy < 0 || nWx < 0 || nPeaks < 1 || H <= 0 || W <= 0) return PIVAMD_ERR_ARG;
    if (work_bytes < pivamd_xcorr_workspace(n_images, nWy, nWx, winY, winX, nPeaks, ensemble))
        return PIVAMD_ERR_ARG;
    if (n_images > 0 && (d_imgA16 ? !d_imgB16 : (!d_imgA || !d_imgB))) return PIVAMD_ERR_ARG;
    XcorrArgs g;
    g.imgA = d_imgA; g.imgB = d_imgB; g.imgA16 = d_imgA16; g.imgB16 = d_imgB16;
    g.mask = d_mask; g.ctrX = d_ctrX; g.ctrY = d_ctrY;
    g.wA = d_wA; g.wB = d_wB; g.pw = d_planeWeight;
    g.pkX = d_pkX; g.pkY = d_pkY; g.pkH = d_pkH; g.sx = d_sx; g.sy = d_sy; g.sxy = d_sxy;
    g.ratio = d_ratio; g.planes = d_planes;
    g.rec = (float *)d_work;
    g.work_planes = (nPeaks > 1 || winY != winX) && d_work ? (float *)d_work + (size_t)n_images * nWy * nWx * kRec
                                                            : nullptr;
    g.total = (long long)n_images * nWy * nWx;
    g.H = H; g.W = W; g.nWy = nWy; g.nWx = nWx; g.nPeaks = nPeaks; g.finder = finder;
    g.impl = impl; g.ensemble = ensemble; g.nblocks = 0;
    int dv = 0;
    if (winY == winX && g_dft_mode == 1 && winY % 16 == 0 && winY <= 64)
        dv = winY == 16 ? kVariantGemm : kVariantMfma;
    else if (winY == winX && winY == 16 && (g_dft_mode == 2 || g_dft_mode == 3))
        dv = kVariantGemm + g_dft_mode - 1;
    g.dft = dv != 0 ? 1 : 0;
    g.variant = dv != 0 ? dv : (winY == winX && winY > 0 && winY <= 128 ? g_variant[winY] : 0);
    if (!ensemble && (!d_pkX || !d_pkY || !d_pkH || !d_sx || !d_sy || !d_sxy || !d_ratio || !d_work))
        return PIVAMD_ERR_ARG;
    if (!pivamd_size_supported(winY, winX)) return PIVAMD_ERR_SIZE;
    if (winY != winX) return launch_rect(winY, winX, g, st);
    return launch_xcorr(winY, g, st);
}

extern "C" int pivamd_xcorr(const float *d_imgA, const float *d_imgB, const float *d_mask, int H, int W,
                            int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                            const float *d_wA, const float *d_wB, const float *d_planeWeight, int ensemble,
                            int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX,
                            float *d_pkY, float *d_pkH, float *d_sx, float *d_sy, float *d_sxy,
                            float *d_ratio, float *d_planes, void *d_work, size_t work_bytes, hipStream_t st)
{
    return xcorr_entry(d_imgA, d_imgB, nullptr, nullptr, d_mask, H, W, n_images, d_ctrX, d_ctrY, nWy, nWx,
                       d_wA, d_wB, d_planeWeight, ensemble, winY, winX, nPeaks, finder, impl, d_pkX, d_pkY,
                       d_pkH, d_sx, d_sy, d_sxy, d_ratio, d_planes, d_work, work_bytes, st);
}

extern "C" int pivamd_xcorr_u16(const unsigned short *d_imgA, const unsigned short *d_imgB, const float *d_mask,
                                int H, int W, int n_images, const float *d_ctrX, const float *d_ctrY, int nWy,
                                int nWx, const float *d_wA, const float *d_wB, const float *d_planeWeight,
                                int ensemble, int winY, int winX, int nPeaks, int finder, int impl, float *d_pkX,
                                float *d_pkY, float *d_pkH, float *d_sx, float *d_sy, float *d_sxy,
                                float *d_ratio, float *d_planes, void *d_work, size_t work_bytes, hipStream_t st)
{
    if (n_images > 0 && (!d_imgA || !d_imgB)) return PIVAMD_ERR_ARG;
    return xcorr_entry(nullptr, nullptr, d_imgA, d_imgB, d_mask, H, W, n_images, d_ctrX, d_ctrY, nWy, nWx,
                       d_wA, d_wB, d_planeWeight, ensemble, winY, winX, nPeaks, finder, impl, d_pkX, d_pkY,
                       d_pkH, d_sx, d_sy, d_sxy, d_ratio, d_planes, d_work, work_bytes, st);
}

// ---- pivamd_xcorr_ragged: several correlation calls of one size, one launch
extern "C" int pivamd_xcorr_job_size(void) { return (int)sizeof(PivamdXcorrJob); }

extern "C" int pivamd_xcorr_ragged(const PivamdXcorrJob *jobs, int n_jobs, int winY, int winX, int nPeaks,
                                   int finder, int impl, hipStream_t st)
{
    if (!jobs || n_jobs < 1 || n_jobs > kMaxRaggedJobs || nPeaks < 1) return PIVAMD_ERR_ARG;
    if (winY != winX || !pivamd_size_supported(winY, winX)) return PIVAMD_ERR_SIZE;
    const int N = winY;
    RaggedArgs ra;
    memset(&ra, 0, sizeof(ra));
    ra.n_jobs = n_jobs;
    for (int j = 0; j < n_jobs; ++j) {
        const PivamdXcorrJob &J = jobs[j];
        if (J.n_images < 0 || J.nWy < 0 || J.nWx < 0 || J.H <= 0 || J.W <= 0) return PIVAMD_ERR_ARG;
        if (J.n_images > 0 && (!J.imgA || !J.imgB)) return PIVAMD_ERR_ARG;
        if (!J.pkX || !J.pkY || !J.pkH || !J.sx || !J.sy || !J.sxy || !J.ratio || !J.work) return PIVAMD_ERR_ARG;
        if (J.work_bytes < pivamd_xcorr_workspace(J.n_images, J.nWy, J.nWx, N, N, nPeaks, 0)) return PIVAMD_ERR_ARG;
        XcorrArgs &g = ra.job[j];
        g.imgA = J.u16 ? nullptr : (const float *)J.imgA;
        g.imgB = J.u16 ? nullptr : (const float *)J.imgB;
        g.imgA16 = J.u16 ? (const unsigned short *)J.imgA : nullptr;
        g.imgB16 = J.u16 ? (const unsigned short *)J.imgB : nullptr;
        g.mask = J.mask; g.ctrX = J.ctrX; g.ctrY = J.ctrY;
        g.wA = J.wA; g.wB = J.wB; g.pw = J.planeWeight;
        g.pkX = J.pkX; g.pkY = J.pkY; g.pkH = J.pkH; g.sx = J.sx; g.sy = J.sy; g.sxy = J.sxy;
        g.ratio = J.ratio; g.planes = J.planes;
        g.rec = (float *)J.work;
        g.work_planes = nPeaks > 1 ? (float *)J.work + (size_t)J.n_images * J.nWy * J.nWx * kRec : nullptr;
        g.total = (long long)J.n_images * J.nWy * J.nWx;
        g.H = J.H; g.W = J.W; g.nWy = J.nWy; g.nWx = J.nWx; g.nPeaks = nPeaks; g.finder = finder;
        g.impl = impl; g.ensemble = 0; g.nblocks = 0; g.dft = 0; g.variant = 0;
    }
    const int variant = g_variant[N];
    int rc;
    switch (N) {
    case 8: rc = launch_ragged_main<8>(ra, variant, st); break;
    case 12: rc = launch_ragged_main<12>(ra, variant, st); break;
    case 16: rc = launch_ragged_main<16>(ra, variant, st); break;
    case 24: rc = launch_ragged_main<24>(ra, variant, st); break;
    case 32: rc = launch_ragged_main<32>(ra, variant, st); break;
    case 48: rc = launch_ragged_main<48>(ra, variant, st); break;
    case 64: rc = launch_ragged_main<64>(ra, variant, st); break;
    case 96: rc = launch_ragged_main<96>(ra, variant, st); break;
    case 128: rc = launch_ragged_main<128>(ra, variant, st); break;
    default: return PIVAMD_ERR_SIZE;
    }
    if (rc != 0) return rc;
    for (int j = 0; j < n_jobs; ++j) {
        const XcorrArgs &g = ra.job[j];
        if (g.total <= 0) continue;
        const int nWin = g.nWy * g.nWx;
        if (nPeaks == 1) {
            const long long fb = (g.total + 255) / 256;
            hipLaunchKernelGGL(fit_kernel, dim3((unsigned)fb), dim3(256), 0, st, g.rec, g.total, nWin, N, N,
                               g.finder, g.impl, g.pkX, g.pkY, g.pkH, g.sx, g.sy, g.sxy);
        } else {
            hipLaunchKernelGGL(multipeak_kernel, dim3((unsigned)g.total), dim3(256), (size_t)N * N * sizeof(float),
                               st, g.work_planes, g.rec, g.total, nWin, N, N, nPeaks, g.finder, g.pw, g);
        }
        if (hipGetLastError() != hipSuccess) return PIVAMD_ERR_LAUNCH;
    }
    return 0;
}

// ---- pivamd_xcorr_noalloc: the correlation with no workspace argument -----
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get the peak records
// (and, for nPeaks > 1, the planes the multi-peak kernel reads) from this
// file-scope device array instead of a caller workspace.  Each GPU the
// library runs on holds its own copy until the process ends.  Every record a
// kernel reads is written earlier in the same call (the validity flag of each
// window included), so nothing in it needs clearing between calls; calls that
// overlap on one GPU (other streams or threads) share it and race.
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

// Device address of the current device's scratch array, looked up once per
// device (hipGetSymbolAddress with that device current).  The array's name
// is only a host-side stand-in: host code passes this address, never the name.
static void *noalloc_scratch()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

extern "C" size_t pivamd_xcorr_noalloc_max_bytes(void) { return kNoallocScratchBytes; }

extern "C" int pivamd_xcorr_noalloc(const float *d_imgA, const float *d_imgB, const float *d_mask, int H, int W,
                                    int n_images, const float *d_ctrX, const float *d_ctrY, int nWy, int nWx,
                                    const float *d_wA, const float *d_wB, const float *d_planeWeight,
                                    int ensemble, int winY, int winX, int nPeaks, int finder, int impl,
                                    float *d_pkX, float *d_pkY, float *d_pkH, float *d_sx, float *d_sy,
                                    float *d_sxy, float *d_ratio, float *d_planes, hipStream_t st)
{
    if (n_images < 0 || nWy < 0 || nWx < 0 || nPeaks < 1) return PIVAMD_ERR_ARG;
    const size_t need = pivamd_xcorr_workspace(n_images, nWy, nWx, winY, winX, nPeaks, ensemble);
    if (need > kNoallocScratchBytes) {
        fprintf(stderr, "pivamd_xcorr_noalloc: %zu bytes of peak records for %d images x %d x %d windows "
                        "exceed the %zu-byte device scratch; split the call or use pivamd_xcorr\n",
                need, n_images, nWy, nWx, kNoallocScratchBytes);
        return PIVAMD_ERR_SIZE;
    }
    void *work = noalloc_scratch();
    if (!work) return PIVAMD_ERR_RUNTIME;
    return xcorr_entry(d_imgA, d_imgB, nullptr, nullptr, d_mask, H, W, n_images, d_ctrX, d_ctrY, nWy, nWx,
                       d_wA, d_wB, d_planeWeight, ensemble, winY, winX, nPeaks, finder, impl, d_pkX, d_pkY,
                       d_pkH, d_sx, d_sy, d_sxy, d_ratio, d_planes, work, kNoallocScratchBytes, st);
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
