// warp.hip -- fused symmetric image warp (gfx950).
//
// Replaces fused_symmetric_warp(_batch) of pivtools_cli/lib/fused_warp.c:
//   phase A  the coarse predictor (nPY x nPX, per image or shared) is
//            upsampled at each pixel by bicubic Keys a=-0.75 with replicate
//            border, its fractional grid index being np.interp of the pixel
//            coordinate over the previous pass's padded centres (clamped);
//   phase B  half = 0.5*d (optionally rounded to whole pixels);
//   phase C  A is sampled at (i - hy, j - hx) and B at (i + hy, j + hx) by
//            bicubic (4x4, floor-1) or Lanczos-3 (6x6, floor-2, 4097-entry
//            LUT with lerp), zero outside the image.
// The arithmetic is the reference's scalar path (impl 0) term for term with
// contraction off, so every output pixel is bit-identical to it.  A second
// bicubic form (mode 2) evaluates the same interpolants separably with fused
// multiply-adds: rows of four taps are summed first and the 4x4 products
// wy*wx are never formed, and B is sampled at the mirror image of A's rounded
// position with A's weights reversed, for about half the instructions; its
// results differ from the reference's in the last bits only.
//
// Mapping: a 256-lane workgroup covers a 64-column x 16-row tile of one
// image (grid.z = image).  Each lane owns one column and four rows.  The
// tile's predictor indices, the predictor cells its stencils touch and the
// predictor y weights and row offsets of each tile row are staged in LDS
// once; image taps take one bounds test per stencil (interior) or one per
// tap (border), and the half-shifts are clamped so that a non-finite
// predictor can never produce an out-of-range read.  The reference-exact
// forms read the taps straight from global memory (wide unaligned loads):
// they are VALU bound at about 400 instructions per pixel, and staging the
// source window in LDS measured slower for them.  The FMA form (mode 2) by
// default stages it: a 512-lane workgroup covers a 64 x 32 tile, converts
// the source rectangles of A and B to float in LDS once, and each tap is
// then an LDS read, which took the form from about 310 to about 190 VALU
// instructions per pixel pair (with B's stencil mirrored from A's).
#include <hip/hip_runtime.h>

#include "pivamd_dev.h"

namespace pivamd {

#pragma clang fp contract(off)

constexpr int kLanczosLut = 4096;
constexpr int kTileX = 64, kTileY = 16, kLanesY = 4;

__device__ __forceinline__ void keys4(float d, float w[4])
{
    const float s0 = 1.0f + d, q2 = 1.0f - d, s3 = 2.0f - d;
    w[1] = (1.25f * d - 2.25f) * d * d + 1.0f;
    w[2] = (1.25f * q2 - 2.25f) * q2 * q2 + 1.0f;
    w[0] = ((-0.75f * s0 + 3.75f) * s0 - 6.0f) * s0 + 3.0f;
    w[3] = ((-0.75f * s3 + 3.75f) * s3 - 6.0f) * s3 + 3.0f;
}

// np.interp(p, ctrs, arange(nC)) with clamp, as the sequential segment walk
// of fused_warp.c:259-307 resolves it for ascending centres.
__device__ float pred_index(float coord, const float *ctrs, int nC)
{
    if (nC == 1) return 0.0f;
    const float c0 = ctrs[0], cl = ctrs[nC - 1];
    if (coord <= c0) return 0.0f;
    if (coord >= cl) return (float)(nC - 1);
    // smallest seg in [0, nC-2] with ctrs[seg+1] >= coord: start from the
    // segment a uniform grid would put it in (exact for the window grids, so
    // two dependent loads instead of a binary search's log2 nC), then walk
    // up and down to it; the walks end for any centres
    const float guess = (coord - c0) / (cl - c0) * (float)(nC - 1);
    int seg = (int)fminf(fmaxf(guess, 0.0f), (float)(nC - 2));   // NaN -> 0
    while (seg < nC - 2 && ctrs[seg + 1] < coord) ++seg;
    while (seg > 0 && ctrs[seg] >= coord) --seg;
    const float den = ctrs[seg + 1] - ctrs[seg];
    if (den < 1e-12f) return (float)seg;
    const float t = (coord - ctrs[seg]) / den;
    // in [0, nC - 1] already for finite ascending centres; the clamp maps
    // anything else (NaN centres) into range so later indexing stays safe
    return fminf(fmaxf((float)seg + t, 0.0f), (float)(nC - 1));
}

// Bicubic tap sum (Keys a=-0.75, 4x4 at floor-1, zero outside the image),
// term order m outer / n inner as fused_warp.c:70-105.  The whole stencil
// is tested once; only border stencils test each tap.  The caller keeps the
// shifts within +-kShiftMax, so every index computation is in range; hm3 =
// max(H - 3, 0) and wm3 = max(W - 3, 0) make the interior test two unsigned
// compares.
__device__ __forceinline__ bool inside(int i0, int lim) { return (unsigned)i0 < (unsigned)lim; }

template <typename TI>
__device__ __forceinline__ float sample_bicubic(const TI *img, float fy, float fx, int H, int W, int hm3,
                                                int wm3)
{
    const float fyf = floorf(fy), fxf = floorf(fx);
    const int iy = (int)fyf - 1, ix = (int)fxf - 1;
    float wy[4], wx[4];
    keys4(fy - fyf, wy);
    keys4(fx - fxf, wx);
    float val = 0.0f;
    if (inside(iy, hm3) && inside(ix, wm3)) {
        const TI *p = img + (iy * W + ix);
#pragma unroll
        for (int m = 0; m < 4; m++) {
#pragma unroll
            for (int n = 0; n < 4; n++) val += wy[m] * wx[n] * (float)p[n];
            p += W;
        }
        return val;
    }
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int row = iy + m;
        if (row < 0 || row >= H) continue;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const int col = ix + n;
            if (col < 0 || col >= W) continue;
            val += wy[m] * wx[n] * (float)img[row * W + col];
        }
    }
    return val;
}

// Keys a = -0.75 weights with fused multiply-adds: the same four cubics
// expanded in powers of d (0 <= d < 1) and evaluated by Horner's rule,
// 11 operations instead of 17:
//   w0 = -0.75 d^3 + 1.5 d^2 - 0.75 d     w1 = 1.25 d^3 - 2.25 d^2 + 1
//   w2 = -1.25 d^3 + 1.5 d^2 + 0.75 d     w3 = 0.75 d^3 - 0.75 d^2
__device__ __forceinline__ void keys4_fma(float d, float w[4])
{
    const float d2 = d * d;
    w[0] = __builtin_fmaf(__builtin_fmaf(-0.75f, d, 1.5f), d, -0.75f) * d;
    w[1] = __builtin_fmaf(__builtin_fmaf(1.25f, d, -2.25f), d2, 1.0f);
    w[2] = __builtin_fmaf(__builtin_fmaf(-1.25f, d, 1.5f), d, 0.75f) * d;
    w[3] = __builtin_fmaf(0.75f, d, -0.75f) * d2;
}

// The separable bicubic sample: sum_m wy[m] * (sum_n wx[n] * img[m][n]),
// zero outside the image.
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));

// A frame as a raw buffer (one 12-byte load per row of taps): the base is
// the frame start rounded down to 4 bytes, so every offset below is aligned.
struct FrameBuf {
    __amdgpu_buffer_rsrc_t rsrc;
    uintptr_t base;        // frame start rounded down to 4 bytes
};

__device__ __forceinline__ FrameBuf frame_buf(const void *frame, size_t bytes)
{
    FrameBuf f;
    f.base = (uintptr_t)frame & ~(uintptr_t)3;
    const int n = (int)(bytes + ((uintptr_t)frame & 3));
    f.rsrc = __builtin_amdgcn_make_buffer_rsrc((void *)f.base, (short)0, n, 0x00020000);
    return f;
}

// The FMA form's stencils of one output pixel (i, j) with half-shifts
// (hy, hx): A's 4 x 4 taps start at (floor(i - hy) - 1, floor(j - hx) - 1)
// with Keys weights wy, wx of the fractional parts.  B is sampled at the
// mirror image of A's position about (i, j), 2 (i, j) - (i - hy, j - hx):
// its taps start at (2i - floor(i - hy) - 2, 2j - floor(j - hx) - 2) and its
// weights are A's in reverse order (the Keys kernel is even, so the weights
// of the fraction 1 - d are those of d reversed, and at d = 0 both give the
// tap at 2i - floor(i - hy)).  The mirror point differs from the rounded
// i + hy by A's own position rounding, so B is as accurate as A, and its
// weights cost nothing.
struct FmaStencil {
    int iyA, ixA, iyB, ixB;
    float wy[4], wx[4];
};

__device__ __forceinline__ FmaStencil fma_stencil(int i, int j, float hy, float hx)
{
    FmaStencil s;
    const float fy = (float)i - hy, fx = (float)j - hx;
    const float fyf = floorf(fy), fxf = floorf(fx);
    keys4_fma(fy - fyf, s.wy);
    keys4_fma(fx - fxf, s.wx);
    s.iyA = (int)fyf - 1;
    s.ixA = (int)fxf - 1;
    s.iyB = 2 * i - (int)fyf - 2;
    s.ixB = 2 * j - (int)fxf - 2;
    return s;
}

// One FMA-form tap sum from global memory: rows iy..iy+3 with weights wy,
// columns ix..ix+3 with weights wx, zero outside the image.
template <typename TI>
__device__ __forceinline__ float sample_bicubic_fast(const TI *img, int iy, int ix, const float (&wy)[4],
                                                     const float (&wx)[4], int H, int W, int hm3, int wm3,
                                                     const FrameBuf &fb)
{
    float val = 0.0f;
    if (inside(iy, hm3) && inside(ix, wm3)) {
        if constexpr (sizeof(TI) == 2) {
            // 16-bit taps: each row's four taps lie in the three aligned
            // dwords from their address rounded down to 4 bytes, at a 0- or
            // 16-bit offset that alignbit takes out.  Aligned reads cost the
            // texture addresser less than 2-byte-aligned 8-byte reads.  The
            // run may extend 2 bytes either side of the taps: it is used only
            // when that stays inside the frame.
            const uintptr_t a0 = (uintptr_t)(img + (iy * W + ix));
            const uintptr_t rowb = (uintptr_t)W * 2u;
            const uintptr_t lo = (uintptr_t)img, hi = (uintptr_t)(img + (size_t)H * W);
            if ((a0 & ~(uintptr_t)3) >= lo && a0 + 3 * rowb + 12 <= hi) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const uintptr_t am = a0 + m * rowb;
                    const int off = (int)((am & ~(uintptr_t)3) - fb.base);
                    const u32x3 v = __builtin_amdgcn_raw_buffer_load_b96(fb.rsrc, off, 0, 0);
                    const uint32_t sh = (uint32_t)(am & 2u) * 8u;
                    const uint32_t d0 = v.x, d1 = v.y, d2 = v.z;
                    const uint32_t t01 = __builtin_amdgcn_alignbit(d1, d0, sh);
                    const uint32_t t23 = __builtin_amdgcn_alignbit(d2, d1, sh);
                    float rs = wx[0] * (float)(t01 & 0xffffu);
                    rs = __builtin_fmaf(wx[1], (float)(t01 >> 16), rs);
                    rs = __builtin_fmaf(wx[2], (float)(t23 & 0xffffu), rs);
                    rs = __builtin_fmaf(wx[3], (float)(t23 >> 16), rs);
                    val = __builtin_fmaf(wy[m], rs, val);
                }
                return val;
            }
        }
        const TI *p = img + (iy * W + ix);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            float rs = wx[0] * (float)p[0];
            rs = __builtin_fmaf(wx[1], (float)p[1], rs);
            rs = __builtin_fmaf(wx[2], (float)p[2], rs);
            rs = __builtin_fmaf(wx[3], (float)p[3], rs);
            val = __builtin_fmaf(wy[m], rs, val);
            p += W;
        }
        return val;
    }
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int row = iy + m;
        if (row < 0 || row >= H) continue;
        float rs = 0.0f;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const int col = ix + n;
            if (col < 0 || col >= W) continue;
            rs = __builtin_fmaf(wx[n], (float)img[row * W + col], rs);
        }
        val = __builtin_fmaf(wy[m], rs, val);
    }
    return val;
}

__device__ __forceinline__ void lanczos_w(float d, const float *lut, float w[6])
{
    const float pos = d * (float)kLanczosLut;
    int idx = (int)pos;
    float fr = pos - (float)idx;
    if (idx >= kLanczosLut) { idx = kLanczosLut - 1; fr = 0.0f; }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const float a = lut[idx * 6 + k], b = lut[(idx + 1) * 6 + k];
        w[k] = a + fr * (b - a);
    }
}

template <typename TI>
__device__ __forceinline__ float sample_lanczos(const TI *img, float fy, float fx, int H, int W, int hm5,
                                                int wm5, const float *lut)
{
    const float fyf = floorf(fy), fxf = floorf(fx);
    const int iy = (int)fyf - 2, ix = (int)fxf - 2;
    float wy[6], wx[6];
    lanczos_w(fy - fyf, lut, wy);
    lanczos_w(fx - fxf, lut, wx);
    float val = 0.0f;
    if (inside(iy, hm5) && inside(ix, wm5)) {
        const TI *p = img + (iy * W + ix);
#pragma unroll
        for (int m = 0; m < 6; m++) {
#pragma unroll
            for (int n = 0; n < 6; n++) val += wy[m] * wx[n] * (float)p[n];
            p += W;
        }
        return val;
    }
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const int row = iy + m;
        if (row < 0 || row >= H) continue;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const int col = ix + n;
            if (col < 0 || col >= W) continue;
            val += wy[m] * wx[n] * (float)img[row * W + col];
        }
    }
    return val;
}

struct WarpArgs {
    const float *imgA, *imgB, *pdy, *pdx, *cy, *cx, *lut;
    const unsigned short *imgA16, *imgB16;   // 16-bit counts instead of imgA/imgB
    float *outA, *outB;
    int H, W, nPY, nPX, predStride, mode, roundShifts;
};

constexpr int kPredMax = 1024;   // predictor cells a tile may stage in LDS
constexpr int kPredMaxStaged = 256;   // the same with the source frames staged too
constexpr int kRowSums = 8;      // predictor rows whose x sums a FAST tile keeps in LDS

// Source staging (the FAST form's default): each frame's source rectangle of
// the tile, converted to float, zero outside the image, kRegH rows of kRegP
// columns.  3*kRegP + 3 = 255 keeps the 4x4 taps of a stencil within the
// 8-bit dword offsets of ds_read2_b32 from one base address.
constexpr int kRegP = 84, kRegH = 48;
// The staged form's tile: 64 x 32 pixels for a 512-lane workgroup (each lane
// one column and four rows), so the per-tile set-up and the source halo are
// shared by twice the pixels of the unstaged 64 x 16 tile.
constexpr int kTileYS = 32, kLanesYS = 8;
// Predictor magnitudes above this (or non-finite) leave the tile unstaged, so
// a staged row lies within kStageRowsMax rows of the frame.
constexpr float kStageShiftMax = 4096.0f;
constexpr int kStageRowsMax = 4300;

template <typename TI>
__device__ __forceinline__ float load_tap(__amdgpu_buffer_rsrc_t r, uint32_t off_bytes)
{
    if constexpr (sizeof(TI) == 2)
        return (float)__builtin_amdgcn_raw_buffer_load_b16(r, off_bytes, 0, 0);
    else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off_bytes, 0, 0));
}

template <typename TI>
__device__ __forceinline__ const TI *warp_src(const WarpArgs &g, bool a)
{
    if constexpr (sizeof(TI) == 2) return a ? g.imgA16 : g.imgB16;
    else return a ? g.imgA : g.imgB;
}

// Half-shifts beyond this put every tap outside any image (zero output);
// clamping them keeps all position and index arithmetic in range for any
// input, non-finite included, and leaves in-range shifts untouched.
constexpr float kShiftMax = 1048576.0f;

// TI: float frames, or 16-bit counts (converted exactly on load).  FAST:
// the separable fused-multiply-add form of mode 2.  STAGE (FAST only): the
// tile's source rectangles of A and B are staged in LDS as floats first, so
// each tap is an LDS read instead of a global load, an unpack and a convert.
// The rectangles are bounded from the staged predictor cells: a Keys bicubic
// stencil has at most 0.4454 negative weight mass (1.375^2 = 1 + 2 * 0.4453),
// so the interpolated predictor stays within half the cells' range beyond
// their extremes.  Zeros stand for pixels outside the image, which gives the
// border stencils' sums exactly; a stencil outside the rectangle (never, for
// finite arithmetic) reads global memory as the unstaged form does.
template <typename TI, bool FAST, bool STAGE = false>
__global__ void __launch_bounds__(STAGE ? 512 : 256, STAGE ? 8 : 1) warp_kernel(WarpArgs g)
{
    static_assert(FAST || !STAGE, "staging is a form of the FMA warp");
    constexpr int kPM = STAGE ? kPredMaxStaged : kPredMax;
    constexpr int TY = STAGE ? kTileYS : kTileY;      // tile rows
    constexpr int LY = STAGE ? kLanesYS : kLanesY;    // lane rows (TY / LY rows per lane)
    constexpr int NT = kTileX * LY;                   // lanes
    __shared__ float s_ly[TY];
    __shared__ float s_lx[kTileX];
    __shared__ float4 s_wy[TY];         // predictor y weights of each tile row
    __shared__ int4 s_ro[TY];           // predictor row offsets of each tile row
    __shared__ float s_pdy[kPM];
    __shared__ float s_pdx[kPM];
    // FAST: x sums sum_n wx[n] P[r][c_n] of each staged predictor row r at
    // each tile column; a pixel then needs only sum_m wy[m] * S[r_m]
    __shared__ float s_sy[FAST ? kRowSums * kTileX : 1];
    __shared__ float s_sx[FAST ? kRowSums * kTileX : 1];
    // STAGE: the source rectangles (A then B) and their origins and heights
    __shared__ float s_reg[STAGE ? 2 * kRegH * kRegP : 1];
    __shared__ int s_rg[8];
    const int tid = (int)threadIdx.x;
    const int tx = tid & (kTileX - 1);
    const int ty = tid / kTileX;
    // Workgroups go to the XCDs (and their shader engines) round robin in
    // launch order, so launch column x of tile row y lands on XCD
    // (x + y * (cols mod 8)) mod 8.  When that pins an XCD to the same tile
    // columns down the whole frame (cols a multiple of 8, as at 1024 and
    // 2048 pixels), the engines holding the border columns fall behind and
    // the warp runs at half speed.  Rotating tile row y by (cols mod 8) + 1
    // columns moves every XCD on by one column per tile row for any width.
    // (The staged form shares halo rows between neighbouring tiles; giving an
    // XCD whole tile rows or contiguous runs of tiles cut its HBM fetch from
    // 3.4 to 1.3-1.9 times the frames but made it 1.7 times slower, so it
    // keeps this rotation too.)
    const int rot = (int)(gridDim.x & 7u) + 1;
    const int bx = (int)((blockIdx.x + blockIdx.y * rot) % gridDim.x);
    const int by = (int)blockIdx.y;
    const int img = (int)blockIdx.z;
    const int x0 = bx * kTileX, y0 = by * TY;
    const int nx_t = min(kTileX, g.W - x0), ny_t = min(TY, g.H - y0);
    if (tid < kTileX) {
        const int x = x0 + min(tid, nx_t - 1);
        s_lx[tid] = pred_index((float)x, g.cx, g.nPX);
    } else if (tid < kTileX + TY) {
        const int y = y0 + min(tid - kTileX, ny_t - 1);
        s_ly[tid - kTileX] = pred_index((float)y, g.cy, g.nPY);
    }
    __syncthreads();
    // The staged predictor rectangle spans the tile's smallest to largest
    // predictor index.  pred_index is non-decreasing in the coordinate for
    // ascending centres, so these sit at the tile ends; one compare per lane
    // verifies it (NaN fails too).  A tile that fails reads the predictor
    // from global memory with clamped indices instead.
    bool bad = false;
    if (tid < kTileX) {
        const float v = s_lx[tid];
        bad = !(v == v) || (tid + 1 < kTileX && !(v <= s_lx[tid + 1]));
    } else if (tid < kTileX + TY) {
        const int k = tid - kTileX;
        const float v = s_ly[k];
        bad = !(v == v) || (k + 1 < TY && !(v <= s_ly[k + 1]));
    }
    const bool mono = !__syncthreads_or(bad);
    int pr0 = 0, pr1 = -1, pc0 = 0, pc1 = -1;
    if (mono) {
        const int fy0 = (int)floorf(s_ly[0]), fy1 = (int)floorf(s_ly[TY - 1]);
        const int fx0 = (int)floorf(s_lx[0]), fx1 = (int)floorf(s_lx[kTileX - 1]);
        pr0 = min(max(fy0 - 1, 0), g.nPY - 1);
        pr1 = min(max(fy1 + 2, 0), g.nPY - 1);
        pc0 = min(max(fx0 - 1, 0), g.nPX - 1);
        pc1 = min(max(fx1 + 2, 0), g.nPX - 1);
    }
    const int pcw = pc1 - pc0 + 1, prh = pr1 - pr0 + 1;
    const bool staged = mono && pr1 >= pr0 && pc1 >= pc0 && pcw * prh <= kPM;
    const float *dyp = g.pdy + (size_t)img * g.predStride;
    const float *dxp = g.pdx + (size_t)img * g.predStride;
    if (staged) {
        for (int e = tid; e < pcw * prh; e += NT) {
            const int r = e / pcw, c = e - r * pcw;
            s_pdy[e] = dyp[(pr0 + r) * g.nPX + pc0 + c];
            s_pdx[e] = dxp[(pr0 + r) * g.nPX + pc0 + c];
        }
    }
    const bool rowsums = FAST && staged && prh <= kRowSums;
    // y weights and row offsets of the predictor stencil, once per tile row
    // (FAST with row sums: the staged row index instead of the offset)
    if (tid < TY) {
        const float fiy = s_ly[tid], fiyf = floorf(fiy);
        const int iyb = (int)fiyf - 1;
        float wy[4];
        keys4(fiy - fiyf, wy);
        int ro[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const int rr = min(max(iyb + m, 0), g.nPY - 1);
            ro[m] = rowsums ? rr - pr0 : (staged ? (rr - pr0) * pcw - pc0 : rr * g.nPX);
        }
        s_wy[tid] = make_float4(wy[0], wy[1], wy[2], wy[3]);
        s_ro[tid] = make_int4(ro[0], ro[1], ro[2], ro[3]);
    }
    __syncthreads();

    if constexpr (FAST) {
        if constexpr (STAGE) {
            // wave 0: bound the tile's half-shifts from the staged predictor
            // cells and place the source rectangles of A and B
            if (tid < 64) {
                int ok = 0, r0a = 0, c0a = 0, ha = 0, r0b = 0, c0b = 0, hb = 0;
                if (staged) {
                    float mny = INFINITY, mxy = -INFINITY, mnx = INFINITY, mxx = -INFINITY;
                    bool bad = false;
                    for (int e = tid; e < pcw * prh; e += 64) {
                        const float vy = s_pdy[e], vx = s_pdx[e];
                        bad = bad || !(fabsf(vy) <= kStageShiftMax) || !(fabsf(vx) <= kStageShiftMax);
                        mny = fminf(mny, vy); mxy = fmaxf(mxy, vy);
                        mnx = fminf(mnx, vx); mxx = fmaxf(mxx, vx);
                    }
#pragma unroll
                    for (int o = 32; o; o >>= 1) {
                        mny = fminf(mny, __shfl_xor(mny, o));
                        mxy = fmaxf(mxy, __shfl_xor(mxy, o));
                        mnx = fminf(mnx, __shfl_xor(mnx, o));
                        mxx = fmaxf(mxx, __shfl_xor(mxx, o));
                    }
                    if (!__any(bad)) {
                        // half the interpolant's bound, widened by one pixel
                        // for rounding and round_shifts
                        const float ry = mxy - mny, rx = mxx - mnx;
                        const float hyl = 0.5f * (mny - 0.5f * ry) - 1.0f, hyh = 0.5f * (mxy + 0.5f * ry) + 1.0f;
                        const float hxl = 0.5f * (mnx - 0.5f * rx) - 1.0f, hxh = 0.5f * (mxx + 0.5f * rx) + 1.0f;
                        const float ylo = (float)y0, yhi = (float)(y0 + ny_t - 1);
                        const float xlo = (float)x0, xhi = (float)(x0 + nx_t - 1);
                        r0a = (int)floorf(ylo - hyh) - 1;
                        c0a = (int)floorf(xlo - hxh) - 1;
                        r0b = (int)floorf(ylo + hyl) - 1;
                        c0b = (int)floorf(xlo + hxl) - 1;
                        ha = (int)floorf(yhi - hyl) + 3 - r0a;
                        hb = (int)floorf(yhi + hyh) + 3 - r0b;
                        const int wa = (int)floorf(xhi - hxl) + 3 - c0a;
                        const int wb = (int)floorf(xhi + hxh) + 3 - c0b;
                        ok = ha <= kRegH && hb <= kRegH && wa <= kRegP && wb <= kRegP;
                    }
                }
                if (tid == 0) {
                    s_rg[0] = ok; s_rg[1] = r0a; s_rg[2] = c0a; s_rg[3] = ha;
                    s_rg[4] = r0b; s_rg[5] = c0b; s_rg[6] = hb; s_rg[7] = 0;
                }
            }
        }
        if (rowsums) {
            const float fxs = s_lx[tx], fxsf = floorf(fxs);
            const int ixs = (int)fxsf - 1;
            float wxs[4];
            keys4_fma(fxs - fxsf, wxs);
            int cs[4];
#pragma unroll
            for (int n = 0; n < 4; n++) cs[n] = min(max(ixs + n, 0), g.nPX - 1) - pc0;
            for (int r = ty; r < prh; r += LY) {
                const float *py = s_pdy + r * pcw, *px = s_pdx + r * pcw;
                float ry = wxs[0] * py[cs[0]], rx = wxs[0] * px[cs[0]];
#pragma unroll
                for (int n = 1; n < 4; n++) {
                    ry = __builtin_fmaf(wxs[n], py[cs[n]], ry);
                    rx = __builtin_fmaf(wxs[n], px[cs[n]], rx);
                }
                s_sy[r * kTileX + tx] = ry;
                s_sx[r * kTileX + tx] = rx;
            }
        }
        if (rowsums || STAGE) __syncthreads();
    }

    const int j = x0 + tx;
    if constexpr (!STAGE) {
        if (j >= g.W) return;
    }
    const size_t plane = (size_t)g.H * g.W;
    const TI *A = warp_src<TI>(g, true) + (size_t)img * plane;
    const TI *B = warp_src<TI>(g, false) + (size_t)img * plane;
    const int hm3 = max(g.H - 3, 0), wm3 = max(g.W - 3, 0);
    const int hm5 = max(g.H - 5, 0), wm5 = max(g.W - 5, 0);
    const FrameBuf fbA = frame_buf(A, plane * sizeof(TI)), fbB = frame_buf(B, plane * sizeof(TI));

    const float fx = s_lx[tx], fxf = floorf(fx);
    const int ix = (int)fxf - 1;
    float wx[4];
    keys4(fx - fxf, wx);
    int cols[4];
#pragma unroll
    for (int n = 0; n < 4; n++) cols[n] = min(max(ix + n, 0), g.nPX - 1);

    // the half-shifts (hy, hx) of tile row ly at this lane's column
    auto half_shifts = [&](int ly, float &hy, float &hx) {
        const float4 wy4 = s_wy[ly];
        const int4 ro4 = s_ro[ly];
        const float wy[4] = {wy4.x, wy4.y, wy4.z, wy4.w};
        const int ro[4] = {ro4.x, ro4.y, ro4.z, ro4.w};
        float va = 0.0f, vb = 0.0f;
        if constexpr (FAST) {
            // separate loops keep the staged reads LDS reads (no flat pointer)
            auto pred_sum = [&](const float *py, const float *px) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    float ry = wx[0] * py[ro[m] + cols[0]], rx = wx[0] * px[ro[m] + cols[0]];
#pragma unroll
                    for (int n = 1; n < 4; n++) {
                        ry = __builtin_fmaf(wx[n], py[ro[m] + cols[n]], ry);
                        rx = __builtin_fmaf(wx[n], px[ro[m] + cols[n]], rx);
                    }
                    va = __builtin_fmaf(wy[m], ry, va);
                    vb = __builtin_fmaf(wy[m], rx, vb);
                }
            };
            if (rowsums) {
                // the same terms in the same order as pred_sum, the x sums read back
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    va = __builtin_fmaf(wy[m], s_sy[ro[m] * kTileX + tx], va);
                    vb = __builtin_fmaf(wy[m], s_sx[ro[m] * kTileX + tx], vb);
                }
            } else if (staged) {
                pred_sum(s_pdy, s_pdx);
            } else {
                pred_sum(dyp, dxp);
            }
        } else if (staged) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
#pragma unroll
                for (int n = 0; n < 4; n++) {
                    const float w = wy[m] * wx[n];
                    va += w * s_pdy[ro[m] + cols[n]];
                    vb += w * s_pdx[ro[m] + cols[n]];
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++) {
#pragma unroll
                for (int n = 0; n < 4; n++) {
                    const float w = wy[m] * wx[n];
                    va += w * dyp[ro[m] + cols[n]];
                    vb += w * dxp[ro[m] + cols[n]];
                }
            }
        }
        hy = 0.5f * va;
        hx = 0.5f * vb;
        if (g.roundShifts) { hy = roundf(hy); hx = roundf(hx); }
        hy = fminf(fmaxf(hy, -kShiftMax), kShiftMax);
        hx = fminf(fmaxf(hx, -kShiftMax), kShiftMax);
    };

    if constexpr (STAGE) {
        const int ok = s_rg[0];
        const int r0a = s_rg[1], c0a = s_rg[2], ha = s_rg[3];
        const int r0b = s_rg[4], c0b = s_rg[5], hb = s_rg[6];
        // lane -> (column c of the rectangle, row phase grp); rows step by
        // kGroups, and every lane stages kSteps rows of each rectangle
        // (kGroups * kSteps = kRegH; rows past its height are never read).
        // Byte offsets are 32-bit: a row above the frame wraps to 2^31 or
        // more and a column outside it is moved there, so both lie beyond
        // the frame's num_records and read 0, as does a row below the frame
        // (the host keeps kStageRowsMax rows either side addressable this
        // way).  All loads of both rectangles are issued before the first
        // LDS write, and the half-shifts are computed while they are in flight.
        constexpr int kGroups = NT / kRegP;
        constexpr int kSteps = kRegH / kGroups;
        static_assert(kGroups * kSteps == kRegH, "staged rows");
        const bool stager = ok && tid < kGroups * kRegP;
        const int grp = tid / kRegP;
        const int c = tid - grp * kRegP;
        const uint32_t rstep = (uint32_t)(kGroups * g.W) * (uint32_t)sizeof(TI);
        float tv[2][kSteps];
        if (stager) {
            const int bytes = (int)(plane * sizeof(TI));
#pragma unroll
            for (int f = 0; f < 2; f++) {
                const int r0 = f ? r0b : r0a, c0 = f ? c0b : c0a;
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc((void *)(f ? B : A), (short)0, bytes, 0x00020000);
                const int cc = c0 + c;
                uint32_t off = (uint32_t)((r0 + grp) * g.W + cc) * (uint32_t)sizeof(TI);
                if ((unsigned)cc >= (unsigned)g.W) off = 0x80000000u;
#pragma unroll
                for (int k = 0; k < kSteps; k++) {
                    tv[f][k] = load_tap<TI>(rs, off);
                    off += rstep;
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        float hys[TY / LY], hxs[TY / LY];
#pragma unroll
        for (int r = 0; r < TY / LY; ++r) half_shifts(ty + r * LY, hys[r], hxs[r]);
        __builtin_amdgcn_sched_barrier(0);
        if (stager) {
#pragma unroll
            for (int f = 0; f < 2; f++) {
                float *dst = s_reg + f * (kRegH * kRegP) + grp * kRegP + c;
#pragma unroll
                for (int k = 0; k < kSteps; k++) dst[k * kGroups * kRegP] = tv[f][k];
            }
        }
        __syncthreads();
        if (j >= g.W) return;
        // one tap sum from the staged rectangle R (origin r0, c0; h rows),
        // the same terms in the same order as sample_bicubic_fast
        auto sample_staged = [&](const float *R, int r0, int c0, int h, const TI *src, const FrameBuf &fb,
                                 int iy, int ix, const float (&wys)[4], const float (&wxs)[4]) -> float {
            const int ry = iy - r0, rx = ix - c0;
            if (ok && (unsigned)ry <= (unsigned)(h - 4) && (unsigned)rx <= (unsigned)(kRegP - 4)) {
                // ry < kRegH and rx < kRegP here: a 24-bit multiply-add
                const float *p = R + (__umul24((unsigned)ry, (unsigned)kRegP) + (unsigned)rx);
                float val = 0.0f;
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    float rs = wxs[0] * p[m * kRegP];
                    rs = __builtin_fmaf(wxs[1], p[m * kRegP + 1], rs);
                    rs = __builtin_fmaf(wxs[2], p[m * kRegP + 2], rs);
                    rs = __builtin_fmaf(wxs[3], p[m * kRegP + 3], rs);
                    val = __builtin_fmaf(wys[m], rs, val);
                }
                return val;
            }
            return sample_bicubic_fast(src, iy, ix, wys, wxs, g.H, g.W, hm3, wm3, fb);
        };
        // output planes as raw buffers based at this image: one 32-bit lane
        // offset, the row step a uniform offset
        const int obytes = (int)(plane * sizeof(float));
        const __amdgpu_buffer_rsrc_t oa =
            __builtin_amdgcn_make_buffer_rsrc((void *)(g.outA + (size_t)img * plane), (short)0, obytes, 0x00020000);
        const __amdgpu_buffer_rsrc_t ob =
            __builtin_amdgcn_make_buffer_rsrc((void *)(g.outB + (size_t)img * plane), (short)0, obytes, 0x00020000);
        const int vo = ((y0 + ty) * g.W + j) * (int)sizeof(float);
#pragma unroll
        for (int r = 0; r < TY / LY; ++r) {
            const int i = y0 + ty + r * LY;
            if (i >= g.H) break;
            const int so = r * LY * g.W * (int)sizeof(float);
            const FmaStencil st = fma_stencil(i, j, hys[r], hxs[r]);
            const float wyr[4] = {st.wy[3], st.wy[2], st.wy[1], st.wy[0]};
            const float wxr[4] = {st.wx[3], st.wx[2], st.wx[1], st.wx[0]};
            const float va = sample_staged(s_reg, r0a, c0a, ha, A, fbA, st.iyA, st.ixA, st.wy, st.wx);
            const float vb = sample_staged(s_reg + kRegH * kRegP, r0b, c0b, hb, B, fbB, st.iyB, st.ixB, wyr, wxr);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, va), oa, vo, so, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, vb), ob, vo, so, 0);
        }
        return;
    }

#pragma unroll
    for (int r = 0; r < kTileY / kLanesY; ++r) {
        const int ly = ty + r * kLanesY;
        const int i = y0 + ly;
        if (i >= g.H) break;
        float hy, hx;
        half_shifts(ly, hy, hx);
        const size_t o = (size_t)img * plane + (size_t)i * g.W + j;
        if constexpr (FAST) {
            const FmaStencil st = fma_stencil(i, j, hy, hx);
            const float wyr[4] = {st.wy[3], st.wy[2], st.wy[1], st.wy[0]};
            const float wxr[4] = {st.wx[3], st.wx[2], st.wx[1], st.wx[0]};
            g.outA[o] = sample_bicubic_fast(A, st.iyA, st.ixA, st.wy, st.wx, g.H, g.W, hm3, wm3, fbA);
            g.outB[o] = sample_bicubic_fast(B, st.iyB, st.ixB, wyr, wxr, g.H, g.W, hm3, wm3, fbB);
        } else if (g.mode == 0) {
            g.outA[o] = sample_bicubic(A, (float)i - hy, (float)j - hx, g.H, g.W, hm3, wm3);
            g.outB[o] = sample_bicubic(B, (float)i + hy, (float)j + hx, g.H, g.W, hm3, wm3);
        } else {
            g.outA[o] = sample_lanczos(A, (float)i - hy, (float)j - hx, g.H, g.W, hm5, wm5, g.lut);
            g.outB[o] = sample_lanczos(B, (float)i + hy, (float)j + hx, g.H, g.W, hm5, wm5, g.lut);
        }
    }
}

#pragma clang fp contract(on)

// Lanczos-3 weight table built on the host with the host libm (the same
// table fused_warp.c:148-169 builds), uploaded once per device.
static float *g_lut_dev[64];

static float lanczos_weight_host(float t)
{
    const float at = fabsf(t);
    if (at < 1e-6f) return 1.0f;
    if (at >= 3.0f) return 0.0f;
    const float pt = (float)M_PI * at;
    const float pta = pt / 3.0f;
    return (sinf(pt) / pt) * (sinf(pta) / pta);
}

static const float *lanczos_lut_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (g_lut_dev[dev]) return g_lut_dev[dev];
    static float host[(kLanczosLut + 1) * 6];
    for (int i = 0; i <= kLanczosLut; i++) {
        const float d = (float)i / (float)kLanczosLut;
        for (int k = 0; k < 6; k++) host[i * 6 + k] = lanczos_weight_host(d - (float)(k - 2));
    }
    float *p = nullptr;
    if (hipMalloc(&p, sizeof(host)) != hipSuccess) return nullptr;
    if (hipMemcpy(p, host, sizeof(host), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    g_lut_dev[dev] = p;
    return p;
}

}  // namespace pivamd

using namespace pivamd;

// Form of the FMA warp (mode 2): 1 stages each tile's source rectangles in
// LDS (the default), 0 reads every tap from global memory.  Process-global,
// like the reference's fused_warp_set_impl selector.
static int g_warp_variant = 1;

extern "C" int pivamd_warp_set_variant(int v)
{
    if (v < 0 || v > 1) return PIVAMD_ERR_ARG;
    g_warp_variant = v;
    return PIVAMD_OK;
}

extern "C" int pivamd_warp_get_variant(void) { return g_warp_variant; }

static int warp_entry(const float *d_imgA, const float *d_imgB, const unsigned short *d_imgA16,
                      const unsigned short *d_imgB16, float *d_outA, float *d_outB, const float *d_predDy,
                      const float *d_predDx, int N, int H, int W, int nPY, int nPX, const float *d_ctrY,
                      const float *d_ctrX, int interp_mode, int shared_predictor, int round_shifts,
                      hipStream_t st)
{
    if (N <= 0 || H <= 0 || W <= 0 || nPY <= 0 || nPX <= 0) return PIVAMD_ERR_ARG;
    if (!(d_imgA16 ? d_imgB16 != nullptr : (d_imgA && d_imgB)) || !d_outA || !d_outB || !d_predDy ||
        !d_predDx || !d_ctrY || !d_ctrX)
        return PIVAMD_ERR_ARG;
    if (N > 65535) return PIVAMD_ERR_ARG;
    WarpArgs g;
    g.imgA = d_imgA; g.imgB = d_imgB; g.imgA16 = d_imgA16; g.imgB16 = d_imgB16;
    g.pdy = d_predDy; g.pdx = d_predDx;
    g.cy = d_ctrY; g.cx = d_ctrX; g.outA = d_outA; g.outB = d_outB;
    g.H = H; g.W = W; g.nPY = nPY; g.nPX = nPX;
    g.predStride = shared_predictor ? 0 : nPY * nPX;
    if (interp_mode < 0 || interp_mode > 2) return PIVAMD_ERR_ARG;
    g.mode = interp_mode;
    g.roundShifts = round_shifts ? 1 : 0;
    g.lut = nullptr;
    if (g.mode == 1) {
        g.lut = lanczos_lut_device();
        if (!g.lut) return PIVAMD_ERR_NOMEM;
    }
    dim3 grid((W + kTileX - 1) / kTileX, (H + kTileY - 1) / kTileY, N);
    // the staged form addresses a frame with 32-bit byte offsets
    // (and the output planes, 4 bytes per pixel, below 2 GiB); rows within
    // kStageRowsMax of the frame stay below 2^31 bytes either side of it
    const size_t row_bytes = (size_t)W * (d_imgA16 ? 2 : 4);
    const bool stage = g_warp_variant == 1 && (size_t)H * W < ((size_t)1 << 29) &&
                       ((size_t)H + kStageRowsMax) * row_bytes < ((size_t)1 << 31);
    if (g.mode == 2 && stage) {
        dim3 grid_s((W + kTileX - 1) / kTileX, (H + kTileYS - 1) / kTileYS, N);
        const int nt = kTileX * kLanesYS;
        if (d_imgA16) hipLaunchKernelGGL((warp_kernel<unsigned short, true, true>), grid_s, dim3(nt), 0, st, g);
        else hipLaunchKernelGGL((warp_kernel<float, true, true>), grid_s, dim3(nt), 0, st, g);
    } else if (g.mode == 2) {
        if (d_imgA16) hipLaunchKernelGGL((warp_kernel<unsigned short, true>), grid, dim3(256), 0, st, g);
        else hipLaunchKernelGGL((warp_kernel<float, true>), grid, dim3(256), 0, st, g);
    } else {
        if (d_imgA16) hipLaunchKernelGGL((warp_kernel<unsigned short, false>), grid, dim3(256), 0, st, g);
        else hipLaunchKernelGGL((warp_kernel<float, false>), grid, dim3(256), 0, st, g);
    }
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

extern "C" int pivamd_warp(const float *d_imgA, const float *d_imgB, float *d_outA, float *d_outB,
                           const float *d_predDy, const float *d_predDx, int N, int H, int W,
                           int nPY, int nPX, const float *d_ctrY, const float *d_ctrX,
                           int interp_mode, int shared_predictor, int round_shifts, hipStream_t st)
{
    return warp_entry(d_imgA, d_imgB, nullptr, nullptr, d_outA, d_outB, d_predDy, d_predDx, N, H, W, nPY,
                      nPX, d_ctrY, d_ctrX, interp_mode, shared_predictor, round_shifts, st);
}

extern "C" int pivamd_warp_u16(const unsigned short *d_imgA, const unsigned short *d_imgB, float *d_outA,
                               float *d_outB, const float *d_predDy, const float *d_predDx, int N, int H,
                               int W, int nPY, int nPX, const float *d_ctrY, const float *d_ctrX,
                               int interp_mode, int shared_predictor, int round_shifts, hipStream_t st)
{
    if (!d_imgA || !d_imgB) return PIVAMD_ERR_ARG;
    return warp_entry(nullptr, nullptr, d_imgA, d_imgB, d_outA, d_outB, d_predDy, d_predDx, N, H, W, nPY,
                      nPX, d_ctrY, d_ctrX, interp_mode, shared_predictor, round_shifts, st);
}
