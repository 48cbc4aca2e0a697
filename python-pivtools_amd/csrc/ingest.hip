// ingest.hip -- camera payloads decoded on the device.
//
// The reference decodes raw camera frames on the host before correlating
// them (pivtools_core/image_handling/readers/set_reader.py:1296-1350
// decode_mono12p; pivraw_reader.py:896-961 _sensor_pixels / _decode_into):
//   Mono12p   PFNC LSB-first packing, two 12-bit pixels in three bytes:
//             pixel j is bits [12j, 12j + 12) of the little-endian byte
//             stream (p0 = b0 | (b1 & 0x0F) << 8, p1 = b1 >> 4 | b2 << 4);
//   dark      max(p, d) - d per photosite in the sensor type, so the unsigned
//             subtraction never wraps;
//   BayerRG8  the 8-bit mosaic read as 2x2 superpixels, each the sum of its
//             four photosites after their own dark subtraction.
// Here the packed payload is what crosses PCIe (1.5 bytes per pixel instead
// of 2) and these kernels expand it in HBM into the uint16 frames that the
// correlation and warp kernels read directly.  Both kernels are HBM-bound
// byte work: each lane moves whole aligned words when the buffers allow it
// and falls back to byte loads otherwise (odd strides, ragged tails).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pivamd_dev.h"

namespace pivamd {

__device__ __forceinline__ uint32_t sub_dark(uint32_t p, uint32_t d) { return p > d ? p - d : 0u; }

// One lane per group of 8 pixels (12 payload bytes).  Frame f of the batch
// is blockIdx.y; its payload starts at raw + f * stride.  FAST: raw and
// stride are 4-byte aligned, out and dark 16-byte aligned and npx % 8 == 0,
// so a group is three dword loads and one 16-byte store.
template <bool FAST, bool DARK>
__global__ void __launch_bounds__(256) mono12p_kernel(const uint8_t *__restrict__ raw, size_t stride, size_t npx,
                                                      const uint16_t *__restrict__ dark, uint16_t *__restrict__ out)
{
    const uint8_t *src = raw + (size_t)blockIdx.y * stride;
    uint16_t *dst = out + (size_t)blockIdx.y * npx;
    const size_t groups = (npx + 7) / 8;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += step) {
        uint32_t p[8];
        if (FAST) {
            const uint32_t *s = (const uint32_t *)(src + 12 * g);
            const uint32_t d0 = s[0], d1 = s[1], d2 = s[2];
            p[0] = d0 & 0xFFFu;
            p[1] = (d0 >> 12) & 0xFFFu;
            p[2] = (d0 >> 24) | ((d1 & 0xFu) << 8);
            p[3] = (d1 >> 4) & 0xFFFu;
            p[4] = (d1 >> 16) & 0xFFFu;
            p[5] = (d1 >> 28) | ((d2 & 0xFFu) << 4);
            p[6] = (d2 >> 8) & 0xFFFu;
            p[7] = d2 >> 20;
            if (DARK) {
                const uint4 dv = *(const uint4 *)(dark + 8 * g);
                const uint32_t dw[4] = {dv.x, dv.y, dv.z, dv.w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    p[2 * k] = sub_dark(p[2 * k], dw[k] & 0xFFFFu);
                    p[2 * k + 1] = sub_dark(p[2 * k + 1], dw[k] >> 16);
                }
            }
            uint4 o;
            o.x = p[0] | (p[1] << 16);
            o.y = p[2] | (p[3] << 16);
            o.z = p[4] | (p[5] << 16);
            o.w = p[6] | (p[7] << 16);
            *(uint4 *)(dst + 8 * g) = o;
        } else {
            // npx is even, so a group holds whole pixel pairs
            const size_t left = npx - 8 * g;
            const int cnt = left < 8 ? (int)left : 8;
            for (int k = 0; k < cnt; k += 2) {
                const uint8_t *b = src + 12 * g + 3 * (k / 2);
                const uint32_t b0 = b[0], b1 = b[1], b2 = b[2];
                uint32_t q0 = b0 | ((b1 & 0xFu) << 8);
                uint32_t q1 = (b1 >> 4) | (b2 << 4);
                if (DARK) {
                    q0 = sub_dark(q0, dark[8 * g + k]);
                    q1 = sub_dark(q1, dark[8 * g + k + 1]);
                }
                dst[8 * g + k] = (uint16_t)q0;
                dst[8 * g + k + 1] = (uint16_t)q1;
            }
        }
    }
}

// One lane per 4 superpixels of one output row.  Sensor frame (H, W) of
// 8-bit photosites at raw + f * stride, output (H/2, W/2) uint16.  FAST: W
// a multiple of 8, raw, stride and dark 8-byte aligned, out 8-byte aligned.
template <bool FAST, bool DARK>
__global__ void __launch_bounds__(256) bayer_rg8_kernel(const uint8_t *__restrict__ raw, size_t stride, int H, int W,
                                                        const uint8_t *__restrict__ dark, uint16_t *__restrict__ out)
{
    const uint8_t *src = raw + (size_t)blockIdx.y * stride;
    const int ho = H / 2, wo = W / 2;
    uint16_t *dst = out + (size_t)blockIdx.y * ho * wo;
    const int per_row = (wo + 3) / 4;
    const size_t items = (size_t)ho * per_row;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t it = (size_t)blockIdx.x * blockDim.x + threadIdx.x; it < items; it += step) {
        const int r = (int)(it / per_row);
        const int c0 = (int)(it % per_row) * 4;
        const size_t top = (size_t)(2 * r) * W, bot = top + W;
        if (FAST) {
            const uint2 t = *(const uint2 *)(src + top + 2 * c0);
            const uint2 u = *(const uint2 *)(src + bot + 2 * c0);
            const uint32_t tw[2] = {t.x, t.y}, uw[2] = {u.x, u.y};
            uint32_t dtw[2] = {0u, 0u}, duw[2] = {0u, 0u};
            if (DARK) {
                const uint2 dt = *(const uint2 *)(dark + top + 2 * c0);
                const uint2 du = *(const uint2 *)(dark + bot + 2 * c0);
                dtw[0] = dt.x;
                dtw[1] = dt.y;
                duw[0] = du.x;
                duw[1] = du.y;
            }
            uint32_t s[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int w = k >> 1, sh = (k & 1) * 16;
                uint32_t a0 = (tw[w] >> sh) & 0xFFu, a1 = (tw[w] >> (sh + 8)) & 0xFFu;
                uint32_t c0v = (uw[w] >> sh) & 0xFFu, c1v = (uw[w] >> (sh + 8)) & 0xFFu;
                if (DARK) {
                    a0 = sub_dark(a0, (dtw[w] >> sh) & 0xFFu);
                    a1 = sub_dark(a1, (dtw[w] >> (sh + 8)) & 0xFFu);
                    c0v = sub_dark(c0v, (duw[w] >> sh) & 0xFFu);
                    c1v = sub_dark(c1v, (duw[w] >> (sh + 8)) & 0xFFu);
                }
                s[k] = a0 + a1 + c0v + c1v;
            }
            uint2 o;
            o.x = s[0] | (s[1] << 16);
            o.y = s[2] | (s[3] << 16);
            *(uint2 *)(dst + (size_t)r * wo + c0) = o;
        } else {
            const int cnt = wo - c0 < 4 ? wo - c0 : 4;
            for (int k = 0; k < cnt; ++k) {
                const size_t x = 2 * (size_t)(c0 + k);
                uint32_t a0 = src[top + x], a1 = src[top + x + 1], b0 = src[bot + x], b1 = src[bot + x + 1];
                if (DARK) {
                    a0 = sub_dark(a0, dark[top + x]);
                    a1 = sub_dark(a1, dark[top + x + 1]);
                    b0 = sub_dark(b0, dark[bot + x]);
                    b1 = sub_dark(b1, dark[bot + x + 1]);
                }
                dst[(size_t)r * wo + c0 + k] = (uint16_t)(a0 + a1 + b0 + b1);
            }
        }
    }
}

static inline bool aligned(const void *p, size_t a) { return ((uintptr_t)p % a) == 0; }

static inline unsigned grid_x(size_t items)
{
    size_t g = (items + 255) / 256;
    return (unsigned)(g > 4096 ? 4096 : (g ? g : 1));
}


// DaVis .set stream decode: the packed pixel encodings of a frame stream
// (set_reader.py:1196-1480), its file-backed corrections in LaVision's order
// (intensity-correction-1, dark-image-subtraction, background-subtraction,
// each clipped: set_reader.py:477-565) and at most one geometry
// (set_reader.py:566-585), fused into one pass.  Each lane writes one output
// pixel: the geometry maps it to its sensor pixel k, which is cut out of the
// two payload bytes that hold it.  Encodings, by sensor pixel k:
//   0 mono-10p       LE16 at 5(k/4) + k%4, >> 2(k%4), & 0x3FF
//   1 mono-10pmsb    BE16 at 5(k/4) + k%4, >> 6 - 2(k%4), & 0x3FF
//   2 mono-12p       LE16 at 3(k/2) + k%2: & 0xFFF, or >> 4
//   3 mono-12pmsb    BE16 at 3(k/2) + k%2: >> 4, or & 0xFFF
//   4 mono-12packed  b0 << 4 | b1 & 0xF, or b2 << 4 | b1 >> 4
//   5 raw-16-bit     LE16 at 2k
//   6 raw-8-bit      byte k
// Geometries (output (i, j) of an H x W sensor frame x):
//   0 none x[i][j]; 1 rotate-180 x[H-1-i][W-1-j]; 2 vertical-mirror x[H-1-i][j];
//   3 horizontal mirror x[i][W-1-j]; 4 rotate-cw-90 x[H-1-j][i];
//   5 rotate-ccw-90 x[j][W-1-i]; 6 rotate-cw-90 + vertical mirror x[H-1-j][W-1-i];
//   7 rotate-ccw-90 + vertical mirror x[j][i].
// The 90-degree geometries give a W x H output.
struct DavisArgs {
    const uint8_t *raw;
    size_t stride;                 // payload bytes from one frame to the next
    const uint16_t *scmos, *dark, *bgnd;   // sensor-shape corrections, nullable
    uint16_t *out;
    int H, W, enc, geo;
};

__device__ __forceinline__ uint32_t le16(const uint8_t *b) { return (uint32_t)b[0] | (uint32_t)b[1] << 8; }
__device__ __forceinline__ uint32_t be16(const uint8_t *b) { return (uint32_t)b[0] << 8 | (uint32_t)b[1]; }

template <int ENC>
__device__ __forceinline__ uint32_t davis_pixel(const uint8_t *src, size_t k)
{
    if constexpr (ENC == 0) {
        const uint32_t m = (uint32_t)(k & 3);
        return (le16(src + 5 * (k >> 2) + m) >> (2 * m)) & 0x3FFu;
    } else if constexpr (ENC == 1) {
        const uint32_t m = (uint32_t)(k & 3);
        return (be16(src + 5 * (k >> 2) + m) >> (6 - 2 * m)) & 0x3FFu;
    } else if constexpr (ENC == 2) {
        const uint32_t w = le16(src + 3 * (k >> 1) + (k & 1));
        return (k & 1) ? w >> 4 : w & 0xFFFu;
    } else if constexpr (ENC == 3) {
        const uint32_t w = be16(src + 3 * (k >> 1) + (k & 1));
        return (k & 1) ? w & 0xFFFu : w >> 4;
    } else if constexpr (ENC == 4) {
        const uint8_t *b = src + 3 * (k >> 1);
        return (k & 1) ? ((uint32_t)b[2] << 4 | (uint32_t)b[1] >> 4) : ((uint32_t)b[0] << 4 | ((uint32_t)b[1] & 0xFu));
    } else if constexpr (ENC == 5) {
        return le16(src + 2 * k);
    } else {
        return src[k];
    }
}

template <int ENC>
__global__ void __launch_bounds__(256) davis_decode_kernel(DavisArgs g)
{
    const bool swap = g.geo >= 4;
    const int Ho = swap ? g.W : g.H, Wo = swap ? g.H : g.W;
    const size_t npx = (size_t)g.H * g.W;
    const uint8_t *src = g.raw + (size_t)blockIdx.y * g.stride;
    uint16_t *dst = g.out + (size_t)blockIdx.y * npx;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < npx; o += step) {
        const int i = (int)(o / (size_t)Wo), j = (int)(o - (size_t)i * Wo);
        int r, c;
        switch (g.geo) {
        case 1: r = g.H - 1 - i; c = g.W - 1 - j; break;
        case 2: r = g.H - 1 - i; c = j; break;
        case 3: r = i; c = g.W - 1 - j; break;
        case 4: r = g.H - 1 - j; c = i; break;
        case 5: r = j; c = g.W - 1 - i; break;
        case 6: r = g.H - 1 - j; c = g.W - 1 - i; break;
        case 7: r = j; c = i; break;
        default: r = i; c = j; break;
        }
        (void)Ho;
        const size_t k = (size_t)r * g.W + c;
        int32_t v = (int32_t)davis_pixel<ENC>(src, k);
        if (g.scmos) v = min(max(v - (int32_t)g.scmos[k] + 300, 0), 65535);
        if (g.dark) v = max(v - (int32_t)g.dark[k], 0);
        if (g.bgnd) v = max(v - (int32_t)g.bgnd[k], 0);
        dst[o] = (uint16_t)v;
    }
}

// Payload bytes of one frame of n pixels in encoding enc (0 when the pixel
// count does not fill whole groups).
__host__ __device__ inline size_t davis_payload_bytes(int enc, size_t n)
{
    switch (enc) {
    case 0: case 1: return n % 4 ? 0 : n / 4 * 5;
    case 2: case 3: case 4: return n % 2 ? 0 : n / 2 * 3;
    case 5: return 2 * n;
    case 6: return n;
    default: return 0;
    }
}

// ---------------------------------------------------------------------------
// TIFF samples to float32 gray frames (generic_readers.py:33-51, 81-122): the
// host decompresses the strips or tiles (and undoes a horizontal predictor);
// here every pixel of n frames of S samples (interleaved, or planar S x H x W
// when PLANAR) becomes one float32 value:
//   S == 1        the sample;
//   S == 3 or 4   BT.601 luma of the first three samples (R, G, B; a fourth
//                 is alpha and ignored), OpenCV's RGB2GRAY rule: for 8- and
//                 16-bit samples (R 9798 + G 19235 + B 3735 + 2^14) >> 15, for
//                 float samples R 0.299f + G 0.587f + B 0.114f in that order;
//   other S       the mean of the samples (numpy's mean: a double sum for
//                 integer samples, a float sum for float ones).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ float gray_luma(T r, T g, T b)
{
    if constexpr (sizeof(T) <= 2 && T(-1) > T(0)) {
        const uint32_t y = (uint32_t)r * 9798u + (uint32_t)g * 19235u + (uint32_t)b * 3735u + 16384u;
        return (float)(y >> 15);
    } else {
        // three rounded products and two rounded sums, as OpenCV's scalar loop
#pragma clang fp contract(off)
        const float pr = (float)r * 0.299f, pg = (float)g * 0.587f, pb = (float)b * 0.114f;
        return (pr + pg) + pb;
    }
}

template <typename T, bool PLANAR>
__global__ void __launch_bounds__(256) gray_kernel(const T *__restrict__ src, size_t npx, int S, int luma,
                                                   float *__restrict__ out)
{
    const size_t f = blockIdx.y;
    const T *fr = src + f * npx * S;
    float *dst = out + f * npx;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx; i += step) {
        auto at = [&](int k) -> T { return PLANAR ? fr[(size_t)k * npx + i] : fr[i * S + k]; };
        float v;
        if (S == 1) {
            v = (float)at(0);
        } else if (luma) {
            v = gray_luma<T>(at(0), at(1), at(2));
        } else if constexpr (sizeof(T) == 4 && T(0.5) != T(0)) {
            float a = at(0);
            for (int k = 1; k < S; ++k) a += at(k);
            v = a / (float)S;
        } else {
            double a = 0.0;
            for (int k = 0; k < S; ++k) a += (double)at(k);
            v = (float)(a / (double)S);
        }
        dst[i] = v;
    }
}

template <typename T>
static int launch_gray(const void *src, int n, size_t npx, int S, int planar, int luma, float *out, hipStream_t st)
{
    const size_t blocks = (npx + 255) / 256;
    const dim3 grid((unsigned)(blocks < 4096 ? blocks : 4096), (unsigned)n);
    if (planar)
        hipLaunchKernelGGL((gray_kernel<T, true>), grid, dim3(256), 0, st, (const T *)src, npx, S, luma, out);
    else
        hipLaunchKernelGGL((gray_kernel<T, false>), grid, dim3(256), 0, st, (const T *)src, npx, S, luma, out);
    return hipGetLastError() == hipSuccess ? PIVAMD_OK : PIVAMD_ERR_LAUNCH;
}

}  // namespace pivamd

using namespace pivamd;

// sample type codes: 0 uint8, 1 uint16, 2 int8, 3 int16, 4 uint32, 5 int32, 6 float32.
// Luma (S of 3 or 4) is defined for uint8, uint16 and float32, as OpenCV's
// conversion is; other types with S of 3 or 4 are refused (1).
extern "C" int pivamd_gray_from_samples(const void *d_src, int n_frames, int H, int W, int S, int planar,
                                        int sample_type, float *d_out, hipStream_t st)
{
    if (n_frames < 0 || H <= 0 || W <= 0 || S < 1 || n_frames > 65535) return PIVAMD_ERR_ARG;
    if (n_frames == 0) return PIVAMD_OK;
    if (!d_src || !d_out) return PIVAMD_ERR_ARG;
    const int luma = (S == 3 || S == 4) ? 1 : 0;
    if (luma && sample_type != 0 && sample_type != 1 && sample_type != 6) return PIVAMD_ERR_ARG;
    const size_t npx = (size_t)H * W;
    switch (sample_type) {
    case 0: return launch_gray<uint8_t>(d_src, n_frames, npx, S, planar, luma, d_out, st);
    case 1: return launch_gray<uint16_t>(d_src, n_frames, npx, S, planar, luma, d_out, st);
    case 2: return launch_gray<int8_t>(d_src, n_frames, npx, S, planar, luma, d_out, st);
    case 3: return launch_gray<int16_t>(d_src, n_frames, npx, S, planar, luma, d_out, st);
    case 4: return launch_gray<uint32_t>(d_src, n_frames, npx, S, planar, luma, d_out, st);
    case 5: return launch_gray<int32_t>(d_src, n_frames, npx, S, planar, luma, d_out, st);
    case 6: return launch_gray<float>(d_src, n_frames, npx, S, planar, luma, d_out, st);
    default: return PIVAMD_ERR_ARG;
    }
}

#define PIVAMD_LAUNCHED() (hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH)

extern "C" int pivamd_unpack_mono12p(const void *d_raw, size_t frame_stride, int n_frames, int H, int W,
                                     const uint16_t *d_dark, uint16_t *d_out, hipStream_t st)
{
    if (n_frames < 0 || H < 0 || W < 0 || n_frames > 65535) return PIVAMD_ERR_ARG;
    const size_t npx = (size_t)H * W;
    if (n_frames == 0 || npx == 0) return 0;
    if (npx % 2) return PIVAMD_ERR_SIZE;   // two pixels per three bytes
    if (frame_stride < npx / 2 * 3 && n_frames > 1) return PIVAMD_ERR_ARG;
    if (!d_raw || !d_out) return PIVAMD_ERR_ARG;
    const uint8_t *raw = (const uint8_t *)d_raw;
    const bool fast = npx % 8 == 0 && aligned(raw, 4) && frame_stride % 4 == 0 && aligned(d_out, 16) &&
                      (!d_dark || aligned(d_dark, 16));
    const dim3 grid(grid_x((npx + 7) / 8), (unsigned)n_frames);
    if (fast) {
        if (d_dark) hipLaunchKernelGGL((mono12p_kernel<true, true>), grid, dim3(256), 0, st, raw, frame_stride, npx, d_dark, d_out);
        else hipLaunchKernelGGL((mono12p_kernel<true, false>), grid, dim3(256), 0, st, raw, frame_stride, npx, d_dark, d_out);
    } else {
        if (d_dark) hipLaunchKernelGGL((mono12p_kernel<false, true>), grid, dim3(256), 0, st, raw, frame_stride, npx, d_dark, d_out);
        else hipLaunchKernelGGL((mono12p_kernel<false, false>), grid, dim3(256), 0, st, raw, frame_stride, npx, d_dark, d_out);
    }
    return PIVAMD_LAUNCHED();
}

extern "C" int pivamd_bayer_rg8_superpixel(const void *d_raw, size_t frame_stride, int n_frames, int H, int W,
                                           const uint8_t *d_dark, uint16_t *d_out, hipStream_t st)
{
    if (n_frames < 0 || H < 0 || W < 0 || n_frames > 65535) return PIVAMD_ERR_ARG;
    if (n_frames == 0 || H == 0 || W == 0) return 0;
    if (H % 2 || W % 2) return PIVAMD_ERR_SIZE;   // 2x2 superpixels
    if (frame_stride < (size_t)H * W && n_frames > 1) return PIVAMD_ERR_ARG;
    if (!d_raw || !d_out) return PIVAMD_ERR_ARG;
    const uint8_t *raw = (const uint8_t *)d_raw;
    const bool fast = W % 8 == 0 && aligned(raw, 8) && frame_stride % 8 == 0 && aligned(d_out, 8) &&
                      (!d_dark || aligned(d_dark, 8));
    const size_t items = (size_t)(H / 2) * ((W / 2 + 3) / 4);
    const dim3 grid(grid_x(items), (unsigned)n_frames);
    if (fast) {
        if (d_dark) hipLaunchKernelGGL((bayer_rg8_kernel<true, true>), grid, dim3(256), 0, st, raw, frame_stride, H, W, d_dark, d_out);
        else hipLaunchKernelGGL((bayer_rg8_kernel<true, false>), grid, dim3(256), 0, st, raw, frame_stride, H, W, d_dark, d_out);
    } else {
        if (d_dark) hipLaunchKernelGGL((bayer_rg8_kernel<false, true>), grid, dim3(256), 0, st, raw, frame_stride, H, W, d_dark, d_out);
        else hipLaunchKernelGGL((bayer_rg8_kernel<false, false>), grid, dim3(256), 0, st, raw, frame_stride, H, W, d_dark, d_out);
    }
    return PIVAMD_LAUNCHED();
}


extern "C" unsigned long pivamd_davis_payload_bytes(int encoding, int H, int W)
{
    if (H <= 0 || W <= 0) return 0;
    return davis_payload_bytes(encoding, (size_t)H * W);
}

extern "C" int pivamd_davis_decode(const void *d_raw, size_t frame_stride, int n_frames, int H, int W,
                                   int encoding, const uint16_t *d_scmos, const uint16_t *d_dark,
                                   const uint16_t *d_bgnd, int geometry, uint16_t *d_out, hipStream_t st)
{
    if (n_frames < 0 || H < 0 || W < 0 || n_frames > 65535) return PIVAMD_ERR_ARG;
    if (encoding < 0 || encoding > 6 || geometry < 0 || geometry > 7) return PIVAMD_ERR_ARG;
    const size_t npx = (size_t)H * W;
    if (n_frames == 0 || npx == 0) return 0;
    const size_t need = davis_payload_bytes(encoding, npx);
    if (need == 0) return PIVAMD_ERR_SIZE;    // the packing needs whole groups
    if (frame_stride < need && n_frames > 1) return PIVAMD_ERR_ARG;
    if (!d_raw || !d_out) return PIVAMD_ERR_ARG;
    DavisArgs g;
    g.raw = (const uint8_t *)d_raw; g.stride = frame_stride;
    g.scmos = d_scmos; g.dark = d_dark; g.bgnd = d_bgnd; g.out = d_out;
    g.H = H; g.W = W; g.enc = encoding; g.geo = geometry;
    const dim3 grid(grid_x(npx), (unsigned)n_frames);
    switch (encoding) {
    case 0: hipLaunchKernelGGL(davis_decode_kernel<0>, grid, dim3(256), 0, st, g); break;
    case 1: hipLaunchKernelGGL(davis_decode_kernel<1>, grid, dim3(256), 0, st, g); break;
    case 2: hipLaunchKernelGGL(davis_decode_kernel<2>, grid, dim3(256), 0, st, g); break;
    case 3: hipLaunchKernelGGL(davis_decode_kernel<3>, grid, dim3(256), 0, st, g); break;
    case 4: hipLaunchKernelGGL(davis_decode_kernel<4>, grid, dim3(256), 0, st, g); break;
    case 5: hipLaunchKernelGGL(davis_decode_kernel<5>, grid, dim3(256), 0, st, g); break;
    default: hipLaunchKernelGGL(davis_decode_kernel<6>, grid, dim3(256), 0, st, g); break;
    }
    return PIVAMD_LAUNCHED();
}
