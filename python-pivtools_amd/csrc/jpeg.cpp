// jpeg.cpp -- baseline JPEG frames, decoded on the host.
//
// The reference reads JPEG through OpenCV's imread (IMREAD_UNCHANGED), which
// decodes with libjpeg-turbo at its defaults, as Pillow does
// (generic_readers.py:125-176; unit-tests/test_generic_readers.py:228-258
// take Pillow's decode as the truth).  This is a decoder of our own for the
// sequential Huffman-coded 8-bit process of ITU-T T.81, with the arithmetic
// those defaults use, so its samples are the same bytes:
//   * Huffman decoding and dequantisation (T.81 F.2.2, restart intervals);
//   * the 8x8 inverse DCT in 32-bit integers: the Loeffler-Ligtenberg-
//     Moschytz flow graph with 13-bit constants and 2 extra bits between
//     the column and row passes, the "islow" method of the IJG library,
//     with its range-limit of the result (values wrap modulo 1024 first);
//   * "fancy" chroma upsampling: a triangle filter (3/4, 1/4) across each
//     doubled axis, with the IJG rounding biases and edge replication;
//   * YCbCr to RGB with 16-bit fixed-point coefficients 1.402, 0.34414,
//     0.71414, 1.772 (JFIF), rounded as the IJG tables round them.
// Progressive, lossless, hierarchical and arithmetic-coded files, 12-bit
// samples, four-component (CMYK / YCCK) files and sampling factors other
// than 1x1, 2x1 and 2x2 relative to the luma are refused with a status code.
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

enum JpegStatus {
    kJOk = 0,
    kJMalformed = 1,
    kJProgressive = 2,
    kJArithmetic = 3,
    kJLossless = 4,
    kJPrecision = 5,
    kJComponents = 6,
    kJSampling = 7,
    kJSmall = 8,
};

struct Huff {
    bool set = false;
    int mincode[17], maxcode[18], valptr[17];
    uint8_t vals[256];
};

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int bw = 0, bh = 0;             // blocks across and down (MCU-padded)
    int cw = 0, ch = 0;             // component samples across and down
    std::vector<int16_t> coef;      // bw * bh blocks of 64 coefficients, natural order
};

struct Reader {
    const uint8_t *d;
    size_t n, pos = 0;
    uint32_t acc = 0;
    int bits = 0;
    bool marker = false;            // hit a marker: feed zeros
    int get_bit()
    {
        if (bits == 0) {
            uint8_t b = 0;
            if (!marker && pos < n) {
                b = d[pos];
                if (b == 0xFF) {
                    const uint8_t nx = pos + 1 < n ? d[pos + 1] : 0;
                    if (nx == 0x00) pos += 2;
                    else { marker = true; b = 0; }
                } else {
                    ++pos;
                }
            }
            acc = b;
            bits = 8;
        }
        --bits;
        return (int)((acc >> bits) & 1u);
    }
    int get_bits(int s)
    {
        int v = 0;
        for (int i = 0; i < s; ++i) v = (v << 1) | get_bit();
        return v;
    }
};

int decode_huff(Reader &r, const Huff &h)
{
    int code = r.get_bit();
    int l = 1;
    while (l <= 16 && code > h.maxcode[l]) {
        code = (code << 1) | r.get_bit();
        ++l;
    }
    if (l > 16) return -1;
    const int k = h.valptr[l] + code - h.mincode[l];
    if (k < 0 || k > 255) return -1;
    return h.vals[k];
}

int extend(int v, int s) { return (s == 0) ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }

const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                         12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                         35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                         58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// islow inverse DCT of one block of coefficients (dequantised on the fly
// with q, both in natural order) into 8x8 samples (row stride `stride`).
inline int descale(int64_t x, int n) { return (int)((x + ((int64_t)1 << (n - 1))) >> n); }

inline uint8_t idct_limit(int x)
{
    const int k = x & 1023;          // the IJG range-limit table, indexed modulo 1024
    if (k < 128) return (uint8_t)(k + 128);
    if (k < 512) return 255;
    if (k < 896) return 0;
    return (uint8_t)(k - 896);
}

void idct_islow(const int16_t *cf, const uint16_t *q, uint8_t *out, int stride)
{
    constexpr int CB = 13, P1 = 2;
    constexpr int64_t F0_298 = 2446, F0_390 = 3196, F0_541 = 4433, F0_765 = 6270, F0_899 = 7373, F1_175 = 9633,
                      F1_501 = 12299, F1_847 = 15137, F1_961 = 16069, F2_053 = 16819, F2_562 = 20995,
                      F3_072 = 25172;
    int ws[64];
    for (int c = 0; c < 8; ++c) {
        const int16_t *raw = cf + c;
        if (!raw[8] && !raw[16] && !raw[24] && !raw[32] && !raw[40] && !raw[48] && !raw[56]) {
            const int dc = (int)raw[0] * q[c] * (1 << P1);
            for (int r = 0; r < 8; ++r) ws[r * 8 + c] = dc;
            continue;
        }
        int64_t col[64];
        for (int r = 0; r < 8; ++r) col[8 * r] = (int64_t)raw[8 * r] * q[8 * r + c];
        int64_t z2 = col[16], z3 = col[48];
        int64_t z1 = (z2 + z3) * F0_541;
        int64_t tmp2 = z1 + z3 * (-F1_847);
        int64_t tmp3 = z1 + z2 * F0_765;
        z2 = col[0];
        z3 = col[32];
        int64_t tmp0 = (z2 + z3) * (1 << CB);
        int64_t tmp1 = (z2 - z3) * (1 << CB);
        const int64_t tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
        tmp0 = col[56];
        tmp1 = col[40];
        tmp2 = col[24];
        tmp3 = col[8];
        z1 = tmp0 + tmp3;
        z2 = tmp1 + tmp2;
        z3 = tmp0 + tmp2;
        int64_t z4 = tmp1 + tmp3;
        const int64_t z5 = (z3 + z4) * F1_175;
        tmp0 *= F0_298;
        tmp1 *= F2_053;
        tmp2 *= F3_072;
        tmp3 *= F1_501;
        z1 *= -F0_899;
        z2 *= -F2_562;
        z3 *= -F1_961;
        z4 *= -F0_390;
        z3 += z5;
        z4 += z5;
        tmp0 += z1 + z3;
        tmp1 += z2 + z4;
        tmp2 += z2 + z3;
        tmp3 += z1 + z4;
        ws[0 * 8 + c] = descale(tmp10 + tmp3, CB - P1);
        ws[7 * 8 + c] = descale(tmp10 - tmp3, CB - P1);
        ws[1 * 8 + c] = descale(tmp11 + tmp2, CB - P1);
        ws[6 * 8 + c] = descale(tmp11 - tmp2, CB - P1);
        ws[2 * 8 + c] = descale(tmp12 + tmp1, CB - P1);
        ws[5 * 8 + c] = descale(tmp12 - tmp1, CB - P1);
        ws[3 * 8 + c] = descale(tmp13 + tmp0, CB - P1);
        ws[4 * 8 + c] = descale(tmp13 - tmp0, CB - P1);
    }
    for (int r = 0; r < 8; ++r) {
        const int *w = ws + r * 8;
        uint8_t *o = out + (size_t)r * stride;
        if (!w[1] && !w[2] && !w[3] && !w[4] && !w[5] && !w[6] && !w[7]) {
            const uint8_t v = idct_limit(descale(w[0], P1 + 3));
            for (int c = 0; c < 8; ++c) o[c] = v;
            continue;
        }
        int64_t z2 = w[2], z3 = w[6];
        int64_t z1 = (z2 + z3) * F0_541;
        int64_t tmp2 = z1 + z3 * (-F1_847);
        int64_t tmp3 = z1 + z2 * F0_765;
        int64_t tmp0 = ((int64_t)w[0] + w[4]) * (1 << CB);
        int64_t tmp1 = ((int64_t)w[0] - w[4]) * (1 << CB);
        const int64_t tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
        tmp0 = w[7];
        tmp1 = w[5];
        tmp2 = w[3];
        tmp3 = w[1];
        z1 = tmp0 + tmp3;
        z2 = tmp1 + tmp2;
        z3 = tmp0 + tmp2;
        int64_t z4 = tmp1 + tmp3;
        const int64_t z5 = (z3 + z4) * F1_175;
        tmp0 *= F0_298;
        tmp1 *= F2_053;
        tmp2 *= F3_072;
        tmp3 *= F1_501;
        z1 *= -F0_899;
        z2 *= -F2_562;
        z3 *= -F1_961;
        z4 *= -F0_390;
        z3 += z5;
        z4 += z5;
        tmp0 += z1 + z3;
        tmp1 += z2 + z4;
        tmp2 += z2 + z3;
        tmp3 += z1 + z4;
        constexpr int S = CB + P1 + 3;
        o[0] = idct_limit(descale(tmp10 + tmp3, S));
        o[7] = idct_limit(descale(tmp10 - tmp3, S));
        o[1] = idct_limit(descale(tmp11 + tmp2, S));
        o[6] = idct_limit(descale(tmp11 - tmp2, S));
        o[2] = idct_limit(descale(tmp12 + tmp1, S));
        o[5] = idct_limit(descale(tmp12 - tmp1, S));
        o[3] = idct_limit(descale(tmp13 + tmp0, S));
        o[4] = idct_limit(descale(tmp13 - tmp0, S));
    }
}

struct Frame {
    int W = 0, H = 0, nc = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0, ri = 0;
    bool adobe = false, jfif = false;
    int adobe_transform = -1;
    Comp comp[3];
    uint16_t q[4][64];              // quantisation tables, natural order
    bool qset[4] = {false, false, false, false};
    Huff dc[4], ac[4];
};

uint16_t be16(const uint8_t *p) { return (uint16_t)((p[0] << 8) | p[1]); }

int build_huff(Huff &h, const uint8_t *counts, const uint8_t *vals, int nv)
{
    int code = 0, k = 0;
    for (int l = 1; l <= 16; ++l) {
        h.valptr[l] = k;
        h.mincode[l] = code;
        code += counts[l - 1];
        k += counts[l - 1];
        h.maxcode[l] = counts[l - 1] ? code - 1 : -1;
        code <<= 1;
    }
    h.maxcode[17] = 0x7fffffff;
    if (k != nv || nv > 256) return kJMalformed;
    memcpy(h.vals, vals, (size_t)nv);
    h.set = true;
    return kJOk;
}

// Decode one scan starting at data[pos] (just past the SOS segment).
int decode_scan(const uint8_t *data, size_t n, size_t &pos, Frame &f, const int *sel, int ns)
{
    Reader r{data, n, pos};
    int pred[3] = {0, 0, 0};
    const bool inter = ns > 1;
    int units_x, units_y;
    if (inter) {
        units_x = f.mcux;
        units_y = f.mcuy;
    } else {
        const Comp &c = f.comp[sel[0]];
        units_x = (c.cw + 7) / 8;
        units_y = (c.ch + 7) / 8;
    }
    const long long total = (long long)units_x * units_y;
    long long done = 0;
    int restarts = 0;
    auto block = [&](int ci, int bx, int by) -> int {
        Comp &c = f.comp[ci];
        const Huff &hd = f.dc[c.td], &ha = f.ac[c.ta];
        if (!hd.set || !ha.set || !f.qset[c.tq]) return kJMalformed;
        int16_t zz[64];
        memset(zz, 0, sizeof zz);
        const int s = decode_huff(r, hd);
        if (s < 0 || s > 11) return kJMalformed;
        int &p = pred[ci];
        p += extend(r.get_bits(s), s);
        if (p < -32768 || p > 32767) return kJMalformed;
        zz[0] = (int16_t)p;
        for (int k = 1; k < 64;) {
            const int rs = decode_huff(r, ha);
            if (rs < 0) return kJMalformed;
            const int run = rs >> 4, sz = rs & 15;
            if (sz == 0) {
                if (run == 15) { k += 16; continue; }
                break;
            }
            k += run;
            if (k > 63) return kJMalformed;
            zz[kZigzag[k]] = (int16_t)extend(r.get_bits(sz), sz);
            ++k;
        }
        if (bx < c.bw && by < c.bh) memcpy(&c.coef[((size_t)by * c.bw + bx) * 64], zz, sizeof zz);
        return kJOk;
    };
    for (long long u = 0; u < total; ++u) {
        if (f.ri && done == f.ri) {
            // expect RSTn: realign to the byte, skip the marker, reset predictions
            r.bits = 0;
            size_t p2 = r.pos;
            while (p2 + 1 < n && !(data[p2] == 0xFF && data[p2 + 1] >= 0xD0 && data[p2 + 1] <= 0xD7)) ++p2;
            if (p2 + 1 >= n) return kJMalformed;
            r.pos = p2 + 2;
            r.marker = false;
            pred[0] = pred[1] = pred[2] = 0;
            done = 0;
            ++restarts;
        }
        const int ux = (int)(u % units_x), uy = (int)(u / units_x);
        if (inter) {
            for (int s = 0; s < ns; ++s) {
                const Comp &c = f.comp[sel[s]];
                for (int by = 0; by < c.v; ++by)
                    for (int bx = 0; bx < c.h; ++bx) {
                        const int rc = block(sel[s], ux * c.h + bx, uy * c.v + by);
                        if (rc) return rc;
                    }
            }
        } else {
            const int rc = block(sel[0], ux, uy);
            if (rc) return rc;
        }
        ++done;
    }
    pos = r.pos;
    return kJOk;
}

int parse(const uint8_t *data, size_t n, Frame &f, bool decode)
{
    if (n < 4 || data[0] != 0xFF || data[1] != 0xD8) return kJMalformed;
    size_t pos = 2;
    bool have_frame = false, scanned = false;
    while (pos + 4 <= n) {
        if (data[pos] != 0xFF) { ++pos; continue; }
        const uint8_t m = data[pos + 1];
        if (m == 0xFF) { ++pos; continue; }
        pos += 2;
        if (m == 0xD9) break;                                   // EOI
        if (m >= 0xD0 && m <= 0xD7) continue;                   // stray RST
        if (pos + 2 > n) return kJMalformed;
        const size_t len = be16(data + pos);
        if (len < 2 || pos + len > n) return kJMalformed;
        const uint8_t *seg = data + pos + 2;
        const size_t sl = len - 2;
        if (m == 0xC0 || m == 0xC1) {
            if (sl < 6) return kJMalformed;
            if (seg[0] != 8) return kJPrecision;
            f.H = be16(seg + 1);
            f.W = be16(seg + 3);
            f.nc = seg[5];
            if (f.nc == 4 || f.nc == 2) return kJComponents;
            if (f.nc != 1 && f.nc != 3) return kJMalformed;
            if (f.W <= 0 || f.H <= 0 || sl < 6 + 3 * (size_t)f.nc) return kJMalformed;
            f.hmax = f.vmax = 1;
            for (int i = 0; i < f.nc; ++i) {
                Comp &c = f.comp[i];
                c.id = seg[6 + 3 * i];
                c.h = seg[7 + 3 * i] >> 4;
                c.v = seg[7 + 3 * i] & 15;
                c.tq = seg[8 + 3 * i];
                if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return kJMalformed;
                f.hmax = c.h > f.hmax ? c.h : f.hmax;
                f.vmax = c.v > f.vmax ? c.v : f.vmax;
            }
            for (int i = 0; i < f.nc; ++i) {
                Comp &c = f.comp[i];
                const int rh = f.hmax / c.h, rv = f.vmax / c.v;
                if (f.hmax % c.h || f.vmax % c.v || rh > 2 || rv > 2 || (rh == 1 && rv == 2)) return kJSampling;
            }
            f.mcux = (f.W + 8 * f.hmax - 1) / (8 * f.hmax);
            f.mcuy = (f.H + 8 * f.vmax - 1) / (8 * f.vmax);
            for (int i = 0; i < f.nc; ++i) {
                Comp &c = f.comp[i];
                c.cw = (int)(((long long)f.W * c.h + f.hmax - 1) / f.hmax);
                c.ch = (int)(((long long)f.H * c.v + f.vmax - 1) / f.vmax);
                c.bw = f.mcux * c.h;
                c.bh = f.mcuy * c.v;
                if (decode) c.coef.assign((size_t)c.bw * c.bh * 64, 0);
            }
            have_frame = true;
        } else if (m == 0xC2 || m == 0xC6 || m == 0xCA || m == 0xCE) {
            return kJProgressive;
        } else if (m == 0xC3 || m == 0xC5 || m == 0xC7 || m == 0xCB || m == 0xCD || m == 0xCF) {
            return kJLossless;
        } else if (m == 0xC9) {
            return kJArithmetic;
        } else if (m == 0xDB) {
            size_t p = 0;
            while (p < sl) {
                const int pq = seg[p] >> 4, tq = seg[p] & 15;
                if (tq > 3 || pq > 1) return kJMalformed;
                if (pq == 1) return kJPrecision;
                if (p + 65 > sl) return kJMalformed;
                for (int k = 0; k < 64; ++k) f.q[tq][kZigzag[k]] = seg[p + 1 + k];
                f.qset[tq] = true;
                p += 65;
            }
        } else if (m == 0xC4) {
            size_t p = 0;
            while (p < sl) {
                if (p + 17 > sl) return kJMalformed;
                const int tc = seg[p] >> 4, th = seg[p] & 15;
                if (tc > 1 || th > 3) return kJMalformed;
                int nv = 0;
                for (int l = 0; l < 16; ++l) nv += seg[p + 1 + l];
                if (p + 17 + nv > sl) return kJMalformed;
                const int rc = build_huff(tc ? f.ac[th] : f.dc[th], seg + p + 1, seg + p + 17, nv);
                if (rc) return rc;
                p += 17 + nv;
            }
        } else if (m == 0xDD) {
            if (sl < 2) return kJMalformed;
            f.ri = be16(seg);
        } else if (m == 0xE0) {
            if (sl >= 5 && memcmp(seg, "JFIF\0", 5) == 0) f.jfif = true;
        } else if (m == 0xEE) {
            if (sl >= 12 && memcmp(seg, "Adobe", 5) == 0) {
                f.adobe = true;
                f.adobe_transform = seg[11];
            }
        } else if (m == 0xDA) {
            if (!have_frame || sl < 1) return kJMalformed;
            const int ns = seg[0];
            if (ns < 1 || ns > f.nc || sl < 1 + 2 * (size_t)ns + 3) return kJMalformed;
            int sel[3] = {0, 0, 0};
            for (int s = 0; s < ns; ++s) {
                const int cid = seg[1 + 2 * s];
                int ci = -1;
                for (int i = 0; i < f.nc; ++i)
                    if (f.comp[i].id == cid) ci = i;
                if (ci < 0) return kJMalformed;
                sel[s] = ci;
                f.comp[ci].td = seg[2 + 2 * s] >> 4;
                f.comp[ci].ta = seg[2 + 2 * s] & 15;
                if (f.comp[ci].td > 3 || f.comp[ci].ta > 3) return kJMalformed;
            }
            const uint8_t ss = seg[1 + 2 * ns], se = seg[2 + 2 * ns], ah = seg[3 + 2 * ns];
            if (ss != 0 || se != 63 || ah != 0) return kJProgressive;
            pos += len;
            if (!decode) return kJOk;
            const int rc = decode_scan(data, n, pos, f, sel, ns);
            if (rc) return rc;
            scanned = true;
            continue;
        }
        pos += len;
    }
    if (!have_frame) return kJMalformed;
    if (decode && !scanned) return kJMalformed;
    return kJOk;
}

// libjpeg's fancy upsampling of one component plane (cw x ch) by (rh, rv).
void upsample(const uint8_t *in, int cw, int ch, int rh, int rv, int W, int H, uint8_t *out)
{
    std::vector<uint8_t> tmp;
    auto row_h = [&](const uint8_t *src, uint8_t *dst) {
        if (rh == 1) {
            memcpy(dst, src, (size_t)cw);
            return;
        }
        if (cw <= 2) {
            for (int x = 0; x < cw; ++x) dst[2 * x] = dst[2 * x + 1] = src[x];
            return;
        }
        int v = src[0];
        dst[0] = (uint8_t)v;
        dst[1] = (uint8_t)((v * 3 + src[1] + 2) >> 2);
        for (int x = 1; x < cw - 1; ++x) {
            v = src[x] * 3;
            dst[2 * x] = (uint8_t)((v + src[x - 1] + 1) >> 2);
            dst[2 * x + 1] = (uint8_t)((v + src[x + 1] + 2) >> 2);
        }
        v = src[cw - 1];
        dst[2 * cw - 2] = (uint8_t)((v * 3 + src[cw - 2] + 1) >> 2);
        dst[2 * cw - 1] = (uint8_t)v;
    };
    const int ow = cw * rh;
    std::vector<uint8_t> full((size_t)ow * ch * rv);
    if (rv == 1) {
        for (int y = 0; y < ch; ++y) row_h(in + (size_t)y * cw, full.data() + (size_t)y * ow);
    } else if (rh == 2 && cw > 2) {
        // h2v2 fancy: column sums 3 * near row + far row, then the h triangle
        for (int y = 0; y < ch; ++y) {
            const uint8_t *r0 = in + (size_t)y * cw;
            for (int v = 0; v < 2; ++v) {
                const int yy = v == 0 ? (y > 0 ? y - 1 : 0) : (y + 1 < ch ? y + 1 : ch - 1);
                const uint8_t *r1 = in + (size_t)yy * cw;
                uint8_t *o = full.data() + (size_t)(2 * y + v) * ow;
                int this_s = r0[0] * 3 + r1[0], next_s = r0[1] * 3 + r1[1], last_s;
                o[0] = (uint8_t)((this_s * 4 + 8) >> 4);
                o[1] = (uint8_t)((this_s * 3 + next_s + 7) >> 4);
                last_s = this_s;
                this_s = next_s;
                for (int x = 2; x < cw; ++x) {
                    next_s = r0[x] * 3 + r1[x];
                    o[2 * x - 2] = (uint8_t)((this_s * 3 + last_s + 8) >> 4);
                    o[2 * x - 1] = (uint8_t)((this_s * 3 + next_s + 7) >> 4);
                    last_s = this_s;
                    this_s = next_s;
                }
                o[2 * cw - 2] = (uint8_t)((this_s * 3 + last_s + 8) >> 4);
                o[2 * cw - 1] = (uint8_t)((this_s * 4 + 7) >> 4);
            }
        }
    } else {
        // rv == 2 with a component too narrow for the filter: replicate
        tmp.resize((size_t)ow);
        for (int y = 0; y < ch; ++y) {
            if (rh == 2) {
                for (int x = 0; x < cw; ++x) tmp[2 * x] = tmp[2 * x + 1] = in[(size_t)y * cw + x];
            } else {
                memcpy(tmp.data(), in + (size_t)y * cw, (size_t)cw);
            }
            memcpy(full.data() + (size_t)(2 * y) * ow, tmp.data(), (size_t)ow);
            memcpy(full.data() + (size_t)(2 * y + 1) * ow, tmp.data(), (size_t)ow);
        }
    }
    for (int y = 0; y < H; ++y) memcpy(out + (size_t)y * W, full.data() + (size_t)y * ow, (size_t)W);
}

}  // namespace

// info[0..2] = width, height, output channels (1 gray, 3 RGB).  Status as
// pivamd_jpeg_decode.
extern "C" int pivamd_jpeg_info(const uint8_t *data, size_t n, int *info)
{
    Frame f;
    const int rc = parse(data, n, f, false);
    if (rc) return rc;
    info[0] = f.W;
    info[1] = f.H;
    info[2] = f.nc;
    return kJOk;
}

// Decode into out (height x width x channels bytes, capacity cap).  Status:
// 0, 1 malformed or truncated, 2 progressive, 3 arithmetic coding, 4
// lossless or hierarchical, 5 12-bit samples, 6 two or four components,
// 7 unsupported sampling factors, 8 out too small.
static int decode_frame(const uint8_t *data, size_t n, uint8_t *out, size_t cap);

extern "C" int pivamd_jpeg_decode(const uint8_t *data, size_t n, uint8_t *out, size_t cap)
{
    {
        // the frame's size first, so a buffer that is too small is refused
        // before any coefficient storage is allocated
        Frame h;
        const int rc = parse(data, n, h, false);
        if (rc) return rc;
        if (cap < (size_t)h.W * h.H * (size_t)h.nc) return kJSmall;
    }
    try {
        return decode_frame(data, n, out, cap);
    } catch (...) {   // allocation failure: no exception crosses the C ABI
        return kJMalformed;
    }
}

static int decode_frame(const uint8_t *data, size_t n, uint8_t *out, size_t cap)
{
    Frame f;
    int rc = parse(data, n, f, true);
    if (rc) return rc;
    const size_t W = (size_t)f.W, H = (size_t)f.H;
    if (cap < W * H * (size_t)f.nc) return kJSmall;
    std::vector<std::vector<uint8_t>> planes((size_t)f.nc);
    for (int i = 0; i < f.nc; ++i) {
        Comp &c = f.comp[i];
        const int pw = c.bw * 8;
        std::vector<uint8_t> pl((size_t)pw * c.bh * 8);
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx)
                idct_islow(&c.coef[((size_t)by * c.bw + bx) * 64], f.q[c.tq], pl.data() + (size_t)by * 8 * pw + bx * 8,
                           pw);
        // crop to the component's own size, then upsample to the frame
        std::vector<uint8_t> cp((size_t)c.cw * c.ch);
        for (int y = 0; y < c.ch; ++y) memcpy(cp.data() + (size_t)y * c.cw, pl.data() + (size_t)y * pw, (size_t)c.cw);
        planes[i].resize(W * H);
        upsample(cp.data(), c.cw, c.ch, f.hmax / c.h, f.vmax / c.v, f.W, f.H, planes[i].data());
    }
    if (f.nc == 1) {
        memcpy(out, planes[0].data(), W * H);
        return kJOk;
    }
    // colour transform as libjpeg decides it: RGB when an Adobe marker says
    // transform 0, or no JFIF/Adobe marker and component ids 'R','G','B'
    const bool rgb = (f.adobe && f.adobe_transform == 0) ||
                     (!f.jfif && !f.adobe && f.comp[0].id == 'R' && f.comp[1].id == 'G' && f.comp[2].id == 'B');
    const uint8_t *Y = planes[0].data(), *Cb = planes[1].data(), *Cr = planes[2].data();
    for (size_t i = 0; i < W * H; ++i) {
        uint8_t *o = out + 3 * i;
        if (rgb) {
            o[0] = Y[i];
            o[1] = Cb[i];
            o[2] = Cr[i];
            continue;
        }
        constexpr int SB = 16;
        constexpr int64_t HALF = (int64_t)1 << (SB - 1);
        constexpr int64_t FR = 91881, FB = 116130, FGR = 46802, FGB = 22554;   // 1.402, 1.772, 0.71414, 0.34414
        const int64_t cb = (int64_t)Cb[i] - 128, cr = (int64_t)Cr[i] - 128;
        const int y = Y[i];
        const int r = y + (int)((FR * cr + HALF) >> SB);
        const int g = y + (int)((-FGB * cb + HALF - FGR * cr) >> SB);
        const int b = y + (int)((FB * cb + HALF) >> SB);
        o[0] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
        o[1] = (uint8_t)(g < 0 ? 0 : (g > 255 ? 255 : g));
        o[2] = (uint8_t)(b < 0 ? 0 : (b > 255 ? 255 : b));
    }
    return kJOk;
}
