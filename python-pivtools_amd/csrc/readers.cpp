// readers.cpp -- the serial decompressors of the camera containers, on the host.
//
// The camera containers the reference reads (pivtools_core/image_handling/
// readers/) store a frame either packed (decoded on the device by
// pivamd_davis_decode) or compressed with a serial code whose token
// boundaries depend on the tokens themselves.  Those are decoded here, on
// the host, into the packed payload that is then copied to the device:
//
//   DaVis token stream  .im7 pack 1 and IMX-compressed .set entries
//                       (davis_tokens.py:1-136): tokens update a running
//                       16-bit value --
//                         int8 byte (not 0x7f/0x80/0x81)  signed delta
//                         0x80 lo hi                      one absolute value
//                         0x7f n (lo hi)*n                n absolute values
//                         0x81 nibbles...                 signed 4-bit deltas,
//                           high nibble first; a low nibble 0x8 ends the run
//                           (a high nibble 0x8 is the delta -8).
//   LZ4 block           .im7 pack 20 and LZ4-compressed .set entries
//                       (im7_reader.py:343-363): the LZ4 block format, one
//                       block with no frame header (a sequence is a token
//                       byte, literal length 15 + 255-run extension, the
//                       literals, a little-endian 16-bit match offset and
//                       match length 4 + nibble + 255-run extension; the last
//                       sequence carries literals only).
//   TIFF LZW            compression 5 (TIFF 6.0 section 13): codes MSB-first,
//                       9 to 12 bits, 256 clear and 257 end of information,
//                       new codes from 258, the width growing one code early
//                       (at 511, 1023 and 2047 entries).
//   PackBits            compression 32773 (TIFF 6.0 section 9): a header byte
//                       n copies n + 1 literal bytes (0..127), repeats the next
//                       byte 1 - n times (-127..-1), or is skipped (-128).
//
//   PNG scanlines       (PNG specification, filter method 0): each row of the
//                       inflated stream is a filter-type byte and the row's
//                       bytes, filtered against the reconstructed left byte a
//                       (bpp bytes back), up byte b and up-left byte c: 0 none,
//                       1 Sub x + a, 2 Up x + b, 3 Average x + floor((a + b)
//                       / 2), 4 Paeth x + the one of a, b, c nearest
//                       p = a + b - c (ties to a, then b), all modulo 256.
//
// The TIFF strips are read by pivtools_amd/tiff.py, as the reference reads
// them through tifffile (generic_readers.py:81-122).
#include <stddef.h>
#include <stdint.h>
#include <string.h>

namespace {

enum TokenStatus { kTokOk = 0, kTokTruncated = 1, kTokOverrun = 2, kTokRange = 3 };

}  // namespace

// Decode nvals running values from data[off:] into out.  prev is the running
// value before the first token.  On return *off_io and *prev_io hold the
// offset after the last token read and the last value, *written the values
// written; the result is 0 or the first failure: 1 the stream ended inside a
// token, 2 a 0x7f run passes nvals, 3 the value left 0..65535.
extern "C" int pivamd_davis_tokens(const uint8_t *data, size_t n, size_t *off_io, uint16_t *out,
                                   size_t nvals, long long *prev_io, size_t *written)
{
    size_t off = *off_io, i = 0;
    long long prev = *prev_io;
    int status = kTokOk;
    while (i < nvals) {
        if (off >= n) { status = kTokTruncated; break; }
        const int b = data[off++];
        if (b == 0x81) {
            bool stop = false;
            while (i < nvals && !stop) {
                if (off >= n) { status = kTokTruncated; break; }
                const int byte = data[off++];
                const int hi = byte >> 4;
                prev += hi >= 8 ? hi - 16 : hi;
                if (prev < 0 || prev > 65535) { status = kTokRange; break; }
                out[i++] = (uint16_t)prev;
                if (i >= nvals) break;
                const int lo = byte & 0xF;
                if (lo == 8) { stop = true; break; }
                prev += lo >= 8 ? lo - 16 : lo;
                if (prev < 0 || prev > 65535) { status = kTokRange; break; }
                out[i++] = (uint16_t)prev;
            }
            if (status != kTokOk) break;
        } else if (b == 0x80) {
            if (off + 2 > n) { status = kTokTruncated; break; }
            prev = (long long)data[off] | ((long long)data[off + 1] << 8);
            off += 2;
            out[i++] = (uint16_t)prev;
        } else if (b == 0x7F) {
            if (off >= n) { status = kTokTruncated; break; }
            const size_t cnt = data[off++];
            if (i + cnt > nvals) { status = kTokOverrun; break; }
            if (off + 2 * cnt > n) { status = kTokTruncated; break; }
            for (size_t k = 0; k < cnt; ++k) {
                prev = (long long)data[off] | ((long long)data[off + 1] << 8);
                off += 2;
                out[i++] = (uint16_t)prev;
            }
        } else {
            prev += b >= 128 ? b - 256 : b;
            if (prev < 0 || prev > 65535) { status = kTokRange; break; }
            out[i++] = (uint16_t)prev;
        }
    }
    *off_io = off;
    *prev_io = prev;
    *written = i;
    return status;
}

// Decompress one LZ4 block of n bytes into dst (capacity cap).  Returns the
// decompressed size, or -1 for a malformed block (a length running past the
// input, a match offset of 0 or before the output start, output past cap).
extern "C" long long pivamd_lz4_block(const uint8_t *src, size_t n, uint8_t *dst, size_t cap)
{
    size_t ip = 0, op = 0;
    while (ip < n) {
        const unsigned token = src[ip++];
        size_t lit = token >> 4;
        if (lit == 15) {
            unsigned char s;
            do {
                if (ip >= n) return -1;
                s = src[ip++];
                lit += s;
            } while (s == 255);
        }
        if (lit > n - ip || lit > cap - op) return -1;
        memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip == n) break;                      // the last sequence: literals only
        if (n - ip < 2) return -1;
        const size_t offset = (size_t)src[ip] | ((size_t)src[ip + 1] << 8);
        ip += 2;
        if (offset == 0 || offset > op) return -1;
        size_t mlen = (token & 15u) + 4;
        if ((token & 15u) == 15) {
            unsigned char s;
            do {
                if (ip >= n) return -1;
                s = src[ip++];
                mlen += s;
            } while (s == 255);
        }
        if (mlen > cap - op) return -1;
        // byte by byte: a match may overlap its own output (offset < mlen)
        const uint8_t *m = dst + op - offset;
        for (size_t k = 0; k < mlen; ++k) dst[op + k] = m[k];
        op += mlen;
    }
    return (long long)op;
}

// Decode one TIFF LZW strip or tile of n bytes into dst (capacity cap; the
// block's expected size).  Returns the bytes written (a stream that ends
// before its end-of-information code keeps what it decoded, as libtiff
// does), or -1 for a code that is not yet in the table or an old-style
// (LSB-first) stream.
extern "C" long long pivamd_tiff_lzw(const uint8_t *src, size_t n, uint8_t *dst, size_t cap)
{
    if (n >= 2 && src[0] == 0 && (src[1] & 1)) return -1;     // old-style LZW, bit-reversed codes
    // table: each code is a prefix code plus a last byte, with its length and
    // first byte, so a string is written back to front without a stack
    static thread_local uint16_t prefix[4096];
    static thread_local uint8_t suffix[4096], first[4096];
    static thread_local uint16_t length[4096];
    for (int c = 0; c < 256; ++c) {
        prefix[c] = 0xFFFF;
        suffix[c] = (uint8_t)c;
        first[c] = (uint8_t)c;
        length[c] = 1;
    }
    int width = 9, next = 258, prev = -1;
    size_t op = 0;
    uint64_t acc = 0;
    int have = 0;
    size_t ip = 0;
    for (;;) {
        while (have < width && ip < n) {
            acc = (acc << 8) | src[ip++];
            have += 8;
        }
        if (have < width) break;                       // input ended without EOI
        const int code = (int)((acc >> (have - width)) & ((1u << width) - 1u));
        have -= width;
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            prev = -1;
            continue;
        }
        int out_code;
        uint8_t head;
        if (code < next) {
            out_code = code;
            head = first[code];
        } else if (code == next && prev >= 0) {
            out_code = -1;                            // KwKwK: prev's string plus its first byte
            head = first[prev];
        } else {
            return -1;
        }
        // add prev + head as a new code
        if (prev >= 0 && next < 4096) {
            prefix[next] = (uint16_t)prev;
            suffix[next] = head;
            first[next] = first[prev];
            length[next] = (uint16_t)(length[prev] + 1);
            ++next;
        }
        const int wcode = out_code >= 0 ? out_code : next - 1;
        const size_t len = length[wcode];
        if (op + len > cap) {
            // write what fits of the string and stop: the block is full
            size_t keep = cap - op;
            int c = wcode;
            for (size_t k = len; k > keep; --k) c = prefix[c];
            for (size_t k = keep; k > 0; --k) {
                dst[op + k - 1] = suffix[c];
                c = prefix[c];
            }
            op = cap;
            break;
        }
        int c = wcode;
        for (size_t k = len; k > 0; --k) {
            dst[op + k - 1] = suffix[c];
            c = prefix[c];
        }
        op += len;
        prev = wcode;
        if (next + 1 >= (1 << width) && width < 12) ++width;
    }
    return (long long)op;
}

// Decode one PackBits strip or tile into dst (capacity cap).  Returns the
// bytes written; a run that passes cap or the input is cut at it.
extern "C" long long pivamd_packbits(const uint8_t *src, size_t n, uint8_t *dst, size_t cap)
{
    size_t ip = 0, op = 0;
    while (ip < n && op < cap) {
        const int h = (int8_t)src[ip++];
        if (h >= 0) {
            size_t cnt = (size_t)h + 1;
            if (cnt > n - ip) cnt = n - ip;
            if (cnt > cap - op) cnt = cap - op;
            memcpy(dst + op, src + ip, cnt);
            ip += cnt;
            op += cnt;
        } else if (h != -128) {
            if (ip >= n) break;
            size_t cnt = (size_t)(1 - h);
            if (cnt > cap - op) cnt = cap - op;
            memset(dst + op, src[ip++], cnt);
            op += cnt;
        }
    }
    return (long long)op;
}

// Undo the PNG row filters of rows scanlines of rowbytes bytes each (src
// holds each row's filter byte then its bytes; dst the reconstructed rows).
// bpp: bytes per complete pixel, at least 1.  Returns 0, or -1 for a filter
// type above 4 (a damaged stream).
extern "C" int pivamd_png_unfilter(const uint8_t *src, size_t rows, size_t rowbytes, int bpp, uint8_t *dst)
{
    if (bpp < 1) return -1;
    const size_t b = (size_t)bpp;
    for (size_t r = 0; r < rows; ++r) {
        const uint8_t *in = src + r * (rowbytes + 1);
        const int type = in[0];
        ++in;
        uint8_t *out = dst + r * rowbytes;
        const uint8_t *up = r ? out - rowbytes : nullptr;
        switch (type) {
        case 0: memcpy(out, in, rowbytes); break;
        case 1:
            for (size_t i = 0; i < rowbytes; ++i) out[i] = (uint8_t)(in[i] + (i >= b ? out[i - b] : 0));
            break;
        case 2:
            for (size_t i = 0; i < rowbytes; ++i) out[i] = (uint8_t)(in[i] + (up ? up[i] : 0));
            break;
        case 3:
            for (size_t i = 0; i < rowbytes; ++i) {
                const unsigned a = i >= b ? out[i - b] : 0, u = up ? up[i] : 0;
                out[i] = (uint8_t)(in[i] + ((a + u) >> 1));
            }
            break;
        case 4:
            for (size_t i = 0; i < rowbytes; ++i) {
                const int a = i >= b ? out[i - b] : 0, u = up ? up[i] : 0, c = (up && i >= b) ? up[i - b] : 0;
                const int p = a + u - c;
                const int pa = p > a ? p - a : a - p, pb = p > u ? p - u : u - p, pc = p > c ? p - c : c - p;
                const int pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? u : c);
                out[i] = (uint8_t)(in[i] + pred);
            }
            break;
        default: return -1;
        }
    }
    return 0;
}
