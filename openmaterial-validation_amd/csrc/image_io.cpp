// image_io.cpp -- saveImage of the reference main (main.cpp:86-120): the
// gamma-corrected float film truncated to 8 bits and written as PNG, BMP,
// TGA or JPEG, or the floats written as PFM, chosen by the file extension,
// with rows written bottom-up (main sets ImageSaver::flipVerticallyOnWrite,
// main.cpp:149).  The reference writes through stb_image_write
// (ImageSaver.h:23-97); the encoders below produce the same bytes:
//   PNG   per-row filter chosen by the smallest sum of |signed residual|
//         (first filter wins ties), zlib stream with fixed Huffman codes and
//         greedy LZ77 over 3-byte hash chains (16384 buckets, chains trimmed
//         to their newest 8 entries when they reach 16, only match starts
//         and literals enter the chains, one-step lazy matching);
//   BMP   24-bit, the RGBA pixels composited over (255, 0, 255) by alpha;
//   TGA   32-bit BGRA, run-length packets per row;
//   JPEG  quality 100, 4:4:4, baseline with the JPEG standard's (Annex K)
//         quantisation and Huffman tables, float AAN DCT (Arai, Agui and
//         Nakajima 1988) with the IJG constants, as Jon Olick's jo_jpeg;
//   PFM   "PF", size, scale -1 (little endian), RGB floats row by row.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "render_core.h"

namespace {

using Bytes = std::vector<uint8_t>;

void put16le(Bytes& o, uint32_t v) { o.push_back(v & 255); o.push_back((v >> 8) & 255); }
void put32le(Bytes& o, uint32_t v) { put16le(o, v & 0xFFFF); put16le(o, v >> 16); }
void put32be(Bytes& o, uint32_t v) { o.push_back(v >> 24); o.push_back((v >> 16) & 255); o.push_back((v >> 8) & 255); o.push_back(v & 255); }

// ---------------------------------------------------------------- zlib (fixed Huffman)
struct BitOut {
    Bytes& out;
    uint32_t buf = 0;
    int cnt = 0;
    explicit BitOut(Bytes& o) : out(o) {}
    void add(uint32_t code, int bits) {
        buf |= code << cnt;
        cnt += bits;
        while (cnt >= 8) { out.push_back(buf & 255); buf >>= 8; cnt -= 8; }
    }
};

uint32_t bitrev(uint32_t code, int bits) {
    uint32_t r = 0;
    while (bits--) { r = (r << 1) | (code & 1); code >>= 1; }
    return r;
}

// RFC 1951 3.2.6 fixed literal/length code
void fixedSymbol(BitOut& w, int n) {
    if (n <= 143) w.add(bitrev(0x30 + n, 8), 8);
    else if (n <= 255) w.add(bitrev(0x190 + n - 144, 9), 9);
    else if (n <= 279) w.add(bitrev(n - 256, 7), 7);
    else w.add(bitrev(0xC0 + n - 280, 8), 8);
}

uint32_t hash3(const uint8_t* d) {
    uint32_t h = d[0] + (d[1] << 8) + (d[2] << 16);
    h ^= h << 3; h += h >> 5; h ^= h << 4; h += h >> 17; h ^= h << 25; h += h >> 6;
    return h;
}

int matchLen(const uint8_t* a, const uint8_t* b, int limit) {
    int i = 0;
    while (i < limit && i < 258 && a[i] == b[i]) ++i;
    return i;
}

Bytes zlibCompress(const uint8_t* data, int n) {
    // RFC 1951 3.2.5 length and distance bases / extra bits
    static const int lenBase[] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258, 259};
    static const int lenExtra[] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
    static const int distBase[] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769, 1025, 1537, 2049, 3073,
                                   4097, 6145, 8193, 12289, 16385, 24577, 32768};
    static const int distExtra[] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
    const int kBuckets = 16384, kKeep = 8;
    Bytes out = {0x78, 0x5E};
    BitOut w(out);
    w.add(1, 1);   // final block
    w.add(1, 2);   // fixed Huffman
    std::vector<std::vector<int>> chain(kBuckets);
    int i = 0;
    while (i < n - 3) {
        std::vector<int>& c = chain[hash3(data + i) & (kBuckets - 1)];
        int best = 3, bestAt = -1;
        for (int p : c)
            if (p > i - 32768) {
                const int d = matchLen(data + p, data + i, n - i);
                if (d >= best) { best = d; bestAt = p; }
            }
        if ((int)c.size() == 2 * kKeep) c.erase(c.begin(), c.begin() + kKeep);
        c.push_back(i);
        if (bestAt >= 0) {   // a longer match one byte later makes this byte a literal
            for (int p : chain[hash3(data + i + 1) & (kBuckets - 1)])
                if (p > i - 32767 && matchLen(data + p, data + i + 1, n - i - 1) > best) { bestAt = -1; break; }
        }
        if (bestAt >= 0) {
            const int dist = i - bestAt;
            int j = 0;
            while (best > lenBase[j + 1] - 1) ++j;
            fixedSymbol(w, j + 257);
            if (lenExtra[j]) w.add(best - lenBase[j], lenExtra[j]);
            j = 0;
            while (dist > distBase[j + 1] - 1) ++j;
            w.add(bitrev(j, 5), 5);
            if (distExtra[j]) w.add(dist - distBase[j], distExtra[j]);
            i += best;
        } else {
            fixedSymbol(w, data[i]);
            ++i;
        }
    }
    for (; i < n; ++i) fixedSymbol(w, data[i]);
    fixedSymbol(w, 256);
    while (w.cnt) w.add(0, 1);
    uint32_t s1 = 1, s2 = 0;   // Adler-32
    for (int k = 0; k < n; ++k) { s1 = (s1 + data[k]) % 65521; s2 = (s2 + s1) % 65521; }
    out.push_back(s2 >> 8); out.push_back(s2 & 255); out.push_back(s1 >> 8); out.push_back(s1 & 255);
    return out;
}

uint32_t crc32(const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t c = k;
            for (int b = 0; b < 8; ++b) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[k] = c;
        }
        init = true;
    }
    uint32_t c = ~0u;
    for (size_t k = 0; k < n; ++k) c = table[(c ^ p[k]) & 255] ^ (c >> 8);
    return ~c;
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// rows are written bottom-up: file row y is memory row h-1-y, the row
// "above" it in the file is memory row h-y
Bytes encodePng(const uint8_t* px, int w, int h) {
    const int n = 4, stride = w * n;
    Bytes filt;
    filt.reserve((size_t)(stride + 1) * h);
    std::vector<int8_t> line(stride), best(stride);
    for (int y = 0; y < h; ++y) {
        const uint8_t* z = px + (size_t)stride * (h - 1 - y);
        const uint8_t* up = y ? z + stride : nullptr;
        int bestType = 0, bestEst = 0x7FFFFFFF;
        for (int t = 0; t < 5; ++t) {
            for (int i = 0; i < stride; ++i) {
                const int a = i >= n ? z[i - n] : 0, b = up ? up[i] : 0, c = (up && i >= n) ? up[i - n] : 0;
                int pred = 0;
                if (t == 1) pred = a;
                else if (t == 2) pred = b;
                else if (t == 3) pred = (a + b) >> 1;
                else if (t == 4) pred = paeth(a, b, c);
                line[i] = (int8_t)(uint8_t)(z[i] - pred);
            }
            int est = 0;
            for (int i = 0; i < stride; ++i) est += std::abs((int)line[i]);
            if (est < bestEst) { bestEst = est; bestType = t; best = line; }
        }
        filt.push_back((uint8_t)bestType);
        filt.insert(filt.end(), (const uint8_t*)best.data(), (const uint8_t*)best.data() + stride);
    }
    const Bytes z = zlibCompress(filt.data(), (int)filt.size());
    Bytes o = {137, 80, 78, 71, 13, 10, 26, 10};
    auto chunk = [&o](const char* tag, const uint8_t* body, size_t len) {
        put32be(o, (uint32_t)len);
        const size_t at = o.size();
        o.insert(o.end(), tag, tag + 4);
        if (len) o.insert(o.end(), body, body + len);
        put32be(o, crc32(o.data() + at, len + 4));
    };
    Bytes ihdr;
    put32be(ihdr, w); put32be(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});
    chunk("IHDR", ihdr.data(), ihdr.size());
    chunk("IDAT", z.data(), z.size());
    chunk("IEND", nullptr, 0);
    return o;
}

Bytes encodeBmp(const uint8_t* px, int w, int h) {
    const int pad = (-w * 3) & 3;
    Bytes o = {'B', 'M'};
    put32le(o, 14 + 40 + (w * 3 + pad) * h);
    put16le(o, 0); put16le(o, 0);
    put32le(o, 14 + 40);
    put32le(o, 40); put32le(o, w); put32le(o, h);
    put16le(o, 1); put16le(o, 24);
    for (int k = 0; k < 6; ++k) put32le(o, 0);
    static const int bg[3] = {255, 0, 255};
    for (int y = 0; y < h; ++y) {          // bottom-up file order with the flip: memory rows in order
        for (int x = 0; x < w; ++x) {
            const uint8_t* d = px + ((size_t)y * w + x) * 4;
            uint8_t c[3];
            for (int k = 0; k < 3; ++k) c[k] = (uint8_t)(bg[k] + ((d[k] - bg[k]) * d[3]) / 255);
            o.push_back(c[2]); o.push_back(c[1]); o.push_back(c[0]);
        }
        for (int k = 0; k < pad; ++k) o.push_back(0);
    }
    return o;
}

Bytes encodeTga(const uint8_t* px, int w, int h) {
    Bytes o = {0, 0, 10};
    put16le(o, 0); put16le(o, 0); o.push_back(0);
    put16le(o, 0); put16le(o, 0); put16le(o, w); put16le(o, h);
    o.push_back(32); o.push_back(8);
    auto pix = [&o](const uint8_t* d) { o.push_back(d[2]); o.push_back(d[1]); o.push_back(d[0]); o.push_back(d[3]); };
    auto same = [](const uint8_t* a, const uint8_t* b) { return std::memcmp(a, b, 4) == 0; };
    for (int y = 0; y < h; ++y) {
        const uint8_t* row = px + (size_t)y * w * 4;
        int len;
        for (int i = 0; i < w; i += len) {
            const uint8_t* begin = row + i * 4;
            bool diff = true;
            len = 1;
            if (i < w - 1) {
                ++len;
                diff = !same(begin, row + (i + 1) * 4);
                if (diff) {   // raw packet: grows while pixel k differs from the pixel at the running cursor
                    const uint8_t* cur = begin;
                    for (int k = i + 2; k < w && len < 128; ++k) {
                        if (!same(cur, row + k * 4)) { cur += 4; ++len; }
                        else { --len; break; }
                    }
                } else {      // run packet
                    for (int k = i + 2; k < w && len < 128; ++k) {
                        if (same(begin, row + k * 4)) ++len;
                        else break;
                    }
                }
            }
            if (diff) {
                o.push_back((uint8_t)(len - 1));
                for (int k = 0; k < len; ++k) pix(begin + k * 4);
            } else {
                o.push_back((uint8_t)(len + 127));
                pix(begin);
            }
        }
    }
    return o;
}

// ---------------------------------------------------------------- JPEG
const uint8_t kZigZag[64] = {0, 1, 5, 6, 14, 15, 27, 28, 2, 4, 7, 13, 16, 26, 29, 42, 3, 8, 12, 17, 25, 30, 41, 43, 9, 11, 18,
                             24, 31, 40, 44, 53, 10, 19, 23, 32, 39, 45, 52, 54, 20, 22, 33, 38, 46, 51, 55, 60, 21, 34, 37, 47, 50, 56, 59, 61,
                             35, 36, 48, 49, 57, 58, 62, 63};

// JPEG standard Annex K.3 table counts (codes of length 1..16) and values
const uint8_t kDcLumCounts[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChrCounts[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcValues[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumCounts[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumValues[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1,
    0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26,
    0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56,
    0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83, 0x84, 0x85,
    0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa,
    0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6,
    0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9,
    0xfa};
const uint8_t kAcChrCounts[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChrValues[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22, 0x32, 0x81, 0x08, 0x14, 0x42,
    0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19,
    0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55,
    0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x82, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8,
    0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4,
    0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9,
    0xfa};
// Annex K.1 quantisation tables (natural order)
const int kLumQ[64] = {16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56, 14, 17, 22, 29, 51, 87,
                       80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92, 49, 64, 78, 87, 103, 121, 120, 101, 72,
                       92, 95, 98, 112, 100, 103, 99};
const int kChrQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99,
                       99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                       99, 99, 99, 99, 99, 99};

struct Code { uint16_t bits = 0, len = 0; };

// canonical Huffman codes from counts per length (JPEG Annex C)
void buildCodes(const uint8_t counts[16], const uint8_t* values, Code table[256]) {
    uint16_t code = 0;
    int k = 0;
    for (int len = 1; len <= 16; ++len) {
        for (int i = 0; i < counts[len - 1]; ++i, ++k) table[values[k]] = {code++, (uint16_t)len};
        code <<= 1;
    }
}

struct JpegBits {
    Bytes& o;
    int buf = 0, cnt = 0;
    explicit JpegBits(Bytes& out) : o(out) {}
    void put(Code c) {
        cnt += c.len;
        buf |= c.bits << (24 - cnt);
        while (cnt >= 8) {
            const uint8_t b = (buf >> 16) & 255;
            o.push_back(b);
            if (b == 255) o.push_back(0);
            buf <<= 8;
            cnt -= 8;
        }
    }
};

// float AAN forward DCT of 8 samples (IJG jfdctflt.c constants)
void dct8(float* d0p, float* d1p, float* d2p, float* d3p, float* d4p, float* d5p, float* d6p, float* d7p) {
    const float d0 = *d0p, d1 = *d1p, d2 = *d2p, d3 = *d3p, d4 = *d4p, d5 = *d5p, d6 = *d6p, d7 = *d7p;
    const float t0 = d0 + d7, t7 = d0 - d7, t1 = d1 + d6, t6 = d1 - d6;
    const float t2 = d2 + d5, t5 = d2 - d5, t3 = d3 + d4, t4 = d3 - d4;
    const float e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
    *d0p = e10 + e11;
    *d4p = e10 - e11;
    const float z1 = (e12 + e13) * 0.707106781f;
    *d2p = e13 + z1;
    *d6p = e13 - z1;
    const float o10 = t4 + t5, o11 = t5 + t6, o12 = t6 + t7;
    const float z5 = (o10 - o12) * 0.382683433f;
    const float z2 = o10 * 0.541196100f + z5;
    const float z4 = o12 * 1.306562965f + z5;
    const float z3 = o11 * 0.707106781f;
    const float z11 = t7 + z3, z13 = t7 - z3;
    *d5p = z13 + z2;
    *d3p = z13 - z2;
    *d1p = z11 + z4;
    *d7p = z11 - z4;
}

Code magnitude(int v) {   // JPEG "additional bits" of a coefficient
    int a = v < 0 ? -v : v;
    const int w = v < 0 ? v - 1 : v;
    Code c;
    c.len = 1;
    while (a >>= 1) ++c.len;
    c.bits = (uint16_t)(w & ((1 << c.len) - 1));
    return c;
}

int encodeBlock(JpegBits& bits, float* blk, const float* scale, int dcPrev, const Code* dc, const Code* ac) {
    for (int r = 0; r < 64; r += 8) dct8(&blk[r], &blk[r + 1], &blk[r + 2], &blk[r + 3], &blk[r + 4], &blk[r + 5], &blk[r + 6], &blk[r + 7]);
    for (int c = 0; c < 8; ++c) dct8(&blk[c], &blk[c + 8], &blk[c + 16], &blk[c + 24], &blk[c + 32], &blk[c + 40], &blk[c + 48], &blk[c + 56]);
    int du[64];
    for (int i = 0; i < 64; ++i) {
        const float v = blk[i] * scale[i];
        du[kZigZag[i]] = (int)(v < 0 ? v - 0.5f : v + 0.5f);
    }
    const int diff = du[0] - dcPrev;
    if (diff == 0) bits.put(dc[0]);
    else { const Code m = magnitude(diff); bits.put(dc[m.len]); bits.put(m); }
    int last = 63;
    while (last > 0 && du[last] == 0) --last;
    if (last == 0) { bits.put(ac[0x00]); return du[0]; }
    for (int i = 1; i <= last; ++i) {
        const int start = i;
        while (du[i] == 0 && i <= last) ++i;
        int zeros = i - start;
        for (int k = 0; k < (zeros >> 4); ++k) bits.put(ac[0xF0]);
        zeros &= 15;
        const Code m = magnitude(du[i]);
        bits.put(ac[(zeros << 4) + m.len]);
        bits.put(m);
    }
    if (last != 63) bits.put(ac[0x00]);
    return du[0];
}

Bytes encodeJpeg(const uint8_t* px, int w, int h, int quality) {
    quality = quality ? quality : 90;
    quality = quality < 1 ? 1 : quality > 100 ? 100 : quality;
    quality = quality < 50 ? 5000 / quality : 200 - quality * 2;
    uint8_t yq[64], cq[64];
    for (int i = 0; i < 64; ++i) {
        const int a = (kLumQ[i] * quality + 50) / 100, b = (kChrQ[i] * quality + 50) / 100;
        yq[kZigZag[i]] = (uint8_t)(a < 1 ? 1 : a > 255 ? 255 : a);
        cq[kZigZag[i]] = (uint8_t)(b < 1 ? 1 : b > 255 ? 255 : b);
    }
    // AAN output scale: 2*sqrt(2) * (1, sqrt(2) cos(k pi / 16) for k = 1..7), IJG constants
    static const float aan[8] = {1.0f * 2.828427125f, 1.387039845f * 2.828427125f, 1.306562965f * 2.828427125f, 1.175875602f * 2.828427125f,
                                 1.0f * 2.828427125f, 0.785694958f * 2.828427125f, 0.541196100f * 2.828427125f, 0.275899379f * 2.828427125f};
    float ys[64], cs[64];
    for (int r = 0, k = 0; r < 8; ++r)
        for (int c = 0; c < 8; ++c, ++k) {
            ys[k] = 1 / (yq[kZigZag[k]] * aan[r] * aan[c]);
            cs[k] = 1 / (cq[kZigZag[k]] * aan[r] * aan[c]);
        }
    Code ydc[256], yac[256], cdc[256], cac[256];
    buildCodes(kDcLumCounts, kDcValues, ydc);
    buildCodes(kAcLumCounts, kAcLumValues, yac);
    buildCodes(kDcChrCounts, kDcValues, cdc);
    buildCodes(kAcChrCounts, kAcChrValues, cac);

    Bytes o = {0xFF, 0xD8, 0xFF, 0xE0, 0, 0x10, 'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0, 0xFF, 0xDB, 0, 0x84, 0};
    o.insert(o.end(), yq, yq + 64);
    o.push_back(1);
    o.insert(o.end(), cq, cq + 64);
    o.insert(o.end(), {0xFF, 0xC0, 0, 0x11, 8, (uint8_t)(h >> 8), (uint8_t)(h & 255), (uint8_t)(w >> 8), (uint8_t)(w & 255),
                       3, 1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1, 0xFF, 0xC4, 0x01, 0xA2, 0});
    o.insert(o.end(), kDcLumCounts, kDcLumCounts + 16);
    o.insert(o.end(), kDcValues, kDcValues + 12);
    o.push_back(0x10);
    o.insert(o.end(), kAcLumCounts, kAcLumCounts + 16);
    o.insert(o.end(), kAcLumValues, kAcLumValues + 162);
    o.push_back(1);
    o.insert(o.end(), kDcChrCounts, kDcChrCounts + 16);
    o.insert(o.end(), kDcValues, kDcValues + 12);
    o.push_back(0x11);
    o.insert(o.end(), kAcChrCounts, kAcChrCounts + 16);
    o.insert(o.end(), kAcChrValues, kAcChrValues + 162);
    o.insert(o.end(), {0xFF, 0xDA, 0, 0xC, 3, 1, 0, 2, 0x11, 3, 0x11, 0, 0x3F, 0});

    JpegBits bits(o);
    int dy = 0, du = 0, dv = 0;
    for (int y = 0; y < h; y += 8)
        for (int x = 0; x < w; x += 8) {
            float Y[64], U[64], V[64];
            for (int row = y, pos = 0; row < y + 8; ++row) {
                const int cr = row < h ? row : h - 1;
                const size_t base = (size_t)(h - 1 - cr) * w * 4;     // bottom-up
                for (int col = x; col < x + 8; ++col, ++pos) {
                    const uint8_t* p = px + base + (size_t)(col < w ? col : w - 1) * 4;
                    const float r = p[0], g = p[1], b = p[2];
                    Y[pos] = +0.29900f * r + 0.58700f * g + 0.11400f * b - 128;
                    U[pos] = -0.16874f * r - 0.33126f * g + 0.50000f * b;
                    V[pos] = +0.50000f * r - 0.41869f * g - 0.08131f * b;
                }
            }
            dy = encodeBlock(bits, Y, ys, dy, ydc, yac);
            du = encodeBlock(bits, U, cs, du, cdc, cac);
            dv = encodeBlock(bits, V, cs, dv, cdc, cac);
        }
    bits.put({0x7F, 7});
    o.push_back(0xFF);
    o.push_back(0xD9);
    return o;
}

Bytes encodePfm(const float* rgba, int w, int h) {
    const std::string head = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.000000\n";
    Bytes o(head.begin(), head.end());
    o.reserve(o.size() + (size_t)w * h * 12);
    for (size_t p = 0; p < (size_t)w * h; ++p) {
        const uint8_t* b = reinterpret_cast<const uint8_t*>(rgba + 4 * p);
        o.insert(o.end(), b, b + 12);
    }
    return o;
}

std::string extensionOf(const std::string& path) {   // utils::path::fileextension
    const size_t slash = path.find_last_of('/');
    const std::string base = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t start = base.find_first_not_of('.'), pos = base.find_last_of('.');
    if (pos == std::string::npos || start >= pos) return "";
    return base.substr(pos + 1);
}

}  // namespace

namespace omv {

int encodeImage(const std::string& ext, const float* rgba, int w, int h, std::vector<uint8_t>& out) {
    if (ext == "pfm") { out = encodePfm(rgba, w, h); return OMV_OK; }
    if (ext != "png" && ext != "bmp" && ext != "tga" && ext != "jpg") {
        setError("This image extension is not supported");
        return OMV_E_ARG;
    }
    std::vector<uint8_t> d((size_t)w * h * 4);
    for (size_t i = 0; i < d.size(); ++i) {   // main.cpp:91: static_cast<unsigned char>(255.0f * f) of [0, 1] values
        const float v = 255.0f * rgba[i];
        d[i] = (uint8_t)(v <= 0.0f ? 0.0f : v >= 255.0f ? 255.0f : v);
    }
    if (ext == "png") out = encodePng(d.data(), w, h);
    else if (ext == "bmp") out = encodeBmp(d.data(), w, h);
    else if (ext == "tga") out = encodeTga(d.data(), w, h);
    else out = encodeJpeg(d.data(), w, h, 100);
    return OMV_OK;
}

int saveImage(const char* filename, const float* rgba, int w, int h) {
    std::vector<uint8_t> bytes;
    const int rc = encodeImage(extensionOf(filename), rgba, w, h, bytes);
    if (rc != OMV_OK) return rc;
    FILE* f = std::fopen(filename, "wb");
    if (!f) { setError(std::string("Cannot open file: ") + filename); return OMV_E_IO; }
    const size_t n = std::fwrite(bytes.data(), 1, bytes.size(), f);
    const int closed = std::fclose(f);
    if (n != bytes.size() || closed != 0) { setError(std::string("Cannot write file: ") + filename); return OMV_E_IO; }
    return OMV_OK;
}

}  // namespace omv
