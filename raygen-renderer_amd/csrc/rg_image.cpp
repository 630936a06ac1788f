// rg_image.cpp — image writers of the render CLI's output stage
// (src/main.cpp:400-410, 479-485: a .hdr path receives the linear HDR image,
// any other extension the tonemapped LDR image; the default output is
// <scene dir>/<scene name>.jpg).  The reference hands both to its
// unvendored image library (saveImage); these are written from the file
// formats themselves:
//   .hdr          Radiance RGBE, flat scanlines (m * 2^(e - 136) per channel)
//   .jpg / .jpeg  baseline sequential JPEG/JFIF, 4:4:4, the example tables of
//                 ITU-T T.81 Annex K scaled by quality as the IJG library does
//   .png          8-bit RGB, zlib stream of stored (uncompressed) deflate blocks
//   .ppm          binary P6
//   .bmp          24-bit bottom-up BI_RGB
// LDR pixels are bytes round(clamp(v, 0, 1) * 255).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rg_internal.h"

namespace rg {
namespace {

uint8_t to_byte(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 1.0f) return 255;
    return (uint8_t)(v * 255.0f + 0.5f);
}

std::string lower_ext(const char *path) {
    const char *dot = std::strrchr(path, '.');
    const char *slash = std::strrchr(path, '/');
    if (!dot || (slash && dot < slash)) return "";
    std::string e(dot + 1);
    for (char &c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

struct File {
    FILE *f;
    explicit File(const char *p) : f(std::fopen(p, "wb")) {}
    ~File() {
        if (f) std::fclose(f);
    }
    bool put(const void *p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
    bool close() {
        const bool ok = std::fclose(f) == 0;
        f = nullptr;
        return ok;
    }
};

// ---- Radiance RGBE ---------------------------------------------------------
bool write_rgbe(const char *path, uint32_t w, uint32_t h, const float *rgba) {
    File F(path);
    if (!F.f) return false;
    char head[128];
    const int n = std::snprintf(head, sizeof head, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
    if (!F.put(head, (size_t)n)) return false;
    std::vector<uint8_t> row((size_t)w * 4);
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++) {
            const float *c = rgba + 4 * ((size_t)y * w + x);
            const float r = std::fmax(c[0], 0.0f), g = std::fmax(c[1], 0.0f), b = std::fmax(c[2], 0.0f);
            const float v = std::fmax(r, std::fmax(g, b));
            uint8_t *o = &row[4 * (size_t)x];
            if (!(v >= 1e-32f) || !std::isfinite(v)) {
                o[0] = o[1] = o[2] = o[3] = 0;
                continue;
            }
            int e;
            std::frexp(v, &e);  // v = f * 2^e, f in [0.5, 1)
            // mantissas m = floor(c * 256 / 2^e) in [0, 255]; decoded as
            // (m + 0.5) * 2^(e - 8) by the common readers, m * 2^(e - 8) by
            // the loader here
            const float s = std::ldexp(1.0f, 8 - e);
            o[0] = (uint8_t)std::fmin(255.0f, r * s);
            o[1] = (uint8_t)std::fmin(255.0f, g * s);
            o[2] = (uint8_t)std::fmin(255.0f, b * s);
            o[3] = (uint8_t)(e + 128);
        }
        if (!F.put(row.data(), row.size())) return false;
    }
    return F.close();
}

// ---- PPM / BMP --------------------------------------------------------------
bool write_ppm(const char *path, uint32_t w, uint32_t h, const float *rgba) {
    File F(path);
    if (!F.f) return false;
    char head[64];
    const int n = std::snprintf(head, sizeof head, "P6\n%u %u\n255\n", w, h);
    if (!F.put(head, (size_t)n)) return false;
    std::vector<uint8_t> row((size_t)w * 3);
    for (uint32_t y = 0; y < h; y++) {
        for (uint32_t x = 0; x < w; x++)
            for (int k = 0; k < 3; k++) row[3 * x + k] = to_byte(rgba[4 * ((size_t)y * w + x) + k]);
        if (!F.put(row.data(), row.size())) return false;
    }
    return F.close();
}

void le32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

bool write_bmp(const char *path, uint32_t w, uint32_t h, const float *rgba) {
    File F(path);
    if (!F.f) return false;
    const uint32_t stride = (3 * w + 3) & ~3u;
    uint8_t hd[54] = {'B', 'M'};
    le32(hd + 2, 54 + stride * h);
    le32(hd + 10, 54);
    le32(hd + 14, 40);
    le32(hd + 18, w);
    le32(hd + 22, h);
    hd[26] = 1;
    hd[28] = 24;
    le32(hd + 34, stride * h);
    le32(hd + 38, 2835);
    le32(hd + 42, 2835);
    if (!F.put(hd, sizeof hd)) return false;
    std::vector<uint8_t> row(stride, 0);
    for (uint32_t yy = 0; yy < h; yy++) {
        const uint32_t y = h - 1 - yy;  // bottom-up
        for (uint32_t x = 0; x < w; x++) {
            const float *c = rgba + 4 * ((size_t)y * w + x);
            row[3 * x] = to_byte(c[2]);
            row[3 * x + 1] = to_byte(c[1]);
            row[3 * x + 2] = to_byte(c[0]);
        }
        if (!F.put(row.data(), row.size())) return false;
    }
    return F.close();
}

// ---- PNG (stored deflate blocks) --------------------------------------------
uint32_t crc32_update(uint32_t c, const uint8_t *p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t v = i;
            for (int k = 0; k < 8; k++) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}

void be32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}

bool png_chunk(File &F, const char *type, const uint8_t *data, uint32_t n) {
    uint8_t hd[8];
    be32(hd, n);
    std::memcpy(hd + 4, type, 4);
    uint32_t c = crc32_update(0xFFFFFFFFu, hd + 4, 4);
    c = crc32_update(c, data, n) ^ 0xFFFFFFFFu;
    uint8_t tail[4];
    be32(tail, c);
    return F.put(hd, 8) && (n == 0 || F.put(data, n)) && F.put(tail, 4);
}

bool write_png(const char *path, uint32_t w, uint32_t h, const float *rgba) {
    // raw scanlines, filter type 0
    const size_t stride = 1 + 3 * (size_t)w;
    std::vector<uint8_t> raw(stride * h);
    for (uint32_t y = 0; y < h; y++) {
        uint8_t *r = &raw[y * stride];
        r[0] = 0;
        for (uint32_t x = 0; x < w; x++)
            for (int k = 0; k < 3; k++) r[1 + 3 * x + k] = to_byte(rgba[4 * ((size_t)y * w + x) + k]);
    }
    // zlib: header, stored blocks of at most 65535 bytes, Adler-32
    std::vector<uint8_t> z;
    z.reserve(raw.size() + raw.size() / 65535 * 5 + 16);
    z.push_back(0x78);
    z.push_back(0x01);
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)len);
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len);
        z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    uint8_t ad[4];
    be32(ad, (b << 16) | a);
    z.insert(z.end(), ad, ad + 4);
    File F(path);
    if (!F.f) return false;
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    uint8_t ihdr[13];
    be32(ihdr, w);
    be32(ihdr + 4, h);
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 2;   // truecolour
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering
    ihdr[12] = 0;  // no interlace
    return F.put(sig, 8) && png_chunk(F, "IHDR", ihdr, 13) && png_chunk(F, "IDAT", z.data(), (uint32_t)z.size()) &&
           png_chunk(F, "IEND", nullptr, 0) && F.close();
}

// ---- baseline JPEG -------------------------------------------------------------
// Quantisation tables of ITU-T T.81 Annex K.1 (natural order) and the
// Huffman tables of K.3 (bits per code length, values).
const uint8_t kQLum[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                           14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                           18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                           49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kQChr[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99,
                           99, 99, 47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                           99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
const uint8_t kZig[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                          41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                          30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
const uint8_t kDcLumBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcLumVal[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kDcChrBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcChrVal[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumVal[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const uint8_t kAcChrBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChrVal[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

struct Huff {
    uint16_t code[256];
    uint8_t len[256];
    void build(const uint8_t *bits, const uint8_t *val) {
        std::memset(len, 0, sizeof len);
        uint16_t c = 0;
        int k = 0;
        for (int l = 1; l <= 16; l++) {
            for (int i = 0; i < bits[l - 1]; i++, k++) {
                code[val[k]] = c++;
                len[val[k]] = (uint8_t)l;
            }
            c <<= 1;
        }
    }
};

struct BitWriter {
    std::vector<uint8_t> &out;
    uint32_t acc = 0;
    int n = 0;
    explicit BitWriter(std::vector<uint8_t> &o) : out(o) {}
    void put(uint32_t v, int bits) {
        for (int i = bits - 1; i >= 0; i--) {
            acc = (acc << 1) | ((v >> i) & 1u);
            if (++n == 8) {
                out.push_back((uint8_t)acc);
                if ((uint8_t)acc == 0xFF) out.push_back(0);  // byte stuffing
                acc = 0;
                n = 0;
            }
        }
    }
    void flush() {
        while (n) put(1, 1);  // pad with ones
    }
};

int magnitude_bits(int v) {
    int a = v < 0 ? -v : v, s = 0;
    while (a) {
        s++;
        a >>= 1;
    }
    return s;
}

void encode_block(BitWriter &bw, const float *blk, const float *qscale, int &pred, const Huff &dc, const Huff &ac) {
    // forward DCT-II, separable, float64
    static double cosv[8][8];
    static bool init = false;
    if (!init) {
        for (int x = 0; x < 8; x++)
            for (int u = 0; u < 8; u++) cosv[x][u] = std::cos((2 * x + 1) * u * M_PI / 16.0);
        init = true;
    }
    double tmp[64], F[64];
    for (int y = 0; y < 8; y++)
        for (int u = 0; u < 8; u++) {
            double s = 0;
            for (int x = 0; x < 8; x++) s += blk[y * 8 + x] * cosv[x][u];
            tmp[y * 8 + u] = s * (u == 0 ? std::sqrt(0.125) : 0.5);
        }
    for (int u = 0; u < 8; u++)
        for (int v = 0; v < 8; v++) {
            double s = 0;
            for (int y = 0; y < 8; y++) s += tmp[y * 8 + u] * cosv[y][v];
            F[v * 8 + u] = s * (v == 0 ? std::sqrt(0.125) : 0.5);
        }
    int q[64];
    for (int i = 0; i < 64; i++) q[i] = (int)std::lround(F[kZig[i]] / qscale[kZig[i]]);
    const int diff = q[0] - pred;
    pred = q[0];
    const int ds = magnitude_bits(diff);
    bw.put(dc.code[ds], dc.len[ds]);
    if (ds) bw.put((uint32_t)(diff < 0 ? diff + (1 << ds) - 1 : diff), ds);
    int run = 0;
    for (int i = 1; i < 64; i++) {
        if (q[i] == 0) {
            run++;
            continue;
        }
        while (run > 15) {
            bw.put(ac.code[0xF0], ac.len[0xF0]);
            run -= 16;
        }
        const int s = magnitude_bits(q[i]);
        const int sym = (run << 4) | s;
        bw.put(ac.code[sym], ac.len[sym]);
        bw.put((uint32_t)(q[i] < 0 ? q[i] + (1 << s) - 1 : q[i]), s);
        run = 0;
    }
    if (run) bw.put(ac.code[0], ac.len[0]);  // end of block
}

void put16(std::vector<uint8_t> &o, uint32_t v) {
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

bool write_jpeg(const char *path, uint32_t w, uint32_t h, const float *rgba, int quality) {
    if (w > 65535 || h > 65535) return false;
    quality = quality < 1 ? 1 : (quality > 100 ? 100 : quality);
    const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;  // IJG quality scaling
    uint8_t ql[64], qc[64];
    for (int i = 0; i < 64; i++) {
        ql[i] = (uint8_t)std::min(255, std::max(1, (kQLum[i] * scale + 50) / 100));
        qc[i] = (uint8_t)std::min(255, std::max(1, (kQChr[i] * scale + 50) / 100));
    }
    float sl[64], sc[64];
    for (int i = 0; i < 64; i++) {
        sl[i] = (float)ql[i];
        sc[i] = (float)qc[i];
    }
    Huff dcl, acl, dcc, acc;
    dcl.build(kDcLumBits, kDcLumVal);
    acl.build(kAcLumBits, kAcLumVal);
    dcc.build(kDcChrBits, kDcChrVal);
    acc.build(kAcChrBits, kAcChrVal);
    std::vector<uint8_t> o;
    o.reserve((size_t)w * h + 1024);
    o.push_back(0xFF);
    o.push_back(0xD8);  // SOI
    static const uint8_t app0[] = {0xFF, 0xE0, 0, 16, 'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0};
    o.insert(o.end(), app0, app0 + sizeof app0);
    for (int t = 0; t < 2; t++) {  // DQT, zig-zag order
        o.push_back(0xFF);
        o.push_back(0xDB);
        put16(o, 67);
        o.push_back((uint8_t)t);
        const uint8_t *q = t ? qc : ql;
        for (int i = 0; i < 64; i++) o.push_back(q[kZig[i]]);
    }
    o.push_back(0xFF);
    o.push_back(0xC0);  // SOF0
    put16(o, 17);
    o.push_back(8);
    put16(o, h);
    put16(o, w);
    o.push_back(3);
    for (int c = 0; c < 3; c++) {
        o.push_back((uint8_t)(c + 1));
        o.push_back(0x11);
        o.push_back(c ? 1 : 0);
    }
    struct T {
        int cls, id;
        const uint8_t *bits, *val;
        int nval;
    } tabs[4] = {{0, 0, kDcLumBits, kDcLumVal, 12},
                 {1, 0, kAcLumBits, kAcLumVal, 162},
                 {0, 1, kDcChrBits, kDcChrVal, 12},
                 {1, 1, kAcChrBits, kAcChrVal, 162}};
    for (const T &t : tabs) {  // DHT
        o.push_back(0xFF);
        o.push_back(0xC4);
        put16(o, (uint32_t)(2 + 1 + 16 + t.nval));
        o.push_back((uint8_t)((t.cls << 4) | t.id));
        o.insert(o.end(), t.bits, t.bits + 16);
        o.insert(o.end(), t.val, t.val + t.nval);
    }
    static const uint8_t sos[] = {0xFF, 0xDA, 0, 12, 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0};
    o.insert(o.end(), sos, sos + sizeof sos);
    BitWriter bw(o);
    int pred[3] = {0, 0, 0};
    float blk[3][64];
    for (uint32_t by = 0; by < h; by += 8)
        for (uint32_t bx = 0; bx < w; bx += 8) {
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++) {
                    // edge blocks repeat the last row / column
                    const uint32_t sx = std::min(bx + x, w - 1), sy = std::min(by + y, h - 1);
                    const float *c = rgba + 4 * ((size_t)sy * w + sx);
                    const float r = to_byte(c[0]), g = to_byte(c[1]), b = to_byte(c[2]);
                    // JFIF YCbCr (full range), level-shifted by 128
                    blk[0][y * 8 + x] = 0.299f * r + 0.587f * g + 0.114f * b - 128.0f;
                    blk[1][y * 8 + x] = -0.168736f * r - 0.331264f * g + 0.5f * b;
                    blk[2][y * 8 + x] = 0.5f * r - 0.418688f * g - 0.081312f * b;
                }
            encode_block(bw, blk[0], sl, pred[0], dcl, acl);
            encode_block(bw, blk[1], sc, pred[1], dcc, acc);
            encode_block(bw, blk[2], sc, pred[2], dcc, acc);
        }
    bw.flush();
    o.push_back(0xFF);
    o.push_back(0xD9);  // EOI
    File F(path);
    if (!F.f) return false;
    return F.put(o.data(), o.size()) && F.close();
}

}  // namespace
}  // namespace rg

using namespace rg;

extern "C" int rg_image_write(const char *path, uint32_t width, uint32_t height, const float *rgba, int quality) {
    if (!path || !rgba || width == 0 || height == 0) return set_error(RG_E_INVALID, "rg_image_write: bad argument");
    const std::string e = lower_ext(path);
    bool ok;
    if (e == "hdr") ok = write_rgbe(path, width, height, rgba);
    else if (e == "jpg" || e == "jpeg") ok = write_jpeg(path, width, height, rgba, quality > 0 ? quality : 90);
    else if (e == "png") ok = write_png(path, width, height, rgba);
    else if (e == "ppm") ok = write_ppm(path, width, height, rgba);
    else if (e == "bmp") ok = write_bmp(path, width, height, rgba);
    else return set_error(RG_E_UNSUPPORTED, "rg_image_write: unknown image extension of " + std::string(path));
    if (!ok) return set_error(RG_E_IO, "rg_image_write: cannot write " + std::string(path));
    return RG_OK;
}
