// rg_jpeg.cpp — baseline JPEG decoder for the material textures of scene
// files (the reference's textures are JFIF JPEGs read by its image library,
// texture.cpp:177-195; that library is not vendored, so this is our own
// decoder of the published format, ITU-T T.81).
//
// Supported: sequential DCT with Huffman coding (SOF0 / SOF1, 8-bit
// samples), one or three components with any sampling factors up to 4,
// interleaved and single-component scans, restart intervals (DRI + RSTn),
// 8- and 16-bit quantisation tables.  Progressive, arithmetic-coded,
// lossless and 12-bit files are refused with an error.
//
// Output: 8-bit RGB, row 0 at the top.  Chroma is upsampled by replication
// and converted with the JFIF YCbCr equations; the inverse DCT is the direct
// separable float transform.  (The reference's library may use a different
// IDCT and chroma filter, so decoded texels can differ from it by a few
// levels: parity unpinned at this step.)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace rg {

namespace {

const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 Annex C) with the decoding procedure of
// F.2.2.3: per code length the largest code and the index of its first value.
struct Huff {
    bool defined = false;
    uint8_t vals[256] = {};
    int maxcode[18] = {};
    int valptr[17] = {};
    int mincode[17] = {};
};

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0;
    int pred = 0;
    int bw = 0, bh = 0;          // blocks per row / column of this component's plane
    std::vector<uint8_t> plane;  // bw*8 x bh*8 samples
};

struct Decoder {
    const uint8_t *p;
    size_t n, pos = 0;
    uint16_t qt[4][64] = {};
    bool qt_defined[4] = {};
    Huff dc[4], ac[4];
    Comp comp[4];
    int ncomp = 0, W = 0, H = 0, hmax = 1, vmax = 1, restart = 0;
    bool frame = false;
    uint32_t bitbuf = 0;
    int bitcnt = 0;
    bool hit_marker = false;
    std::string err;

    bool fail(const std::string &m) {
        if (err.empty()) err = m;
        return false;
    }
    int u8() { return pos < n ? p[pos++] : -1; }
    int u16() {
        const int a = u8(), b = u8();
        return (a < 0 || b < 0) ? -1 : (a << 8) | b;
    }

    bool read_dqt(size_t end) {
        while (pos < end) {
            const int pq = u8();
            if (pq < 0) return fail("truncated DQT");
            const int t = pq & 15, prec = pq >> 4;
            if (t > 3) return fail("bad DQT table id");
            for (int k = 0; k < 64; k++) {
                const int v = prec ? u16() : u8();
                if (v < 0) return fail("truncated DQT");
                qt[t][kZigzag[k]] = (uint16_t)v;
            }
            qt_defined[t] = true;
        }
        return true;
    }

    bool read_dht(size_t end) {
        while (pos < end) {
            const int tc = u8();
            if (tc < 0) return fail("truncated DHT");
            const int cls = tc >> 4, id = tc & 15;
            if (cls > 1 || id > 3) return fail("bad DHT table id");
            Huff &hf = cls ? ac[id] : dc[id];
            int bits[17] = {0}, total = 0;
            for (int l = 1; l <= 16; l++) {
                bits[l] = u8();
                if (bits[l] < 0) return fail("truncated DHT");
                total += bits[l];
            }
            if (total > 256) return fail("DHT with more than 256 codes");
            for (int k = 0; k < total; k++) {
                const int v = u8();
                if (v < 0) return fail("truncated DHT");
                hf.vals[k] = (uint8_t)v;
            }
            int code = 0, k = 0;
            for (int l = 1; l <= 16; l++) {
                hf.valptr[l] = k;
                hf.mincode[l] = code;
                code += bits[l];
                k += bits[l];
                hf.maxcode[l] = bits[l] ? code - 1 : -1;
                code <<= 1;
            }
            hf.maxcode[17] = 0x7FFFFFFF;
            hf.defined = true;
        }
        return true;
    }

    bool read_sof(size_t end, int marker) {
        if (marker != 0xC0 && marker != 0xC1)
            return fail(marker == 0xC2 ? "progressive JPEG is not supported" : "unsupported JPEG coding process");
        const int prec = u8();
        H = u16();
        W = u16();
        ncomp = u8();
        if (prec != 8) return fail("only 8-bit JPEG samples are supported");
        if (W <= 0 || H <= 0 || W > 32768 || H > 32768) return fail("bad JPEG size");
        if (ncomp != 1 && ncomp != 3) return fail("only 1- or 3-component JPEGs are supported");
        for (int c = 0; c < ncomp; c++) {
            comp[c].id = u8();
            const int hv = u8();
            comp[c].tq = u8();
            comp[c].h = hv >> 4;
            comp[c].v = hv & 15;
            if (comp[c].h < 1 || comp[c].h > 4 || comp[c].v < 1 || comp[c].v > 4 || comp[c].tq > 3 || comp[c].tq < 0)
                return fail("bad JPEG component");
            hmax = std::max(hmax, comp[c].h);
            vmax = std::max(vmax, comp[c].v);
        }
        if (pos > end) return fail("truncated SOF");
        const int mx = (W + 8 * hmax - 1) / (8 * hmax), my = (H + 8 * vmax - 1) / (8 * vmax);
        for (int c = 0; c < ncomp; c++) {
            comp[c].bw = mx * comp[c].h;
            comp[c].bh = my * comp[c].v;
            comp[c].plane.assign((size_t)comp[c].bw * 8 * comp[c].bh * 8, 0);
        }
        frame = true;
        return true;
    }

    // entropy-coded segment bits (byte stuffing removed; a marker stops input)
    int bit() {
        if (bitcnt == 0) {
            int b = 0;
            if (!hit_marker && pos < n) {
                b = p[pos];
                if (b == 0xFF) {
                    const int nx = pos + 1 < n ? p[pos + 1] : 0;
                    if (nx == 0x00) pos += 2;
                    else {
                        hit_marker = true;  // feed zeros until the restart / end
                        b = 0;
                    }
                } else {
                    pos++;
                }
            }
            bitbuf = (uint32_t)b;
            bitcnt = 8;
        }
        bitcnt--;
        return (int)((bitbuf >> bitcnt) & 1u);
    }
    int bits(int s) {
        int v = 0;
        for (int k = 0; k < s; k++) v = (v << 1) | bit();
        return v;
    }
    int decode(const Huff &hf) {
        int code = bit(), l = 1;
        while (l <= 16 && code > hf.maxcode[l]) {
            code = (code << 1) | bit();
            l++;
        }
        if (l > 16) return -1;
        return hf.vals[hf.valptr[l] + code - hf.mincode[l]];
    }
    static int extend(int v, int s) { return (s > 0 && v < (1 << (s - 1))) ? v - (1 << s) + 1 : v; }

    bool block(Comp &c, int bx, int by, const float (&cosk)[8][8]) {
        int coef[64] = {0};
        const int t = decode(dc[c.td]);
        if (t < 0 || t > 11) return fail("bad JPEG DC code");
        c.pred += extend(bits(t), t);
        coef[0] = c.pred * qt[c.tq][0];
        for (int k = 1; k < 64;) {
            const int rs = decode(ac[c.ta]);
            if (rs < 0) return fail("bad JPEG AC code");
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break;  // end of block
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) return fail("JPEG AC run past the block");
            coef[kZigzag[k]] = extend(bits(s), s) * qt[c.tq][kZigzag[k]];
            k++;
        }
        // separable inverse DCT: rows then columns
        float tmp[64];
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) {
                float a = 0.0f;
                for (int u = 0; u < 8; u++) a += cosk[x][u] * (float)coef[y * 8 + u];
                tmp[y * 8 + x] = a;
            }
        uint8_t *dst = c.plane.data() + (size_t)by * 8 * (c.bw * 8) + (size_t)bx * 8;
        for (int x = 0; x < 8; x++)
            for (int y = 0; y < 8; y++) {
                float a = 0.0f;
                for (int v = 0; v < 8; v++) a += cosk[y][v] * tmp[v * 8 + x];
                const float s = std::round(a * 0.25f + 128.0f);
                dst[(size_t)y * (c.bw * 8) + x] = (uint8_t)(s < 0.0f ? 0.0f : (s > 255.0f ? 255.0f : s));
            }
        return true;
    }

    bool restart_marker() {
        // skip to the RSTn marker, reset the bit reader and the predictors
        bitcnt = 0;
        hit_marker = false;
        while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)) pos++;
        if (pos + 1 >= n) return fail("missing JPEG restart marker");
        pos += 2;
        for (int c = 0; c < ncomp; c++) comp[c].pred = 0;
        return true;
    }

    bool read_sos() {
        if (!frame) return fail("JPEG scan before frame header");
        const int len = u16();
        const size_t end = pos + (size_t)len - 2;
        const int ns = u8();
        if (ns < 1 || ns > ncomp) return fail("bad JPEG scan");
        int idx[4];
        for (int k = 0; k < ns; k++) {
            const int cid = u8(), tt = u8();
            idx[k] = -1;
            for (int c = 0; c < ncomp; c++)
                if (comp[c].id == cid) idx[k] = c;
            if (idx[k] < 0 || tt < 0) return fail("JPEG scan names an unknown component");
            comp[idx[k]].td = tt >> 4;
            comp[idx[k]].ta = tt & 15;
            if (comp[idx[k]].td > 3 || comp[idx[k]].ta > 3 || !dc[comp[idx[k]].td].defined ||
                !ac[comp[idx[k]].ta].defined || !qt_defined[comp[idx[k]].tq])
                return fail("JPEG scan uses an undefined table");
        }
        pos = end;  // Ss, Se, Ah/Al: fixed for sequential DCT
        float cosk[8][8];
        for (int x = 0; x < 8; x++)
            for (int u = 0; u < 8; u++)
                cosk[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * M_PI / 16.0));
        for (int c = 0; c < ncomp; c++) comp[c].pred = 0;
        bitcnt = 0;
        hit_marker = false;
        int mcu = 0;
        if (ns == 1) {
            // non-interleaved: the component's own block grid over the image
            Comp &c = comp[idx[0]];
            const int bw = ((W * c.h + hmax - 1) / hmax + 7) / 8, bh = ((H * c.v + vmax - 1) / vmax + 7) / 8;
            for (int by = 0; by < bh; by++)
                for (int bx = 0; bx < bw; bx++) {
                    if (restart && mcu > 0 && mcu % restart == 0 && !restart_marker()) return false;
                    if (!block(c, bx, by, cosk)) return false;
                    mcu++;
                }
        } else {
            const int mx = (W + 8 * hmax - 1) / (8 * hmax), my = (H + 8 * vmax - 1) / (8 * vmax);
            for (int y = 0; y < my; y++)
                for (int x = 0; x < mx; x++) {
                    if (restart && mcu > 0 && mcu % restart == 0 && !restart_marker()) return false;
                    for (int k = 0; k < ns; k++) {
                        Comp &c = comp[idx[k]];
                        for (int v = 0; v < c.v; v++)
                            for (int h = 0; h < c.h; h++)
                                if (!block(c, x * c.h + h, y * c.v + v, cosk)) return false;
                    }
                    mcu++;
                }
        }
        // continue after the entropy-coded data
        bitcnt = 0;
        while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] != 0x00 && !(p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)))
            pos++;
        return true;
    }

    bool run(std::vector<uint8_t> &rgb) {
        if (n < 4 || p[0] != 0xFF || p[1] != 0xD8) return fail("not a JPEG file");
        pos = 2;
        bool scanned = false;
        while (pos < n) {
            if (p[pos] != 0xFF) {
                pos++;
                continue;
            }
            while (pos < n && p[pos] == 0xFF) pos++;
            if (pos >= n) break;
            const int m = p[pos++];
            if (m == 0xD9) break;                      // EOI
            if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
            if (m == 0xDA) {
                if (!read_sos()) return false;
                scanned = true;
                continue;
            }
            const int len = u16();
            if (len < 2 || pos + (size_t)len - 2 > n) return fail("truncated JPEG segment");
            const size_t end = pos + (size_t)len - 2;
            bool ok = true;
            if (m == 0xDB) ok = read_dqt(end);
            else if (m == 0xC4) ok = read_dht(end);
            else if (m == 0xDD) restart = u16();
            else if (m >= 0xC0 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) ok = read_sof(end, m);
            if (!ok) return false;
            pos = end;
        }
        if (!scanned) return fail("JPEG without image data");
        // upsample (replication) and convert to RGB
        rgb.assign((size_t)W * H * 3, 0);
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                float s[3] = {0, 0, 0};
                for (int c = 0; c < ncomp; c++) {
                    const Comp &cc = comp[c];
                    const int sx = x * cc.h / hmax, sy = y * cc.v / vmax;
                    s[c] = (float)cc.plane[(size_t)sy * (cc.bw * 8) + sx];
                }
                uint8_t *o = &rgb[3 * ((size_t)y * W + x)];
                if (ncomp == 1) {
                    o[0] = o[1] = o[2] = (uint8_t)s[0];
                    continue;
                }
                const float Y = s[0], cb = s[1] - 128.0f, cr = s[2] - 128.0f;
                const float c3[3] = {Y + 1.402f * cr, Y - 0.344136f * cb - 0.714136f * cr, Y + 1.772f * cb};
                for (int k = 0; k < 3; k++) {
                    const float r = std::round(c3[k]);
                    o[k] = (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r));
                }
            }
        return true;
    }
};

}  // namespace

bool decode_jpeg(const uint8_t *data, size_t size, uint32_t &w, uint32_t &h, std::vector<uint8_t> &rgb,
                 std::string &err) {
    Decoder d;
    d.p = data;
    d.n = size;
    if (!d.run(rgb)) {
        err = d.err.empty() ? "JPEG decode failed" : d.err;
        return false;
    }
    w = (uint32_t)d.W;
    h = (uint32_t)d.H;
    return true;
}

}  // namespace rg
