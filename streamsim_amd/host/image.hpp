// image.hpp -- 8-bit images for reflectance-variance textures: the first
// three bands of a JPEG or PNG file as a GDAL byte read gives them.
//
// JPEG: baseline and extended sequential Huffman files (ITU-T T.81), 8-bit
// precision, one or three components, any sampling factors, restart
// intervals, interleaved or single-component scans.  Samples follow the
// IJG/libjpeg-turbo decoder's defaults bit for bit: the "islow" integer
// inverse DCT (Loeffler-Ligtenberg-Moschytz with 13-bit constants, 2 extra
// bits between the passes), fancy (triangle) upsampling of 2:1 components,
// replication otherwise, and the fixed-point YCbCr -> RGB conversion with
// 16-bit tables.  Progressive and arithmetic-coded files are refused.
//
// PNG: non-interlaced images of any colour type at 1 to 16 bits, inflated by
// zlib; palette images give their indices (one band), 16-bit samples are
// clamped to 255, as a byte read of the band does.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

namespace facetimg {

struct Image {
    int width = 0, height = 0, channels = 0;  // channels interleaved per pixel
    std::vector<uint8_t> data;
};

namespace jpeg_detail {

// zig-zag position -> natural (row-major) index of an 8x8 block
constexpr std::array<uint8_t, 64> kNatural = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
    30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    bool defined = false;
    int maxcode[18] = {};  // largest code of each length, -1 if none
    int valptr[17] = {};   // index of the first value of each length
    int mincode[17] = {};
    uint8_t values[256] = {};
    void build(const uint8_t counts[16], const uint8_t* vals, int n) {
        std::copy(vals, vals + n, values);
        int code = 0, k = 0;
        for (int len = 1; len <= 16; ++len) {
            valptr[len] = k;
            mincode[len] = code;
            code += counts[len - 1];
            k += counts[len - 1];
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        defined = true;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;     // Huffman tables of the current scan
    int dcPred = 0;
    int width = 0, height = 0;          // downsampled size
    int blocksX = 0, blocksY = 0;       // padded block grid (whole MCUs)
    std::vector<int16_t> coef;          // blocksX * blocksY * 64, natural order
    std::vector<uint8_t> plane;         // (8 blocksX) x (8 blocksY) samples
};

class BitReader {
public:
    BitReader(const std::vector<uint8_t>& b, size_t pos) : b_(b), pos_(pos) {}
    int bit() {
        if (count_ == 0) fill();
        --count_;
        return int((acc_ >> count_) & 1u);
    }
    int bits(int n) {  // n <= 16
        uint32_t v = 0;
        for (int i = 0; i < n; ++i) v = (v << 1) | uint32_t(bit());
        return int(v);
    }
    // discards the partial byte and expects RSTn
    void restart() {
        count_ = 0;
        marker_ = false;
        while (pos_ + 1 < b_.size() && !(b_[pos_] == 0xFF && b_[pos_ + 1] >= 0xD0 && b_[pos_ + 1] <= 0xD7)) ++pos_;
        if (pos_ + 1 >= b_.size()) throw std::runtime_error("JPEG restart marker missing");
        pos_ += 2;
    }
    size_t position() const { return pos_; }

private:
    void fill() {
        uint8_t x = 0;
        if (!marker_ && pos_ < b_.size()) {
            x = b_[pos_];
            if (x == 0xFF) {
                const uint8_t next = pos_ + 1 < b_.size() ? b_[pos_ + 1] : 0;
                if (next == 0x00) {
                    pos_ += 2;
                } else {  // a marker: the entropy-coded segment ended; feed zeros
                    marker_ = true;
                    x = 0;
                }
            } else {
                ++pos_;
            }
        }
        acc_ = x;
        count_ = 8;
    }
    const std::vector<uint8_t>& b_;
    size_t pos_;
    uint32_t acc_ = 0;
    int count_ = 0;
    bool marker_ = false;
};

inline int decode(BitReader& br, const Huffman& h) {
    if (!h.defined) throw std::runtime_error("JPEG Huffman table missing");
    int code = br.bit();
    int len = 1;
    while (code > h.maxcode[len]) {
        code = (code << 1) | br.bit();
        if (++len > 16) throw std::runtime_error("corrupt JPEG Huffman code");
    }
    return h.values[h.valptr[len] + code - h.mincode[len]];
}

inline int extend(int v, int t) { return t == 0 ? 0 : (v < (1 << (t - 1)) ? v - (1 << t) + 1 : v); }

// islow inverse DCT of one dequantised block into 8x8 samples (+128, clamped)
inline void idct_islow(const int16_t* in, const uint16_t* q, uint8_t* out, int stride) {
    constexpr int kConst = 13, kPass1 = 2;
    constexpr int64_t c0298 = 2446, c0390 = 3196, c0541 = 4433, c0765 = 6270, c0899 = 7373, c1175 = 9633,
                      c1501 = 12299, c1847 = 15137, c1961 = 16069, c2053 = 16819, c2562 = 20995, c3072 = 25172;
    const auto descale = [](int64_t x, int n) { return (x + (int64_t(1) << (n - 1))) >> n; };
    int64_t ws[64];
    for (int col = 0; col < 8; ++col) {
        int64_t d[8];
        for (int r = 0; r < 8; ++r) d[r] = int64_t(in[r * 8 + col]) * q[r * 8 + col];
        int64_t z1 = (d[2] + d[6]) * c0541;
        const int64_t t2 = z1 - d[6] * c1847, t3 = z1 + d[2] * c0765;
        const int64_t t0 = (d[0] + d[4]) * (int64_t(1) << kConst), t1 = (d[0] - d[4]) * (int64_t(1) << kConst);
        const int64_t e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
        int64_t o0 = d[7], o1 = d[5], o2 = d[3], o3 = d[1];
        z1 = o0 + o3;
        int64_t z2 = o1 + o2, z3 = o0 + o2, z4 = o1 + o3;
        const int64_t z5 = (z3 + z4) * c1175;
        o0 *= c0298;
        o1 *= c2053;
        o2 *= c3072;
        o3 *= c1501;
        z1 *= -c0899;
        z2 *= -c2562;
        z3 = z3 * -c1961 + z5;
        z4 = z4 * -c0390 + z5;
        o0 += z1 + z3;
        o1 += z2 + z4;
        o2 += z2 + z3;
        o3 += z1 + z4;
        const int n = kConst - kPass1;
        ws[0 * 8 + col] = descale(e10 + o3, n);
        ws[7 * 8 + col] = descale(e10 - o3, n);
        ws[1 * 8 + col] = descale(e11 + o2, n);
        ws[6 * 8 + col] = descale(e11 - o2, n);
        ws[2 * 8 + col] = descale(e12 + o1, n);
        ws[5 * 8 + col] = descale(e12 - o1, n);
        ws[3 * 8 + col] = descale(e13 + o0, n);
        ws[4 * 8 + col] = descale(e13 - o0, n);
    }
    for (int row = 0; row < 8; ++row) {
        const int64_t* d = ws + row * 8;
        int64_t z1 = (d[2] + d[6]) * c0541;
        const int64_t t2 = z1 - d[6] * c1847, t3 = z1 + d[2] * c0765;
        const int64_t t0 = (d[0] + d[4]) * (int64_t(1) << kConst), t1 = (d[0] - d[4]) * (int64_t(1) << kConst);
        const int64_t e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
        int64_t o0 = d[7], o1 = d[5], o2 = d[3], o3 = d[1];
        z1 = o0 + o3;
        int64_t z2 = o1 + o2, z3 = o0 + o2, z4 = o1 + o3;
        const int64_t z5 = (z3 + z4) * c1175;
        o0 *= c0298;
        o1 *= c2053;
        o2 *= c3072;
        o3 *= c1501;
        z1 *= -c0899;
        z2 *= -c2562;
        z3 = z3 * -c1961 + z5;
        z4 = z4 * -c0390 + z5;
        o0 += z1 + z3;
        o1 += z2 + z4;
        o2 += z2 + z3;
        o3 += z1 + z4;
        const int n = kConst + kPass1 + 3;
        const int64_t v[8] = {e10 + o3, e11 + o2, e12 + o1, e13 + o0, e13 - o0, e12 - o1, e11 - o2, e10 - o3};
        for (int k = 0; k < 8; ++k)
            out[row * stride + k] = uint8_t(std::clamp<int64_t>(descale(v[k], n) + 128, 0, 255));
    }
}

// one component plane upsampled by (fh, fv) to (outW, outH) samples
inline std::vector<uint8_t> upsample(const Component& c, int fh, int fv, int outW, int outH) {
    const int stride = c.blocksX * 8;
    const auto at = [&](int x, int y) {
        x = std::clamp(x, 0, c.width - 1);
        y = std::clamp(y, 0, c.height - 1);
        return int(c.plane[size_t(y) * stride + x]);
    };
    std::vector<uint8_t> out(size_t(outW) * outH);
    const int W = c.width;
    if (fh == 1 && fv == 1) {
        for (int y = 0; y < outH; ++y)
            for (int x = 0; x < outW; ++x) out[size_t(y) * outW + x] = uint8_t(at(x, y));
    } else if (fh == 2 && fv == 1 && W > 2) {
        for (int y = 0; y < outH; ++y)
            for (int x = 0; x < outW; ++x) {
                const int i = x >> 1;
                int v;
                if ((x & 1) == 0) v = i == 0 ? at(0, y) : (3 * at(i, y) + at(i - 1, y) + 1) >> 2;
                else v = i == W - 1 ? at(i, y) : (3 * at(i, y) + at(i + 1, y) + 2) >> 2;
                out[size_t(y) * outW + x] = uint8_t(v);
            }
    } else if (fh == 2 && fv == 2 && W > 2) {
        for (int y = 0; y < outH; ++y) {
            const int j = y >> 1, jn = (y & 1) ? j + 1 : j - 1;  // nearest and next-nearest rows
            const auto colsum = [&](int i) { return 3 * at(i, j) + at(i, jn); };
            for (int x = 0; x < outW; ++x) {
                const int i = x >> 1;
                int v;
                if ((x & 1) == 0) v = i == 0 ? (colsum(0) * 4 + 8) >> 4 : (colsum(i) * 3 + colsum(i - 1) + 8) >> 4;
                else v = i == W - 1 ? (colsum(i) * 4 + 7) >> 4 : (colsum(i) * 3 + colsum(i + 1) + 7) >> 4;
                out[size_t(y) * outW + x] = uint8_t(v);
            }
        }
    } else if (fh == 1 && fv == 2) {
        for (int y = 0; y < outH; ++y) {
            const int j = y >> 1, jn = (y & 1) ? j + 1 : j - 1, bias = (y & 1) ? 2 : 1;
            for (int x = 0; x < outW; ++x) out[size_t(y) * outW + x] = uint8_t((3 * at(x, j) + at(x, jn) + bias) >> 2);
        }
    } else {  // replication
        for (int y = 0; y < outH; ++y)
            for (int x = 0; x < outW; ++x) out[size_t(y) * outW + x] = uint8_t(at(x / fh, y / fv));
    }
    return out;
}

}  // namespace jpeg_detail

inline Image decodeJpeg(const std::vector<uint8_t>& b) {
    using namespace jpeg_detail;
    const auto u16 = [&](size_t o) {
        if (o + 2 > b.size()) throw std::runtime_error("truncated JPEG");
        return int(b[o]) << 8 | int(b[o + 1]);
    };
    if (b.size() < 4 || b[0] != 0xFF || b[1] != 0xD8) throw std::runtime_error("not a JPEG file");
    uint16_t quant[4][64] = {};
    Huffman dc[4], ac[4];
    std::vector<Component> comps;
    int width = 0, height = 0, hmax = 1, vmax = 1, restartInterval = 0;
    bool jfif = false, adobe = false, frame = false;
    int adobeTransform = -1;
    int mcusX = 0, mcusY = 0;
    size_t pos = 2;
    while (pos + 4 <= b.size()) {
        if (b[pos] != 0xFF) {
            ++pos;
            continue;
        }
        const int m = b[pos + 1];
        if (m == 0xFF) {
            ++pos;
            continue;
        }
        if (m == 0xD9) break;
        if ((m >= 0xD0 && m <= 0xD7) || m == 0x01) {
            pos += 2;
            continue;
        }
        const int len = u16(pos + 2);
        const size_t seg = pos + 4, end = pos + 2 + size_t(len);
        if (len < 2 || end > b.size()) throw std::runtime_error("truncated JPEG segment");
        if (m == 0xE0 && len >= 7 && std::equal(b.begin() + long(seg), b.begin() + long(seg) + 5, "JFIF")) {
            jfif = true;
        } else if (m == 0xEE && len >= 14 && std::equal(b.begin() + long(seg), b.begin() + long(seg) + 5, "Adobe")) {
            adobe = true;
            adobeTransform = b[seg + 11];
        } else if (m == 0xDB) {
            size_t p = seg;
            while (p < end) {
                const int pq = b[p] >> 4, tq = b[p] & 15;
                if (tq > 3 || p + 1 + (pq ? 128u : 64u) > end) throw std::runtime_error("bad JPEG quantisation table");
                ++p;
                for (int k = 0; k < 64; ++k) {
                    const int v = pq ? u16(p + 2 * size_t(k)) : b.at(p + size_t(k));
                    quant[tq][kNatural[size_t(k)]] = uint16_t(v);
                }
                p += pq ? 128 : 64;
            }
        } else if (m == 0xC4) {
            size_t p = seg;
            while (p < end) {
                const int tc = b[p] >> 4, th = b[p] & 15;
                if (th > 3 || tc > 1 || p + 17 > end) throw std::runtime_error("bad JPEG Huffman table");
                uint8_t counts[16];
                int n = 0;
                for (int i = 0; i < 16; ++i) n += counts[i] = b[p + 1 + size_t(i)];
                if (n > 256 || p + 17 + size_t(n) > end) throw std::runtime_error("bad JPEG Huffman table");
                (tc ? ac[th] : dc[th]).build(counts, &b[p + 17], n);
                p += 17 + size_t(n);
            }
        } else if (m == 0xDD) {
            if (end < seg + 2) throw std::runtime_error("bad JPEG restart interval");
            restartInterval = u16(seg);
        } else if (m == 0xC0 || m == 0xC1) {
            if (end < seg + 6 || end < seg + 6 + 3 * size_t(b[seg + 5])) throw std::runtime_error("bad JPEG frame");
            if (b[seg] != 8) throw std::runtime_error("only 8-bit JPEG samples are read");
            height = u16(seg + 1);
            width = u16(seg + 3);
            const int nc = b[seg + 5];
            if (width < 1 || height < 1 || (nc != 1 && nc != 3) || size_t(width) * size_t(height) > (size_t(1) << 28))
                throw std::runtime_error("unsupported JPEG frame");
            comps.assign(size_t(nc), Component{});
            for (int i = 0; i < nc; ++i) {
                Component& c = comps[size_t(i)];
                c.id = b[seg + 6 + 3 * size_t(i)];
                c.h = b[seg + 7 + 3 * size_t(i)] >> 4;
                c.v = b[seg + 7 + 3 * size_t(i)] & 15;
                c.tq = b[seg + 8 + 3 * size_t(i)] & 3;
                if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) throw std::runtime_error("bad JPEG sampling factors");
                hmax = std::max(hmax, c.h);
                vmax = std::max(vmax, c.v);
            }
            mcusX = (width + 8 * hmax - 1) / (8 * hmax);
            mcusY = (height + 8 * vmax - 1) / (8 * vmax);
            for (Component& c : comps) {
                c.width = (width * c.h + hmax - 1) / hmax;
                c.height = (height * c.v + vmax - 1) / vmax;
                c.blocksX = mcusX * c.h;
                c.blocksY = mcusY * c.v;
                c.coef.assign(size_t(c.blocksX) * c.blocksY * 64, 0);
            }
            frame = true;
        } else if (m >= 0xC2 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) {
            throw std::runtime_error("progressive, lossless and arithmetic-coded JPEG files are not read");
        } else if (m == 0xDA) {
            if (!frame) throw std::runtime_error("JPEG scan before its frame");
            const int ns = b[seg];
            if (ns < 1 || end < seg + 1 + 2 * size_t(ns)) throw std::runtime_error("bad JPEG scan header");
            std::vector<Component*> sc;
            for (int i = 0; i < ns; ++i) {
                const int id = b[seg + 1 + 2 * size_t(i)], t = b[seg + 2 + 2 * size_t(i)];
                Component* c = nullptr;
                for (Component& x : comps)
                    if (x.id == id) c = &x;
                if (!c) throw std::runtime_error("JPEG scan names an unknown component");
                c->td = t >> 4;
                c->ta = t & 15;
                c->dcPred = 0;
                sc.push_back(c);
            }
            BitReader br(b, end);
            const auto block = [&](Component& c, int bx, int by) {
                int16_t* z = &c.coef[(size_t(by) * c.blocksX + bx) * 64];
                const int t = decode(br, dc[c.td & 3]);
                if (t > 16) throw std::runtime_error("corrupt JPEG DC code");
                c.dcPred += t ? extend(br.bits(t), t) : 0;
                z[0] = int16_t(c.dcPred);
                for (int k = 1; k < 64;) {
                    const int rs = decode(br, ac[c.ta & 3]);
                    const int r = rs >> 4, s = rs & 15;
                    if (s == 0) {
                        if (r != 15) break;
                        k += 16;
                        continue;
                    }
                    k += r;
                    if (k > 63) throw std::runtime_error("corrupt JPEG block");
                    z[kNatural[size_t(k)]] = int16_t(extend(br.bits(s), s));
                    ++k;
                }
            };
            const auto restart = [&](long done, long total) {
                if (restartInterval && done % restartInterval == 0 && done < total) {
                    br.restart();
                    for (Component* c : sc) c->dcPred = 0;
                }
            };
            if (ns == 1) {  // non-interleaved: the component's own blocks
                Component& c = *sc[0];
                const int bw = (c.width + 7) / 8, bh = (c.height + 7) / 8;
                long n = 0;
                for (int by = 0; by < bh; ++by)
                    for (int bx = 0; bx < bw; ++bx) {
                        block(c, bx, by);
                        restart(++n, long(bw) * bh);
                    }
            } else {
                long n = 0;
                for (int my = 0; my < mcusY; ++my)
                    for (int mx = 0; mx < mcusX; ++mx) {
                        for (Component* c : sc)
                            for (int v = 0; v < c->v; ++v)
                                for (int h = 0; h < c->h; ++h) block(*c, mx * c->h + h, my * c->v + v);
                        restart(++n, long(mcusX) * mcusY);
                    }
            }
            // continue after the entropy-coded data
            size_t p = br.position();
            while (p + 1 < b.size() && !(b[p] == 0xFF && b[p + 1] != 0x00 && !(b[p + 1] >= 0xD0 && b[p + 1] <= 0xD7))) ++p;
            pos = p;
            continue;
        }
        pos = end;
    }
    if (!frame) throw std::runtime_error("JPEG file has no frame");
    for (Component& c : comps) {
        const int stride = c.blocksX * 8;
        c.plane.assign(size_t(stride) * c.blocksY * 8, 0);
        for (int by = 0; by < c.blocksY; ++by)
            for (int bx = 0; bx < c.blocksX; ++bx)
                idct_islow(&c.coef[(size_t(by) * c.blocksX + bx) * 64], quant[c.tq],
                           &c.plane[size_t(by) * 8 * stride + size_t(bx) * 8], stride);
        c.coef.clear();
        c.coef.shrink_to_fit();
    }
    Image img;
    img.width = width;
    img.height = height;
    img.channels = int(comps.size());
    img.data.resize(size_t(width) * height * comps.size());
    std::vector<std::vector<uint8_t>> full;
    for (const Component& c : comps) full.push_back(upsample(c, hmax / c.h, vmax / c.v, width, height));
    const size_t np = size_t(width) * height;
    if (comps.size() == 1) {
        std::copy(full[0].begin(), full[0].end(), img.data.begin());
        return img;
    }
    // colour space as libjpeg infers it: JFIF -> YCbCr; Adobe transform 0 ->
    // RGB; component ids 'R','G','B' -> RGB; otherwise YCbCr
    bool ycc = true;
    if (!jfif) {
        if (adobe) ycc = adobeTransform != 0;
        else if (comps[0].id == 'R' && comps[1].id == 'G' && comps[2].id == 'B') ycc = false;
    }
    if (!ycc) {
        for (size_t p = 0; p < np; ++p)
            for (int k = 0; k < 3; ++k) img.data[3 * p + size_t(k)] = full[size_t(k)][p];
        return img;
    }
    // fixed-point YCbCr -> RGB, 16 fraction bits, tables of the chroma terms
    constexpr int kBits = 16;
    constexpr int64_t kHalf = int64_t(1) << (kBits - 1);
    const auto fix = [](double x) { return int64_t(x * double(int64_t(1) << kBits) + 0.5); };
    int crR[256], cbB[256];
    int64_t crG[256], cbG[256];
    for (int i = 0; i < 256; ++i) {
        const int64_t x = i - 128;
        crR[i] = int((fix(1.40200) * x + kHalf) >> kBits);
        cbB[i] = int((fix(1.77200) * x + kHalf) >> kBits);
        crG[i] = -fix(0.71414) * x;
        cbG[i] = -fix(0.34414) * x + kHalf;
    }
    for (size_t p = 0; p < np; ++p) {
        const int y = full[0][p], cb = full[1][p], cr = full[2][p];
        img.data[3 * p + 0] = uint8_t(std::clamp(y + crR[cr], 0, 255));
        img.data[3 * p + 1] = uint8_t(std::clamp(y + int((cbG[cb] + crG[cr]) >> kBits), 0, 255));
        img.data[3 * p + 2] = uint8_t(std::clamp(y + cbB[cb], 0, 255));
    }
    return img;
}

inline Image decodePng(const std::vector<uint8_t>& b) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (b.size() < 8 || !std::equal(sig, sig + 8, b.begin())) throw std::runtime_error("not a PNG file");
    const auto u32 = [&](size_t o) {
        if (o + 4 > b.size()) throw std::runtime_error("truncated PNG");
        return uint32_t(b[o]) << 24 | uint32_t(b[o + 1]) << 16 | uint32_t(b[o + 2]) << 8 | uint32_t(b[o + 3]);
    };
    uint32_t width = 0, height = 0;
    int depth = 0, type = -1, interlace = 0;
    std::vector<uint8_t> z;
    size_t pos = 8;
    while (pos + 8 <= b.size()) {
        const uint32_t len = u32(pos);
        const std::string kind(b.begin() + long(pos) + 4, b.begin() + long(pos) + 8);
        const size_t d = pos + 8;
        if (d + len + 4 > b.size()) throw std::runtime_error("truncated PNG chunk");
        if (kind == "IHDR") {
            if (len < 13) throw std::runtime_error("bad PNG header");
            width = u32(d);
            height = u32(d + 4);
            depth = b[d + 8];
            type = b[d + 9];
            interlace = b[d + 12];
        } else if (kind == "IDAT") {
            z.insert(z.end(), b.begin() + long(d), b.begin() + long(d + len));
        } else if (kind == "IEND") {
            break;
        }
        pos = d + len + 4;
    }
    const int samples = type == 0 ? 1 : type == 2 ? 3 : type == 3 ? 1 : type == 4 ? 2 : type == 6 ? 4 : 0;
    if (!samples || width < 1 || height < 1 || !(depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16))
        throw std::runtime_error("unsupported PNG image");
    if (interlace) throw std::runtime_error("interlaced PNG images are not read");
    if (size_t(width) * height > (size_t(1) << 28)) throw std::runtime_error("PNG image too large");
    const size_t bitsPerPixel = size_t(samples) * depth, rowBytes = (size_t(width) * bitsPerPixel + 7) / 8;
    const size_t bpp = std::max<size_t>(1, bitsPerPixel / 8);
    std::vector<uint8_t> raw((rowBytes + 1) * height);
    z_stream s{};
    if (inflateInit(&s) != Z_OK) throw std::runtime_error("zlib initialisation failed");
    s.next_in = z.data();
    s.avail_in = uInt(z.size());
    s.next_out = raw.data();
    s.avail_out = uInt(raw.size());
    const int rc = inflate(&s, Z_FINISH);
    inflateEnd(&s);
    if (rc != Z_STREAM_END || s.avail_out != 0) throw std::runtime_error("corrupt PNG image data");
    // undo the row filters (none, sub, up, average, Paeth)
    std::vector<uint8_t> prev(rowBytes, 0), cur(rowBytes);
    Image img;
    img.width = int(width);
    img.height = int(height);
    img.channels = std::min(samples, 3);
    img.data.resize(size_t(width) * height * size_t(img.channels));
    for (uint32_t y = 0; y < height; ++y) {
        const uint8_t* row = &raw[size_t(y) * (rowBytes + 1)];
        const int filter = row[0];
        for (size_t i = 0; i < rowBytes; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int pred = 0;
            switch (filter) {
            case 0: pred = 0; break;
            case 1: pred = a; break;
            case 2: pred = up; break;
            case 3: pred = (a + up) >> 1; break;
            case 4: {
                const int p = a + up - c, pa = std::abs(p - a), pb = std::abs(p - up), pc = std::abs(p - c);
                pred = (pa <= pb && pa <= pc) ? a : pb <= pc ? up : c;
                break;
            }
            default: throw std::runtime_error("corrupt PNG row filter");
            }
            cur[i] = uint8_t(row[1 + i] + pred);
        }
        for (uint32_t x = 0; x < width; ++x)
            for (int k = 0; k < img.channels; ++k) {
                int v;
                if (depth == 16) {
                    const size_t o = (size_t(x) * samples + size_t(k)) * 2;
                    v = std::min(255, int(cur[o]) << 8 | int(cur[o + 1]));
                } else if (depth == 8) {
                    v = cur[size_t(x) * samples + size_t(k)];
                } else {  // 1, 2, 4 bits: one sample per pixel (grey or palette index)
                    const size_t bit = size_t(x) * depth;
                    v = (cur[bit / 8] >> (8 - depth - int(bit % 8))) & ((1 << depth) - 1);
                }
                img.data[(size_t(y) * width + x) * size_t(img.channels) + size_t(k)] = uint8_t(v);
            }
        std::swap(prev, cur);
    }
    return img;
}

inline Image readImage(const std::filesystem::path& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open image: " + path.string());
    const std::vector<uint8_t> b((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (b.size() >= 2 && b[0] == 0xFF && b[1] == 0xD8) return decodeJpeg(b);
    if (b.size() >= 8 && b[0] == 0x89 && b[1] == 'P') return decodePng(b);
    throw std::runtime_error("unsupported image format (JPEG or PNG expected): " + path.string());
}

}  // namespace facetimg
