// Image.h - the reference's Image (Utils/Image.h:22-95, Utils/Image.cpp), header
// only, without tinyexr or stb: planar float channels (mData[c][x + y * width]),
// pixel access, and the file formats the reference reads and writes:
//   writeEXR  single-part scanline OpenEXR, no compression, HALF channels in
//             B, G, R order from channels 2, 1, 0 (Image.cpp:49-106);
//   readEXR   scanline OpenEXR with NONE, ZIPS or ZIP compression and HALF,
//             FLOAT or UINT channels; R, G, B (by the first letter of the
//             channel name) go to channels 0-2 and the image keeps the file's
//             channel count, extra channels zero (Image.cpp:108-180);
//   readMisc  8-bit PNG (grey, grey + alpha, RGB, RGBA; not interlaced),
//             rows flipped bottom-up as the reference loads them, /255 and,
//             with toLinear, ^2.2 (Image.cpp:219-259);
//   writePNG  8-bit PNG of every channel, with gammaCorrect v^(1/2.2) and a
//             clamp to [0, 1] for linear images (Image.cpp:182-217);
//   load      .exr through readEXR, anything else through readMisc.
// Link with -lz.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../crayhip.h"

namespace CRay {

class RuntimeError : public std::runtime_error {
   public:
    RuntimeError(int code, const std::string& what)
        : std::runtime_error(what), mCode(code) {}
    int code() const { return mCode; }

   private:
    int mCode;
};

enum class ColorSpace { Linear, Gamma2_2 };

namespace imgdetail {

[[noreturn]] inline void fail(const std::string& m) { throw RuntimeError(CRH_ERR_INVALID_ARG, m); }

inline std::string readFile(const std::filesystem::path& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) fail("cannot open " + p.string());
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}
inline void writeFile(const std::filesystem::path& p, const std::string& b) {
    std::ofstream f(p, std::ios::binary);
    f.write(b.data(), std::streamsize(b.size()));
    if (!f) fail("cannot write " + p.string());
}

// float -> IEEE half, round to nearest even (NaN stays NaN).
inline uint16_t floatToHalf(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t exp = (x >> 23) & 0xFFu;
    uint32_t mant = x & 0x007FFFFFu;
    if (exp == 0xFFu) {
        uint32_t m = mant >> 13;
        return uint16_t(sign | 0x7C00u | (mant ? (m ? m : 1u) : 0u));
    }
    const int e = int(exp) - 127 + 15;
    if (e >= 0x1F) return uint16_t(sign | 0x7C00u);
    if (e <= 0) {  // subnormal half (or zero)
        if (exp == 0) return uint16_t(sign);
        const uint32_t m = mant | 0x00800000u;
        const uint32_t s = uint32_t(126 - int(exp));  // 14 .. 126
        if (s > 25) return uint16_t(sign);
        uint32_t h = m >> s;
        const uint32_t r = m & ((1u << s) - 1u), half = 1u << (s - 1);
        if (r > half || (r == half && (h & 1u))) ++h;
        return uint16_t(sign | h);
    }
    uint32_t h = sign | (uint32_t(e) << 10) | (mant >> 13);
    const uint32_t r = mant & 0x1FFFu;
    if (r > 0x1000u || (r == 0x1000u && (h & 1u))) ++h;
    return uint16_t(h);
}

inline float halfToFloat(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1Fu, mant = h & 0x3FFu, bits;
    if (exp == 0x1Fu) {
        bits = sign | 0x7F800000u | (mant << 13);
    } else if (exp == 0) {
        if (mant == 0) {
            bits = sign;
        } else {  // subnormal: normalise
            int e = -1;
            do {
                ++e;
                mant <<= 1;
            } while (!(mant & 0x400u));
            bits = sign | (uint32_t(127 - 15 - e) << 23) | ((mant & 0x3FFu) << 13);
        }
    } else {
        bits = sign | ((exp + 112u) << 23) | (mant << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

inline void put32(std::string& b, int32_t v) { b.append(reinterpret_cast<const char*>(&v), 4); }
inline void putF(std::string& b, float v) { b.append(reinterpret_cast<const char*>(&v), 4); }
inline void putAttr(std::string& b, const char* name, const char* type, const std::string& v) {
    b += name;
    b += '\0';
    b += type;
    b += '\0';
    put32(b, int32_t(v.size()));
    b += v;
}

struct Reader {
    const std::string& s;
    size_t p = 0;
    void need(size_t n) const {
        if (p + n > s.size()) fail("truncated image file");
    }
    template <class T>
    T get() {
        need(sizeof(T));
        T v;
        std::memcpy(&v, s.data() + p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    std::string cstr() {
        size_t e = s.find('\0', p);
        if (e == std::string::npos) fail("truncated image file");
        std::string r = s.substr(p, e - p);
        p = e + 1;
        return r;
    }
};

inline std::string inflateAll(const char* src, size_t n, size_t expect) {
    std::string out(expect, '\0');
    uLongf len = uLongf(expect);
    if (uncompress(reinterpret_cast<Bytef*>(&out[0]), &len, reinterpret_cast<const Bytef*>(src),
                   uLong(n)) != Z_OK || len != expect)
        fail("corrupt compressed image data");
    return out;
}

inline uint32_t be32(const unsigned char* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3];
}
inline void putBe32(std::string& b, uint32_t v) {
    const char c[4] = {char(v >> 24), char(v >> 16), char(v >> 8), char(v)};
    b.append(c, 4);
}

}  // namespace imgdetail

class Image {
   public:
    using Ref = std::shared_ptr<Image>;
    Image(size_t width, size_t height, int channels, ColorSpace colorSpace = ColorSpace::Linear)
        : mWidth(int(width)),
          mHeight(int(height)),
          mChannels(channels),
          mColorSpace(colorSpace),
          mData(size_t(channels), std::vector<float>(width * height, 0.f)) {}

    void resize(size_t width, size_t height) {
        mWidth = int(width);
        mHeight = int(height);
        for (auto& layer : mData) layer.resize(width * height, 0.f);
    }
    // The channel count; kept channels keep their data, new ones are 0.
    void resizeChannels(int channels) {
        mChannels = channels;
        mData.resize(size_t(channels), std::vector<float>(size_t(mWidth) * mHeight, 0.f));
    }

    int getWidth() const { return mWidth; }
    int getHeight() const { return mHeight; }
    int getChannels() const { return mChannels; }
    int getArea() const { return mWidth * mHeight; }
    ColorSpace getColorSpace() const { return mColorSpace; }
    const std::vector<std::vector<float>>& getRawData() const { return mData; }

    float getPixel(int x, int y, int c) const { return mData[c][size_t(x) + size_t(y) * mWidth]; }
    float getPixel(int index, int c) const { return mData[c][index]; }
    float& getPixel(int x, int y, int c) { return mData[c][size_t(x) + size_t(y) * mWidth]; }
    float& getPixel(int index, int c) { return mData[c][index]; }
    void setPixel(int x, int y, int c, float v) { setPixel(x + y * mWidth, c, v); }
    void setPixel(int index, int c, float v) {
        if (index < 0 || index >= getArea() || c < 0 || c >= mChannels)
            imgdetail::fail("pixel index out of range");
        mData[c][index] = v;
    }

    void writeEXR(const std::filesystem::path& filename) const {
        using namespace imgdetail;
        if (filename.extension() != ".exr") fail("Output file must be .exr");
        if (mChannels < 3) fail("writeEXR needs at least 3 channels");
        const char* names[3] = {"B", "G", "R"};
        const int src[3] = {2, 1, 0};
        std::string chl;
        for (const char* n : names) {
            chl += n;
            chl += '\0';
            put32(chl, 1);                // HALF
            chl += std::string(4, '\0');  // pLinear + reserved
            put32(chl, 1);
            put32(chl, 1);
        }
        chl += '\0';
        std::string box;
        put32(box, 0);
        put32(box, 0);
        put32(box, mWidth - 1);
        put32(box, mHeight - 1);
        std::string f1, v2;
        putF(f1, 1.0f);
        putF(v2, 0.0f);
        putF(v2, 0.0f);
        std::string out;
        put32(out, 20000630);
        put32(out, 2);
        putAttr(out, "channels", "chlist", chl);
        putAttr(out, "compression", "compression", std::string(1, '\0'));
        putAttr(out, "dataWindow", "box2i", box);
        putAttr(out, "displayWindow", "box2i", box);
        putAttr(out, "lineOrder", "lineOrder", std::string(1, '\0'));
        putAttr(out, "pixelAspectRatio", "float", f1);
        putAttr(out, "screenWindowCenter", "v2f", v2);
        putAttr(out, "screenWindowWidth", "float", f1);
        out += '\0';
        const uint64_t lineBytes = uint64_t(mWidth) * 2 * 3;
        const uint64_t first = out.size() + 8 * uint64_t(mHeight);
        for (int y = 0; y < mHeight; ++y) {
            uint64_t off = first + uint64_t(y) * (8 + lineBytes);
            out.append(reinterpret_cast<const char*>(&off), 8);
        }
        for (int y = 0; y < mHeight; ++y) {
            put32(out, y);
            put32(out, int32_t(lineBytes));
            for (int c = 0; c < 3; ++c)
                for (int x = 0; x < mWidth; ++x) {
                    uint16_t h = floatToHalf(getPixel(x, y, src[c]));
                    out.append(reinterpret_cast<const char*>(&h), 2);
                }
        }
        writeFile(filename, out);
    }

    void readEXR(const std::filesystem::path& filename) {
        using namespace imgdetail;
        if (filename.extension() != ".exr") fail("Output file must be .exr");
        const std::string s = readFile(filename);
        Reader r{s};
        if (r.get<int32_t>() != 20000630) fail("not an OpenEXR file: " + filename.string());
        const int32_t version = r.get<int32_t>();
        if ((version & 0xFF) != 2 || (version & 0x1E00))  // tiled / deep / multi-part
            fail("only single-part scanline OpenEXR files are read");
        struct Ch {
            std::string name;
            int32_t type;
        };
        std::vector<Ch> chans;
        int32_t comp = -1, x0 = 0, y0 = 0, x1 = -1, y1 = -1;
        for (;;) {
            std::string name = r.cstr();
            if (name.empty()) break;
            std::string type = r.cstr();
            const int32_t size = r.get<int32_t>();
            r.need(size_t(size));
            const size_t end = r.p + size_t(size);
            if (name == "channels") {
                for (;;) {
                    std::string cn = r.cstr();
                    if (cn.empty()) break;
                    const int32_t t = r.get<int32_t>();
                    r.p += 12;  // pLinear, reserved, xSampling, ySampling
                    chans.push_back({cn, t});
                }
            } else if (name == "compression") {
                comp = uint8_t(s[r.p]);
            } else if (name == "dataWindow") {
                x0 = r.get<int32_t>();
                y0 = r.get<int32_t>();
                x1 = r.get<int32_t>();
                y1 = r.get<int32_t>();
            }
            r.p = end;
        }
        if (chans.empty() || x1 < x0 || y1 < y0) fail("EXR header without channels or window");
        if (comp != 0 && comp != 2 && comp != 3)
            fail("EXR compression " + std::to_string(comp) + " is not read (NONE, ZIPS, ZIP)");
        const int W = x1 - x0 + 1, H = y1 - y0 + 1;
        const int lines = comp == 3 ? 16 : 1;
        size_t pixBytes = 0;
        for (const Ch& c : chans) pixBytes += c.type == 1 ? 2 : 4;
        const size_t lineBytes = pixBytes * size_t(W);
        const int nChunks = (H + lines - 1) / lines;
        std::vector<uint64_t> offsets(static_cast<size_t>(nChunks));
        for (auto& o : offsets) o = r.get<uint64_t>();
        int rgb[3] = {-1, -1, -1};
        for (int c = 0; c < 3; ++c)
            for (size_t j = 0; j < chans.size(); ++j)
                if (!chans[j].name.empty() && chans[j].name[0] == "RGB"[c]) {
                    rgb[c] = int(j);
                    break;
                }
        if (rgb[0] < 0 || rgb[1] < 0 || rgb[2] < 0) fail("EXR file without R, G and B channels");
        resize(size_t(W), size_t(H));
        resizeChannels(int(chans.size()));
        for (auto& layer : mData) std::fill(layer.begin(), layer.end(), 0.f);
        for (int k = 0; k < nChunks; ++k) {
            Reader cr{s, size_t(offsets[size_t(k)])};
            const int32_t y = cr.get<int32_t>() - y0;
            const int32_t size = cr.get<int32_t>();
            cr.need(size_t(size));
            const int n = std::min(lines, H - y);
            if (y < 0 || n <= 0) fail("EXR scanline out of the data window");
            const size_t raw = lineBytes * size_t(n);
            std::string block;
            if (comp == 0 || size_t(size) == raw) {
                block.assign(s.data() + cr.p, raw);
            } else {  // zlib, then the predictor and the byte split undone
                std::string t = inflateAll(s.data() + cr.p, size_t(size), raw);
                for (size_t i = 1; i < t.size(); ++i)
                    t[i] = char(uint8_t(t[i - 1]) + uint8_t(t[i]) - 128);
                block.resize(raw);
                const size_t half = (raw + 1) / 2;
                for (size_t i = 0; i < raw; ++i)
                    block[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
            }
            for (int l = 0; l < n; ++l) {
                const char* line = block.data() + size_t(l) * lineBytes;
                size_t off = 0;
                for (size_t j = 0; j < chans.size(); ++j) {
                    const int bytes = chans[j].type == 1 ? 2 : 4;
                    for (int x = 0; x < W; ++x) {
                        const char* q = line + off + size_t(x) * bytes;
                        float v;
                        if (chans[j].type == 1) {
                            uint16_t h;
                            std::memcpy(&h, q, 2);
                            v = halfToFloat(h);
                        } else if (chans[j].type == 2) {
                            std::memcpy(&v, q, 4);
                        } else {
                            uint32_t u;
                            std::memcpy(&u, q, 4);
                            v = float(u);
                        }
                        for (int c = 0; c < 3; ++c)
                            if (rgb[c] == int(j)) getPixel(x, y + l, c) = v;
                    }
                    off += size_t(bytes) * W;
                }
            }
        }
        mColorSpace = ColorSpace::Linear;
    }

    void writePNG(const std::filesystem::path& filename, bool gammaCorrect) const {
        using namespace imgdetail;
        if (filename.extension() != ".png") fail("Output file must be .png");
        if (mChannels < 1 || mChannels > 4) fail("writePNG needs 1 to 4 channels");
        static const int colorType[5] = {0, 0, 4, 2, 6};
        std::string raw;
        raw.reserve(size_t(getArea()) * mChannels + mHeight);
        for (int y = 0; y < mHeight; ++y) {
            raw += '\0';  // filter: none
            for (int x = 0; x < mWidth; ++x)
                for (int c = 0; c < mChannels; ++c) {
                    float v = getPixel(x, y, c);
                    if (gammaCorrect) {
                        if (mColorSpace == ColorSpace::Linear) v = std::pow(v, 1 / 2.2f);
                        v = std::min(std::max(v, 0.f), 1.f);
                    }
                    raw += char(static_cast<unsigned char>(v * 255.f));
                }
        }
        uLongf zlen = compressBound(uLong(raw.size()));
        std::string z(zlen, '\0');
        if (compress2(reinterpret_cast<Bytef*>(&z[0]), &zlen,
                      reinterpret_cast<const Bytef*>(raw.data()), uLong(raw.size()), 6) != Z_OK)
            fail("PNG compression failed");
        z.resize(zlen);
        auto chunk = [](std::string& out, const char* type, const std::string& data) {
            putBe32(out, uint32_t(data.size()));
            std::string td = std::string(type, 4) + data;
            out += td;
            putBe32(out, uint32_t(crc32(0, reinterpret_cast<const Bytef*>(td.data()),
                                        uInt(td.size()))));
        };
        std::string ihdr;
        putBe32(ihdr, uint32_t(mWidth));
        putBe32(ihdr, uint32_t(mHeight));
        ihdr += char(8);
        ihdr += char(colorType[mChannels]);
        ihdr += std::string(3, '\0');
        std::string out("\x89PNG\r\n\x1a\n", 8);
        chunk(out, "IHDR", ihdr);
        chunk(out, "IDAT", z);
        chunk(out, "IEND", std::string());
        writeFile(filename, out);
    }

    void readMisc(const std::filesystem::path& filename, bool toLinear) {
        using namespace imgdetail;
        const std::string s = readFile(filename);
        if (s.size() < 8 || s.compare(0, 8, std::string("\x89PNG\r\n\x1a\n", 8)) != 0)
            fail("unsupported image format (PNG and EXR are read): " + filename.string());
        const auto* u = reinterpret_cast<const unsigned char*>(s.data());
        size_t p = 8;
        uint32_t W = 0, H = 0;
        int depth = 0, ctype = -1, interlace = 0;
        std::string idat;
        while (p + 12 <= s.size()) {
            const uint32_t len = be32(u + p);
            if (p + 12 + size_t(len) > s.size()) fail("truncated PNG file");
            const std::string type = s.substr(p + 4, 4);
            const unsigned char* d = u + p + 8;
            if (type == "IHDR") {
                W = be32(d);
                H = be32(d + 4);
                depth = d[8];
                ctype = d[9];
                interlace = d[12];
            } else if (type == "IDAT") {
                idat.append(reinterpret_cast<const char*>(d), len);
            } else if (type == "IEND") {
                break;
            }
            p += 12 + size_t(len);
        }
        int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
        if (depth != 8 || ch == 0 || interlace != 0)
            fail("PNG must be 8-bit grey, grey + alpha, RGB or RGBA, not interlaced");
        if (ch == 2) fail("Unsupported color type: 2");  // the reference takes 1, 3 or 4
        const size_t stride = size_t(W) * ch;
        std::string img = inflateAll(idat.data(), idat.size(), (stride + 1) * H);
        std::vector<unsigned char> prev(stride, 0), cur(stride);
        std::vector<unsigned char> pix(stride * H);
        for (uint32_t y = 0; y < H; ++y) {
            const auto* row = reinterpret_cast<const unsigned char*>(img.data()) + y * (stride + 1);
            const int f = row[0];
            for (size_t i = 0; i < stride; ++i) {
                const int a = i >= size_t(ch) ? cur[i - ch] : 0, b = prev[i],
                          c = i >= size_t(ch) ? prev[i - ch] : 0;
                int v = row[1 + i];
                if (f == 1) v += a;
                else if (f == 2) v += b;
                else if (f == 3) v += (a + b) / 2;
                else if (f == 4) {
                    const int pa = std::abs(b - c), pb = std::abs(a - c),
                              pc = std::abs(a + b - 2 * c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                } else if (f != 0) {
                    fail("corrupt PNG filter");
                }
                cur[i] = static_cast<unsigned char>(v);
            }
            std::copy(cur.begin(), cur.end(), pix.begin() + (H - 1 - y) * stride);  // bottom-up
            std::swap(prev, cur);
        }
        resizeChannels(ch);
        resize(W, H);
        for (int c = 0; c < ch; ++c)
            for (size_t i = 0; i < size_t(W) * H; ++i) {
                float v = float(pix[i * ch + c]) / 255.f;
                if (toLinear) v = float(std::pow(double(v), 2.2));  // pow<float>(v, 2.2)
                mData[c][i] = v;
            }
        mColorSpace = ColorSpace::Linear;
    }

    static Image load(const std::filesystem::path& filename, bool toLinear = true) {
        Image image(0, 0, 3);
        if (filename.extension() == ".exr")
            image.readEXR(filename);
        else
            image.readMisc(filename, toLinear);
        return image;
    }

   private:
    int mWidth;
    int mHeight;
    int mChannels;
    ColorSpace mColorSpace = ColorSpace::Linear;
    std::vector<std::vector<float>> mData;  // mData[channel][x + y * width]
};

}  // namespace CRay
