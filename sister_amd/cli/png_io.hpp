// png_io.hpp -- minimal PNG / PGM / PPM reading and writing for compute_disp
// (zlib inflate/deflate; non-interlaced PNG, 8- or 16-bit, gray, gray+alpha,
// RGB, RGBA or palette).  Color images come back in BGR order, as
// cv::imread returns them.
#pragma once
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sister/SisterMultiviewDisparities.hpp"

namespace pngio {

inline uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

inline std::vector<uint8_t> read_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), {});
}

inline uint8_t paeth(int a, int b, int c)
{
    int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    return (uint8_t)((pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c));
}

inline sister::Mat8 read_png(const std::vector<uint8_t>& f)
{
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8)) throw std::runtime_error("not a PNG");
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, pal;
    for (size_t p = 8; p + 8 <= f.size();) {
        uint32_t len = be32(&f[p]);
        std::string type((const char*)&f[p + 4], 4);
        if (p + 12 + len > f.size()) throw std::runtime_error("truncated PNG");
        const uint8_t* d = &f[p + 8];
        if (type == "IHDR") {
            W = be32(d); H = be32(d + 4); depth = d[8]; ctype = d[9]; interlace = d[12];
        } else if (type == "PLTE") {
            pal.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        p += 12 + len;
    }
    if (interlace) throw std::runtime_error("interlaced PNG not supported");
    if (depth != 8 && depth != 16) throw std::runtime_error("PNG bit depth must be 8 or 16");
    const int spp = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4;
    const int bpp = spp * depth / 8;
    const size_t stride = (size_t)W * bpp;
    std::vector<uint8_t> raw(H * (stride + 1));
    uLongf n = raw.size();
    if (uncompress(raw.data(), &n, idat.data(), idat.size()) != Z_OK || n != raw.size())
        throw std::runtime_error("corrupt PNG data");
    std::vector<uint8_t> img(H * stride);
    for (uint32_t y = 0; y < H; y++) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t* s = &raw[y * (stride + 1) + 1];
        uint8_t* o = &img[y * stride];
        const uint8_t* up = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t x = 0; x < stride; x++) {
            int a = x >= (size_t)bpp ? o[x - bpp] : 0, b = up ? up[x] : 0, c = (up && x >= (size_t)bpp) ? up[x - bpp] : 0;
            int v = s[x];
            switch (ft) {
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: v += paeth(a, b, c); break;
            default: break;
            }
            o[x] = (uint8_t)v;
        }
    }
    const int och = (ctype == 0 || ctype == 4) ? 1 : 3;
    sister::Mat8 out(H, W, och);
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const uint8_t* px = &img[y * stride + x * bpp];
            auto smp = [&](int k) { return px[k * depth / 8]; };   // high byte of 16-bit samples
            uint8_t* q = out.ptr(y) + x * och;
            if (ctype == 0 || ctype == 4) q[0] = smp(0);
            else if (ctype == 3) { int i = px[0]; q[0] = pal[3 * i + 2]; q[1] = pal[3 * i + 1]; q[2] = pal[3 * i]; }
            else { q[0] = smp(2); q[1] = smp(1); q[2] = smp(0); }   // RGB -> BGR
        }
    return out;
}

// binary PGM (P5) or PPM (P6), 8-bit
inline sister::Mat8 read_pnm(const std::vector<uint8_t>& f)
{
    size_t p = 2;
    auto next = [&]() {
        while (p < f.size()) {
            if (f[p] == '#') { while (p < f.size() && f[p] != '\n') p++; }
            else if (isspace(f[p])) p++;
            else break;
        }
        int v = 0;
        while (p < f.size() && isdigit(f[p])) v = v * 10 + (f[p++] - '0');
        return v;
    };
    const bool color = f[1] == '6';
    int W = next(), H = next(), mx = next();
    p++;
    if (mx > 255) throw std::runtime_error("16-bit PNM not supported");
    const int ch = color ? 3 : 1;
    if (p + (size_t)W * H * ch > f.size()) throw std::runtime_error("truncated PNM");
    sister::Mat8 out(H, W, ch);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const uint8_t* s = &f[p + ((size_t)y * W + x) * ch];
            uint8_t* q = out.ptr(y) + x * ch;
            if (color) { q[0] = s[2]; q[1] = s[1]; q[2] = s[0]; } else q[0] = s[0];
        }
    return out;
}

inline sister::Mat8 imread(const std::string& path)
{
    std::vector<uint8_t> f = read_file(path);
    if (f.size() >= 2 && f[0] == 'P' && (f[1] == '5' || f[1] == '6')) return read_pnm(f);
    return read_png(f);
}

inline void put32(std::vector<uint8_t>& v, uint32_t x)
{
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}

inline void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data)
{
    put32(out, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    put32(out, (uint32_t)crc32(0, td.data(), td.size()));
}

// PNG from filtered-free scanlines (filter byte 0 already in `raw`)
inline void write_png_raw(const std::string& path, const std::vector<uint8_t>& raw, int W, int H, int depth,
                          int color_type)
{
    uLongf n = compressBound(raw.size());
    std::vector<uint8_t> z(n);
    if (compress2(z.data(), &n, raw.data(), raw.size(), 6) != Z_OK) throw std::runtime_error("deflate failed");
    z.resize(n);
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ih;
    put32(ih, W);
    put32(ih, H);
    ih.push_back((uint8_t)depth); ih.push_back((uint8_t)color_type); ih.push_back(0); ih.push_back(0); ih.push_back(0);
    chunk(out, "IHDR", ih);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write((const char*)out.data(), out.size());
}

// gray PNG, 8 or 16 bits (samples given as 16-bit; 8-bit keeps the low byte)
inline void write_png(const std::string& path, const uint16_t* px, int W, int H, int depth)
{
    const int bps = depth / 8;
    std::vector<uint8_t> raw;
    raw.reserve((size_t)H * (W * bps + 1));
    for (int y = 0; y < H; y++) {
        raw.push_back(0);
        for (int x = 0; x < W; x++) {
            uint16_t v = px[(size_t)y * W + x];
            if (bps == 2) raw.push_back((uint8_t)(v >> 8));
            raw.push_back((uint8_t)v);
        }
    }
    write_png_raw(path, raw, W, H, depth, 0);
}

// 8-bit colour PNG of a BGR image (stored as RGB)
inline void write_png_bgr(const std::string& path, const sister::Mat8& img)
{
    if (img.channels != 3) throw std::runtime_error("write_png_bgr: expected 3 channels");
    std::vector<uint8_t> raw;
    raw.reserve((size_t)img.rows * (img.cols * 3 + 1));
    for (int y = 0; y < img.rows; y++) {
        raw.push_back(0);
        const uint8_t* r = img.ptr(y);
        for (int x = 0; x < img.cols; x++) {
            raw.push_back(r[3 * x + 2]);
            raw.push_back(r[3 * x + 1]);
            raw.push_back(r[3 * x]);
        }
    }
    write_png_raw(path, raw, img.cols, img.rows, 8, 2);
}

}  // namespace pngio
