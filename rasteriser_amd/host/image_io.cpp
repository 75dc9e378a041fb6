// image_io.cpp — 8-bit image files for the host layer: PNG (zlib inflate/deflate,
// written here) and binary PPM/PGM.  Covers what the reference reads and writes on
// its frame path: the material texture loaded by Material's constructor
// (material.h:20-23) and frame.png / depth.png (renderer.cpp:92-93).
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "rz_host.hpp"

namespace rz {
namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put_be32(std::string& s, uint32_t v) {
    s.push_back((char)(v >> 24)); s.push_back((char)(v >> 16)); s.push_back((char)(v >> 8)); s.push_back((char)v);
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

Image8 decode_png(const std::string& data, const std::string& path) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (data.size() < 8 || std::memcmp(data.data(), sig, 8) != 0) throw std::runtime_error(path + ": not a PNG");
    const uint8_t* p = (const uint8_t*)data.data();
    size_t pos = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::string idat;
    while (pos + 12 <= data.size()) {
        const uint32_t len = be32(p + pos);
        const char* type = (const char*)p + pos + 4;
        if (pos + 12 + (size_t)len > data.size()) throw std::runtime_error(path + ": truncated PNG chunk");
        const uint8_t* body = p + pos + 8;
        if (!std::memcmp(type, "IHDR", 4)) {
            W = be32(body); H = be32(body + 4);
            depth = body[8]; ctype = body[9]; interlace = body[12];
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.append((const char*)body, len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    int C = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (depth != 8 || C == 0 || interlace != 0 || W == 0 || H == 0)
        throw std::runtime_error(path + ": only 8-bit non-interlaced grey/RGB/RGBA PNG is supported");
    const size_t stride = (size_t)W * C;
    std::string raw((stride + 1) * H, '\0');
    uLongf rawlen = (uLongf)raw.size();
    if (uncompress((Bytef*)&raw[0], &rawlen, (const Bytef*)idat.data(), (uLong)idat.size()) != Z_OK ||
        rawlen != raw.size())
        throw std::runtime_error(path + ": corrupt PNG image data");
    Image8 img;
    img.width = (int)W; img.height = (int)H; img.channels = C;
    img.data.resize(stride * H);
    const uint8_t* src = (const uint8_t*)raw.data();
    for (uint32_t y = 0; y < H; ++y) {
        const int filter = src[y * (stride + 1)];
        const uint8_t* in = src + y * (stride + 1) + 1;
        uint8_t* out = img.data.data() + y * stride;
        const uint8_t* prev = y ? out - stride : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)C ? out[i - C] : 0, b = prev ? prev[i] : 0;
            const int c = (prev && i >= (size_t)C) ? prev[i - C] : 0;
            int v = in[i];
            switch (filter) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: throw std::runtime_error(path + ": bad PNG filter");
            }
            out[i] = (uint8_t)v;
        }
    }
    return img;
}

Image8 decode_pnm(const std::string& data, const std::string& path) {
    std::istringstream in(data);
    std::string magic;
    in >> magic;
    const int C = magic == "P6" ? 3 : magic == "P5" ? 1 : 0;
    if (!C) throw std::runtime_error(path + ": only binary P5/P6 PNM is supported");
    int W = 0, H = 0, maxv = 0;
    auto next = [&](int& v) {
        in >> std::ws;
        while (in.peek() == '#') { std::string line; std::getline(in, line); in >> std::ws; }
        in >> v;
    };
    next(W); next(H); next(maxv);
    in.get();
    if (W <= 0 || H <= 0 || maxv != 255) throw std::runtime_error(path + ": unsupported PNM header");
    Image8 img;
    img.width = W; img.height = H; img.channels = C;
    img.data.resize((size_t)W * H * C);
    in.read((char*)img.data.data(), (std::streamsize)img.data.size());
    if ((size_t)in.gcount() != img.data.size()) throw std::runtime_error(path + ": truncated PNM");
    return img;
}

}  // namespace

Image8 load_image(const std::string& path) {
    const std::string data = read_file(path);
    if (data.size() >= 8 && (uint8_t)data[0] == 137 && data[1] == 'P') return decode_png(data, path);
    if (data.size() >= 2 && data[0] == 'P') return decode_pnm(data, path);
    throw std::runtime_error(path + ": unknown image format (PNG or binary PNM expected)");
}

void save_png(const std::string& path, const uint8_t* interleaved, int W, int H, int C) {
    const size_t stride = (size_t)W * C;
    std::string raw;
    raw.reserve((stride + 1) * H);
    for (int y = 0; y < H; ++y) {
        raw.push_back('\0');
        raw.append((const char*)interleaved + y * stride, stride);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::string z(zlen, '\0');
    if (compress2((Bytef*)&z[0], &zlen, (const Bytef*)raw.data(), (uLong)raw.size(), 6) != Z_OK)
        throw std::runtime_error("deflate failed");
    z.resize(zlen);
    std::string out("\x89PNG\r\n\x1a\n", 8);
    auto chunk = [&](const char* type, const std::string& body) {
        put_be32(out, (uint32_t)body.size());
        std::string tb(type, 4);
        tb += body;
        out += tb;
        put_be32(out, (uint32_t)crc32(0, (const Bytef*)tb.data(), (uInt)tb.size()));
    };
    std::string ihdr;
    put_be32(ihdr, (uint32_t)W);
    put_be32(ihdr, (uint32_t)H);
    ihdr.push_back(8);
    ihdr.push_back((char)(C == 1 ? 0 : C == 3 ? 2 : C == 4 ? 6 : 0));
    ihdr.append(3, '\0');
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", "");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write(out.data(), (std::streamsize)out.size());
}

// frame.png: the planar [3][H][W] frame written interleaved (CImg::save writes the
// same pixels).
void save_frame_png(const std::string& path, const uint8_t* planar, int W, int H) {
    std::vector<uint8_t> rgb((size_t)W * H * 3);
    const size_t plane = (size_t)W * H;
    for (size_t i = 0; i < plane; ++i)
        for (int c = 0; c < 3; ++c) rgb[3 * i + c] = planar[i + c * plane];
    save_png(path, rgb.data(), W, H, 3);
}

// depth.png: depth_buffer.normalize(0,255) (renderer.cpp:93, CImg.h:26786-26794) in
// binary32, then the (unsigned char) conversion CImg's writer applies to a float image
// (truncation, CImg.h:52409-52410).
void save_depth_png(const std::string& path, const float* depth, int W, int H) {
    const size_t n = (size_t)W * H;
    float mn = depth[0], mx = depth[0];
    for (size_t i = 0; i < n; ++i) {
        if (depth[i] < mn) mn = depth[i];
        if (depth[i] > mx) mx = depth[i];
    }
    std::vector<uint8_t> g(n);
    for (size_t i = 0; i < n; ++i) {
        float v = 0.f;
        if (mn != mx) v = (depth[i] - mn) / (mx - mn) * (255.f - 0.f) + 0.f;
        g[i] = (uint8_t)(v <= 0.f ? 0 : v >= 255.f ? 255 : (int)v);
    }
    save_png(path, g.data(), W, H, 1);
}

}  // namespace rz

// C entry points of the image codec (used by the host layer's tests).
extern "C" int rzh_load_image(const char* path, int* W, int* H, int* C, uint8_t* out) {
    try {
        rz::Image8 img = rz::load_image(path);
        *W = img.width; *H = img.height; *C = img.channels;
        if (out) std::memcpy(out, img.data.data(), img.data.size());
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

extern "C" int rzh_save_png(const char* path, const uint8_t* data, int W, int H, int C) {
    try {
        rz::save_png(path, data, W, H, C);
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

extern "C" int rzh_save_depth_png(const char* path, const float* depth, int W, int H) {
    try {
        rz::save_depth_png(path, depth, W, H);
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}
