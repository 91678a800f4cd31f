#include "images.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace mc3d {

namespace {

// ParaView's blue-white-red map in 33 steps, as the reference tabulates it
// (src/PngWriter.hpp:284-300): red, green, blue in [0, 1].
const double kBwrR[33] = {0.07514311,  0.247872569, 0.339526309, 0.409505078, 0.468487184, 0.520796675, 0.568724526,
                          0.613686735, 0.656658579, 0.698372844, 0.739424025, 0.780330104, 0.821573924, 0.863634967,
                          0.907017747, 0.936129275, 0.943467973, 0.990146732, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                          0.961891484, 0.916482116};
const double kBwrG[33] = {0.468049805, 0.498782363, 0.528909511, 0.558608486, 0.588057293, 0.617435078, 0.646924167,
                          0.676713218, 0.707001303, 0.738002964, 0.769954435, 0.803121429, 0.837809045, 0.874374691,
                          0.913245283, 0.938743558, 0.943498599, 0.928791426, 0.88332677,  0.833985467, 0.788626485,
                          0.746206642, 0.70590052,  0.667019783, 0.6289553,   0.591130233, 0.552955184, 0.513776083,
                          0.472800903, 0.428977855, 0.380759558, 0.313155629, 0.236630659};
const double kBwrB[33] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0.983038586, 0.943398095, 0.917447482,
                          0.861943246, 0.803839606, 0.750707739, 0.701389973, 0.654994046, 0.610806959, 0.568237474,
                          0.526775617, 0.485962266, 0.445364274, 0.404551679, 0.363073592, 0.320428137, 0.265499262,
                          0.209939162};

// piecewise-linear map over equal chunks of width w, truncated to int
void linear_map(const int *r, const int *g, const int *b, double w, double f, int &R, int &G, int &B) {
    const int c = (int)std::floor(f / w);
    const double t = (f - w * c) / w;
    R = (int)((1.0 - t) * r[c] + t * r[c + 1]);
    G = (int)((1.0 - t) * g[c] + t * g[c + 1]);
    B = (int)((1.0 - t) * b[c] + t * b[c + 1]);
}

void put_u32(std::vector<unsigned char> &v, unsigned x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((unsigned char)(x >> s));
}

void chunk(std::vector<unsigned char> &out, const char *type, const unsigned char *data, size_t n) {
    put_u32(out, (unsigned)n);
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + n);
    put_u32(out, (unsigned)crc32(0L, out.data() + at, (uInt)(n + 4)));
}

}  // namespace

void colormap_rgb(int cm, double f, int &r, int &g, int &b) {
    if (cm == RAINBOW) {
        // blue, cyan, green, yellow, red (src/PngWriter.hpp:270-281)
        static const int R[5] = {0, 0, 0, 255, 255}, G[5] = {0, 255, 255, 255, 0}, B[5] = {255, 255, 0, 0, 0};
        linear_map(R, G, B, 0.2500001, f, r, g, b);
    } else if (cm == BWR) {
        static int R[33], G[33], B[33];
        static const bool init = [] {
            for (int i = 0; i < 33; ++i) {
                R[i] = (int)(255.0 * kBwrR[i]);
                G[i] = (int)(255.0 * kBwrG[i]);
                B[i] = (int)(255.0 * kBwrB[i]);
            }
            return true;
        }();
        (void)init;
        linear_map(R, G, B, 0.031250001, f, r, g, b);
    } else {
        r = g = b = (int)(f * 255.0);
    }
}

std::vector<unsigned char> render_plane(const double *plane, int n1, int n2, const ImageOutput &img) {
    const int W = img.width, H = img.height;
    const double d = std::max((n1 - 1) / (W + 1.0), (n2 - 1) / (H + 1.0));
    std::vector<long> src((size_t)W * H, -1);
    for (int jp = 0; jp < H; ++jp)
        for (int ip = 0; ip < W; ++ip) {
            const double s1 = d / 2 + d * ip, s2 = d / 2 + d * jp;
            const long i1 = (long)std::floor(s1 + 0.5), i2 = (long)std::floor(s2 + 0.5);
            if (i1 <= n1 - 1 && i2 <= n2 - 1) src[(size_t)jp * W + ip] = i2 * n1 + i1;
        }
    // colour range: fixed, or the extremes of the sampled values
    // (src/CurvilinearCSolver_Core.cpp:1686-1703)
    double lo = 100000000.0, hi = -100000000.0;
    if (img.has_range) {
        lo = img.vmin;
        hi = img.vmax;
    } else {
        for (long s : src)
            if (s >= 0) {
                lo = std::fmin(lo, plane[s]);
                hi = std::fmax(hi, plane[s]);
            }
    }
    std::vector<unsigned char> rgb((size_t)W * H * 3);
    for (int jp = 0; jp < H; ++jp)
        for (int ip = 0; ip < W; ++ip) {
            const long s = src[(size_t)jp * W + ip];
            int r = 73, g = 175, b = 205;
            if (s >= 0) {
                double f = (plane[s] - lo) / (hi - lo);
                if (hi - lo < 1e-6) f = 0.0;
                f = std::min(1.0, std::max(0.0, f));
                colormap_rgb(img.colormap, f, r, g, b);
            }
            // PNG rows run from the top: jp = 0 is the bottom row
            unsigned char *px = &rgb[((size_t)(H - 1 - jp) * W + ip) * 3];
            px[0] = (unsigned char)r;
            px[1] = (unsigned char)g;
            px[2] = (unsigned char)b;
        }
    return rgb;
}

bool write_png(const std::string &path, const unsigned char *rgb, int width, int height) {
    std::vector<unsigned char> raw;
    raw.reserve((size_t)height * (3 * (size_t)width + 1));
    for (int y = 0; y < height; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), rgb + (size_t)y * width * 3, rgb + (size_t)(y + 1) * width * 3);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) return false;
    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    put_u32(ihdr, (unsigned)width);
    put_u32(ihdr, (unsigned)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, filter 0, no interlace
    chunk(out, "IHDR", ihdr.data(), ihdr.size());
    chunk(out, "IDAT", z.data(), zn);
    chunk(out, "IEND", nullptr, 0);
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(out.data(), 1, out.size(), fp) == out.size();
    return std::fclose(fp) == 0 && ok;
}

}  // namespace mc3d
