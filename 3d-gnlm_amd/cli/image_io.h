// image_io.h -- 8-bit image files for the nl-means command line (host code).
//
// The reference's drivers read their input with OpenCV imread(path, 0) and
// write 8-bit slices with imwrite (3DNLM/nlm.cpp:9-15,31;
// "2D NLM/11x11/nlm.cpp:8,33,43").  This header replaces that dependency
// with the two formats those drivers use in practice:
//   * BMP: uncompressed 8-bit palettised, 24-bit BGR or 32-bit BGRA, bottom-up
//     or top-down rows.  Colour pixels and palette entries become gray by the
//     ITU-R BT.601 luma weights in 14-bit fixed point, rounded half up:
//     gray = (4899 R + 9617 G + 1868 B + 8192) >> 14, the conversion OpenCV
//     applies when it decodes a BMP as grayscale (a gray palette maps to its
//     own indices, so the reference's 8-bit data files read unchanged).
//   * PGM: binary "P5" with maxval <= 255.
// Writing produces 8-bit BMPs with a gray palette (the layout of the
// reference's data files).
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace gnlm_cli {

struct Gray8 {
    int rows = 0, cols = 0;
    std::vector<unsigned char> px;  // row-major, rows[0] is the top row
};

inline unsigned char luma601(unsigned r, unsigned g, unsigned b)
{
    return (unsigned char)((4899u * r + 9617u * g + 1868u * b + 8192u) >> 14);
}

namespace detail {

inline bool read_file(const std::string &path, std::vector<unsigned char> &buf)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return false;
    }
    buf.resize((size_t)n);
    const bool ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
    std::fclose(f);
    return ok;
}

inline uint32_t le32(const unsigned char *p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }
inline uint16_t le16(const unsigned char *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
inline void put32(unsigned char *p, uint32_t v)
{
    p[0] = (unsigned char)v;
    p[1] = (unsigned char)(v >> 8);
    p[2] = (unsigned char)(v >> 16);
    p[3] = (unsigned char)(v >> 24);
}
inline void put16(unsigned char *p, uint16_t v)
{
    p[0] = (unsigned char)v;
    p[1] = (unsigned char)(v >> 8);
}

inline bool decode_bmp(const std::vector<unsigned char> &b, Gray8 &img, std::string &err)
{
    if (b.size() < 54 || b[0] != 'B' || b[1] != 'M') return err = "not a BMP file", false;
    const uint32_t off = le32(&b[10]), hsz = le32(&b[14]);
    if (hsz < 40) return err = "unsupported BMP header (OS/2 core header)", false;
    const int32_t w = (int32_t)le32(&b[18]), h = (int32_t)le32(&b[22]);
    const int bpp = le16(&b[28]);
    const uint32_t comp = le32(&b[30]);
    uint32_t ncolors = le32(&b[46]);
    if (comp != 0) return err = "compressed BMPs are not supported", false;
    if (w <= 0 || h == 0 || h == INT32_MIN) return err = "bad BMP dimensions", false;
    if (bpp != 8 && bpp != 24 && bpp != 32) return err = "BMP must be 8, 24 or 32 bits per pixel", false;
    const int rows = h < 0 ? -h : h, cols = w;
    const bool bottom_up = h > 0;
    const size_t stride = ((size_t)cols * (bpp / 8) + 3) / 4 * 4;
    if ((size_t)off + stride * (size_t)rows > b.size()) return err = "truncated BMP pixel data", false;
    unsigned char lut[256];
    if (bpp == 8) {
        if (ncolors == 0 || ncolors > 256) ncolors = 256;
        const size_t pal = 14 + (size_t)hsz;
        if (pal + 4 * (size_t)ncolors > off) return err = "bad BMP palette", false;
        for (int i = 0; i < 256; ++i) lut[i] = 0;
        for (uint32_t i = 0; i < ncolors; ++i) {
            const unsigned char *q = &b[pal + 4 * i];  // B, G, R, reserved
            lut[i] = luma601(q[2], q[1], q[0]);
        }
    }
    img.rows = rows;
    img.cols = cols;
    img.px.assign((size_t)rows * cols, 0);
    for (int r = 0; r < rows; ++r) {
        const unsigned char *src = &b[off + stride * (size_t)(bottom_up ? rows - 1 - r : r)];
        unsigned char *dst = &img.px[(size_t)r * cols];
        for (int c = 0; c < cols; ++c) {
            if (bpp == 8) dst[c] = lut[src[c]];
            else {
                const unsigned char *q = src + (size_t)c * (bpp / 8);
                dst[c] = luma601(q[2], q[1], q[0]);
            }
        }
    }
    return true;
}

inline bool decode_pgm(const std::vector<unsigned char> &b, Gray8 &img, std::string &err)
{
    size_t i = 2;
    int vals[3];
    for (int k = 0; k < 3; ++k) {
        for (;;) {  // skip white space and comments
            while (i < b.size() && (b[i] == ' ' || b[i] == '\t' || b[i] == '\r' || b[i] == '\n')) ++i;
            if (i < b.size() && b[i] == '#') {
                while (i < b.size() && b[i] != '\n') ++i;
                continue;
            }
            break;
        }
        long v = 0;
        size_t j = i;
        while (j < b.size() && b[j] >= '0' && b[j] <= '9' && v < (1L << 24)) v = v * 10 + (b[j++] - '0');
        if (j == i) return err = "bad PGM header", false;
        vals[k] = (int)v;
        i = j;
    }
    ++i;  // the single white-space byte before the raster
    const int cols = vals[0], rows = vals[1], maxval = vals[2];
    if (cols < 1 || rows < 1 || maxval < 1 || maxval > 255) return err = "PGM must be 8-bit (maxval <= 255)", false;
    if (i + (size_t)rows * cols > b.size()) return err = "truncated PGM raster", false;
    img.rows = rows;
    img.cols = cols;
    img.px.assign(b.begin() + (long)i, b.begin() + (long)(i + (size_t)rows * cols));
    return true;
}

}  // namespace detail

// Reads an 8-bit grayscale image (BMP or binary PGM, by content).
inline bool read_gray8(const std::string &path, Gray8 &img, std::string &err)
{
    std::vector<unsigned char> b;
    if (!detail::read_file(path, b)) return err = "cannot read " + path, false;
    if (b.size() >= 2 && b[0] == 'B' && b[1] == 'M') return detail::decode_bmp(b, img, err);
    if (b.size() >= 2 && b[0] == 'P' && b[1] == '5') return detail::decode_pgm(b, img, err);
    return err = path + ": not a BMP or binary PGM image", false;
}

// Writes rows x cols 8-bit pixels (row-major, top row first) as an 8-bit
// bottom-up BMP with a gray palette.
inline bool write_bmp_gray8(const std::string &path, const unsigned char *px, int rows, int cols)
{
    const uint32_t stride = ((uint32_t)cols + 3) / 4 * 4;
    const uint32_t off = 14 + 40 + 1024, size = off + stride * (uint32_t)rows;
    std::vector<unsigned char> b(size, 0);
    b[0] = 'B';
    b[1] = 'M';
    detail::put32(&b[2], size);
    detail::put32(&b[10], off);
    detail::put32(&b[14], 40);
    detail::put32(&b[18], (uint32_t)cols);
    detail::put32(&b[22], (uint32_t)rows);
    detail::put16(&b[26], 1);
    detail::put16(&b[28], 8);
    detail::put32(&b[34], stride * (uint32_t)rows);
    detail::put32(&b[46], 256);
    for (int i = 0; i < 256; ++i) {
        unsigned char *q = &b[54 + 4 * i];
        q[0] = q[1] = q[2] = (unsigned char)i;
    }
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) b[off + stride * (uint32_t)(rows - 1 - r) + (uint32_t)c] = px[(size_t)r * cols + c];
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace gnlm_cli
