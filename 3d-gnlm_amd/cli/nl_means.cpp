// nl_means.cpp -- the nl-means command line, on the MI355X library.
//
// 3D mode mirrors the reference's driver 3DNLM/nlm.cpp:27-101:
//   read an 8-bit image, reflect-101 pad it by (SS+NS-2)/2 (:33-40), stack
//   depth + SS + NS - 2 copies of the padded slice (:41,53-59), run the 3D
//   filter through the drop-in's host-pointer entry point (:63), truncate
//   the float result to 8 bits (:66-70), write output.bin = int depth,
//   int rows, int cols, then the bytes in [slice][row][col] order (:17-24,
//   72-73), and optionally each slice as i_<k>_result.bmp (:9-15,75-76).
//   SS, NS and depth come from flags whose defaults are the reference's
//   config.h (SS 13, NS 17, depth 128; 3DNLM/config.h:4-9); h defaults to
//   PARAMETER = 10 on the 0-255 scale (3DNLM/main.cu:13).
// 2D mode mirrors "2D NLM/11x11/nlm.cpp:6-60": image, search size and
//   neighbourhood size as positional arguments; the padded input and the
//   output go to BMP files (the output rounded and clamped to 8 bits, as an
//   8-bit image writer stores a float image).
// Differences from the reference, by design: images need not be square (the
// reference indexes its padded copy with the row and column counts swapped,
// which is only right for square images), and the slice folder is written
// only when asked for.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "gnlm.h"
#include "image_io.h"

namespace {

int reflect101(int i, int n)
{
    if (n == 1) return 0;
    const int p = 2 * (n - 1);
    int m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - m;
}

void usage(const char *argv0)
{
    std::fprintf(stderr,
                 "usage:\n"
                 "  %s <image> [--ss 13] [--ns 17] [--depth 128] [--h 10] [--out output.bin] [--slices DIR]\n"
                 "      3D: the image stacked depth times, as 3DNLM's nl-means\n"
                 "  %s --2d <image> <search_size> <neighbor_size> [--h 10] [--out output.bmp]\n"
                 "      [--padded input.bmp]\n"
                 "      2D, as \"2D NLM/11x11\"'s driver\n"
                 "images: 8-bit BMP (palettised, 24- or 32-bit colour is converted to gray) or binary PGM\n",
                 argv0, argv0);
}

bool parse_int(const char *s, int &v)
{
    char *e = nullptr;
    const long x = std::strtol(s, &e, 10);
    if (!e || *e || x < -(1L << 30) || x > (1L << 30)) return false;
    v = (int)x;
    return true;
}

bool parse_float(const char *s, float &v)
{
    char *e = nullptr;
    v = std::strtof(s, &e);
    return e && !*e;
}

bool save_bin(const std::string &path, const std::vector<unsigned char> &v, int dim, int rows, int cols)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const int hdr[3] = {dim, rows, cols};
    bool ok = std::fwrite(hdr, sizeof(int), 3, f) == 3;
    ok = ok && std::fwrite(v.data(), 1, v.size(), f) == v.size();
    return std::fclose(f) == 0 && ok;
}

int fail(const std::string &msg)
{
    std::fprintf(stderr, "nl-means: %s\n", msg.c_str());
    return 1;
}

int run3d(const gnlm_cli::Gray8 &img, int ss, int ns, int depth, float h, const std::string &out,
          const std::string &slices)
{
    const int rows = img.rows, cols = img.cols;
    const int pad = (ss + ns - 2) / 2;
    const long long W = cols + 2LL * pad, H = rows + 2LL * pad, D = depth + 2LL * pad;
    std::vector<unsigned char> in((size_t)(W * H * D));
    // one padded slice, then depth + SS + NS - 2 copies of it
    for (long long y = 0; y < H; ++y)
        for (long long x = 0; x < W; ++x)
            in[(size_t)(y * W + x)] =
                img.px[(size_t)reflect101((int)y - pad, rows) * cols + (size_t)reflect101((int)x - pad, cols)];
    for (long long z = 1; z < D; ++z) std::memcpy(&in[(size_t)(z * W * H)], in.data(), (size_t)(W * H));
    std::vector<float> res((size_t)cols * rows * depth);
    const int st = gnlm_nl_means_3d_u8(in.data(), res.data(), cols, rows, depth, ss, ns, h);
    if (st != GNLM_OK) return fail(std::string("gnlm_nl_means_3d_u8: ") + gnlm_status_string(st));
    // float -> 8 bits by truncation, as the reference's (unsigned char) cast;
    // the result is a convex combination of 0..255 inputs
    std::vector<unsigned char> r8(res.size());
    for (size_t i = 0; i < res.size(); ++i) {
        const float v = res[i] < 0.0f ? 0.0f : (res[i] > 255.0f ? 255.0f : res[i]);
        r8[i] = (unsigned char)v;
    }
    if (!save_bin(out, r8, depth, rows, cols)) return fail("cannot write " + out);
    if (!slices.empty()) {
        for (int k = 0; k < depth; ++k) {
            const std::string p = slices + "/i_" + std::to_string(k) + "_result.bmp";
            if (!gnlm_cli::write_bmp_gray8(p, &r8[(size_t)k * rows * cols], rows, cols)) return fail("cannot write " + p);
        }
    }
    std::printf("nl-means: %d x %d x %d, SS %d, NS %d, h %g -> %s\n", cols, rows, depth, ss, ns, (double)h, out.c_str());
    return 0;
}

int run2d(const gnlm_cli::Gray8 &img0, int sr, int nb, float h, const std::string &out, const std::string &padded)
{
    // The drop-in's 2D entry point takes rows no wider than the image is tall
    // (its output stride is the row count); a wider image runs transposed,
    // which gives the same filter (the search and patch windows are squares).
    const bool tr = img0.cols > img0.rows;
    gnlm_cli::Gray8 img = img0;
    if (tr) {
        img.rows = img0.cols;
        img.cols = img0.rows;
        for (int r = 0; r < img0.rows; ++r)
            for (int c = 0; c < img0.cols; ++c) img.px[(size_t)c * img.cols + r] = img0.px[(size_t)r * img0.cols + c];
    }
    const int rows = img.rows, cols = img.cols, pad = (sr + nb - 2) / 2;
    const int W = cols + 2 * pad, H = rows + 2 * pad;
    std::vector<unsigned char> in((size_t)W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            in[(size_t)y * W + x] = img.px[(size_t)reflect101(y - pad, rows) * cols + (size_t)reflect101(x - pad, cols)];
    if (!padded.empty()) {
        // the padded input in the original orientation
        std::vector<unsigned char> p(in.size());
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) p[tr ? (size_t)x * H + y : (size_t)y * W + x] = in[(size_t)y * W + x];
        if (!gnlm_cli::write_bmp_gray8(padded, p.data(), tr ? W : H, tr ? H : W)) return fail("cannot write " + padded);
    }
    std::vector<float> res((size_t)rows * rows);  // rows of stride IMAGE_SIZE_Y = rows
    const int st = gnlm_nl_means_2d_u8(in.data(), res.data(), W, H, sr, nb, cols, rows, h);
    if (st != GNLM_OK) return fail(std::string("gnlm_nl_means_2d_u8: ") + gnlm_status_string(st));
    std::vector<unsigned char> o((size_t)img0.rows * img0.cols);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            float v = std::nearbyint(res[(size_t)r * rows + c]);
            v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
            o[tr ? (size_t)c * img0.cols + r : (size_t)r * cols + c] = (unsigned char)v;
        }
    if (!gnlm_cli::write_bmp_gray8(out, o.data(), img0.rows, img0.cols)) return fail("cannot write " + out);
    std::printf("nl-means 2D: %d x %d, search %d, neighbourhood %d, h %g -> %s\n", img0.cols, img0.rows, sr, nb,
                (double)h, out.c_str());
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    bool two_d = false;
    int ss = 13, ns = 17, depth = 128;
    float h = 10.0f;
    std::string out, slices, padded;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](int &v) { return i + 1 < argc && parse_int(argv[++i], v); };
        if (a == "--2d") two_d = true;
        else if (a == "--ss") {
            if (!need(ss)) return usage(argv[0]), 2;
        } else if (a == "--ns") {
            if (!need(ns)) return usage(argv[0]), 2;
        } else if (a == "--depth") {
            if (!need(depth)) return usage(argv[0]), 2;
        } else if (a == "--h") {
            if (!(i + 1 < argc && parse_float(argv[++i], h))) return usage(argv[0]), 2;
        } else if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--slices" && i + 1 < argc) slices = argv[++i];
        else if (a == "--padded" && i + 1 < argc) padded = argv[++i];
        else if (a == "-h" || a == "--help") return usage(argv[0]), 0;
        else if (!a.empty() && a[0] == '-') return usage(argv[0]), 2;
        else pos.push_back(a);
    }
    if (pos.size() != (two_d ? 3u : 1u)) return usage(argv[0]), 2;
    gnlm_cli::Gray8 img;
    std::string err;
    if (!gnlm_cli::read_gray8(pos[0], img, err)) return fail(err);
    if (!(h > 0.0f) || !std::isfinite(h)) return fail("h must be positive");
    if (two_d) {
        int sr = 0, nb = 0;
        if (!parse_int(pos[1].c_str(), sr) || !parse_int(pos[2].c_str(), nb) || sr < 1 || nb < 1 || sr % 2 == 0 ||
            nb % 2 == 0)
            return fail("search and neighbourhood sizes must be odd positive integers");
        return run2d(img, sr, nb, h, out.empty() ? "output.bmp" : out, padded);
    }
    if (ss < 1 || ns < 1 || ss % 2 == 0 || ns % 2 == 0) return fail("SS and NS must be odd positive integers");
    if (depth < 1) return fail("depth must be positive");
    return run3d(img, ss, ns, depth, h, out.empty() ? "output.bin" : out, slices);
}
