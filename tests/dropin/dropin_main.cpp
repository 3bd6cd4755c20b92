// Drop-in check: drives NLMeansProcessor::NL_Means exactly as the
// reference's drivers do (3DNLM/nlm.cpp:27-85, "2D NLM/11x11/nlm.cpp:6-48"),
// with a seeded synthetic 8-bit image instead of an OpenCV BMP read:
//   3D: reflect-101 pad by (SS+NS-2)/2, replicate the padded slice along z,
//       call NL_Means(in, out);
//   2D: reflect-101 pad by (sr+nb-2)/2, call NL_Means(in, out, W, H, sr, nb, X, Y).
// Writes the unpadded 8-bit image and both outputs as raw little-endian files
// into the directory given as argv[1]; the pytest compares them with the
// CPU oracle.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "config.h"
#include "gnlm_main.h"

static int refl(int i, int n)
{
    if (n <= 1) return 0;
    const int p = 2 * (n - 1);
    int m = i % p;
    if (m < 0) m += p;
    return m < n ? m : p - m;
}

static bool dump(const std::string &path, const void *p, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, bytes, f) == bytes;
    std::fclose(f);
    return ok;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <out_dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int X = IMAGE_SIZE_X, Y = IMAGE_SIZE_Y;
    std::vector<unsigned char> img((size_t)X * Y);
    unsigned s = 12345u;
    for (auto &v : img) {
        s = s * 1664525u + 1013904223u;
        v = (unsigned char)(s >> 24);
    }
    // ---- 3D, as 3DNLM/nlm.cpp:33-65 ----
    const int pad = (SS + NS - 2) / 2;
    const int W = X + 2 * pad, H = Y + 2 * pad, D = depth + 2 * pad;
    std::vector<unsigned char> in3((size_t)W * H * D);
    for (int d = 0; d < D; ++d)
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j)
                in3[(size_t)d * W * H + (size_t)i * W + j] = img[(size_t)refl(i - pad, Y) * X + refl(j - pad, X)];
    std::vector<float> out3((size_t)X * Y * depth);
    NLMeansProcessor::NL_Means(in3.data(), out3.data());
    // ---- 2D, as "2D NLM/11x11/nlm.cpp:14-36" (square image) ----
    const int n2 = 32, sr = 5, nb = 3, pad2 = (sr + nb - 2) / 2, W2 = n2 + 2 * pad2;
    std::vector<unsigned char> img2((size_t)n2 * n2), in2((size_t)W2 * W2);
    for (auto &v : img2) {
        s = s * 1664525u + 1013904223u;
        v = (unsigned char)(s >> 24);
    }
    for (int i = 0; i < W2; ++i)
        for (int j = 0; j < W2; ++j) in2[(size_t)i * W2 + j] = img2[(size_t)refl(i - pad2, n2) * n2 + refl(j - pad2, n2)];
    std::vector<float> out2((size_t)n2 * n2);
    NLMeansProcessor::NL_Means(in2.data(), out2.data(), W2, W2, sr, nb, n2, n2);
    const bool ok = dump(dir + "/img3.u8", img.data(), img.size()) && dump(dir + "/out3.f32", out3.data(), out3.size() * 4) &&
                    dump(dir + "/img2.u8", img2.data(), img2.size()) && dump(dir + "/out2.f32", out2.data(), out2.size() * 4);
    std::printf("dropin: wrote %s\n", ok ? "ok" : "FAILED");
    return ok ? 0 : 1;
}
