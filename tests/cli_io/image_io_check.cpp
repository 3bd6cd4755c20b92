// image_io_check -- CPU-only driver of image_io.h for the tests:
//   image_io_check <in image> <out raw> [<out bmp>]
// decodes <in image> to 8-bit gray, writes "rows cols\n" + the raw bytes to
// <out raw>, and re-encodes it as an 8-bit BMP to <out bmp> if given.
#include <cstdio>
#include "image_io.h"

int main(int argc, char **argv)
{
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: %s <in image> <out raw> [<out bmp>]\n", argv[0]);
        return 2;
    }
    gnlm_cli::Gray8 img;
    std::string err;
    if (!gnlm_cli::read_gray8(argv[1], img, err)) {
        std::fprintf(stderr, "image_io_check: %s\n", err.c_str());
        return 1;
    }
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    std::fprintf(f, "%d %d\n", img.rows, img.cols);
    const bool ok = std::fwrite(img.px.data(), 1, img.px.size(), f) == img.px.size();
    if (std::fclose(f) != 0 || !ok) return 1;
    if (argc == 4 && !gnlm_cli::write_bmp_gray8(argv[3], img.px.data(), img.rows, img.cols)) return 1;
    return 0;
}
