// gnlm_api.cpp -- the C-ABI of libgnlm.so: argument checking, path choice,
// and the host-pointer entry points that back the C++ drop-in.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include "gnlm_internal.h"

namespace gnlm {

double weight_scale(int P, float h)
{
    // exp(-(S/P)/h^2) = exp2(S * k),  k = -log2(e) / (P h^2)
    const double hd = (double)h;
    return -1.4426950408889634 / ((double)P * hd * hd);
}

namespace {

inline bool bad_h(float h) { return !(h > 0.0f) || !std::isfinite(h); }

inline int hip_status(hipError_t e)
{
    if (e == hipSuccess) return GNLM_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return GNLM_ENOMEM;
    if (e == hipErrorInvalidValue) return GNLM_EINVAL;
    if (e == hipErrorNotSupported) return GNLM_EUNSUPPORTED;
    return GNLM_EHIP;
}

int check_padded(int in_pad, long long &sy, long long &sz, int nx, int ny, int R)
{
    if (in_pad < 0) return GNLM_EINVAL;
    if (in_pad == 0) return GNLM_OK;
    if (in_pad < R) return GNLM_EINVAL;  // the padding must cover the halo
    if (sy == 0) sy = (long long)nx + 2LL * in_pad;
    if (sz == 0) sz = sy * ((long long)ny + 2LL * in_pad);
    if (sy < (long long)nx + 2LL * in_pad) return GNLM_EINVAL;
    return GNLM_OK;
}

}  // namespace
}  // namespace gnlm

using namespace gnlm;

extern "C" {

const char *gnlm_status_string(int status)
{
    switch (status) {
    case GNLM_OK: return "ok";
    case GNLM_EINVAL: return "invalid argument";
    case GNLM_ENOMEM: return "out of memory";
    case GNLM_EUNSUPPORTED: return "unsupported parameter combination";
    case GNLM_EHIP: return "HIP runtime error";
    case GNLM_ESTATE: return "object in wrong state";
    default: return "unknown status";
    }
}

int gnlm_version(void) { return 100; }

int gnlm3d_forward_ex(const float *d_in, int in_pad, long long in_sy, long long in_sz, float *d_out,
                      float *d_Z, int nx, int ny, int nz, int r_p, int r_s, float h, int path, void *stream)
{
    if (!d_in || !d_out) return GNLM_EINVAL;
    if (bad_h(h)) return GNLM_EINVAL;
    int kernel = 0;
    int st = gnlm3d_forward_route(nx, ny, nz, r_p, r_s, path, &kernel);
    if (st != GNLM_OK) return st;
    const int R = r_p + r_s;
    st = check_padded(in_pad, in_sy, in_sz, nx, ny, R);
    if (st != GNLM_OK) return st;
    InDesc in{d_in, nx, ny, nz, in_pad, in_sy, in_sz};
    const int P = (2 * r_p + 1) * (2 * r_p + 1) * (2 * r_p + 1);
    const float k = (float)weight_scale(P, h);
    hipStream_t s = (hipStream_t)stream;
    if (kernel != GNLM_KERNEL_GENERIC) {
        const int shape = fwd3d_pick_shape(in, d_out, d_Z, r_p, r_s, k, kernel, s, true);
        return hip_status(fwd3d_tiled_launch_variant(in, d_out, d_Z, r_p, r_s, k, kernel, s, nullptr, shape));
    }
    return hip_status(fwd_generic_launch(in, d_out, d_Z, r_p, r_s, r_p, r_s, k, s));
}

int gnlm3d_forward_route(int nx, int ny, int nz, int r_p, int r_s, int path, int *kernel)
{
    if (!kernel) return GNLM_EINVAL;
    if (nx < 1 || ny < 1 || nz < 1 || r_p < 0 || r_s < 0) return GNLM_EINVAL;
    if (path < GNLM_PATH_AUTO || path > GNLM_PATH_GENERIC) return GNLM_EINVAL;
    // 2^40 voxels (4 TiB per FP32 buffer) exceed any device's memory
    if ((long long)nx * ny * nz > (1LL << 40)) return GNLM_EUNSUPPORTED;
    return fwd3d_route(nx, ny, nz, r_p, r_s, path, kernel);
}

int gnlm3d_backward_route(int nx, int ny, int nz, int r_p, int r_s, int *kernel)
{
    if (!kernel) return GNLM_EINVAL;
    if (nx < 1 || ny < 1 || nz < 1 || r_p < 0 || r_s < 0) return GNLM_EINVAL;
    if ((long long)nx * ny * nz > (1LL << 40)) return GNLM_EUNSUPPORTED;
    return bwd3d_route(nx, ny, nz, r_p, r_s, kernel);
}

int gnlm3d_forward(const float *d_in, float *d_out, float *d_Z, int nx, int ny, int nz, int r_p, int r_s,
                   float h, void *stream)
{
    return gnlm3d_forward_ex(d_in, 0, 0, 0, d_out, d_Z, nx, ny, nz, r_p, r_s, h, GNLM_PATH_AUTO, stream);
}

int gnlm2d_forward_ex(const float *d_in, int in_pad, long long in_sy, float *d_out, float *d_Z, int nx, int ny,
                      int r_p, int r_s, float h, int path, void *stream)
{
    if (!d_in || !d_out) return GNLM_EINVAL;
    if (nx < 1 || ny < 1 || r_p < 0 || r_s < 0 || bad_h(h)) return GNLM_EINVAL;
    if (path < GNLM_PATH_AUTO || path > GNLM_PATH_GENERIC) return GNLM_EINVAL;
    const int R = r_p + r_s;
    long long in_sz = 0;
    int st = check_padded(in_pad, in_sy, in_sz, nx, ny, R);
    if (st != GNLM_OK) return st;
    // An image has no z padding: the single plane sits at z = 0 of the
    // buffer, so the plane stride must not contribute (load_in adds
    // sz * (z + pad) with z = 0).
    InDesc in{d_in, nx, ny, 1, in_pad, in_sy, 0};
    const int P = (2 * r_p + 1) * (2 * r_p + 1);
    const float k = (float)weight_scale(P, h);
    hipStream_t s = (hipStream_t)stream;
    const bool tiled_ok = fwd2d_tiled_supported(r_p, r_s);
    if (path == GNLM_PATH_TILED && !tiled_ok) return GNLM_EUNSUPPORTED;
    if (path != GNLM_PATH_GENERIC && tiled_ok) return hip_status(fwd2d_tiled_launch(in, d_out, d_Z, r_p, r_s, k, s));
    return hip_status(fwd_generic_launch(in, d_out, d_Z, r_p, r_s, 0, 0, k, s));
}

int gnlm2d_forward(const float *d_in, float *d_out, int nx, int ny, int r_p, int r_s, float h, void *stream)
{
    return gnlm2d_forward_ex(d_in, 0, 0, d_out, nullptr, nx, ny, r_p, r_s, h, GNLM_PATH_AUTO, stream);
}

int gnlm2d_forward_u8(const unsigned char *d_in, int in_pad, long long in_sy, float *d_out, float *d_Z, int nx,
                      int ny, int r_p, int r_s, float h, void *stream)
{
    if (!d_in || !d_out) return GNLM_EINVAL;
    if (nx < 1 || ny < 1 || r_p < 0 || r_s < 0 || bad_h(h)) return GNLM_EINVAL;
    const int R = r_p + r_s;
    long long in_sz = 0;
    int st = check_padded(in_pad, in_sy, in_sz, nx, ny, R);
    if (st != GNLM_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    if (fwd2d_tiled_supported(r_p, r_s)) {
        InDesc in{nullptr, nx, ny, 1, in_pad, in_sy, 0};
        const float k = (float)weight_scale((2 * r_p + 1) * (2 * r_p + 1), h);
        return hip_status(fwd2d_tiled_u8_launch(in, d_in, d_out, d_Z, r_p, r_s, k, s));
    }
    // other radii: the FP32 path on a float copy (a padded input is converted
    // up to its last sample)
    const long long n = in_pad == 0 ? (long long)nx * ny
                                    : (long long)nx + 2LL * in_pad + in_sy * ((long long)ny + 2LL * in_pad - 1);
    float *d_f = nullptr;
    hipError_t e = hipMallocAsync((void **)&d_f, sizeof(float) * (size_t)n, s);
    if (e != hipSuccess) return hip_status(e);
    e = u8_to_f32_any_launch(d_in, d_f, n, s);
    st = e == hipSuccess ? gnlm2d_forward_ex(d_f, in_pad, in_sy, d_out, d_Z, nx, ny, r_p, r_s, h, GNLM_PATH_AUTO,
                                             stream)
                         : hip_status(e);
    const hipError_t e2 = hipFreeAsync(d_f, s);
    return st != GNLM_OK ? st : hip_status(e2);
}

int gnlm3d_backward(const float *d_in, const float *d_out, const float *d_Z, const float *d_gout, int nx, int ny,
                    int nz, int r_p, int r_s, float h, float *d_gin, float *d_gh, void *stream)
{
    if (!d_in || !d_out || !d_Z || !d_gout || !d_gin) return GNLM_EINVAL;
    if (bad_h(h)) return GNLM_EINVAL;
    int kernel = 0;
    const int st = gnlm3d_backward_route(nx, ny, nz, r_p, r_s, &kernel);
    if (st != GNLM_OK) return st;
    return hip_status(bwd3d_launch(d_in, d_out, d_Z, d_gout, nx, ny, nz, r_p, r_s, h, d_gin, d_gh,
                                   (hipStream_t)stream));
}

int gnlm_u8_to_f32(const unsigned char *d_in, float *d_out, long long n, void *stream)
{
    if (!d_in || !d_out || n < 0) return GNLM_EINVAL;
    if (n == 0) return GNLM_OK;
    return hip_status(u8_to_f32_launch(d_in, d_out, n, (hipStream_t)stream));
}

int gnlm3d_forward_u8_route(int nx, int ny, int nz, int r_p, int r_s, int *kernel)
{
    const int st = gnlm3d_forward_route(nx, ny, nz, r_p, r_s, GNLM_PATH_AUTO, kernel);
    if (st == GNLM_OK && fwd3d_u8_fits(nx, ny, nz, r_p, r_s)) *kernel = GNLM_KERNEL_U8_INTEGER;
    return st;
}

int gnlm3d_forward_u8(const unsigned char *d_in, int in_pad, long long in_sy, long long in_sz, float *d_out,
                      float *d_Z, int nx, int ny, int nz, int r_p, int r_s, float h, void *stream)
{
    if (!d_in || !d_out) return GNLM_EINVAL;
    if (bad_h(h)) return GNLM_EINVAL;
    int kernel = 0;
    int st = gnlm3d_forward_u8_route(nx, ny, nz, r_p, r_s, &kernel);
    if (st != GNLM_OK) return st;
    const int R = r_p + r_s;
    st = check_padded(in_pad, in_sy, in_sz, nx, ny, R);
    if (st != GNLM_OK) return st;
    const int P = (2 * r_p + 1) * (2 * r_p + 1) * (2 * r_p + 1);
    const float k = (float)weight_scale(P, h);
    hipStream_t s = (hipStream_t)stream;
    if (kernel == GNLM_KERNEL_U8_INTEGER)
        return hip_status(fwd3d_u8_launch(d_in, in_pad, in_sy, in_sz, d_out, d_Z, nx, ny, nz, r_p, r_s, k, s));
    // other radii and sizes: the float kernels on a float copy of the input
    // (same layout; a padded input is converted up to its last sample)
    const long long n = in_pad == 0 ? (long long)nx * ny * nz
                                    : (long long)nx + 2LL * in_pad + in_sy * ((long long)ny + 2LL * in_pad - 1) +
                                          in_sz * ((long long)nz + 2LL * in_pad - 1);
    float *d_f = nullptr;
    hipError_t e = hipMallocAsync((void **)&d_f, sizeof(float) * (size_t)n, s);
    if (e != hipSuccess) return hip_status(e);
    e = u8_to_f32_any_launch(d_in, d_f, n, s);
    st = e == hipSuccess ? gnlm3d_forward_ex(d_f, in_pad, in_sy, in_sz, d_out, d_Z, nx, ny, nz, r_p, r_s, h,
                                             GNLM_PATH_AUTO, stream)
                         : hip_status(e);
    const hipError_t e2 = hipFreeAsync(d_f, s);
    return st != GNLM_OK ? st : hip_status(e2);
}

// ---------------------------------------------------------------------------
// host-pointer entry points (drop-in backing)
// ---------------------------------------------------------------------------

int gnlm_nl_means_3d_u8(const unsigned char *h_in, float *h_out, int nx, int ny, int nz, int ss, int ns,
                        float h)
{
    if (!h_in || !h_out || nx < 1 || ny < 1 || nz < 1) return GNLM_EINVAL;
    if (ss < 1 || ns < 1 || (ss % 2) == 0 || (ns % 2) == 0) return GNLM_EINVAL;
    const int r_s = (ss - 1) / 2, r_p = (ns - 1) / 2;
    const int pad = (ss + ns - 2) / 2;  // = r_s + r_p, 3DNLM/nlm.cpp:33-36
    const long long W = nx + 2LL * pad, H = ny + 2LL * pad, D = nz + 2LL * pad;
    const long long nin = W * H * D, nout = (long long)nx * ny * nz;
    unsigned char *d_u8 = nullptr;
    float *d_o = nullptr;
    int st = GNLM_OK;
    hipError_t e = hipMalloc((void **)&d_u8, (size_t)nin);
    if (e == hipSuccess) e = hipMalloc((void **)&d_o, sizeof(float) * (size_t)nout);
    if (e == hipSuccess) e = hipMemcpy(d_u8, h_in, (size_t)nin, hipMemcpyHostToDevice);
    if (e != hipSuccess) st = hip_status(e);
    if (st == GNLM_OK) st = gnlm3d_forward_u8(d_u8, pad, W, W * H, d_o, nullptr, nx, ny, nz, r_p, r_s, h, nullptr);
    if (st == GNLM_OK) {
        e = hipMemcpy(h_out, d_o, sizeof(float) * (size_t)nout, hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = hip_status(e);
    }
    if (d_u8) (void)hipFree(d_u8);
    if (d_o) (void)hipFree(d_o);
    return st;
}

int gnlm_nl_means_2d_u8(const unsigned char *h_in, float *h_out, int W, int H, int sr_size, int nb_size,
                        int image_size_x, int image_size_y, float h)
{
    if (!h_in || !h_out || W < 1 || H < 1 || image_size_x < 1 || image_size_y < 1) return GNLM_EINVAL;
    if (sr_size < 1 || nb_size < 1 || (sr_size % 2) == 0 || (nb_size % 2) == 0) return GNLM_EINVAL;
    const int r_s = (sr_size - 1) / 2, r_p = (nb_size - 1) / 2;
    const int pad = (sr_size + nb_size - 2) / 2;  // "2D NLM/11x11/nlm.cpp:15-18"
    const int nx = image_size_x, ny = image_size_y;
    if (W < nx + 2 * pad || H < ny + 2 * pad) return GNLM_EINVAL;
    // Output rows have stride IMAGE_SIZE_Y in the reference, so wider rows
    // would overlap; the reference's driver only produces X <= Y (square).
    if (image_size_x > image_size_y) return GNLM_EINVAL;
    const long long nin = (long long)W * H;
    // The reference writes rows of stride IMAGE_SIZE_Y ("2D NLM/11x11/main.cu:64").
    const long long nout = (long long)image_size_y * image_size_y;
    unsigned char *d_u8 = nullptr;
    float *d_o = nullptr;
    int st = GNLM_OK;
    hipError_t e = hipMalloc((void **)&d_u8, (size_t)nin);
    if (e == hipSuccess) e = hipMalloc((void **)&d_o, sizeof(float) * (size_t)nx * ny);
    if (e == hipSuccess) e = hipMemcpy(d_u8, h_in, (size_t)nin, hipMemcpyHostToDevice);
    if (e != hipSuccess) st = hip_status(e);
    if (st == GNLM_OK) st = gnlm2d_forward_u8(d_u8, pad, W, d_o, nullptr, nx, ny, r_p, r_s, h, nullptr);
    if (st == GNLM_OK) {
        if (image_size_x == image_size_y) {
            e = hipMemcpy(h_out, d_o, sizeof(float) * (size_t)nout, hipMemcpyDeviceToHost);
        } else {
            e = hipMemcpy2D(h_out, sizeof(float) * (size_t)image_size_y, d_o, sizeof(float) * (size_t)nx,
                            sizeof(float) * (size_t)nx, (size_t)ny, hipMemcpyDeviceToHost);
        }
        if (e != hipSuccess) st = hip_status(e);
    }
    if (d_u8) (void)hipFree(d_u8);
    if (d_o) (void)hipFree(d_o);
    return st;
}

}  // extern "C"
