// nlm3d_rows.hip -- per-volume launch of the "rows" forward variant
// (nlm3d_rows.h): a padding pre-pass into a stream-ordered workspace, then
// one workgroup per tile.
#include <hip/hip_runtime.h>
#include "gnlm_internal.h"
#include "nlm3d_rows.h"

namespace gnlm {
namespace {

struct RowsArgs {
    PadVol v;
    float *out;
    float *Zacc;
    int nx, ny, nz;
    float k;
    int ntx, nty, ntz;
};

template <int RP, int RS, bool SHORT>
__global__ __launch_bounds__(kRowsThreads) __attribute__((amdgpu_waves_per_eu(SHORT ? 4 : (RP == 1 ? 3 : 2)))) void nlm3d_fwd_rows(RowsArgs a)
{
    extern __shared__ __align__(16) float lds[];
    using C = RowsCfg<RP, RS, SHORT>;
    const int t = xcd_tile(blockIdx.x, a.ntx * a.nty * a.ntz);
    RowsTile tl;
    tl.v = a.v;
    tl.out = a.out;
    tl.Zacc = a.Zacc;
    tl.nx = a.nx;
    tl.ny = a.ny;
    tl.nz = a.nz;
    tl.k = a.k;
    tl.x0 = (t % a.ntx) * C::TX;
    tl.y0 = ((t / a.ntx) % a.nty) * C::TY;
    tl.z0 = (t / (a.ntx * a.nty)) * C::TZ;
    rows_tile<C>(tl, lds);
}

// Fill the padded copy: reflect-101 samples on [-R, n + R) in every axis
// (or the caller's own halo when the input arrives padded), 0 beyond.
__global__ void rows_pad(InDesc in, float *dst, int px, int py, int pz, int lp, int r)
{
    const long long n = (long long)px * py * pz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int ix = (int)(i % px);
        const long long q = i / px;
        const int x = ix - lp;
        const int y = (int)(q % py) - r;
        const int z = (int)(q / py) - r;
        float v = 0.0f;
        if (x >= -r && x < in.nx + r && y >= -r && y < in.ny + r && z >= -r && z < in.nz + r) v = load_in(in, x, y, z);
        dst[i] = v;
    }
}

template <int RP, int RS, bool SHORT>
constexpr size_t rows_lds()
{
    return RowsCfg<RP, RS, SHORT>::LDS_FLOATS * sizeof(float);
}

template <int RP, int RS, bool SHORT>
hipError_t rows_launch_t(const InDesc &in, float *out, float *Zout, float k, hipStream_t s)
{
    using C = RowsCfg<RP, RS, SHORT>;
    RowsArgs a;
    a.nx = in.nx;
    a.ny = in.ny;
    a.nz = in.nz;
    a.k = k;
    a.out = out;
    a.ntx = (in.nx + C::TX - 1) / C::TX;
    a.nty = (in.ny + C::TY - 1) / C::TY;
    a.ntz = (in.nz + C::TZ - 1) / C::TZ;
    const int px = a.ntx * C::TX + 2 * C::LP;
    const int py = a.nty * C::TY + 2 * C::R;
    const int pz = a.ntz * C::TZ + 2 * C::R;
    const long long npad = (long long)px * py * pz;
    const long long nv = (long long)in.nx * in.ny * in.nz;
    if (npad >= (1LL << 30) || nv >= (1LL << 31)) return hipErrorInvalidValue;
    const size_t bytes = sizeof(float) * (size_t)(npad + (Zout ? 0 : nv));
    float *ws = nullptr;
    hipError_t e = hipMallocAsync((void **)&ws, bytes, s);
    if (e != hipSuccess) return e;
    a.v.p = ws;
    a.v.px = px;
    a.v.pxy = px * py;
    a.v.bytes = (unsigned)(sizeof(float) * npad);
    a.Zacc = Zout ? Zout : ws + npad;
    long long g = (npad + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(rows_pad, dim3((unsigned)g), dim3(256), 0, s, in, ws, px, py, pz, C::LP, C::R);
    e = hipGetLastError();
    if (e == hipSuccess) {
        const size_t lds = rows_lds<RP, RS, SHORT>();
        e = hipFuncSetAttribute((const void *)nlm3d_fwd_rows<RP, RS, SHORT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess) {
            hipLaunchKernelGGL((nlm3d_fwd_rows<RP, RS, SHORT>), dim3(a.ntx * a.nty * a.ntz), dim3(kRowsThreads), lds,
                               s, a);
            e = hipGetLastError();
        }
    }
    const hipError_t e2 = hipFreeAsync(ws, s);
    return e != hipSuccess ? e : e2;
}

// Instantiated radii: r_p in {1, 2}; r_s up to 7 (r_p = 1) or 6 (r_p = 2),
// so that the loaded rows stay at <= 32 samples; short tiles for r_p = 2.
#define GNLM_ROWS_CASES(X)                                                                                       \
    X(1, 1, false) X(1, 2, false) X(1, 3, false) X(1, 4, false) X(1, 5, false) X(1, 6, false) X(1, 7, false)     \
    X(2, 1, false) X(2, 2, false) X(2, 3, false) X(2, 4, false) X(2, 5, false) X(2, 6, false) X(2, 1, true)      \
    X(2, 2, true) X(2, 3, true) X(2, 4, true) X(2, 5, true) X(2, 6, true)

}  // namespace

bool fwd3d_rows_supported(int r_p, int r_s, bool short_tiles)
{
#define GNLM_ROWS_SUP(RP, RS, SH) \
    if (r_p == RP && r_s == RS && short_tiles == SH) return rows_lds<RP, RS, SH>() <= kLdsLimitBytes;
    GNLM_ROWS_CASES(GNLM_ROWS_SUP)
#undef GNLM_ROWS_SUP
    return false;
}

size_t fwd3d_rows_lds_bytes(int r_p, int r_s, bool short_tiles)
{
#define GNLM_ROWS_LDS(RP, RS, SH) \
    if (r_p == RP && r_s == RS && short_tiles == SH) return rows_lds<RP, RS, SH>();
    GNLM_ROWS_CASES(GNLM_ROWS_LDS)
#undef GNLM_ROWS_LDS
    return (size_t)-1;
}

hipError_t fwd3d_rows_launch(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k, bool short_tiles,
                             hipStream_t s)
{
    // a caller-padded input must carry the whole halo the copy needs
    if (in.pad > 0 && in.pad < r_p + r_s) return hipErrorInvalidValue;
#define GNLM_ROWS_RUN(RP, RS, SH) \
    if (r_p == RP && r_s == RS && short_tiles == SH) return rows_launch_t<RP, RS, SH>(in, out, Zout, k, s);
    GNLM_ROWS_CASES(GNLM_ROWS_RUN)
#undef GNLM_ROWS_RUN
    return hipErrorInvalidValue;
}

}  // namespace gnlm
