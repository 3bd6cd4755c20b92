// gnlm_device.h -- device-side helpers shared by the NLM kernels (gfx950).
//
// Volumes are x-fastest: element (x, y, z) of an nx*ny*nz volume lives at
// x + nx*(y + ny*z), as the reference lays them out (3DNLM/nlm.cpp:55-61,
// 3DNLM/main.cu:125).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gnlm {

// Input descriptor.  pad == 0: `data` is the unpadded volume and samples
// outside it follow the reflect-101 rule of OpenCV BORDER_REFLECT_101, which
// the reference's driver applies on the host (3DNLM/nlm.cpp:36-41).
// pad > 0: `data` is already padded by `pad` voxels per side (the reference's
// drop-in input, 3DNLM/main.cu:7613-7620) and is read as given; sy/sz are its
// row and plane strides in elements.
struct InDesc {
    const float *data;
    int nx, ny, nz;
    int pad;
    long long sy, sz;
};

__device__ __forceinline__ int reflect101(int i, int n)
{
    if ((unsigned)i < (unsigned)n) return i;  // inside: no division
    if (n <= 1) return 0;
    const int period = 2 * (n - 1);
    int m = i % period;
    m += (m < 0) ? period : 0;
    return (m < n) ? m : period - m;
}

// Sample u~ at (x, y, z) in volume coordinates (may lie outside the volume by
// up to the padding radius in padded mode, by any amount in reflect mode),
// from data laid out as d describes (load_at: samples of another type, e.g.
// an 8-bit image, with d's sizes, padding and strides in elements).
template <class T>
__device__ __forceinline__ T load_at(const T *data, const InDesc &d, int x, int y, int z)
{
    if (d.pad == 0) {
        x = reflect101(x, d.nx);
        y = reflect101(y, d.ny);
        z = reflect101(z, d.nz);
        return data[(long long)x + (long long)d.nx * ((long long)y + (long long)d.ny * z)];
    }
    return data[(long long)(x + d.pad) + d.sy * (long long)(y + d.pad) + d.sz * (long long)(z + d.pad)];
}

__device__ __forceinline__ float load_in(const InDesc &d, int x, int y, int z) { return load_at(d.data, d, x, y, z); }

// v_exp_f32: 2^x.  The NLM weight exp(-S/(P h^2)) is evaluated as
// exp2(S * k) with k = -log2(e)/(P h^2) folded on the host.
__device__ __forceinline__ float exp2_fast(float x) { return __builtin_amdgcn_exp2f(x); }

// blockIdx -> tile index, grouping consecutive tiles on one XCD.  The
// dispatcher hands workgroup b to XCD (b % 8); tiles that share halo data
// are adjacent in tile order, so giving each XCD a contiguous run of tiles
// keeps their shared input in that XCD's L2.  The first nmain tiles (rounded
// down to a multiple of 8) are split into 8 such runs; the tiles after them
// keep their order and are dealt round robin -- a launch passes the count of
// its full-cost tiles, so that cheaper tiles at the end of the tile order (a
// partial last layer) are spread over all XCDs and fill the last round
// instead of all landing in one XCD's run.
__device__ __forceinline__ int xcd_tile(int b, int nmain)
{
    constexpr int NXCD = 8;
    const int q = nmain / NXCD;
    return b < NXCD * q ? (b % NXCD) * q + (b / NXCD) : b;
}

}  // namespace gnlm
