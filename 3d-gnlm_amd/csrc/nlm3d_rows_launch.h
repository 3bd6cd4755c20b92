// nlm3d_rows_launch.h -- per-volume launch of the "rows" forward tiles
// (nlm3d_rows.h): a padding pre-pass into a stream-ordered workspace, then one
// workgroup per tile.  Included by the translation units that instantiate
// the tile shapes (nlm3d_rows.hip, nlm3d_rows_big.hip); everything here has
// internal linkage.
#pragma once
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

// The tile of workgroup blockIdx.x (tiles of one XCD are contiguous).  Tiles
// run z-layer by z-layer; a partial last layer (nz not a multiple of TZ) skips
// the work of its missing planes, so it is dealt over the XCDs last.
template <class C>
__device__ __forceinline__ void rows_kernel_body(const RowsArgs &a, float *lds)
{
    const int nfull = a.ntx * a.nty * (a.ntz - (a.nz % C::TZ != 0 ? 1 : 0));
    const int t = xcd_tile(blockIdx.x, nfull);
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

// dz-split launch: workgroup b handles tile b % ntiles over the dz range of
// part b / ntiles (nsplit near-equal ranges of the 2 r_s + 1 planes of
// offsets), writing per-dz partial sums into slots (RowsSplit).
struct RowsSplitArgs {
    float *slots;
    int nsplit;
};

template <class C>
__device__ __forceinline__ void rows_kernel_body_split(const RowsArgs &a, const RowsSplitArgs &sa, float *lds)
{
    const int ntiles = a.ntx * a.nty * a.ntz;
    const int part = blockIdx.x / ntiles, t = blockIdx.x % ntiles;
    constexpr int ND = 2 * C::RS + 1;
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
    RowsSplit sp;
    sp.slots = sa.slots;
    sp.dzlo = -C::RS + part * ND / sa.nsplit;
    sp.dzhi = -C::RS + (part + 1) * ND / sa.nsplit - 1;
    rows_tile<C, true>(tl, lds, sp);
}

// The split launch's totals: for each voxel 0 + slot 0 + slot 1 + ... in dz
// order (the additions of the single-workgroup tile), then y = Y / Z.
__global__ void rows_split_combine(const float *slots, int nd, long long nvox, float *out, float *Zout)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nvox;
         i += (long long)gridDim.x * blockDim.x) {
        float Z = 0.0f, Y = 0.0f;
        for (int j = 0; j < nd; ++j) {
            Z = Z + slots[(2LL * j + 1) * nvox + i];
            Y = Y + slots[2LL * j * nvox + i];
        }
        if (Zout) Zout[i] = Z;
        out[i] = Y / Z;
    }
}

// Fill the padded copy: reflect-101 samples on [-R, n + R) in every axis
// (or the caller's own halo when the input arrives padded), 0 beyond.
// Grid rows walk the planes, so the index arithmetic within a plane is
// 32-bit (the copy is below 2^30 floats).
__global__ void rows_pad(InDesc in, float *dst, int px, int py, int pz, int lp, int r)
{
    const int nxy = px * py;
    for (int zp = blockIdx.y; zp < pz; zp += gridDim.y) {
        const int z = zp - r;
        float *plane = dst + (long long)zp * nxy;
        const bool zin = z >= -r && z < in.nz + r;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nxy; i += gridDim.x * blockDim.x) {
            const int x = i % px - lp;
            const int y = i / px - r;
            float v = 0.0f;
            if (zin && x >= -r && x < in.nx + r && y >= -r && y < in.ny + r) v = load_in(in, x, y, z);
            plane[i] = v;
        }
    }
}

template <class C>
constexpr size_t rows_lds_cfg()
{
    return C::LDS_FLOATS * sizeof(float);
}

// Workspace = padded copy (+ the Z accumulator when the caller wants no Z
// and the tiles keep their totals in the output buffers), allocated and
// released on the stream around the two launches -- or, when the caller
// hands one in (gws), taken from it: nothing is allocated then, and a
// workspace too small is hipErrorOutOfMemory.  Every float of the padded copy
// is written by rows_pad and the Z accumulator is written before it is read
// (first dz), so a reused workspace needs no clearing.  Tiles with register
// totals and no Z wanted get Zacc = NULL and write y alone.
template <class C>
hipError_t rows_launch_cfg(const InDesc &in, float *out, float *Zout, float k, hipStream_t s, void (*kernel)(RowsArgs),
                           const Workspace *gws = nullptr)
{
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
    const bool zws = !Zout && !C::REG_TOTALS;  // Z accumulator in the workspace
    const size_t bytes = sizeof(float) * (size_t)(npad + (zws ? nv : 0));
    float *ws = nullptr;
    hipError_t e = hipSuccess;
    if (gws) {
        if (gws->bytes < bytes || !gws->p) return hipErrorOutOfMemory;
        ws = static_cast<float *>(gws->p);
    } else {
        e = hipMallocAsync((void **)&ws, bytes, s);
        if (e != hipSuccess) return e;
    }
    a.v.p = ws;
    a.v.px = px;
    a.v.pxy = px * py;
    a.v.bytes = (unsigned)(sizeof(float) * npad);
    a.Zacc = Zout ? Zout : (zws ? ws + npad : nullptr);
    const int gx = (px * py + 255) / 256 < 1024 ? (px * py + 255) / 256 : 1024;
    const int gy = pz < 65535 ? pz : 65535;
    hipLaunchKernelGGL(rows_pad, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, in, ws, px, py, pz, C::LP, C::R);
    e = hipGetLastError();
    if (e == hipSuccess) {
        const size_t lds = rows_lds_cfg<C>();
        e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(kernel, dim3(a.ntx * a.nty * a.ntz), dim3(C::NT), lds, s, a);
            e = hipGetLastError();
        }
    }
    const hipError_t e2 = gws ? hipSuccess : hipFreeAsync(ws, s);
    return e != hipSuccess ? e : e2;
}

// The dz-split form of rows_launch_cfg (register-total tiles): padded copy
// and 2 (2 r_s + 1) slots per voxel in one stream-ordered workspace, the
// split kernel on ntiles x nsplit workgroups, then the combine.
template <class C>
hipError_t rows_launch_split_cfg(const InDesc &in, float *out, float *Zout, float k, hipStream_t s,
                                 void (*kernel)(RowsArgs, RowsSplitArgs), int nsplit)
{
    static_assert(C::REG_TOTALS, "register-total tiles");
    constexpr int ND = 2 * C::RS + 1;
    if (nsplit < 2 || nsplit > ND) return hipErrorInvalidValue;
    RowsArgs a;
    a.nx = in.nx;
    a.ny = in.ny;
    a.nz = in.nz;
    a.k = k;
    a.out = out;
    a.Zacc = Zout;
    a.ntx = (in.nx + C::TX - 1) / C::TX;
    a.nty = (in.ny + C::TY - 1) / C::TY;
    a.ntz = (in.nz + C::TZ - 1) / C::TZ;
    const int px = a.ntx * C::TX + 2 * C::LP;
    const int py = a.nty * C::TY + 2 * C::R;
    const int pz = a.ntz * C::TZ + 2 * C::R;
    const long long npad = (long long)px * py * pz;
    const long long nv = (long long)in.nx * in.ny * in.nz;
    const long long ntiles = (long long)a.ntx * a.nty * a.ntz;
    if (npad >= (1LL << 30) || nv >= (1LL << 31) || ntiles * nsplit >= (1LL << 31)) return hipErrorInvalidValue;
    // the slots start on a 256-byte boundary after the padded copy
    const long long soff = (npad + 63) / 64 * 64;
    const size_t bytes = sizeof(float) * (size_t)(soff + 2LL * ND * nv);
    float *ws = nullptr;
    hipError_t e = hipMallocAsync((void **)&ws, bytes, s);
    if (e != hipSuccess) return e;
    a.v.p = ws;
    a.v.px = px;
    a.v.pxy = px * py;
    a.v.bytes = (unsigned)(sizeof(float) * npad);
    RowsSplitArgs sa;
    sa.slots = ws + soff;
    sa.nsplit = nsplit;
    const int gx = (px * py + 255) / 256 < 1024 ? (px * py + 255) / 256 : 1024;
    const int gy = pz < 65535 ? pz : 65535;
    hipLaunchKernelGGL(rows_pad, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, in, ws, px, py, pz, C::LP, C::R);
    e = hipGetLastError();
    if (e == hipSuccess) {
        const size_t lds = rows_lds_cfg<C>();
        e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(kernel, dim3((unsigned)(ntiles * nsplit)), dim3(C::NT), lds, s, a, sa);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) {
        const long long g = (nv + 255) / 256 < 4096 ? (nv + 255) / 256 : 4096;
        hipLaunchKernelGGL(rows_split_combine, dim3((unsigned)g), dim3(256), 0, s, (const float *)sa.slots, ND, nv, out,
                           Zout);
        e = hipGetLastError();
    }
    const hipError_t e2 = hipFreeAsync(ws, s);
    return e != hipSuccess ? e : e2;
}

}  // namespace
}  // namespace gnlm
