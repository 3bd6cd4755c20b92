// nlm_generic.hip -- per-offset NLM forward for any radii (2D and 3D).
//
// Handles what the tiled kernels do not: large patches such as the
// reference's default configuration (SS = 13, NS = 17 in 3DNLM/config.h:4-5)
// or search windows whose halo would not fit a tile in LDS.  For each offset
// d the separable box sum is formed in three streaming passes over HBM
// scratch volumes, then the weight is accumulated:
//   pass X : s_x(x,y,z)  = sum_i fma(e,e)  on  x in [0,nx), y,z extended by r_p
//   pass Y : s_xy(x,y,z) = sum_j s_x         on  y in [0,ny), z extended
//   pass ZW: S = sum_k s_xy;  w = exp2(S k);  Zd += w;  Yd = fma(w, u~(x+d), Yd)
//   after each dz (3D) / dy (2D): Z += Zd, Y += Yd (two-level accumulation)
// The arithmetic order is the canonical one of the CPU oracle and of the 3D
// direct-sum tiles (variants 0 and 1), which agree with this path bit for bit
// on the patch distances; the rows tiles and the 2D tile sum in pairwise
// trees and agree to rounding (bit for bit on 8-bit integer-valued input while
// P * 255^2 < 2^24, where every partial sum is an exact FP32 integer).
// The passes are HBM/L2-bound (about 30 B per voxel per offset) and each
// offset costs three launches; the path trades speed for generality.
#include "gnlm_internal.h"

namespace gnlm {
namespace {

struct GenArgs {
    InDesc in;
    int rp, rpz;      // patch radii (xy, z)
    int dx, dy, dz;   // current offset
    float *sx;        // nx * (ny + 2rp) * (nz + 2rpz)
    float *sxy;       // nx * ny * (nz + 2rpz)
    float *Z, *Y;     // nx * ny * nz accumulators
    float *Zd, *Yd;   // per-block partials (block = one dz in 3D, one dy in 2D)
    float k;
};

__global__ void gen_pass_x(GenArgs a)
{
    const int nx = a.in.nx, ey = a.in.ny + 2 * a.rp, ez = a.in.nz + 2 * a.rpz;
    const long long n = (long long)nx * ey * ez;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const long long r = i / nx;
        const int y = (int)(r % ey) - a.rp;
        const int z = (int)(r / ey) - a.rpz;
        float acc = 0.0f;
        for (int px = -a.rp; px <= a.rp; ++px) {
            const float e = load_in(a.in, x + px, y, z) - load_in(a.in, x + px + a.dx, y + a.dy, z + a.dz);
            acc = fmaf(e, e, acc);
        }
        a.sx[i] = acc;
    }
}

__global__ void gen_pass_y(GenArgs a)
{
    const int nx = a.in.nx, ny = a.in.ny, ey = ny + 2 * a.rp, ez = a.in.nz + 2 * a.rpz;
    const long long n = (long long)nx * ny * ez;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const long long r = i / nx;
        const int y = (int)(r % ny);
        const int z = (int)(r / ny);
        const float *col = a.sx + x + (long long)nx * (y + (long long)ey * z);
        float s = 0.0f;
        for (int j = 0; j <= 2 * a.rp; ++j) s = s + col[(long long)j * nx];
        a.sxy[i] = s;
    }
}

__global__ void gen_pass_zw(GenArgs a)
{
    const int nx = a.in.nx, ny = a.in.ny, nz = a.in.nz;
    const long long n = (long long)nx * ny * nz;
    const long long plane = (long long)nx * ny;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const long long r = i / nx;
        const int y = (int)(r % ny);
        const int z = (int)(r / ny);
        const float *col = a.sxy + i;  // plane z of the extended stack is z - rpz + rpz
        float S = 0.0f;
        for (int k = 0; k <= 2 * a.rpz; ++k) S = S + col[(long long)k * plane];
        const float w = exp2_fast(S * a.k);
        a.Zd[i] = a.Zd[i] + w;
        a.Yd[i] = fmaf(w, load_in(a.in, x + a.dx, y + a.dy, z + a.dz), a.Yd[i]);
    }
}

// Close a block of the two-level accumulation: totals += partials, partials = 0.
__global__ void gen_fold_block(GenArgs a, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        a.Z[i] = a.Z[i] + a.Zd[i];
        a.Y[i] = a.Y[i] + a.Yd[i];
        a.Zd[i] = 0.0f;
        a.Yd[i] = 0.0f;
    }
}

__global__ void gen_fill0(float *p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void gen_finish(const float *Z, const float *Y, float *out, float *Zout, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        out[i] = Y[i] / Z[i];
        if (Zout) Zout[i] = Z[i];
    }
}

inline int grid_for(long long n)
{
    long long g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

hipError_t fwd_generic_launch(const InDesc &in, float *out, float *Zout, int rp, int rs, int rpz, int rsz,
                              float k, hipStream_t s)
{
    const long long nx = in.nx, ny = in.ny, nz = in.nz;
    const long long nsx = nx * (ny + 2 * rp) * (nz + 2 * rpz);
    const long long nsxy = nx * ny * (nz + 2 * rpz);
    const long long nv = nx * ny * nz;
    float *ws = nullptr;
    hipError_t e = hipMalloc((void **)&ws, sizeof(float) * (size_t)(nsx + nsxy + 4 * nv));
    if (e != hipSuccess) return e;
    GenArgs a;
    a.in = in;
    a.rp = rp;
    a.rpz = rpz;
    a.sx = ws;
    a.sxy = ws + nsx;
    a.Z = a.sxy + nsxy;
    a.Y = a.Z + nv;
    a.Zd = a.Y + nv;
    a.Yd = a.Zd + nv;
    a.k = k;
    hipLaunchKernelGGL(gen_fill0, dim3(grid_for(4 * nv)), dim3(256), 0, s, a.Z, 4 * nv);
    // two-level accumulation: blocks are dz planes of offsets in 3D, dy rows in 2D
    const bool is3d = rsz > 0;
    for (int dz = -rsz; dz <= rsz; ++dz) {
        for (int dy = -rs; dy <= rs; ++dy) {
            for (int dx = -rs; dx <= rs; ++dx) {
                a.dx = dx;
                a.dy = dy;
                a.dz = dz;
                hipLaunchKernelGGL(gen_pass_x, dim3(grid_for(nsx)), dim3(256), 0, s, a);
                hipLaunchKernelGGL(gen_pass_y, dim3(grid_for(nsxy)), dim3(256), 0, s, a);
                hipLaunchKernelGGL(gen_pass_zw, dim3(grid_for(nv)), dim3(256), 0, s, a);
            }
            if (!is3d) hipLaunchKernelGGL(gen_fold_block, dim3(grid_for(nv)), dim3(256), 0, s, a, nv);
        }
        if (is3d) hipLaunchKernelGGL(gen_fold_block, dim3(grid_for(nv)), dim3(256), 0, s, a, nv);
    }
    hipLaunchKernelGGL(gen_finish, dim3(grid_for(nv)), dim3(256), 0, s, a.Z, a.Y, out, Zout, nv);
    e = hipGetLastError();
    // The workspace is released once the stream has drained the passes.
    hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(ws);
    return e != hipSuccess ? e : e2;
}

}  // namespace gnlm
