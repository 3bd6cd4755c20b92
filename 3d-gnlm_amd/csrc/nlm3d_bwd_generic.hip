// nlm3d_bwd_generic.hip -- adjoint of the 3D NLM forward for any radii.
//
// The rows-organised adjoints (nlm3d_bwd_rows.hip, r_p <= 2) and the LDS-block
// tile (nlm3d_bwd.hip, r_p <= 3 with a small search window) hold a tile's
// halo on chip; larger patches -- the reference's own default geometry,
// NS = 17 and SS = 13 (3DNLM/config.h:4-5) -- do not fit.  This path streams
// the same gather form through HBM scratch volumes, six passes per offset d
// (the forward being differentiated is 3DNLM/main.cu:52-68,107-111,125):
//   X   s_x  = sum_px (u~(x+p) - u~(x+d+p))^2            x in D, y,z in E
//   Y   s_xy = sum_py s_x                                  x,y in D, z in E
//   ZW  S = sum_pz s_xy;  w = exp2(S k);
//       E = -2c w [G(x) u~(x+d) - GY(x) + G(x+d) u~(x) - GY(x+d)]   on D
//       dh(x) += w S (G(x) u~(x+d) - GY(x))              x in the volume
//   EX, EY  box sums of E along x, then y                 (E = 0 outside D)
//   EZG bE = box sum along z;
//       grad(z) += G(z+d) w_d(z) + bE (u~(z) - u~(z+d))    z in E
// with E = [-R, n+R) the extended domain (R = r_p + r_s) on which the
// gradient of the reflect-101 extension lives, D = [-r_s, n+r_s) where E_d
// can be nonzero (x or x+d in the volume), c = 1/(P h^2), G = g/Z and GY =
// g y/Z (zero outside the volume).  Afterwards the gradient is folded onto
// the volume through the reflect-101 preimages (bwd3d_fold_launch) and the
// dL/dh partials are reduced in a fixed order, scaled by 2c/h.  Every sum is
// direct (left to right) and every position has one owner, so the result
// is deterministic.  Cost: about 6 (n + 2R)^3 element-passes per offset; it
// is the path of last resort, not a fast one.
#include <algorithm>
#include "gnlm_internal.h"

namespace gnlm {
namespace {

struct GenBwdArgs {
    InDesc in;          // unpadded input, reflect-101 on read
    const float2 *GG;   // (G, GY) per volume voxel
    int rp, rs, R;
    int dx, dy, dz;     // current offset
    float k, negc2;     // w = exp2(S k); E carries -2c
    int ex, ey, ez;     // extended domain E (n + 2R per axis)
    int dxn, dyn, dzn;  // domain D (n + 2 r_s per axis)
    float *sx, *sxy;    // X and Y pass outputs
    float *w, *E;       // on D
    float *bx, *bxy;    // EX and EY pass outputs
    float *gext;        // gradient on E
    float *dh;          // per-voxel dL/dh partials (volume)
};

__device__ __forceinline__ float2 gg_at(const GenBwdArgs &a, int x, int y, int z)
{
    if ((unsigned)x >= (unsigned)a.in.nx || (unsigned)y >= (unsigned)a.in.ny || (unsigned)z >= (unsigned)a.in.nz)
        return make_float2(0.0f, 0.0f);
    return a.GG[x + (long long)a.in.nx * (y + (long long)a.in.ny * z)];
}

// linear index i over an (n0, n1, n2) box -> (i0, i1, i2)
__device__ __forceinline__ void unrank3(long long i, int n0, int n1, int &i0, int &i1, int &i2)
{
    i0 = (int)(i % n0);
    const long long r = i / n0;
    i1 = (int)(r % n1);
    i2 = (int)(r / n1);
}

#define GRID_LOOP(n) \
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (n); i += (long long)gridDim.x * blockDim.x)

__global__ void gb_make_gg(const float *g, const float *y, const float *Z, float2 *GG, long long n)
{
    GRID_LOOP(n)
    {
        const float G = g[i] / Z[i];
        GG[i] = make_float2(G, G * y[i]);
    }
}

__global__ void gb_fill0(float *p, long long n)
{
    GRID_LOOP(n) p[i] = 0.0f;
}

// X: s_x on x in D, y,z in E
__global__ void gb_pass_x(GenBwdArgs a)
{
    const long long n = (long long)a.dxn * a.ey * a.ez;
    GRID_LOOP(n)
    {
        int xi, yi, zi;
        unrank3(i, a.dxn, a.ey, xi, yi, zi);
        const int x = xi - a.rs, y = yi - a.R, z = zi - a.R;
        float acc = 0.0f;
        for (int px = -a.rp; px <= a.rp; ++px) {
            const float e = load_in(a.in, x + px, y, z) - load_in(a.in, x + px + a.dx, y + a.dy, z + a.dz);
            acc = fmaf(e, e, acc);
        }
        a.sx[i] = acc;
    }
}

// Y: s_xy on x,y in D, z in E
__global__ void gb_pass_y(GenBwdArgs a)
{
    const long long n = (long long)a.dxn * a.dyn * a.ez;
    GRID_LOOP(n)
    {
        int xi, yi, zi;
        unrank3(i, a.dxn, a.dyn, xi, yi, zi);
        // s_x row y (in D) + py sits at E index yi + (R - rs) + py
        const float *col = a.sx + xi + (long long)a.dxn * ((yi + a.R - a.rs - a.rp) + (long long)a.ey * zi);
        float s = 0.0f;
        for (int j = 0; j <= 2 * a.rp; ++j) s = s + col[(long long)j * a.dxn];
        a.sxy[i] = s;
    }
}

// ZW: S, w, E on D; dL/dh partials on the volume
__global__ void gb_pass_zw(GenBwdArgs a)
{
    const long long n = (long long)a.dxn * a.dyn * a.dzn;
    const long long plane = (long long)a.dxn * a.dyn;
    GRID_LOOP(n)
    {
        int xi, yi, zi;
        unrank3(i, a.dxn, a.dyn, xi, yi, zi);
        const int x = xi - a.rs, y = yi - a.rs, z = zi - a.rs;
        const float *col = a.sxy + xi + (long long)a.dxn * yi + plane * (zi + a.R - a.rs - a.rp);
        float S = 0.0f;
        for (int k = 0; k <= 2 * a.rp; ++k) S = S + col[(long long)k * plane];
        const float w = exp2_fast(S * a.k);
        const float2 g0 = gg_at(a, x, y, z), g1 = gg_at(a, x + a.dx, y + a.dy, z + a.dz);
        const float u0 = load_in(a.in, x, y, z), u1 = load_in(a.in, x + a.dx, y + a.dy, z + a.dz);
        const float t1 = fmaf(g0.x, u1, -g0.y);
        const float t2 = fmaf(g1.x, u0, -g1.y);
        a.w[i] = w;
        a.E[i] = (a.negc2 * w) * (t1 + t2);
        if ((unsigned)x < (unsigned)a.in.nx && (unsigned)y < (unsigned)a.in.ny && (unsigned)z < (unsigned)a.in.nz) {
            float *dh = a.dh + x + (long long)a.in.nx * (y + (long long)a.in.ny * z);
            *dh = fmaf(w * S, t1, *dh);
        }
    }
}

// EX: box of E along x, on x in E, y,z in D (E = 0 outside D)
__global__ void gb_pass_ex(GenBwdArgs a)
{
    const long long n = (long long)a.ex * a.dyn * a.dzn;
    GRID_LOOP(n)
    {
        int xi, yi, zi;
        unrank3(i, a.ex, a.dyn, xi, yi, zi);
        const float *row = a.E + (long long)a.dxn * (yi + (long long)a.dyn * zi);
        float s = 0.0f;
        for (int px = -a.rp; px <= a.rp; ++px) {
            const int xd = xi - (a.R - a.rs) + px;  // D index of x + px
            if ((unsigned)xd < (unsigned)a.dxn) s = s + row[xd];
        }
        a.bx[i] = s;
    }
}

// EY: box along y, on x,y in E, z in D
__global__ void gb_pass_ey(GenBwdArgs a)
{
    const long long n = (long long)a.ex * a.ey * a.dzn;
    GRID_LOOP(n)
    {
        int xi, yi, zi;
        unrank3(i, a.ex, a.ey, xi, yi, zi);
        float s = 0.0f;
        for (int py = -a.rp; py <= a.rp; ++py) {
            const int yd = yi - (a.R - a.rs) + py;
            if ((unsigned)yd < (unsigned)a.dyn) s = s + a.bx[xi + (long long)a.ex * (yd + (long long)a.dyn * zi)];
        }
        a.bxy[i] = s;
    }
}

// EZG: box along z and the gradient update, on E
__global__ void gb_pass_ezg(GenBwdArgs a)
{
    const long long n = (long long)a.ex * a.ey * a.ez;
    GRID_LOOP(n)
    {
        int xi, yi, zi;
        unrank3(i, a.ex, a.ey, xi, yi, zi);
        const int x = xi - a.R, y = yi - a.R, z = zi - a.R;
        float bE = 0.0f;
        for (int pz = -a.rp; pz <= a.rp; ++pz) {
            const int zd = zi - (a.R - a.rs) + pz;
            if ((unsigned)zd < (unsigned)a.dzn) bE = bE + a.bxy[xi + (long long)a.ex * (yi + (long long)a.ey * zd)];
        }
        float gsum = bE * (load_in(a.in, x, y, z) - load_in(a.in, x + a.dx, y + a.dy, z + a.dz));
        const int xd = x + a.rs, yd = y + a.rs, zd = z + a.rs;
        if ((unsigned)xd < (unsigned)a.dxn && (unsigned)yd < (unsigned)a.dyn && (unsigned)zd < (unsigned)a.dzn) {
            const float2 g1 = gg_at(a, x + a.dx, y + a.dy, z + a.dz);
            gsum = fmaf(g1.x, a.w[xd + (long long)a.dxn * (yd + (long long)a.dyn * zd)], gsum);
        }
        a.gext[i] = a.gext[i] + gsum;
    }
}

inline unsigned grid_of(long long n)
{
    long long g = (n + 255) / 256;
    return (unsigned)std::max<long long>(1, std::min<long long>(g, 65536));
}

}  // namespace

hipError_t bwd3d_generic_launch(const float *u, const float *y, const float *Z, const float *g, int nx, int ny,
                                int nz, int rp, int rs, float h, float *gu, float *gh, hipStream_t s)
{
    GenBwdArgs a;
    a.in = InDesc{u, nx, ny, nz, 0, 0, 0};
    a.rp = rp;
    a.rs = rs;
    a.R = rp + rs;
    a.ex = nx + 2 * a.R;
    a.ey = ny + 2 * a.R;
    a.ez = nz + 2 * a.R;
    a.dxn = nx + 2 * rs;
    a.dyn = ny + 2 * rs;
    a.dzn = nz + 2 * rs;
    const int P = (2 * rp + 1) * (2 * rp + 1) * (2 * rp + 1);
    const double c = 1.0 / ((double)P * (double)h * (double)h);
    a.k = (float)weight_scale(P, h);
    a.negc2 = (float)(-2.0 * c);
    const long long nv = (long long)nx * ny * nz;
    const long long nE = (long long)a.ex * a.ey * a.ez, nD = (long long)a.dxn * a.dyn * a.dzn;
    const long long nsx = (long long)a.dxn * a.ey * a.ez, nsxy = (long long)a.dxn * a.dyn * a.ez;
    const long long nbx = (long long)a.ex * a.dyn * a.dzn, nbxy = (long long)a.ex * a.ey * a.dzn;
    // (G, GY) | dh | gext | s_x | s_xy | w | E | b_x | b_xy   (float2 first: 8-byte aligned)
    const long long nfl = 2 * nv + nv + nE + nsx + nsxy + 2 * nD + nbx + nbxy;
    float *ws = nullptr;
    hipError_t e = hipMallocAsync((void **)&ws, sizeof(float) * (size_t)nfl, s);
    if (e != hipSuccess) return e;
    float2 *GG = reinterpret_cast<float2 *>(ws);
    a.GG = GG;
    a.dh = ws + 2 * nv;
    a.gext = a.dh + nv;
    a.sx = a.gext + nE;
    a.sxy = a.sx + nsx;
    a.w = a.sxy + nsxy;
    a.E = a.w + nD;
    a.bx = a.E + nD;
    a.bxy = a.bx + nbx;
    hipLaunchKernelGGL(gb_make_gg, dim3(grid_of(nv)), dim3(256), 0, s, g, y, Z, GG, nv);
    hipLaunchKernelGGL(gb_fill0, dim3(grid_of(nv + nE)), dim3(256), 0, s, a.dh, nv + nE);
    e = hipGetLastError();
    for (int dz = -rs; dz <= rs && e == hipSuccess; ++dz)
        for (int dy = -rs; dy <= rs && e == hipSuccess; ++dy)
            for (int dx = -rs; dx <= rs && e == hipSuccess; ++dx) {
                a.dx = dx;
                a.dy = dy;
                a.dz = dz;
                hipLaunchKernelGGL(gb_pass_x, dim3(grid_of(nsx)), dim3(256), 0, s, a);
                hipLaunchKernelGGL(gb_pass_y, dim3(grid_of(nsxy)), dim3(256), 0, s, a);
                hipLaunchKernelGGL(gb_pass_zw, dim3(grid_of(nD)), dim3(256), 0, s, a);
                hipLaunchKernelGGL(gb_pass_ex, dim3(grid_of(nbx)), dim3(256), 0, s, a);
                hipLaunchKernelGGL(gb_pass_ey, dim3(grid_of(nbxy)), dim3(256), 0, s, a);
                hipLaunchKernelGGL(gb_pass_ezg, dim3(grid_of(nE)), dim3(256), 0, s, a);
                e = hipGetLastError();
            }
    if (e == hipSuccess) e = bwd3d_fold_launch(a.gext, gu, nx, ny, nz, a.R, s);
    if (e == hipSuccess && gh) {
        if (nv < (1LL << 31))
            e = bwd3d_reduce_dh_launch(a.dh, (int)nv, (float)(2.0 * c / (double)h), gh, s);
        else
            e = hipErrorNotSupported;
    }
    const hipError_t e2 = hipFreeAsync(ws, s);
    return e != hipSuccess ? e : e2;
}

}  // namespace gnlm
