// nlm3d_fwd.hip -- per-volume tiled 3D NLM forward (one launch per volume).
// The tile computation itself lives in nlm3d_tile.h.
#include <cmath>
#include "gnlm_internal.h"
#include "nlm3d_tile.h"

namespace gnlm {
namespace {

struct Fwd3Args {
    InDesc in;
    float *out;
    float *Zout;
    int rs;
    float k;
    int ntx, nty, ntz;
};

template <int RP, bool WIDE>
__global__ __launch_bounds__(kFwdThreads) void nlm3d_fwd_slab(Fwd3Args a, float sigma)
{
    extern __shared__ float lds[];
    using C = SlabCfg<RP>;
    const int nblocks = a.ntx * a.nty * a.ntz;
    const int t = xcd_tile(blockIdx.x, nblocks);
    Fwd3Tile tl;
    tl.in = a.in;
    tl.out = a.out;
    tl.Zout = a.Zout;
    tl.rs = a.rs;
    tl.k = a.k;
    tl.sigma = sigma;
    tl.x0 = (t % a.ntx) * C::TX;
    tl.y0 = ((t / a.ntx) % a.nty) * C::TY;
    tl.z0 = (t / (a.ntx * a.nty)) * C::TZ;
    if constexpr (WIDE)
        fwd3_slab2_tile<RP>(tl, lds, sigma);
    else
        fwd3_slab_tile<RP>(tl, lds, sigma);
}

template <int RP, bool WIDE>
hipError_t launch_slab(const InDesc &in, float *out, float *Zout, int rs, float k, hipStream_t s)
{
    using C = SlabCfg<RP>;
    Fwd3Args a;
    a.in = in;
    a.out = out;
    a.Zout = Zout;
    a.rs = rs;
    a.k = k;
    a.ntx = (in.nx + C::TX - 1) / C::TX;
    a.nty = (in.ny + C::TY - 1) / C::TY;
    a.ntz = (in.nz + C::TZ - 1) / C::TZ;
    const size_t lds = (WIDE ? slab2_lds_floats(RP, rs) : slab_lds_floats(RP, rs)) * sizeof(float);
    hipError_t e = hipFuncSetAttribute((const void *)nlm3d_fwd_slab<RP, WIDE>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const float sigma = std::sqrt(-k);
    hipLaunchKernelGGL((nlm3d_fwd_slab<RP, WIDE>), dim3(a.ntx * a.nty * a.ntz), dim3(kFwdThreads), lds, s, a,
                       sigma);
    return hipGetLastError();
}

template <class C>
__global__ __launch_bounds__(kFwdThreads) void nlm3d_fwd_tiled(Fwd3Args a)
{
    extern __shared__ float lds[];
    const int nblocks = a.ntx * a.nty * a.ntz;
    const int t = xcd_tile(blockIdx.x, nblocks);
    Fwd3Tile tl;
    tl.in = a.in;
    tl.out = a.out;
    tl.Zout = a.Zout;
    tl.rs = a.rs;
    tl.k = a.k;
    tl.sigma = 0.0f;
    tl.x0 = (t % a.ntx) * C::TX;
    tl.y0 = ((t / a.ntx) % a.nty) * C::TY;
    tl.z0 = (t / (a.ntx * a.nty)) * C::TZ;
    fwd3_tile<C>(tl, lds);
}

template <class C>
hipError_t launch_cfg(const InDesc &in, float *out, float *Zout, int rs, float k, hipStream_t s)
{
    Fwd3Args a;
    a.in = in;
    a.out = out;
    a.Zout = Zout;
    a.rs = rs;
    a.k = k;
    a.ntx = (in.nx + C::TX - 1) / C::TX;
    a.nty = (in.ny + C::TY - 1) / C::TY;
    a.ntz = (in.nz + C::TZ - 1) / C::TZ;
    const size_t lds = fwd3_lds_floats<C>(rs) * sizeof(float);
    hipError_t e = hipFuncSetAttribute((const void *)nlm3d_fwd_tiled<C>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nlm3d_fwd_tiled<C>, dim3(a.ntx * a.nty * a.ntz), dim3(kFwdThreads), lds, s, a);
    return hipGetLastError();
}

template <int RP>
hipError_t launch_rp(const InDesc &in, float *out, float *Zout, int rs, float k, int variant, hipStream_t s,
                     const Workspace *ws, int shape)
{
    if (fwd3_is_rows(variant))
        return fwd3d_rows_launch(in, out, Zout, RP, rs, k, variant == kVRowsShort, s, ws, shape);
    if (variant == kVSlab) return launch_slab<RP, false>(in, out, Zout, rs, k, s);
    if (variant == kVSlab2) return launch_slab<RP, true>(in, out, Zout, rs, k, s);
    if (variant == kV16x16x8) return launch_cfg<Fwd3Cfg<RP, 16, 16, 8>>(in, out, Zout, rs, k, s);
    return launch_cfg<Fwd3Cfg<RP, 16, 16, 16>>(in, out, Zout, rs, k, s);
}

template <int RP>
size_t lds_rp(int rs, int variant)
{
    if (fwd3_is_rows(variant)) return fwd3d_rows_lds_bytes(RP, rs, variant == kVRowsShort);
    if (variant == kVSlab) return slab_lds_floats(RP, rs) * sizeof(float);
    if (variant == kVSlab2) return slab2_lds_floats(RP, rs) * sizeof(float);
    if (variant == kV16x16x8) return fwd3_lds_floats<Fwd3Cfg<RP, 16, 16, 8>>(rs) * sizeof(float);
    return fwd3_lds_floats<Fwd3Cfg<RP, 16, 16, 16>>(rs) * sizeof(float);
}

}  // namespace

size_t fwd3d_lds_bytes(int r_p, int r_s, int variant)
{
    if (variant == kVRowsBig) return fwd3d_rows_big_lds_bytes(r_p, r_s);
    if (variant == kVRows && r_p == 4) return fwd3d_rows_lds_bytes(4, r_s, false);
    switch (r_p) {
    case 0: return lds_rp<0>(r_s, variant);
    case 1: return lds_rp<1>(r_s, variant);
    case 2: return lds_rp<2>(r_s, variant);
    case 3: return lds_rp<3>(r_s, variant);
    default: return (size_t)-1;
    }
}

bool fwd3d_variant_supported(int r_p, int r_s, int variant)
{
    if (variant == kVRowsBig) return fwd3d_rows_big_supported(r_p, r_s);
    // r_p = 4 has rows tiles (16 x 8 x 8) besides the big ones
    if (variant == kVRows && r_p == 4) return fwd3d_rows_supported(4, r_s, false);
    if (r_s < 0 || r_p < 0 || r_p > 3 || variant < 0 || variant >= kNumFwd3Variants) return false;
    if (fwd3_is_rows(variant)) return fwd3d_rows_supported(r_p, r_s, variant == kVRowsShort);
    return fwd3d_lds_bytes(r_p, r_s, variant) <= kLdsLimitBytes;
}

bool fwd3d_tiled_supported(int r_p, int r_s)
{
    for (int v = 0; v < kNumFwd3Variants; ++v)
        if (fwd3d_variant_supported(r_p, r_s, v)) return true;
    return false;
}

int fwd3d_num_shapes(int r_p, int r_s, int variant)
{
    if (!fwd3d_variant_supported(r_p, r_s, variant)) return 0;
    // r_p = 4: the 16 x 8 x 8 rows tiles (and their deep 16 x 8 x 16 form) and
    // the 16 x 24 x 24 big tiles pair their tree outputs from even run starts
    // alike, so they give the same bits.  Shapes: 0 = the variant itself,
    // 1 = the other of rows / big, 2 = deep rows.
    if (r_p == 4 && (variant == kVRows || variant == kVRowsBig))
        return fwd3d_variant_supported(4, r_s, kVRows) && fwd3d_variant_supported(4, r_s, kVRowsBig) ? 3 : 1;
    if (fwd3_is_rows(variant)) return fwd3d_rows_num_shapes(r_p, variant == kVRowsShort);
    return 1;
}

bool fwd3d_shape_distinct(const InDesc &in, int r_p, int r_s, int variant, int shape)
{
    // the last shape of the r_p 1-3 rows families is the dz split
    if (shape > 0 && fwd3_is_rows(variant) && r_p >= 1 && r_p <= 3 &&
        shape == fwd3d_rows_num_shapes(r_p, variant == kVRowsShort) - 1)
        return fwd3d_rows_split_applies(in, r_p, r_s, variant == kVRowsShort);
    return true;
}

hipError_t fwd3d_tiled_launch_variant(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k,
                                      int variant, hipStream_t s, const Workspace *ws, int shape)
{
    if (!fwd3d_variant_supported(r_p, r_s, variant)) return hipErrorInvalidValue;
    if (shape >= fwd3d_num_shapes(r_p, r_s, variant)) return hipErrorInvalidValue;
    if (r_p == 4 && (variant == kVRows || variant == kVRowsBig) && shape >= 0) {
        // shape 0 = the variant itself, 1 = the other r_p = 4 tile, 2 = deep rows (same bits)
        if (shape == 2) return fwd3d_rows_launch(in, out, Zout, 4, r_s, k, false, s, ws, 1);
        const bool big = (variant == kVRowsBig) != (shape == 1);
        return big ? fwd3d_rows_big_launch(in, out, Zout, r_p, r_s, k, s, ws)
                   : fwd3d_rows_launch(in, out, Zout, 4, r_s, k, false, s, ws);
    }
    if (variant == kVRowsBig) return fwd3d_rows_big_launch(in, out, Zout, r_p, r_s, k, s, ws);
    if (variant == kVRows && r_p == 4) return fwd3d_rows_launch(in, out, Zout, 4, r_s, k, false, s, ws);
    switch (r_p) {
    case 0: return launch_rp<0>(in, out, Zout, r_s, k, variant, s, ws, shape);
    case 1: return launch_rp<1>(in, out, Zout, r_s, k, variant, s, ws, shape);
    case 2: return launch_rp<2>(in, out, Zout, r_s, k, variant, s, ws, shape);
    case 3: return launch_rp<3>(in, out, Zout, r_s, k, variant, s, ws, shape);
    default: return hipErrorInvalidValue;
    }
}

bool fwd3d_rows_fits(int nx, int ny, int nz, int r_p, int r_s, int variant)
{
    // the padded copy as rows_launch_cfg lays it out: whole tiles plus the
    // halo R = r_p + r_s (x: LP = 4 ceil(R / 4)), below 2^30 floats, and the
    // volume below 2^31 voxels (32-bit offsets in the kernels)
    // of every launch shape the variant may run: the deep shapes (rows tiles
    // of r_p 1 and 4, short r_p 2) double the tile depth, and r_p = 4 also
    // runs the other of its rows / big tiles
    if ((long long)nx * ny * nz >= (1LL << 31)) return false;
    const long long R = r_p + r_s, LP = 4 * ((R + 3) / 4);
    auto fits = [&](long long tx, long long ty, long long tz) {
        const long long px = (nx + tx - 1) / tx * tx + 2 * LP, py = (ny + ty - 1) / ty * ty + 2 * R,
                        pz = (nz + tz - 1) / tz * tz + 2 * R;
        return px * py * pz < (1LL << 30);
    };
    if (r_p == 4 && (variant == kVRows || variant == kVRowsBig))
        return fits(fwd3_tx(kVRows, 4), fwd3_ty(kVRows, 4), 2 * fwd3_tz(kVRows, 4)) &&
               fits(fwd3_tx(kVRowsBig, 4), fwd3_ty(kVRowsBig, 4), fwd3_tz(kVRowsBig, 4));
    const int deep = (variant == kVRows && r_p == 1) || (r_p == 2 && variant == kVRowsShort) ? 2 : 1;
    return fits(fwd3_tx(variant, r_p), fwd3_ty(variant, r_p), deep * fwd3_tz(variant, r_p));
}

int fwd3d_route(int nx, int ny, int nz, int r_p, int r_s, int path, int *kernel)
{
    if (path == GNLM_PATH_GENERIC || !fwd3d_tiled_supported(r_p, r_s)) {
        if (path == GNLM_PATH_TILED) return GNLM_EUNSUPPORTED;
        *kernel = GNLM_KERNEL_GENERIC;
        return GNLM_OK;
    }
    int v = fwd3d_pick_variant(r_p, r_s);
    if ((fwd3_is_rows(v) || v == kVRowsBig) && !fwd3d_rows_fits(nx, ny, nz, r_p, r_s, v)) {
        // beyond the padded copy's 32-bit offsets: the tiles that read the
        // input in place and index in 64 bits, else the per-offset passes
        v = -1;
        for (int c : {(int)kVSlab2, (int)kVSlab, (int)kV16x16x8, (int)kV16x16x16})
            if (fwd3d_variant_supported(r_p, r_s, c)) {
                v = c;
                break;
            }
        if (v < 0) {
            if (path == GNLM_PATH_TILED) return GNLM_EUNSUPPORTED;
            *kernel = GNLM_KERNEL_GENERIC;
            return GNLM_OK;
        }
    }
    *kernel = v;
    return GNLM_OK;
}

}  // namespace gnlm
