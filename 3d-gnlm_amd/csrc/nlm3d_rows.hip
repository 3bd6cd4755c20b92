// nlm3d_rows.hip -- per-volume launch of the "rows" forward variants 4 and 5
// (r_p 1-4 and 2; tile bodies in nlm3d_rows.h, launch in nlm3d_rows_launch.h).
#include <atomic>
#include "nlm3d_rows_launch.h"

namespace gnlm {
namespace {

// Waves per SIMD each tile shape is compiled for (its register budget).
template <int RP, bool SHORT>
constexpr int rows_waves()
{
    return SHORT ? 4 : (RP == 2 ? 2 : (RP >= 3 ? 4 : 3));
}

template <int RP, int RS, bool SHORT, int WAVES>
__global__ __launch_bounds__(kRowsThreads) __attribute__((amdgpu_waves_per_eu(WAVES))) void nlm3d_fwd_rows(RowsArgs a)
{
    extern __shared__ __align__(16) float lds[];
    rows_kernel_body<RowsCfg<RP, RS, SHORT>>(a, lds);
}

// The dz-split form (small volumes; same tile, same bits).
template <int RP, int RS, bool SHORT, int WAVES>
__global__ __launch_bounds__(kRowsThreads) __attribute__((amdgpu_waves_per_eu(WAVES))) void nlm3d_fwd_rows_split(
    RowsArgs a, RowsSplitArgs sa)
{
    extern __shared__ __align__(16) float lds[];
    rows_kernel_body_split<RowsCfg<RP, RS, SHORT>>(a, sa, lds);
}

// Deep tiles (RowsCfg ZM = 2): 512 threads, two workgroups per CU at four
// waves per SIMD (128 VGPRs).
template <int RP, int RS, bool SHORT>
__global__ __launch_bounds__(2 * kRowsThreads) __attribute__((amdgpu_waves_per_eu(4))) void nlm3d_fwd_rows_deep(
    RowsArgs a)
{
    extern __shared__ __align__(16) float lds[];
    rows_kernel_body<RowsCfg<RP, RS, SHORT, 2>>(a, lds);
}

template <int RP, int RS, bool SHORT>
constexpr size_t rows_lds()
{
    return rows_lds_cfg<RowsCfg<RP, RS, SHORT>>();
}

// r_p = 1 tiles come in two register budgets: 168 VGPRs (3 tiles per CU) and
// 128 VGPRs (4 per CU; the spills fall in the per-dz flush, outside the offset
// loop).  Per CU the two process tiles at about the same rate, so the choice
// is the number of rounds of tiles: a launch of t tiles takes about
// ceil(t / 4c) rounds of 4/3 the length at 4 per CU, against ceil(t / 3c)
// at 3 per CU (c compute units).  At 128^3 (800 tiles, 256 CUs) that is one
// round against two: 1.52 -> 1.36 ms for r_s = 4.  Past four rounds the
// quantisation no longer matters and the 168-VGPR build stays.
bool rows_rp1_four_per_cu(long long tiles)
{
    const long long c = device_cus(), r3 = (tiles + 3 * c - 1) / (3 * c), r4 = (tiles + 4 * c - 1) / (4 * c);
    return r3 <= 4 && 4 * r4 < 3 * r3;
}

// Workgroups of a dz-split launch per tile (1: no split).  The split is for
// volumes with fewer tiles than the GPU holds at once (the 96 x 96 x 48
// configuration: 168 tiles for 768 workgroup slots): each tile is spread
// over up to 2 r_s + 1 workgroups, as many as fill the slots in one round.
// Not with a caller's workspace (it has no room for the slots), nor when the
// slots would pass 256 MiB.
std::atomic<bool> g_dz_split{true};  // gnlm_set_dz_split

template <class C>
int rows_split_parts(const InDesc &in, const Workspace *ws, int waves)
{
    if (ws || !g_dz_split.load(std::memory_order_relaxed)) return 1;
    constexpr int ND = 2 * C::RS + 1;
    const long long nv = (long long)in.nx * in.ny * in.nz;
    if (8LL * ND * nv > (256LL << 20)) return 1;
    const long long tiles = (long long)((in.nx + C::TX - 1) / C::TX) * ((in.ny + C::TY - 1) / C::TY) *
                            ((in.nz + C::TZ - 1) / C::TZ);
    const long long by_lds = (long long)kLdsLimitBytes / (long long)rows_lds_cfg<C>();
    const long long per_cu = by_lds < waves ? by_lds : waves;
    long long n = (long long)device_cus() * per_cu / tiles;
    if (n > ND) n = ND;
    return n < 2 ? 1 : (int)n;
}

// Launch shapes of one tile family: builds and tile depths that give the
// same bits (the autotuner times them at a volume's own size).
//   r_p = 1:        0 = 168 VGPRs, 3 tiles per CU; 1 = 128 VGPRs, 4 per CU;
//                   2 = deep 16 x 14 x 28 tiles, 2 per CU
//   r_p = 2 short:  0 = 16 x 12 x 12, 4 per CU; 1 = deep 16 x 12 x 24, 2 per CU
//   r_p = 3:        0 = 128 VGPRs (4 per CU, spills outside the offset loop);
//                   1 = 168 VGPRs (3 per CU, no spills)
//   r_p = 4:        0 = 16 x 8 x 8, 4 per CU; 1 = deep 16 x 8 x 16, 2 per CU
//   r_p 1-3, last:  the dz split of shape 0's build where rows_split_parts
//                   allows one (small volumes), else shape 0 itself; the
//                   default rule (no recorded shape) takes it wherever
//                   rows_split_parts allows it and r_s >= 3: the split at
//                   least doubles the workgroups of a launch that fills
//                   under half the GPU
// (deep 16 x 10 x 20 tiles for r_p = 3 give the same bits too but ran 2-8%
// slower than both builds at 96^3-256^3, so they are not a candidate)
// (r_p = 4's big 16 x 24 x 24 tile is launched from the variant dispatch.)
// shape < 0: the default rule below.
int rows_num_shapes(int r_p, bool short_tiles)
{
    if (r_p == 1) return 4;
    if (r_p == 2 && short_tiles) return 3;
    if (r_p == 2) return 2;
    if (r_p == 3) return 3;
    if (r_p == 4) return 2;
    return 1;
}

template <int RP, int RS, bool SHORT>
hipError_t rows_launch_t(const InDesc &in, float *out, float *Zout, float k, hipStream_t s, const Workspace *ws,
                         int shape)
{
    using C = RowsCfg<RP, RS, SHORT>;
    if constexpr (RP <= 3) {
        // the fixed rule takes the split from r_s = 3 on (343 offsets and
        // more per voxel: the slots' traffic and the combine stay small
        // against the work); the tuner times it wherever it applies
        if ((shape < 0 && RS >= 3) || shape == rows_num_shapes(RP, SHORT) - 1) {
            constexpr int W0 = rows_waves<RP, SHORT>();
            const int ns = rows_split_parts<C>(in, ws, W0);
            if (ns >= 2)
                return rows_launch_split_cfg<C>(in, out, Zout, k, s, nlm3d_fwd_rows_split<RP, RS, SHORT, W0>, ns);
            if (shape >= 0) shape = 0;
        }
    }
    if constexpr (RP == 1) {
        if (shape < 0) {
            const long long tiles = (long long)((in.nx + C::TX - 1) / C::TX) * ((in.ny + C::TY - 1) / C::TY) *
                                    ((in.nz + C::TZ - 1) / C::TZ);
            shape = rows_rp1_four_per_cu(tiles) ? 1 : 0;
        }
        if (shape == 1) return rows_launch_cfg<C>(in, out, Zout, k, s, nlm3d_fwd_rows<RP, RS, SHORT, 4>, ws);
        if (shape == 2)
            return rows_launch_cfg<RowsCfg<RP, RS, SHORT, 2>>(in, out, Zout, k, s, nlm3d_fwd_rows_deep<RP, RS, SHORT>,
                                                               ws);
    } else if constexpr (RP == 2 && SHORT) {
        if (shape == 1)
            return rows_launch_cfg<RowsCfg<RP, RS, SHORT, 2>>(in, out, Zout, k, s, nlm3d_fwd_rows_deep<RP, RS, SHORT>,
                                                               ws);
    } else if constexpr (RP == 3) {
        if (shape == 1) return rows_launch_cfg<C>(in, out, Zout, k, s, nlm3d_fwd_rows<RP, RS, SHORT, 3>, ws);
    } else if constexpr (RP == 4) {
        if (shape == 1)
            return rows_launch_cfg<RowsCfg<RP, RS, SHORT, 2>>(in, out, Zout, k, s, nlm3d_fwd_rows_deep<RP, RS, SHORT>,
                                                               ws);
    }
    return rows_launch_cfg<C>(in, out, Zout, k, s, nlm3d_fwd_rows<RP, RS, SHORT, rows_waves<RP, SHORT>()>, ws);
}

// Instantiated radii: r_p in {1, 2, 3, 4} with r_s up to 8 (r_p = 1) or 7
// (r_p 2-4): the loaded rows hold 16 + 2 LP <= 40 samples (LP = 4 ceil((r_p +
// r_s) / 4)); short tiles for r_p = 2.
#define GNLM_ROWS_CASES(X)                                                                                       \
    X(1, 1, false) X(1, 2, false) X(1, 3, false) X(1, 4, false) X(1, 5, false) X(1, 6, false) X(1, 7, false)     \
    X(1, 8, false) X(2, 1, false) X(2, 2, false) X(2, 3, false) X(2, 4, false) X(2, 5, false) X(2, 6, false)     \
    X(2, 7, false) X(2, 1, true) X(2, 2, true) X(2, 3, true) X(2, 4, true) X(2, 5, true) X(2, 6, true)           \
    X(2, 7, true) X(3, 1, false) X(3, 2, false) X(3, 3, false) X(3, 4, false) X(3, 5, false) X(3, 6, false)      \
    X(3, 7, false) X(4, 1, false) X(4, 2, false) X(4, 3, false) X(4, 4, false) X(4, 5, false) X(4, 6, false)     \
    X(4, 7, false)

}  // namespace

int device_cus()
{
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cache[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

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

int fwd3d_rows_num_shapes(int r_p, bool short_tiles) { return rows_num_shapes(r_p, short_tiles); }

void fwd3d_set_dz_split(bool on) { g_dz_split.store(on, std::memory_order_relaxed); }

bool fwd3d_rows_split_applies(const InDesc &in, int r_p, int r_s, bool short_tiles)
{
#define GNLM_ROWS_SPLIT(RP, RS, SH)                                                              \
    if (r_p == RP && r_s == RS && short_tiles == SH) {                                             \
        if constexpr (RP <= 3) return rows_split_parts<RowsCfg<RP, RS, SH>>(in, nullptr, rows_waves<RP, SH>()) >= 2; \
        return false;                                                                            \
    }
    GNLM_ROWS_CASES(GNLM_ROWS_SPLIT)
#undef GNLM_ROWS_SPLIT
    return false;
}

hipError_t fwd3d_rows_launch(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k, bool short_tiles,
                             hipStream_t s, const Workspace *ws, int shape)
{
    // a caller-padded input must carry the whole halo the copy needs
    if (in.pad > 0 && in.pad < r_p + r_s) return hipErrorInvalidValue;
    if (shape >= rows_num_shapes(r_p, short_tiles)) return hipErrorInvalidValue;
#define GNLM_ROWS_RUN(RP, RS, SH) \
    if (r_p == RP && r_s == RS && short_tiles == SH) return rows_launch_t<RP, RS, SH>(in, out, Zout, k, s, ws, shape);
    GNLM_ROWS_CASES(GNLM_ROWS_RUN)
#undef GNLM_ROWS_RUN
    return hipErrorInvalidValue;
}

}  // namespace gnlm
