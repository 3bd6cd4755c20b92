// nlm3d_rows.hip -- per-volume launch of the "rows" forward variants 4 and 5
// (r_p 1-4 and 2; tile bodies in nlm3d_rows.h, launch in nlm3d_rows_launch.h).
#include "nlm3d_rows_launch.h"

namespace gnlm {
namespace {

template <int RP, int RS, bool SHORT>
__global__ __launch_bounds__(kRowsThreads) __attribute__((amdgpu_waves_per_eu(SHORT ? 4 : (RP == 2 ? 2 : (RP >= 3 ? 4 : 3))))) void nlm3d_fwd_rows(RowsArgs a)
{
    extern __shared__ __align__(16) float lds[];
    rows_kernel_body<RowsCfg<RP, RS, SHORT>>(a, lds);
}

template <int RP, int RS, bool SHORT>
constexpr size_t rows_lds()
{
    return rows_lds_cfg<RowsCfg<RP, RS, SHORT>>();
}

template <int RP, int RS, bool SHORT>
hipError_t rows_launch_t(const InDesc &in, float *out, float *Zout, float k, hipStream_t s)
{
    return rows_launch_cfg<RowsCfg<RP, RS, SHORT>>(in, out, Zout, k, s, nlm3d_fwd_rows<RP, RS, SHORT>);
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
