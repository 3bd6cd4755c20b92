// nlm3d_rows_big.hip -- the rows forward on large patches (variant 6: r_p 4-8,
// the reference's default NS = 17 and the NS = 13-17 of its sweeps,
// 3DNLM/config.h:4-5, 3DNLM/run.sh:25-34): 16 x (32 - 2 r_p)^2 tiles on 1024
// threads (tile body nlm3d_rows.h, launch nlm3d_rows_launch.h).
#include "nlm3d_rows_launch.h"

namespace gnlm {
namespace {

template <int RP, int RS>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) void nlm3d_fwd_rows_big(RowsArgs a)
{
    extern __shared__ __align__(16) float lds[];
    rows_kernel_body<RowsBigCfg<RP, RS>>(a, lds);
}

// Instantiated radii: r_p 4-8 with r_s 1-7.
#define GNLM_ROWS_BIG_CASES(X)                                                          \
    X(4, 1) X(4, 2) X(4, 3) X(4, 4) X(4, 5) X(4, 6) X(4, 7) X(5, 1) X(5, 2) X(5, 3)     \
    X(5, 4) X(5, 5) X(5, 6) X(5, 7) X(6, 1) X(6, 2) X(6, 3) X(6, 4) X(6, 5) X(6, 6)     \
    X(6, 7) X(7, 1) X(7, 2) X(7, 3) X(7, 4) X(7, 5) X(7, 6) X(7, 7) X(8, 1) X(8, 2)     \
    X(8, 3) X(8, 4) X(8, 5) X(8, 6) X(8, 7)

}  // namespace

bool fwd3d_rows_big_supported(int r_p, int r_s)
{
#define GNLM_RB_SUP(RP, RS) \
    if (r_p == RP && r_s == RS) return rows_lds_cfg<RowsBigCfg<RP, RS>>() <= kLdsLimitBytes;
    GNLM_ROWS_BIG_CASES(GNLM_RB_SUP)
#undef GNLM_RB_SUP
    return false;
}

size_t fwd3d_rows_big_lds_bytes(int r_p, int r_s)
{
#define GNLM_RB_LDS(RP, RS) \
    if (r_p == RP && r_s == RS) return rows_lds_cfg<RowsBigCfg<RP, RS>>();
    GNLM_ROWS_BIG_CASES(GNLM_RB_LDS)
#undef GNLM_RB_LDS
    return (size_t)-1;
}

hipError_t fwd3d_rows_big_launch(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k, hipStream_t s,
                                 const Workspace *ws)
{
    // a caller-padded input must carry the whole halo the copy needs
    if (in.pad > 0 && in.pad < r_p + r_s) return hipErrorInvalidValue;
#define GNLM_RB_RUN(RP, RS)                                                                       \
    if (r_p == RP && r_s == RS)                                                                   \
        return rows_launch_cfg<RowsBigCfg<RP, RS>>(in, out, Zout, k, s, nlm3d_fwd_rows_big<RP, RS>, ws);
    GNLM_ROWS_BIG_CASES(GNLM_RB_RUN)
#undef GNLM_RB_RUN
    return hipErrorInvalidValue;
}

}  // namespace gnlm
