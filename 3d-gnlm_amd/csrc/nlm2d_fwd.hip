// nlm2d_fwd.hip -- tiled 2D NLM forward for gfx950.
//
// The 2D variant of the reference ("2D NLM/11x11/main.cu:36-63": patch SSD
// over an nb x nb patch for every offset of the sr x sr window, weight
// exp(-(S/nb^2)/PARAMETER^2), weighted mean of the candidate centres), with
// the patch radius generic instead of hard-wired to nb = 11.
//
// One workgroup (256 threads) per 32 x 64 output tile (x by y; taller than
// wide, so the y halo of the phase-1 rows is a smaller share):
//   * the tile's input with its r_p + r_s halo is staged into LDS with
//     reflect-101 applied on the fly;
//   * per offset, phase 1 forms x-direction patch-row sums (s_x) for the
//     tile's 64 + 2 r_p rows in segments of 16 outputs; each thread keeps its
//     segment's centre samples in registers for the whole offset loop and
//     reads only the shifted samples from LDS.  The sums are shared pairwise
//     trees over the differences (half of the squares folded into fused
//     multiply-adds), about 3 operations per output for any nb instead of
//     nb.  s_x is stored column-major with an odd pitch so both its writes
//     and the column reads are conflict-free; it is double-buffered, so one
//     barrier per offset suffices;
//   * phase 2: a thread owns 8 consecutive rows of one column, loads the
//     8 + 2 r_p s_x values it needs once, forms the 8 window sums as shared
//     trees in registers, and accumulates Z and Y in registers.
// The weights are summed per dy, then over dy, as in the CPU oracle and the
// generic path.  The patch sums add the same non-negative terms in a
// different order (trees instead of left to right), so results agree with
// those to rounding (~1e-7 relative in S); for integer-valued input (the
// uint8 images of the reference) every S is an exact integer in FP32 and the
// results are identical bit for bit.
//
// dy split (SPLIT, images with fewer tiles than half the workgroup slots):
// workgroup b computes tile b % ntiles over the dy range of part b / ntiles
// and writes each dy's partial sums (Y, Z) into slot dy + r_s; a combine
// pass forms 0 + partial(-r_s) + ... + partial(r_s) in dy order and y = Y/Z,
// the additions the unsplit tile makes in registers: the same bits.
//
// 8-bit build (INT8): the tile is staged from the caller's 8-bit image, and
// since every partial patch sum of squared 8-bit differences is an integer
// below 2^24 (11 x 11 x 255^2 < 2^23), both box sums run as sliding windows
// (nlm_box.h:box_run*), 2 operations per output instead of about 3: the same
// S, so the same bits as the FP32 build on the converted image.
#include <atomic>
#include <type_traits>
#include "gnlm_internal.h"
#include "nlm_box.h"

namespace gnlm {
namespace {

constexpr int NT = 256;

template <int RP>
struct G2 {
    static constexpr int TX = 32, TY = 64;
    static constexpr int SEG = 16;                  // outputs per phase-1 segment
    static constexpr int NSEG = TX / SEG;           // segments per row
    static constexpr int EY = TY + 2 * RP;          // phase-1 rows
    static constexpr int EYP = EY | 1;              // odd column pitch of s_x
    static constexpr int NR = EY * NSEG;            // phase-1 work items
    static constexpr int ROWS = (NR + NT - 1) / NT;
    static constexpr int ES = SEG + 2 * RP;         // samples per segment
    static constexpr int YR = 8;                    // rows per thread in phase 2
    static constexpr int SXN = TX * EYP;
};

struct Fwd2Args {
    InDesc in;                 // layout (and, FP32 build, the samples)
    const unsigned char *in8;  // the samples, 8-bit build
    float *out;
    float *Zout;
    int rs;
    float k;
    int ntx, nty;
    float *slots;  // dy split: 2 (2 r_s + 1) floats per pixel (slot j: Y at 2 j npix, Z at (2 j + 1) npix)
    int nsplit;    // dy split: parts per tile
};

__host__ __device__ constexpr int lds2_x(int TX, int R) { return (TX + 2 * R) | 1; }

template <int RP>
size_t fwd2_lds_bytes(int rs)
{
    using G = G2<RP>;
    const int R = RP + rs;
    return ((size_t)lds2_x(G::TX, R) * (G::TY + 2 * R) + 2 * (size_t)G::SXN) * sizeof(float);
}

// Workgroup barrier for LDS traffic only (the offset loop shares nothing else).
// As inline asm it also keeps the compiler from interleaving the arithmetic of
// consecutive unrolled offsets across it, which would otherwise keep several
// offsets' values live at once.
__device__ __forceinline__ void lds_barrier2()
{
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// An LDS address the compiler cannot see through: loads from it take their
// offsets as immediates (up to 1020 bytes for paired 32-bit reads) from this
// one base, instead of one address add per read pair, which it emits when the
// base is a known constant plus a lane offset.
typedef const __attribute__((address_space(3))) float lds_cfloat;
typedef __attribute__((address_space(3))) float lds_float;
__device__ __forceinline__ lds_cfloat *lds_opaque(const float *p)
{
    unsigned a = (unsigned)(size_t)(lds_cfloat *)p;
    asm volatile("" : "+v"(a));
    return (lds_cfloat *)(size_t)a;
}
__device__ __forceinline__ lds_float *lds_opaque_w(float *p)
{
    unsigned a = (unsigned)(size_t)(lds_float *)p;
    asm volatile("" : "+v"(a));
    return (lds_float *)(size_t)a;
}

// compile-time loop: f(std::integral_constant<int, I>) for I = I0 .. I1
template <int I, int I1, class F>
__device__ __forceinline__ void static_for2(F &&f)
{
    if constexpr (I <= I1) {
        f(std::integral_constant<int, I>{});
        static_for2<I + 1, I1>(f);
    }
}

// RSC = 0: search radius a.rs at run time; the shifted samples of a segment
// are read from LDS for every offset.  RSC > 0: the search radius is RSC
// (a.rs must equal it); a segment's shifted samples for all dx of one dy
// (ES + 2 RSC of them) are read from LDS once per dy into registers, and dx
// is unrolled so that each offset's window is a compile-time register range
// -- a twentieth of the phase-1 LDS reads at r_s = 10.  Same arithmetic, same
// bits.
// Register budget of the compile-time-radius builds: the waves per SIMD that
// their LDS image allows anyway (patch 11: 3 workgroups per CU; smaller: 4),
// so that the unrolled offsets do not take the whole register file.
template <int RP, int RSC>
constexpr int waves2() { return RSC == 0 ? 1 : (RP >= 4 ? 3 : 4); }

template <int RP, int RSC, bool INT8, bool SPLIT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(waves2<RP, RSC>()))) void nlm2d_fwd_tiled(
    Fwd2Args a)
{
    using G = G2<RP>;
    // the sliding windows are exact: a patch sum plus one x-row sum stays
    // below 2^24
    static_assert(!INT8 || (2 * RP + 2) * (2 * RP + 1) * 65025 < (1 << 24), "8-bit sums must stay exact");
    extern __shared__ float lds[];
    const int R = RP + (RSC > 0 ? RSC : a.rs);
    const int LXr = G::TX + 2 * R;
    const int LX = lds2_x(G::TX, R);
    const int LY = G::TY + 2 * R;
    float *U = lds;
    float *SXB = lds + LX * LY;

    const int ntiles = a.ntx * a.nty;
    const int t = SPLIT ? (int)(blockIdx.x % ntiles) : xcd_tile(blockIdx.x, ntiles);
    const int x0 = (t % a.ntx) * G::TX;
    const int y0 = (t / a.ntx) * G::TY;
    const int tid = threadIdx.x;

    for (int i = tid; i < LX * LY; i += NT) {
        const int lx = i % LX, ly = i / LX;
        float v = 0.0f;
        if (lx < LXr) {
            if constexpr (INT8) v = (float)load_at(a.in8, a.in, x0 - R + lx, y0 - R + ly, 0);
            else v = load_in(a.in, x0 - R + lx, y0 - R + ly, 0);
        }
        U[i] = v;
    }
    __syncthreads();

    // phase-1 items: (row ry, segment sg), ry fastest.  Waves with no item
    // skip phase 1 (a scalar branch); in a partly filled wave the lanes past
    // the last item repeat it and only their LDS stores are masked, so that
    // phase 1 has no divergent control flow around its register arrays.
    float cen[G::ROWS][G::ES];
    int base[G::ROWS];
    bool wav[G::ROWS], own[G::ROWS];
#pragma unroll
    for (int q = 0; q < G::ROWS; ++q) {
        const int r = tid + q * NT;
        own[q] = r < G::NR;
        wav[q] = __builtin_amdgcn_readfirstlane(r) < G::NR;
        const int rc = own[q] ? r : G::NR - 1;
        const int ry = rc % G::EY, sg = rc / G::EY;
        base[q] = (R - RP + ry) * LX + (R - RP) + sg * G::SEG;
        if (wav[q]) {
#pragma unroll
            for (int i = 0; i < G::ES; ++i) cen[q][i] = U[base[q] + i];
        }
    }

    // phase-2 ownership: column cx, rows cy0 .. cy0+YR-1
    const int cx = tid % G::TX;
    const int cy0 = (tid / G::TX) * G::YR;
    float Zacc[G::YR], Yacc[G::YR];
#pragma unroll
    for (int m = 0; m < G::YR; ++m) {
        Zacc[m] = 0.0f;
        Yacc[m] = 0.0f;
    }
    const int candbase = (R + cy0) * LX + (R + cx);

    const int rs = RSC > 0 ? RSC : a.rs;
    int dylo = -rs, dyhi = rs;
    if constexpr (SPLIT) {
        const int part = blockIdx.x / ntiles, nd = 2 * rs + 1;
        dylo = -rs + part * nd / a.nsplit;
        dyhi = -rs + (part + 1) * nd / a.nsplit - 1;
    }
    const int gx = x0 + cx;
    const long long npix = (long long)a.in.nx * a.in.ny;
    int it = 0;
    for (int dy = dylo; dy <= dyhi; ++dy) {
        float Zd[G::YR], Yd[G::YR];  // per-dy partial sums (two-level accumulation)
#pragma unroll
        for (int m = 0; m < G::YR; ++m) {
            Zd[m] = 0.0f;
            Yd[m] = 0.0f;
        }
        // the s_x buffer of the dy's first offset; the offsets alternate between
        // the two (compile-time-radius builds: by the parity of dx, so that every
        // LDS address of an offset is a per-dy base plus an immediate)
        float *const SXe = SXB + (it & 1) * G::SXN, *const SXo = SXB + ((it & 1) ^ 1) * G::SXN;
        lds_cfloat *const colE = lds_opaque(SXe + cx * G::EYP + cy0), *const colO = lds_opaque(SXo + cx * G::EYP + cy0);
        lds_float *dstE[G::ROWS], *dstO[G::ROWS];
#pragma unroll
        for (int q = 0; q < G::ROWS; ++q) {
            const int r = tid + q * NT;
            const int rc = own[q] ? r : G::NR - 1;
            const int o = (rc / G::EY) * G::SEG * G::EYP + rc % G::EY;
            dstE[q] = lds_opaque_w(SXe + o);
            dstO[q] = lds_opaque_w(SXo + o);
        }
        // one offset: phase 1 with shifted sample i of row item q from shv(q, i),
        // the barrier, phase 2; odd: the offset takes the other buffer than the
        // dy's first offset
        auto offset = [&](int dx, bool odd, auto &&shv) {
            const int doff = dy * LX + dx;
#pragma unroll
            for (int q = 0; q < G::ROWS; ++q) {
                if (wav[q]) {
                    float e[G::ES];
#pragma unroll
                    for (int i = 0; i < G::ES; ++i) e[i] = cen[q][i] - shv(q, i);
                    float sx[G::SEG];
                    if constexpr (INT8) box_run_sq<2 * RP + 1, G::SEG>(e, sx);
                    else box_tree_sq<2 * RP + 1, G::SEG>(e, sx);
                    lds_float *dst = odd ? dstO[q] : dstE[q];
                    if (own[q]) {
#pragma unroll
                        for (int j = 0; j < G::SEG; ++j) dst[j * G::EYP] = sx[j];
                    }
                }
            }
            lds_barrier2();

            float col[G::YR + 2 * RP];
            lds_cfloat *src = odd ? colO : colE;
#pragma unroll
            for (int j = 0; j < G::YR + 2 * RP; ++j) col[j] = src[j];
            const float *cand = U + candbase + doff;
            float S[G::YR];
            if constexpr (INT8) box_run<2 * RP + 1, G::YR>(col, S);
            else box_tree<2 * RP + 1, G::YR>(col, S);
#pragma unroll
            for (int m = 0; m < G::YR; ++m) {
                const float w = exp2_fast(S[m] * a.k);
                Zd[m] = Zd[m] + w;
                Yd[m] = fmaf(w, cand[m * LX], Yd[m]);
            }
            // The partial sums of this offset are final here: without this the
            // compiler sinks the unrolled offsets' weights below the dy loop's
            // last barrier and keeps every offset's LDS reads live (hundreds of
            // spilled registers).  The empty volatile asm keeps them in order.
            if constexpr (RSC > 0) {
#pragma unroll
                for (int m = 0; m < G::YR; ++m) asm volatile("" : "+v"(Zd[m]), "+v"(Yd[m]));
            }
            ++it;
        };
        if constexpr (RSC > 0) {
            // the shifted samples of every dx of this dy: window[j] = sample at x offset j - RSC
            constexpr int NWIN = G::ES + 2 * RSC;
            float win[G::ROWS][NWIN];
#pragma unroll
            for (int q = 0; q < G::ROWS; ++q) {
                if (wav[q]) {
                    const float *sh = U + base[q] + dy * LX - RSC;
#pragma unroll
                    for (int j = 0; j < NWIN; ++j) win[q][j] = sh[j];
                }
            }
            static_for2<-RSC, RSC>([&](auto dxc) {
                constexpr int DX = decltype(dxc)::value;
                offset(DX, ((DX + RSC) & 1) != 0, [&](int q, int i) { return win[q][i + DX + RSC]; });
            });
        } else {
            for (int dx = -rs; dx <= rs; ++dx)
                offset(dx, ((dx + rs) & 1) != 0, [&](int q, int i) { return U[base[q] + dy * LX + dx + i]; });
        }
        if constexpr (SPLIT) {
            if (gx < a.in.nx) {
                float *ys = a.slots + 2LL * (dy + rs) * npix, *zs = ys + npix;
#pragma unroll
                for (int m = 0; m < G::YR; ++m) {
                    const int gy = y0 + cy0 + m;
                    if (gy < a.in.ny) {
                        ys[gx + (long long)a.in.nx * gy] = Yd[m];
                        zs[gx + (long long)a.in.nx * gy] = Zd[m];
                    }
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < G::YR; ++m) {
                Zacc[m] = Zacc[m] + Zd[m];
                Yacc[m] = Yacc[m] + Yd[m];
            }
        }
    }

    if (!SPLIT && gx < a.in.nx) {
#pragma unroll
        for (int m = 0; m < G::YR; ++m) {
            const int gy = y0 + cy0 + m;
            if (gy < a.in.ny) {
                const long long o = gx + (long long)a.in.nx * gy;
                a.out[o] = Yacc[m] / Zacc[m];
                if (a.Zout) a.Zout[o] = Zacc[m];
            }
        }
    }
}

// the dy split's totals: 0 + slot 0 + slot 1 + ... in dy order, then y = Y / Z
__global__ void nlm2d_split_combine(const float *slots, int nd, long long npix, float *out, float *Zout)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npix;
         i += (long long)gridDim.x * blockDim.x) {
        float Z = 0.0f, Y = 0.0f;
        for (int j = 0; j < nd; ++j) {
            Z = Z + slots[(2LL * j + 1) * npix + i];
            Y = Y + slots[2LL * j * npix + i];
        }
        if (Zout) Zout[i] = Z;
        out[i] = Y / Z;
    }
}

std::atomic<bool> g_dy_split{true};  // gnlm_set_dz_split covers the 2D dy split too

// Parts per tile of the dy split (1: none): as many parts (up to 2 r_s + 1)
// as fill the workgroup slots in one round, for images whose tiles fill at
// most a quarter of them, at search windows of 21 x 21 or more; the slots
// stay within 256 MiB.  The split costs an allocation, a combine launch and
// 16 (2 r_s + 1) bytes of traffic per pixel, so it pays only where a round of
// tiles is long: at patch 11 / search 21, 256^2 0.088 against 0.238 ms and
// 512^2 0.18 against 0.24 ms, but 768^2 (two parts) 0.46 against 0.30 ms;
// at patch 5 / search 11 it was never faster (512^2 0.091 against 0.073 ms).
template <int RP, int RSC>
int split2_parts(long long tiles, long long npix, int rs)
{
    if (!g_dy_split.load(std::memory_order_relaxed)) return 1;
    const long long nd = 2LL * rs + 1;
    if (nd * nd < 441 || 8 * nd * npix > (256LL << 20)) return 1;
    const long long by_lds = (long long)kLdsLimitBytes / (long long)fwd2_lds_bytes<RP>(rs);
    const long long per_cu = by_lds < waves2<RP, RSC>() ? by_lds : waves2<RP, RSC>();
    long long n = (long long)device_cus() * per_cu / tiles;
    if (n > nd) n = nd;
    return n < 4 ? 1 : (int)n;
}

template <int RP, int RSC, bool INT8>
hipError_t launch2_k(const InDesc &in, const unsigned char *in8, float *out, float *Zout, int rs, float k,
                     hipStream_t s)
{
    using G = G2<RP>;
    Fwd2Args a;
    a.in = in;
    a.in8 = in8;
    a.out = out;
    a.Zout = Zout;
    a.rs = rs;
    a.k = k;
    a.ntx = (in.nx + G::TX - 1) / G::TX;
    a.nty = (in.ny + G::TY - 1) / G::TY;
    a.slots = nullptr;
    a.nsplit = 1;
    const long long tiles = (long long)a.ntx * a.nty;
    const long long npix = (long long)in.nx * in.ny;
    const size_t lds = fwd2_lds_bytes<RP>(rs);
    const int ns = split2_parts<RP, RSC>(tiles, npix, rs);
    if (ns < 2) {
        hipError_t e = hipFuncSetAttribute((const void *)nlm2d_fwd_tiled<RP, RSC, INT8, false>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((nlm2d_fwd_tiled<RP, RSC, INT8, false>), dim3((unsigned)tiles), dim3(NT), lds, s, a);
        return hipGetLastError();
    }
    const int nd = 2 * rs + 1;
    hipError_t e = hipMallocAsync((void **)&a.slots, sizeof(float) * 2 * (size_t)nd * (size_t)npix, s);
    if (e != hipSuccess) return e;
    a.nsplit = ns;
    e = hipFuncSetAttribute((const void *)nlm2d_fwd_tiled<RP, RSC, INT8, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        hipLaunchKernelGGL((nlm2d_fwd_tiled<RP, RSC, INT8, true>), dim3((unsigned)(tiles * ns)), dim3(NT), lds, s, a);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        const long long g = (npix + 255) / 256 < 4096 ? (npix + 255) / 256 : 4096;
        hipLaunchKernelGGL(nlm2d_split_combine, dim3((unsigned)g), dim3(256), 0, s, (const float *)a.slots, nd, npix,
                           out, Zout);
        e = hipGetLastError();
    }
    const hipError_t e2 = hipFreeAsync(a.slots, s);
    return e != hipSuccess ? e : e2;
}

// Search radii with a register-window build (the reference's own 2D geometry,
// patch 11 / search 21, and patch 5 / search 11); the others run the
// run-time-radius kernel.
template <int RP, bool INT8>
hipError_t launch2_rp(const InDesc &in, const unsigned char *in8, float *out, float *Zout, int rs, float k,
                      hipStream_t s)
{
    if constexpr (RP == 5) {
        if (rs == 10) return launch2_k<RP, 10, INT8>(in, in8, out, Zout, rs, k, s);
    } else if constexpr (RP == 2) {
        if (rs == 5) return launch2_k<RP, 5, INT8>(in, in8, out, Zout, rs, k, s);
    }
    return launch2_k<RP, 0, INT8>(in, in8, out, Zout, rs, k, s);
}

template <bool INT8>
hipError_t launch2(const InDesc &in, const unsigned char *in8, float *out, float *Zout, int r_p, int r_s, float k,
                   hipStream_t s)
{
    switch (r_p) {
    case 0: return launch2_rp<0, INT8>(in, in8, out, Zout, r_s, k, s);
    case 1: return launch2_rp<1, INT8>(in, in8, out, Zout, r_s, k, s);
    case 2: return launch2_rp<2, INT8>(in, in8, out, Zout, r_s, k, s);
    case 3: return launch2_rp<3, INT8>(in, in8, out, Zout, r_s, k, s);
    case 4: return launch2_rp<4, INT8>(in, in8, out, Zout, r_s, k, s);
    case 5: return launch2_rp<5, INT8>(in, in8, out, Zout, r_s, k, s);
    default: return hipErrorInvalidValue;
    }
}

constexpr size_t kLdsLimit = 160 * 1024;

}  // namespace

bool fwd2d_tiled_supported(int r_p, int r_s)
{
    if (r_s < 0) return false;
    switch (r_p) {
    case 0: return fwd2_lds_bytes<0>(r_s) <= kLdsLimit;
    case 1: return fwd2_lds_bytes<1>(r_s) <= kLdsLimit;
    case 2: return fwd2_lds_bytes<2>(r_s) <= kLdsLimit;
    case 3: return fwd2_lds_bytes<3>(r_s) <= kLdsLimit;
    case 4: return fwd2_lds_bytes<4>(r_s) <= kLdsLimit;
    case 5: return fwd2_lds_bytes<5>(r_s) <= kLdsLimit;
    default: return false;
    }
}

hipError_t fwd2d_tiled_launch(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k,
                              hipStream_t s)
{
    return launch2<false>(in, nullptr, out, Zout, r_p, r_s, k, s);
}

hipError_t fwd2d_tiled_u8_launch(const InDesc &in, const unsigned char *in8, float *out, float *Zout, int r_p,
                                 int r_s, float k, hipStream_t s)
{
    return launch2<true>(in, in8, out, Zout, r_p, r_s, k, s);
}

void fwd2d_set_dy_split(bool on) { g_dy_split.store(on, std::memory_order_relaxed); }

}  // namespace gnlm
