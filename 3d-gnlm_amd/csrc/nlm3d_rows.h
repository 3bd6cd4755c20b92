// nlm3d_rows.h -- "rows" forward variant of the 3D NLM (device code), shared
// by the per-volume launch (nlm3d_fwd.hip) and the persistent batched launch
// (nlm3d_batch.hip).
//
// Same mathematics as the other variants (3DNLM/main.cu:52-68 patch SSD over
// the full offset cube, :107-111 weights, :125 normalisation), organised so
// that the input never sits in LDS:
//
//   * A pre-pass copies the volume with its reflect-101 border (the
//     reference pads on the host, 3DNLM/nlm.cpp:36-41) into an aligned HBM
//     buffer: rows start on 16-byte boundaries and every sample a tile can
//     touch exists, so the kernel reads rows with 128-bit global loads and
//     no border logic.  The tile's working set (~100 KiB) stays in L2.
//   * One workgroup (256 threads; 1024 for the large-patch tiles) per
//     16 x TY x TZ tile; LDS holds only two partial-sum buffers (34 KiB for
//     the short r_p = 2 tiles, up to 135 KiB for the large-patch ones), so up
//     to four workgroups share a CU and hide each other's barriers.
//   * Offsets are visited dz, dy, dx.  Phase-1 threads own whole rows
//     (y_ext, z_ext) of the tile grown by r_p: the centre samples of a row
//     stay in registers for the whole kernel, and the shifted row for one
//     (dy, dz) is loaded once (16 + 2 LP samples) and serves all 2 r_s + 1
//     dx offsets from registers -- dx is unrolled at compile time, so the
//     shift is a static register index.
//   * phase 1: differences, x box sums of their squares (shared pair sums,
//              half of the squares folded into fused multiply-adds) -> s_x
//              rows (four 128-bit LDS stores per row);
//     phase 2: (x vector, z_ext, run of y) items: y box sums -> s_xy;
//     phase 3: (group of V3 = 1, 2 or 4 x columns, y, run of z) items: z box
//              sums, w = exp2(S k), per-dz partial sums in registers; the
//              candidate samples u~(x+d) come straight from the padded copy
//              (L1/L2 hits), one 32/64/128-bit load per group and plane.
//     Two barriers per offset; the phase-1 and phase-2 patterns are
//     conflict-free for 64 banks (row pitches of 20 floats, plane paddings
//     of 16 floats).
//   * Weights accumulate in two levels (partial over the (dy, dx) offsets of
//     one dz, then the total over dz) as in the other variants.  The r_p 1-4
//     tiles keep the totals in registers and write y and Z once per voxel;
//     the big tiles (1024 threads, 128 VGPRs) keep them in the output
//     buffers between dz steps (one read-modify-write of two floats per
//     voxel per dz).  Either way each total is (old total + partial), in dz
//     order, so both give the same bits.
// Box sums add non-negative terms only (no running add/subtract windows), in
// shared pairwise trees instead of the CPU oracle's left-to-right chain, so the
// results agree with the oracle to rounding (~1e-7 relative in S) rather
// than bit for bit.
#pragma once
#include <type_traits>
#include "gnlm_device.h"
#include "nlm_box.h"

namespace gnlm {

constexpr int kRowsThreads = 256;

// Padded, aligned copy of a volume.  Sample (x, y, z) of the reflect-101
// extension is p[(z + R) * pxy + (y + R) * px + (x + LP)] for x in
// [-LP, nxr + LP), y and z in [-R, n?r + R), where nxr, nyr, nzr round the
// volume up to whole tiles; samples outside [-R, n + R) hold 0 and only feed
// outputs that are discarded.  The copy is smaller than 4 GiB: the kernel
// addresses it with 32-bit byte offsets.
struct PadVol {
    const float *p;
    int px;          // row pitch in floats (multiple of 4)
    int pxy;         // plane pitch in floats
    unsigned bytes;  // size of the copy (< 4 GiB: offsets are 32-bit)
};

// Tiles: 16 x (16 - 2 r_p) x TZ, so that the grown rows of one z_ext plane
// fill 16 lanes: r_p = 1: 16 x 14 x 14; r_p = 2: 16 x 12 x 28 (two rows per
// thread) or, SHORT, 16 x 12 x 12; r_p = 3: 16 x 10 x 10 (phase-3 items of x
// pairs with 5-plane runs, 160 items); r_p = 4: 16 x 8 x 8 (single columns,
// 4-plane runs).
// SHORT (r_p = 2 only): one phase-1 row per thread and 3-plane phase-3 runs
// -- a larger halo share than the 28-plane tiles, but half the live
// registers: the per-volume kernel runs four waves per SIMD (128 VGPRs; its
// few spills sit outside the offset loops) instead of two.
// ZM = 2 ("deep" tiles, r_p 1, short r_p 2 and r_p 4): the same tile twice
// as deep in z on 512 threads.  The halo share of phases 1 and 2 drops (r_p =
// 2: phase-1 rows 448 for 288 tile rows instead of 256 for 144; r_p = 4: 384
// for 128 instead of 256 for 64), and every voxel gets exactly
// the arithmetic of the ZM = 1 tile -- x and y runs are unchanged, and the
// phase-3 z runs still start at multiples of RZ in volume coordinates, since
// TZ stays a multiple of RZ -- so the outputs are bit-identical.  Two such
// workgroups share a CU (four waves per SIMD); rows, phase-2 items and
// phase-3 items beyond the tile's count are idle (whole waves).
template <int RP_, int RS_, bool SHORT_ = false, int ZM_ = 1>
struct RowsCfg {
    static constexpr int RP = RP_, RS = RS_, ZM = ZM_;
    static constexpr bool SHORT = SHORT_;
    // Y, Z totals over dz stay in registers: one write of y and Z per voxel
    static constexpr bool REG_TOTALS = true;
    static_assert(!SHORT || RP == 2, "short tiles are for r_p = 2");
    static_assert(ZM == 1 || (ZM == 2 && (RP != 2 || SHORT)), "deep tiles: r_p 1, 3, 4 or short r_p 2");
    static constexpr int NT = kRowsThreads * ZM;         // threads per workgroup
    static constexpr int W = 2 * RP + 1;                 // patch width
    static constexpr int R = RP + RS;                    // halo of the copy in y, z
    static constexpr int LP = 4 * ((R + 3) / 4);         // halo of the copy in x (aligned)
    static constexpr int TX = 16;
    static constexpr int NS = TX + 2 * LP;               // samples per loaded row
    static_assert(RP >= 1 && RP <= 4, "rows tiles are for r_p 1-4");
    static constexpr int TY = 16 - 2 * RP;
    static constexpr int TZ = ZM * (RP == 1 ? 14 : (RP == 3 ? 10 : (RP == 4 ? 8 : (SHORT ? 12 : 28))));
    static constexpr int EX = TX + 2 * RP, EY = TY + 2 * RP, EZ = TZ + 2 * RP;
    static constexpr int NROWS = EY * EZ;
    static constexpr int ROWS = (NROWS + NT - 1) / NT;   // phase-1 rows per thread
    // rows past NROWS (deep tiles) are masked per lane in phase 1
    static_assert(EY == 16 && (NROWS % NT == 0 || ROWS == 1), "16-lane groups hold one z_ext plane of rows");
    // phase 2 items: (x pair, z_ext, run of RY y), P2R per thread
    static constexpr int VX = 2;
    static constexpr int RY = RP == 1 ? 7 : (RP == 2 ? 6 : (RP == 3 ? 5 : 4));
    static constexpr int NXV = TX / VX;
    static constexpr int P2N = NXV * EZ * (TY / RY);
    static constexpr int P2R = (P2N + NT - 1) / NT;
    static_assert(TY % RY == 0 && (P2N % NT == 0 || P2R == 1), "whole phase-2 rounds");
    // phase 3 items: (group of V3 x columns, y, run of RZ z); the short tiles
    // read their columns four at a time (128-bit LDS reads: the phase's LDS
    // time halves, and 192 items leave a fourth wave free)
    static constexpr int RZ = SHORT ? 3 : (RP == 3 ? 5 : (RP == 4 ? 4 : 7));
    static constexpr int V3 = SHORT ? 4 : ((RP == 2 || RP == 4) ? 1 : 2);
    static constexpr int P3N = (TX / V3) * TY * (TZ / RZ);
    static constexpr int P3R = (P3N + NT - 1) / NT;
    static_assert(TZ % RZ == 0, "whole z runs");
    // LDS: s_x [z_ext][y_ext][x] (row pitch 20, plane pitch 20*16+16),
    //      s_xy [z_ext][y][x] (plane pitch 16*TY+16)
    static constexpr int SXR = 20;
    static constexpr int SXP = SXR * EY + 16;
    static constexpr int SXN = SXP * EZ;
    static constexpr int SYP = 16 * TY + 16;
    static constexpr int SYN = SYP * EZ;
    static constexpr int LDS_FLOATS = SXN + SYN;
};

// BIG (r_p 4-8: the large patches of the reference's default configuration,
// NS = 17, and of its sweeps, NS = 13-17): 16 x (32 - 2 r_p) x (32 - 2 r_p)
// tiles on 1024 threads, one phase-1 row of the 32 x 32 grown rows per thread
// (at most 128 VGPRs: four waves per SIMD); phase-2 items are single columns
// (x, z_ext, half of the y range).
template <int RP_, int RS_>
struct RowsBigCfg {
    static constexpr int RP = RP_, RS = RS_;
    static_assert(RP >= 4 && RP <= 8, "big tiles are for r_p 4-8");
    static constexpr bool SHORT = false;
    // 1024 threads leave 128 VGPRs: the totals live in the output buffers
    // between dz steps (registers for them cost more in spills than the
    // read-modify-write of two floats per voxel and dz)
    static constexpr bool REG_TOTALS = false;
    static constexpr int NT = 1024;
    static constexpr int W = 2 * RP + 1;
    static constexpr int R = RP + RS;
    static constexpr int LP = 4 * ((R + 3) / 4);
    static constexpr int TX = 16;
    static constexpr int NS = TX + 2 * LP;
    static constexpr int EY = 32, EZ = 32;
    static constexpr int TY = EY - 2 * RP, TZ = EZ - 2 * RP;
    static constexpr int EX = TX + 2 * RP;
    static constexpr int NROWS = EY * EZ;
    static constexpr int ROWS = NROWS / NT;
    static_assert(ROWS == 1, "one phase-1 row per thread");
    static constexpr int VX = 1;
    static constexpr int RY = TY / 2;
    static constexpr int NXV = TX / VX;
    static constexpr int P2N = NXV * EZ * (TY / RY);
    static constexpr int P2R = P2N / NT;
    static_assert(TY % RY == 0 && P2N % NT == 0, "whole phase-2 rounds");
    // z runs: (x, y, run) items fill the 1024 threads as far as TZ allows
    static constexpr int RZ = RP == 4 ? 6 : (RP == 5 ? 11 : (RP == 6 ? 4 : (RP == 7 ? 6 : 4)));
    static constexpr int V3 = 1;
    static constexpr int P3N = TX * TY * (TZ / RZ);
    static constexpr int P3R = (P3N + NT - 1) / NT;
    static_assert(TZ % RZ == 0, "whole z runs");
    // LDS: s_x [z_ext][y_ext][x] (row pitch 20, plane pitch 656 = 16 mod 64),
    //      s_xy [z_ext][y][x] (plane pitch = 16 mod 64)
    static constexpr int SXR = 20;
    static constexpr int SXP = SXR * EY + 16;
    static constexpr int SXN = SXP * EZ;
    static constexpr int SYP = (16 * TY - 16 + 63) / 64 * 64 + 16;
    static constexpr int SYN = SYP * EZ;
    static constexpr int LDS_FLOATS = SXN + SYN;
};

// Global-address-space view of the output buffers.  When the tile body is
// a called function (the persistent batched kernel) its pointers arrive
// generic and memory operations through them become flat operations, which
// count against the LDS wait counter as well (every lds_barrier would wait
// for them); address space 1 keeps them global.
typedef __attribute__((address_space(1))) float gfloat;

// The padded copy is read with buffer loads: a scalar resource (base, size),
// a 32-bit per-lane byte offset that is fixed for the whole tile, and a
// uniform (scalar) byte offset for the current offset d -- no vector address
// arithmetic in the loops.  Accesses are in bounds by construction; loads
// past the resource's size would return 0.
struct PadRsrc {
    __amdgpu_buffer_rsrc_t r;
};

__device__ __forceinline__ PadRsrc pad_rsrc(const PadVol &v)
{
    // the base may arrive in vector registers (called tile bodies): it is
    // uniform, so take it from the first lane to keep the resource scalar
    const unsigned long long b = (unsigned long long)v.p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    const unsigned nb = __builtin_amdgcn_readfirstlane(v.bytes);
    void *p = (void *)(((unsigned long long)hi << 32) | lo);
    return PadRsrc{__builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)nb, 0x00020000)};
}

// A resource over any buffer of `bytes` bytes (uniform base and size).
__device__ __forceinline__ PadRsrc buf_rsrc(const void *base, unsigned bytes)
{
    const unsigned long long b = (unsigned long long)base;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    void *p = (void *)(((unsigned long long)hi << 32) | lo);
    return PadRsrc{__builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes),
                                                     0x00020000)};
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// V consecutive floats (V = 2 or 4) through a buffer resource, as raw bits.
template <int V>
__device__ __forceinline__ void buf_load_v(const PadRsrc &r, unsigned voff, int soff, unsigned *dst)
{
    if constexpr (V == 4) {
        const u32x4 t = __builtin_amdgcn_raw_buffer_load_b128(r.r, (int)voff, soff, 0);
        dst[0] = t[0];
        dst[1] = t[1];
        dst[2] = t[2];
        dst[3] = t[3];
    } else {
        const u32x2 t = __builtin_amdgcn_raw_buffer_load_b64(r.r, (int)voff, soff, 0);
        dst[0] = t[0];
        dst[1] = t[1];
    }
}
template <int V>
__device__ __forceinline__ void buf_store_v(const PadRsrc &r, unsigned voff, const unsigned *src)
{
    if constexpr (V == 4) {
        u32x4 t;
        t[0] = src[0];
        t[1] = src[1];
        t[2] = src[2];
        t[3] = src[3];
        __builtin_amdgcn_raw_buffer_store_b128(t, r.r, (int)voff, 0, 0);
    } else {
        u32x2 t;
        t[0] = src[0];
        t[1] = src[1];
        __builtin_amdgcn_raw_buffer_store_b64(t, r.r, (int)voff, 0, 0);
    }
}

__device__ __forceinline__ float pad_load(const PadRsrc &r, unsigned voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r.r, (int)voff, soff, 0));
}

template <int NS>
__device__ __forceinline__ void pad_load_row(const PadRsrc &r, unsigned voff, int soff, float *dst)
{
#pragma unroll
    for (int j = 0; j < NS / 4; ++j) {
        const auto t = __builtin_amdgcn_raw_buffer_load_b128(r.r, (int)(voff + 16u * j), soff, 0);
        dst[4 * j + 0] = __uint_as_float(t[0]);
        dst[4 * j + 1] = __uint_as_float(t[1]);
        dst[4 * j + 2] = __uint_as_float(t[2]);
        dst[4 * j + 3] = __uint_as_float(t[3]);
    }
}

// Workgroup barrier that waits for LDS traffic only: global loads in flight
// (the next rows, candidate samples) stay in flight across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

struct RowsTile {
    PadVol v;
    float *out;    // y (the big tiles: the running Y total between dz steps)
    float *Zacc;   // Z, or NULL for no Z (the big tiles: the running Z total
                   // between dz steps, never NULL)
    int nx, ny, nz;
    float k;       // w = exp2(S * k), k = -log2(e) / (P h^2)
    int x0, y0, z0;
};

// A dz-split launch (small volumes: several workgroups per tile, one per
// range of dz): the workgroup visits dz in [dzlo, dzhi] only and writes each
// dz's partial sums into slot dz + r_s of `slots` (slot j: Y at slots + 2 j
// nvox, Z at slots + (2 j + 1) nvox, voxel-indexed); a combine pass then
// forms 0 + partial(-r_s) + partial(-r_s + 1) + ... in dz order -- the
// additions a single workgroup makes -- so the output bits are the same.
struct RowsSplit {
    float *slots;
    int dzlo, dzhi;
};

// phase 1 for the offset whose first shifted sample is sh[SH]
template <class C, int SH>
__device__ __forceinline__ void rows_phase1(const float (&cen)[C::ROWS][C::EX], const float (&sh)[C::ROWS][C::NS],
                                            float *SX, const int (&sxw)[C::ROWS], const bool (&ok)[C::ROWS])
{
#pragma unroll
    for (int q = 0; q < C::ROWS; ++q) {
        if (!ok[q]) continue;
        float e[C::EX];
#pragma unroll
        for (int i = 0; i < C::EX; ++i) e[i] = cen[q][i] - sh[q][SH + i];
        float s[C::TX];
        box_tree_sq<C::W, C::TX>(e, s);
        float4 *dst = reinterpret_cast<float4 *>(SX + sxw[q]);
#pragma unroll
        for (int j = 0; j < C::TX / 4; ++j) dst[j] = make_float4(s[4 * j], s[4 * j + 1], s[4 * j + 2], s[4 * j + 3]);
    }
}

// compile-time loop: f(std::integral_constant<int, I>) for I = I0 .. I1
template <int I, int I1, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I <= I1) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, I1>(f);
    }
}

template <class C>
__device__ __forceinline__ void rows_phase2(const float *src, float *dst)
{
    constexpr int NIN = C::RY + 2 * C::RP;
    float v[C::VX][NIN];
#pragma unroll
    for (int j = 0; j < NIN; ++j) {
        if constexpr (C::VX == 1) {
            v[0][j] = src[j * C::SXR];
        } else if constexpr (C::VX == 4) {
            const float4 t = *reinterpret_cast<const float4 *>(src + j * C::SXR);
            v[0][j] = t.x;
            v[1][j] = t.y;
            v[2][j] = t.z;
            v[3][j] = t.w;
        } else {
            const float2 t = *reinterpret_cast<const float2 *>(src + j * C::SXR);
            v[0][j] = t.x;
            v[1][j] = t.y;
        }
    }
    float s[C::VX][C::RY];
#pragma unroll
    for (int c = 0; c < C::VX; ++c) box_tree<C::W, C::RY>(v[c], s[c]);
#pragma unroll
    for (int o = 0; o < C::RY; ++o) {
        if constexpr (C::VX == 1)
            dst[o * 16] = s[0][o];
        else if constexpr (C::VX == 4)
            *reinterpret_cast<float4 *>(dst + o * 16) = make_float4(s[0][o], s[1][o], s[2][o], s[3][o]);
        else
            *reinterpret_cast<float2 *>(dst + o * 16) = make_float2(s[0][o], s[1][o]);
    }
}

template <class C, bool SPLIT = false>
__device__ __forceinline__ void rows_tile(const RowsTile &ta, float *lds, RowsSplit sp = RowsSplit{nullptr, 0, 0})
{
    static_assert(!SPLIT || C::REG_TOTALS, "dz-split launches are for the register-total tiles");
    const RowsTile a = ta;  // registers: no reloads through a generic pointer in the loops
    constexpr int RP = C::RP, RS = C::RS, RZ = C::RZ, NT = C::NT;
    float *SX = lds;
    float *SXY = lds + C::SXN;
    const int tid = threadIdx.x;
    const PadRsrc rs = pad_rsrc(a.v);
    gfloat *out = (gfloat *)a.out;
    gfloat *Zacc = (gfloat *)a.Zacc;
    const int px = __builtin_amdgcn_readfirstlane(a.v.px), pxy = __builtin_amdgcn_readfirstlane(a.v.pxy);
    const float kw = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(a.k)));

    // z planes of this tile inside the volume: a partial tile at the far z
    // end skips the rows, items and loads that only feed discarded planes
    // (wave-uniform: each wave covers 4 consecutive z_ext values in phase 1)
    const int tzv = min(C::TZ, a.nz - a.z0);

    // ---- phase-1 rows: y_ext = r % 16, z_ext = r / 16 ----
    float cen[C::ROWS][C::EX];
    unsigned rbase[C::ROWS];  // byte offsets: loads take a uniform base + 32-bit lane offset
    int sxw[C::ROWS];
    bool p1ok[C::ROWS];
#pragma unroll
    for (int q = 0; q < C::ROWS; ++q) {
        const int r = tid + q * NT;
        const int yp = r % C::EY, zp = r / C::EY;
        // z_ext grows with the lane index within a wave: the first lane decides
        // for the whole wave, which keeps the branch scalar
        p1ok[q] = __builtin_amdgcn_readfirstlane(zp < tzv + 2 * RP) && (C::NROWS % NT == 0 || r < C::NROWS);
        // at offset (dy, dz) = (-RS, -RS): every scalar offset added below is >= 0
        rbase[q] = 4u * (unsigned)((a.z0 - RP + zp + C::R - RS) * pxy + (a.y0 - RP + yp + C::R - RS) * px + a.x0);
        sxw[q] = zp * C::SXP + yp * C::SXR;
        float row[C::NS];
        pad_load_row<C::NS>(rs, rbase[q], 4 * (RS * pxy + RS * px), row);
#pragma unroll
        for (int i = 0; i < C::EX; ++i) cen[q][i] = p1ok[q] ? row[C::LP - RP + i] : 0.0f;
    }

    // ---- phase-2 items: lanes run x pair fastest, then z_ext ----
    int p2src[C::P2R], p2dst[C::P2R];
    bool p2ok[C::P2R];
#pragma unroll
    for (int m = 0; m < C::P2R; ++m) {
        const int it = tid + m * NT;
        const int xv = it % C::NXV;
        const int zp2 = (it / C::NXV) % C::EZ;
        const int yr2 = it / (C::NXV * C::EZ);
        // A wave holds 64 / NXV consecutive z_ext values of one y run when EZ
        // is a multiple of that count, and the first lane can decide for the
        // wave; otherwise (deep tiles: EZ = 28, 30) a wave can wrap into the
        // next y run, whose low z_ext rows are needed, so each lane decides.
        // Items past P2N (deep tiles) skip the phase.
        if constexpr (C::EZ % (64 / C::NXV) == 0)
            p2ok[m] = __builtin_amdgcn_readfirstlane(zp2 < tzv + 2 * RP) && (C::P2N % NT == 0 || it < C::P2N);
        else
            p2ok[m] = zp2 < tzv + 2 * RP && it < C::P2N;
        p2src[m] = zp2 * C::SXP + yr2 * C::RY * C::SXR + xv * C::VX;
        p2dst[m] = zp2 * C::SYP + yr2 * C::RY * 16 + xv * C::VX;
    }

    // ---- phase-3 items: (x group of V3 columns, y, z run) ----
    constexpr int V3 = C::V3, NXG = 16 / V3;
    int p3r[C::P3R];
    unsigned cb[C::P3R];
    bool p3ok[C::P3R];
#pragma unroll
    for (int k = 0; k < C::P3R; ++k) {
        const int it = tid + k * NT;
        const int itc = ((C::P3N % NT == 0) || it < C::P3N) ? it : C::P3N - 1;
        const int x = (itc % NXG) * V3, y = (itc / NXG) % C::TY, zr = itc / (NXG * C::TY);
        // items past P3N (whole waves for the short tiles) skip the phase
        p3ok[k] = __builtin_amdgcn_readfirstlane(((C::P3N % NT == 0) || it < C::P3N) && zr * RZ < tzv);
        p3r[k] = zr * RZ * C::SYP + y * 16 + x;
        // candidate u~(x + d) at d = (-RS, -RS, -RS)
        cb[k] = 4u * (unsigned)((a.z0 + zr * RZ + C::R - RS) * pxy + (a.y0 + y + C::R - RS) * px + a.x0 + x + C::LP -
                                RS);
    }

    // V3 > 1: the flush of whole groups uses 64/128-bit buffer accesses
    const long long nvox = (long long)a.nx * a.ny * a.nz;
    const bool vec_ok = V3 > 1 && (((unsigned long long)a.out | (unsigned long long)a.Zacc) & (4 * V3 - 1)) == 0 &&
                        (a.nx % V3) == 0 && nvox < (1LL << 30);
    const unsigned vbytes = vec_ok ? (unsigned)(4 * nvox) : 0u;
    const PadRsrc ro = buf_rsrc(a.out, vbytes), rz = buf_rsrc(a.Zacc, vbytes);

    // Write (Z, Y) sums of this thread's items: with acc, added to the totals
    // already in the output buffers (Zacc, out); with last, y = Y / Z is stored.
    // No Z buffer (register totals only): y alone.
    const bool wz = !C::REG_TOTALS || a.Zacc != nullptr;
    auto store = [&](float(&Zs)[C::P3R][V3][RZ], float(&Ys)[C::P3R][V3][RZ], bool acc, bool last) {
#pragma unroll
        for (int k = 0; k < C::P3R; ++k) {
            const int it = tid + k * NT;
            const int x = (it % NXG) * V3, y = (it / NXG) % C::TY, zr = it / (NXG * C::TY);
            const int gy = a.y0 + y, gz = a.z0 + zr * RZ;
            if (!p3ok[k] || it >= C::P3N || gy >= a.ny) continue;
            if constexpr (V3 > 1) {
                // a whole group at an aligned position: one 64/128-bit access per
                // buffer and plane (vec_ok: aligned buffers, rows of a multiple of
                // V3 samples, byte offsets within 32 bits)
                if (vec_ok && a.x0 + x + V3 - 1 < a.nx) {
#pragma unroll
                    for (int o = 0; o < RZ; ++o) {
                        if (gz + o >= a.nz) continue;
                        const unsigned boff = 4u * (unsigned)(a.x0 + x + a.nx * (gy + a.ny * (gz + o)));
                        float Z[V3], Y[V3];
#pragma unroll
                        for (int c = 0; c < V3; ++c) {
                            Z[c] = Zs[k][c][o];
                            Y[c] = Ys[k][c][o];
                        }
                        if (acc) {
                            unsigned zt[V3], yt[V3];
                            buf_load_v<V3>(rz, boff, 0, zt);
                            buf_load_v<V3>(ro, boff, 0, yt);
#pragma unroll
                            for (int c = 0; c < V3; ++c) {
                                Z[c] = __uint_as_float(zt[c]) + Z[c];
                                Y[c] = __uint_as_float(yt[c]) + Y[c];
                            }
                        }
                        unsigned zs[V3], ys[V3];
#pragma unroll
                        for (int c = 0; c < V3; ++c) {
                            zs[c] = __float_as_uint(Z[c]);
                            ys[c] = __float_as_uint(last ? Y[c] / Z[c] : Y[c]);
                        }
                        if (wz) buf_store_v<V3>(rz, boff, zs);
                        buf_store_v<V3>(ro, boff, ys);
                    }
                    continue;
                }
            }
#pragma unroll
            for (int c = 0; c < V3; ++c) {
                const int gx = a.x0 + x + c;
                if (gx >= a.nx) continue;
#pragma unroll
                for (int o = 0; o < RZ; ++o) {
                    if (gz + o >= a.nz) continue;
                    const int oi = gx + a.nx * (gy + a.ny * (gz + o));
                    float Z = Zs[k][c][o], Y = Ys[k][c][o];
                    if (acc) {
                        Z = Zacc[oi] + Z;
                        Y = out[oi] + Y;
                    }
                    if (wz) Zacc[oi] = Z;
                    out[oi] = last ? Y / Z : Y;
                }
            }
        }
    };

    // totals over dz, in registers when the tile keeps them there
    float Zt[C::P3R][V3][RZ], Yt[C::P3R][V3][RZ];
#pragma unroll
    for (int k = 0; k < C::P3R; ++k)
#pragma unroll
        for (int c = 0; c < V3; ++c)
#pragma unroll
            for (int o = 0; o < RZ; ++o) {
                Zt[k][c][o] = 0.0f;
                Yt[k][c][o] = 0.0f;
            }
    const int dzlo = SPLIT ? sp.dzlo : -RS, dzhi = SPLIT ? sp.dzhi : RS;
    for (int dz = dzlo; dz <= dzhi; ++dz) {
        float Zd[C::P3R][V3][RZ], Yd[C::P3R][V3][RZ];
#pragma unroll
        for (int k = 0; k < C::P3R; ++k)
#pragma unroll
            for (int c = 0; c < V3; ++c)
#pragma unroll
                for (int o = 0; o < RZ; ++o) {
                    Zd[k][c][o] = 0.0f;
                    Yd[k][c][o] = 0.0f;
                }
        for (int dy = -RS; dy <= RS; ++dy) {
            // scalar offsets are unsigned in the buffer address: count from d = -RS
            const int soffp = (dz + RS) * pxy + (dy + RS) * px;
            float sh[C::ROWS][C::NS];
#pragma unroll
            for (int q = 0; q < C::ROWS; ++q) pad_load_row<C::NS>(rs, rbase[q], 4 * soffp, sh[q]);

            // dx unrolled: the shift into sh is a compile-time register index
            static_for<-RS, RS>([&](auto dxc) {
                constexpr int DX = decltype(dxc)::value;
                rows_phase1<C, C::LP - RP + DX>(cen, sh, SX, sxw, p1ok);
                // candidate samples u~(x + d), in flight across the barriers
                float cand[C::P3R][V3][RZ];
#pragma unroll
                for (int o = 0; o < RZ; ++o)
#pragma unroll
                    for (int k = 0; k < C::P3R; ++k) {
                        if constexpr (V3 > 1) {
                            // the group's samples in one load (dword-aligned: the
                            // shift dx moves the group by single samples)
                            unsigned t[V3];
                            buf_load_v<V3>(rs, cb[k], 4 * (soffp + DX + RS + o * pxy), t);
#pragma unroll
                            for (int c = 0; c < V3; ++c) cand[k][c][o] = __uint_as_float(t[c]);
                        } else {
#pragma unroll
                            for (int c = 0; c < V3; ++c)
                                cand[k][c][o] = pad_load(rs, cb[k] + 4u * c, 4 * (soffp + DX + RS + o * pxy));
                        }
                    }
                lds_barrier();
#pragma unroll
                for (int m = 0; m < C::P2R; ++m)
                    if (p2ok[m]) rows_phase2<C>(SX + p2src[m], SXY + p2dst[m]);
                lds_barrier();
#pragma unroll
                for (int k = 0; k < C::P3R; ++k) {
                    if (p3ok[k]) {
                        float v[V3][RZ + 2 * RP];
#pragma unroll
                        for (int j = 0; j < RZ + 2 * RP; ++j) {
                            const float *src = SXY + p3r[k] + j * C::SYP;
                            if constexpr (V3 == 4) {
                                const float4 t = *reinterpret_cast<const float4 *>(src);
                                v[0][j] = t.x;
                                v[1][j] = t.y;
                                v[2][j] = t.z;
                                v[3][j] = t.w;
                            } else if constexpr (V3 == 2) {
                                const float2 t = *reinterpret_cast<const float2 *>(src);
                                v[0][j] = t.x;
                                v[1][j] = t.y;
                            } else {
                                v[0][j] = src[0];
                            }
                        }
#pragma unroll
                        for (int c = 0; c < V3; ++c) {
                            float S[RZ];
                            box_tree<C::W, RZ>(v[c], S);
#pragma unroll
                            for (int o = 0; o < RZ; ++o) {
                                const float w = exp2_fast(S[o] * kw);
                                Zd[k][c][o] = Zd[k][c][o] + w;
                                Yd[k][c][o] = fmaf(w, cand[k][c][o], Yd[k][c][o]);
                            }
                        }
                    }
                }
            });
        }
        // close the dz block: totals += partials (old total + partial, in
        // this order, in registers or in the output buffers) -- or, split,
        // the partials into their slot
        if constexpr (SPLIT) {
            const long long nvox = (long long)a.nx * a.ny * a.nz;
            gfloat *ys = (gfloat *)sp.slots + 2LL * (dz + RS) * nvox, *zs = ys + nvox;
#pragma unroll
            for (int k = 0; k < C::P3R; ++k) {
                const int it = tid + k * NT;
                const int x = (it % NXG) * V3, y = (it / NXG) % C::TY, zr = it / (NXG * C::TY);
                const int gy = a.y0 + y, gz = a.z0 + zr * RZ;
                if (!p3ok[k] || it >= C::P3N || gy >= a.ny) continue;
#pragma unroll
                for (int c = 0; c < V3; ++c) {
                    const int gx = a.x0 + x + c;
                    if (gx >= a.nx) continue;
#pragma unroll
                    for (int o = 0; o < RZ; ++o) {
                        if (gz + o >= a.nz) continue;
                        const long long oi = gx + (long long)a.nx * (gy + (long long)a.ny * (gz + o));
                        ys[oi] = Yd[k][c][o];
                        zs[oi] = Zd[k][c][o];
                    }
                }
            }
        } else if constexpr (C::REG_TOTALS) {
#pragma unroll
            for (int k = 0; k < C::P3R; ++k)
#pragma unroll
                for (int c = 0; c < V3; ++c)
#pragma unroll
                    for (int o = 0; o < RZ; ++o) {
                        Zt[k][c][o] = Zt[k][c][o] + Zd[k][c][o];
                        Yt[k][c][o] = Yt[k][c][o] + Yd[k][c][o];
                    }
        } else {
            store(Zd, Yd, dz != -RS, dz == RS);
        }
    }
    if constexpr (C::REG_TOTALS && !SPLIT) store(Zt, Yt, false, true);
}

}  // namespace gnlm
