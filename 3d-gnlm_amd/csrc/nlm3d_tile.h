// nlm3d_tile.h -- the 3D NLM forward for one output tile (device code),
// shared by the per-volume launch (nlm3d_fwd.hip) and the persistent
// batched launch (nlm3d_batch.hip), so that both produce identical bits.
//
// Computes, for every voxel x of the tile (reference math:
// 3DNLM/main.cu:52-68 patch SSD over the full offset cube, :107-111 weights,
// :125 normalisation):
//   S_d(x) = sum_{p in [-r_p,r_p]^3} (u~(x+p) - u~(x+d+p))^2
//   w_d(x) = exp(-(S_d(x)/P)/h^2) = exp2(S_d(x) * k)
//   y(x)   = sum_d w_d(x) u~(x+d) / sum_d w_d(x),   d in [-r_s,r_s]^3
//
// One workgroup of 256 threads (4 wave64) per TX x TY x TZ tile, TX*TY = 256:
//   * the tile's input block with its r_p+r_s halo is staged once into LDS,
//     applying reflect-101 on the fly (no padded copy in HBM);
//   * for each offset d the patch box sum is separable:
//       phase 1  rows of s_x (x sums) over the tile grown by r_p in y and z.
//                A thread owns whole rows; the rows' centre samples never
//                change across offsets and stay in registers, only the
//                shifted row is read from LDS.  s_x goes to an LDS buffer
//                laid out x-major with an odd pitch, so the row-wise writes
//                and the column-wise reads of phase 2 are conflict-free;
//       phase 2  each thread owns one (x, y) column: s_xy for the tile's
//                z-extent plus halo, summed from LDS into registers;
//       phase 3  z sums in registers, exp2 weight, Z/Y accumulation in
//                registers (2 TZ accumulators per thread).
//     The s_x buffer is double-buffered: one barrier per offset.
//   * Every partial sum is a direct sum in a fixed order (x, then y, then z,
//     each left to right; squared differences fused with fma), the order the
//     CPU oracle uses: patch distances agree bit for bit with it and with the
//     generic path, and no result depends on the tile shape.
//   * Weights accumulate in two levels: a partial over the 121 (dy, dx)
//     offsets of each dz, then the total over dz.  One running total over all
//     1331 offsets loses ~6e-6 relative to rounding as it grows; the
//     two-level sum is ~10x more accurate (the oracle uses the same order).
#pragma once
#include "gnlm_device.h"

namespace gnlm {

constexpr int kFwdThreads = 256;

template <int RP_, int TX_, int TY_, int TZ_>
struct Fwd3Cfg {
    static constexpr int RP = RP_, TX = TX_, TY = TY_, TZ = TZ_;
    static_assert(TX * TY == kFwdThreads, "phase 2 maps one thread per (x, y) column");
    static constexpr int EX = TX + 2 * RP;  // extended row length (samples)
    static constexpr int EY = TY + 2 * RP;  // rows per plane in phase 1
    static constexpr int EZ = TZ + 2 * RP;  // planes in phase 1 / phase 2
    static constexpr int NR = EY * EZ;      // phase-1 rows per offset
    static constexpr int NRP = NR | 1;      // odd pitch of the s_x buffer
    static constexpr int ROWS = (NR + kFwdThreads - 1) / kFwdThreads;
    static constexpr int SXN = TX * NRP;    // floats per s_x buffer
};

__host__ __device__ constexpr int fwd3_lds_x(int TX, int R) { return (TX + 2 * R) | 1; }

template <class C>
__host__ __device__ constexpr size_t fwd3_lds_floats(int rs)
{
    return (size_t)fwd3_lds_x(C::TX, C::RP + rs) * (C::TY + 2 * (C::RP + rs)) * (C::TZ + 2 * (C::RP + rs)) +
           2 * (size_t)C::SXN;
}

// One tile of one volume.
struct Fwd3Tile {
    InDesc in;
    float *out;
    float *Zout;
    int rs;
    float k;      // exp2 scale (direct-sum variants): w = exp2(S * k)
    float sigma;  // input pre-scale (slab variant): sigma = sqrt(-k)
    int x0, y0, z0;
};

template <class C>
__device__ __forceinline__ void fwd3_tile(const Fwd3Tile &a, float *lds)
{
    constexpr int RP = C::RP;
    constexpr int NT = kFwdThreads;
    const int R = RP + a.rs;
    const int LXr = C::TX + 2 * R;
    const int LX = fwd3_lds_x(C::TX, R);
    const int LY = C::TY + 2 * R;
    const int LZ = C::TZ + 2 * R;
    const int LXY = LX * LY;
    float *U = lds;
    float *SXB = lds + LXY * LZ;
    const int x0 = a.x0, y0 = a.y0, z0 = a.z0;
    const int tid = threadIdx.x;

    // ---- stage the input block (tile + halo R) into LDS ----
    for (int i = tid; i < LXY * LZ; i += NT) {
        const int lx = i % LX;
        const int r = i / LX;
        const int ly = r % LY;
        const int lz = r / LY;
        float v = 0.0f;
        if (lx < LXr) v = load_in(a.in, x0 - R + lx, y0 - R + ly, z0 - R + lz);
        U[i] = v;
    }
    __syncthreads();

    // ---- phase-1 rows owned by this thread; centre samples in registers ----
    float cen[C::ROWS][C::EX];
    int rowbase[C::ROWS];
#pragma unroll
    for (int q = 0; q < C::ROWS; ++q) {
        const int r = tid + q * NT;
        const int ry = r % C::EY, rz = r / C::EY;
        rowbase[q] = (R - RP + rz) * LXY + (R - RP + ry) * LX + (R - RP);
        if (r < C::NR) {
#pragma unroll
            for (int i = 0; i < C::EX; ++i) cen[q][i] = U[rowbase[q] + i];
        }
    }

    // ---- phase-2/3 ownership: column (cx, cy), all TZ planes ----
    const int cx = tid % C::TX;
    const int cy = tid / C::TX;
    float Zacc[C::TZ], Yacc[C::TZ];
#pragma unroll
    for (int z = 0; z < C::TZ; ++z) {
        Zacc[z] = 0.0f;
        Yacc[z] = 0.0f;
    }
    const int candbase = R * LXY + (R + cy) * LX + (R + cx);

    int it = 0;
    for (int dz = -a.rs; dz <= a.rs; ++dz) {
        // per-dz partial sums (two-level accumulation, see the file header)
        float Zd[C::TZ], Yd[C::TZ];
#pragma unroll
        for (int z = 0; z < C::TZ; ++z) {
            Zd[z] = 0.0f;
            Yd[z] = 0.0f;
        }
        for (int dy = -a.rs; dy <= a.rs; ++dy)
            for (int dx = -a.rs; dx <= a.rs; ++dx, ++it) {
                float *SX = SXB + (it & 1) * C::SXN;
                const int doff = dz * LXY + dy * LX + dx;

                // phase 1: s_x rows
#pragma unroll
                for (int q = 0; q < C::ROWS; ++q) {
                    const int r = tid + q * NT;
                    if (r < C::NR) {
                        const float *sh = U + rowbase[q] + doff;
                        float e[C::EX];
#pragma unroll
                        for (int i = 0; i < C::EX; ++i) e[i] = cen[q][i] - sh[i];
#pragma unroll
                        for (int j = 0; j < C::TX; ++j) {
                            float acc = 0.0f;
#pragma unroll
                            for (int i = 0; i <= 2 * RP; ++i) acc = fmaf(e[j + i], e[j + i], acc);
                            SX[j * C::NRP + r] = acc;
                        }
                    }
                }
                __syncthreads();

                // phase 2: s_xy for planes pz (column-wise y sums)
                float sxy[C::EZ];
#pragma unroll
                for (int pz = 0; pz < C::EZ; ++pz) {
                    const float *col = SX + cx * C::NRP + pz * C::EY + cy;
                    float s = 0.0f;
#pragma unroll
                    for (int j = 0; j <= 2 * RP; ++j) s = s + col[j];
                    sxy[pz] = s;
                }

                // phase 3: z sums, weights, accumulation
                const float *cand = U + candbase + doff;
#pragma unroll
                for (int z = 0; z < C::TZ; ++z) {
                    float S = 0.0f;
#pragma unroll
                    for (int kk = 0; kk <= 2 * RP; ++kk) S = S + sxy[z + kk];
                    const float w = exp2_fast(S * a.k);
                    Zd[z] = Zd[z] + w;
                    Yd[z] = fmaf(w, cand[z * LXY], Yd[z]);
                }
            }
#pragma unroll
        for (int z = 0; z < C::TZ; ++z) {
            Zacc[z] = Zacc[z] + Zd[z];
            Yacc[z] = Yacc[z] + Yd[z];
        }
    }

    // ---- store ----
    const int gx = x0 + cx, gy = y0 + cy;
    if (gx < a.in.nx && gy < a.in.ny) {
#pragma unroll
        for (int z = 0; z < C::TZ; ++z) {
            const int gz = z0 + z;
            if (gz < a.in.nz) {
                const long long o = gx + (long long)a.in.nx * (gy + (long long)a.in.ny * gz);
                a.out[o] = Yacc[z] / Zacc[z];
                if (a.Zout) a.Zout[o] = Zacc[z];
            }
        }
    }
}

// ===========================================================================
// Variant "slab": DPP y sums, per-dz input slabs.
//
// Tile 16 x (16 - 2 r_p) x (32 - 2 r_p) voxels, 256 threads.
//   * The input is pre-scaled by sigma = sqrt(log2(e) / (P h^2)) while it is
//     staged, so S' = sigma^2 S and w = exp2(-S') needs no multiply; the
//     candidate values carry the same factor and the final ratio divides it
//     out (y = Y / (sigma Z)).
//   * Offsets are visited dz-major.  For each dz only the 32 input planes the
//     shifted patches touch are staged in LDS (a "slab"), which leaves room
//     for a 16 x 12 x 28 tile with a 7-voxel halo in 160 KiB.  The centre
//     samples of each phase-1 row are loaded once into registers.
//   * phase 1: thread <-> row (y_ext, z_ext), 16 y_ext x 32 z_ext rows, two
//     per thread, z_ext fastest across lanes (conflict-free LDS rows); each
//     x sum is a direct fma chain over the window in registers.
//   * phase 2: thread <-> (x, y_ext) with y_ext = lane % 16, so one 16-lane
//     DPP row holds a whole extended column: the y box sum is formed with
//     row shifts (3 VALU ops for a 5-wide window) and reads no LDS; its
//     input column of 32 planes comes in with eight 128-bit LDS reads.
//   * phase 3: direct z sums over the window in registers; exp2, Z, Y.
//   The y_ext lanes of the halo (r_p at each end of a row) carry no output.
// Every partial sum adds non-negative terms directly (no running add/subtract
// windows): a running window that subtracts a large term that has left it
// keeps an absolute error of that term's ulp, which at high contrast and
// small h moved outputs by 4e-4 relative in testing.  The DPP y sum groups
// its terms differently from the CPU oracle's left-to-right order and the
// input is pre-scaled, so results agree with the oracle to rounding rather
// than bit for bit; they do not depend on where tiles start.
// ===========================================================================

// DPP row shift (within 16-lane rows, zero fill at the row ends).
template <int CTRL>
__device__ __forceinline__ float dpp_row(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
constexpr int kRowShl1 = 0x101, kRowShl2 = 0x102, kRowShl3 = 0x103;
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr3 = 0x113;

// Sum of v over the 2RP+1 lanes centred on this one within its 16-lane row.
// Written so the result does not depend on which way row_shl/row_shr move
// data: every term and its mirror are both included.
template <int RP>
__device__ __forceinline__ float dpp_box(float v)
{
    if constexpr (RP == 0) {
        return v;
    } else if constexpr (RP == 1) {
        return (v + dpp_row<kRowShl1>(v)) + dpp_row<kRowShr1>(v);
    } else if constexpr (RP == 2) {
        const float p = v + dpp_row<kRowShl1>(v);   // v(i) + v(i+s)
        const float q = p + dpp_row<kRowShr2>(p);   // + v(i-2s) + v(i-s)
        return q + dpp_row<kRowShl2>(v);            // + v(i+2s)
    } else {
        float s = v;
        s = s + dpp_row<kRowShl1>(v);
        s = s + dpp_row<kRowShr1>(v);
        s = s + dpp_row<kRowShl2>(v);
        s = s + dpp_row<kRowShr2>(v);
        s = s + dpp_row<kRowShl3>(v);
        s = s + dpp_row<kRowShr3>(v);
        return s;
    }
}

template <int RP_>
struct SlabCfg {
    static constexpr int RP = RP_;
    static constexpr int TX = 16;
    static constexpr int EY = 16;            // one DPP row of lanes
    static constexpr int TY = EY - 2 * RP;
    static constexpr int EZ = 32;            // planes per slab / per phase-1 column
    static constexpr int TZ = EZ - 2 * RP;
    static constexpr int EX = TX + 2 * RP;   // samples per phase-1 row
    static constexpr int SYP = 36;           // y pitch of s_x (== 4 mod 32: b128 reads conflict-free)
    static constexpr int SXP = EY * SYP;     // x pitch of s_x
    static constexpr int SXN = TX * SXP;     // floats in the s_x buffer
    static constexpr int ROWS = EY * EZ / kFwdThreads;  // phase-1 rows per thread (2)
};

__host__ __device__ constexpr size_t slab_lds_floats(int rp, int rs)
{
    return (size_t)(((16 + 2 * (rp + rs)) * (16 + 2 * rs)) | 1) * 32 + (size_t)16 * 16 * 36;
}

template <int RP>
__device__ __forceinline__ void fwd3_slab_tile(const Fwd3Tile &a, float *lds, float sigma)
{
    using C = SlabCfg<RP>;
    constexpr int NT = kFwdThreads;
    const int rs = a.rs;
    const int LX = C::TX + 2 * (RP + rs);
    const int LY = C::EY + 2 * rs;
    const int PP = (LX * LY) | 1;            // odd plane pitch: conflict-free phase-1 rows
    float *SL = lds;                          // slab: 32 planes x LY x LX (pitch PP)
    float *SX = lds + PP * C::EZ;             // s_x[x][y_ext][z_ext]
    const int x0 = a.x0, y0 = a.y0, z0 = a.z0;
    const int tid = threadIdx.x;

    // phase-1 rows: z_ext = tid % 32, y_ext = tid / 32 + 8 q
    const int r_z = tid % C::EZ;
    float cen[C::ROWS][C::EX];
#pragma unroll
    for (int q = 0; q < C::ROWS; ++q) {
        const int r_y = tid / C::EZ + 8 * q;
#pragma unroll
        for (int i = 0; i < C::EX; ++i)
            cen[q][i] = sigma * load_in(a.in, x0 - RP + i, y0 - RP + r_y, z0 - RP + r_z);
    }
    // phase-2/3 ownership: column x = tid / 16, y_ext = tid % 16
    const int cx = tid / C::EY;
    const int cye = tid % C::EY;
    float Zacc[C::TZ], Yacc[C::TZ];
#pragma unroll
    for (int t = 0; t < C::TZ; ++t) {
        Zacc[t] = 0.0f;
        Yacc[t] = 0.0f;
    }

    for (int dz = -rs; dz <= rs; ++dz) {
        float Zd[C::TZ], Yd[C::TZ];  // per-dz partial sums
#pragma unroll
        for (int t = 0; t < C::TZ; ++t) {
            Zd[t] = 0.0f;
            Yd[t] = 0.0f;
        }
        // ---- stage the slab for this dz: planes z0 - RP + dz + [0, 32) ----
        __syncthreads();  // previous slab no longer read
        const int xs = x0 - RP - rs, ys = y0 - RP - rs, zs = z0 - RP + dz;
        for (int i = tid; i < LX * LY * C::EZ; i += NT) {
            const int lx = i % LX;
            const int r = i / LX;
            const int ly = r % LY;
            const int lz = r / LY;
            SL[lx + LX * ly + PP * lz] = sigma * load_in(a.in, xs + lx, ys + ly, zs + lz);
        }
        __syncthreads();

        for (int dy = -rs; dy <= rs; ++dy)
            for (int dx = -rs; dx <= rs; ++dx) {
                // ---- phase 1: direct x sums of squared differences ----
#pragma unroll
                for (int q = 0; q < C::ROWS; ++q) {
                    const int r_y = tid / C::EZ + 8 * q;
                    const float *sh = SL + (rs + dx) + LX * (r_y + dy + rs) + PP * r_z;
                    float e[C::EX];
#pragma unroll
                    for (int i = 0; i < C::EX; ++i) e[i] = cen[q][i] - sh[i];
                    float *dst = SX + C::SYP * r_y + r_z;
#pragma unroll
                    for (int j = 0; j < C::TX; ++j) {
                        float s = e[j] * e[j];
#pragma unroll
                        for (int i = 1; i <= 2 * RP; ++i) s = fmaf(e[j + i], e[j + i], s);
                        dst[j * C::SXP] = s;
                    }
                }
                __syncthreads();

                // ---- phase 2: column of s_x (8 x b128), DPP y box ----
                float sxy[C::EZ];
                const float4 *col = reinterpret_cast<const float4 *>(SX + cx * C::SXP + cye * C::SYP);
#pragma unroll
                for (int m = 0; m < C::EZ / 4; ++m) {
                    const float4 v = col[m];
                    sxy[4 * m + 0] = dpp_box<RP>(v.x);
                    sxy[4 * m + 1] = dpp_box<RP>(v.y);
                    sxy[4 * m + 2] = dpp_box<RP>(v.z);
                    sxy[4 * m + 3] = dpp_box<RP>(v.w);
                }

                // ---- phase 3: direct z sums, weights, accumulation ----
                const float *cand = SL + (cx + dx + RP + rs) + LX * (cye + dy + rs) + PP * RP;
#pragma unroll
                for (int t = 0; t < C::TZ; ++t) {
                    float S = sxy[t];
#pragma unroll
                    for (int k = 1; k <= 2 * RP; ++k) S = S + sxy[t + k];
                    const float w = exp2_fast(-S);
                    Zd[t] = Zd[t] + w;
                    Yd[t] = fmaf(w, cand[t * PP], Yd[t]);
                }
                __syncthreads();  // s_x is rewritten by the next offset's phase 1
            }
#pragma unroll
        for (int t = 0; t < C::TZ; ++t) {
            Zacc[t] = Zacc[t] + Zd[t];
            Yacc[t] = Yacc[t] + Yd[t];
        }
    }

    // ---- store (lanes of the y halo carry no output) ----
    const int gx = x0 + cx, gy = y0 + cye - RP;
    if (cye >= RP && cye < C::EY - RP && gx < a.in.nx && gy < a.in.ny) {
        const float inv_sigma = 1.0f / sigma;
#pragma unroll
        for (int t = 0; t < C::TZ; ++t) {
            const int gz = z0 + t;
            if (gz < a.in.nz) {
                const long long o = gx + (long long)a.in.nx * (gy + (long long)a.in.ny * gz);
                a.out[o] = (Yacc[t] / Zacc[t]) * inv_sigma;
                if (a.Zout) a.Zout[o] = Zacc[t];
            }
        }
    }
}

// ===========================================================================
// Variant "slab2": the slab scheme with the LDS traffic cut by wide accesses.
// Same tile, same arithmetic as "slab" (identical bits), but
//   * the slab is stored z-fastest ([y][x][z], x pitch 36 = 4 mod 32, y pitch
//     = 4 mod 32), staged with one 128-bit write per 4 planes;
//   * a phase-1 thread owns a PAIR of rows (y_ext, z_ext = 2p, 2p+1): every
//     shifted sample pair is one 64-bit read, every s_x pair one 64-bit write;
//   * the 28 candidate values of a phase-3 column are eight 128-bit reads.
// Per tile and offset this is ~210 LDS instructions instead of ~430.
// ===========================================================================
__host__ __device__ constexpr int slab2_yp(int lx) { return 36 * lx + ((4 - 4 * lx) % 32 + 32) % 32; }

__host__ __device__ constexpr size_t slab2_lds_floats(int rp, int rs)
{
    return (size_t)slab2_yp(16 + 2 * (rp + rs)) * (16 + 2 * rs) + (size_t)16 * 16 * 36;
}

template <int RP>
__device__ __forceinline__ void fwd3_slab2_tile(const Fwd3Tile &a, float *lds, float sigma)
{
    using C = SlabCfg<RP>;
    constexpr int NT = kFwdThreads;
    constexpr int XP = 36;                    // x pitch of the slab (floats)
    const int rs = a.rs;
    const int LX = C::TX + 2 * (RP + rs);
    const int LY = C::EY + 2 * rs;
    const int YP = slab2_yp(LX);              // y pitch of the slab, == 4 mod 32
    float *SL = lds;                          // slab [ly][lx][lz], lz in [0, 32)
    float *SX = lds + YP * LY;                // s_x[x][y_ext][z_ext]
    const int x0 = a.x0, y0 = a.y0, z0 = a.z0;
    const int tid = threadIdx.x;

    // phase-1 rows: y_ext = tid / 16, z_ext = 2 * (tid % 16) + {0, 1}
    const int p_y = tid / 16;
    const int p_z = 2 * (tid % 16);
    float cen0[C::EX], cen1[C::EX];
#pragma unroll
    for (int i = 0; i < C::EX; ++i) {
        cen0[i] = sigma * load_in(a.in, x0 - RP + i, y0 - RP + p_y, z0 - RP + p_z);
        cen1[i] = sigma * load_in(a.in, x0 - RP + i, y0 - RP + p_y, z0 - RP + p_z + 1);
    }
    // phase-2/3 ownership: column x = tid / 16, y_ext = tid % 16
    const int cx = tid / C::EY;
    const int cye = tid % C::EY;
    float Zacc[C::TZ], Yacc[C::TZ];
#pragma unroll
    for (int t = 0; t < C::TZ; ++t) {
        Zacc[t] = 0.0f;
        Yacc[t] = 0.0f;
    }

    for (int dz = -rs; dz <= rs; ++dz) {
        float Zd[C::TZ], Yd[C::TZ];  // per-dz partial sums
#pragma unroll
        for (int t = 0; t < C::TZ; ++t) {
            Zd[t] = 0.0f;
            Yd[t] = 0.0f;
        }
        // ---- stage the slab: one 128-bit LDS write per (x, y, 4 planes) ----
        __syncthreads();
        const int xs = x0 - RP - rs, ys = y0 - RP - rs, zs = z0 - RP + dz;
        for (int i = tid; i < LX * LY * (C::EZ / 4); i += NT) {
            const int lx = i % LX;
            const int r = i / LX;
            const int ly = r % LY;
            const int q = r / LY;
            float4 v;
            v.x = sigma * load_in(a.in, xs + lx, ys + ly, zs + 4 * q + 0);
            v.y = sigma * load_in(a.in, xs + lx, ys + ly, zs + 4 * q + 1);
            v.z = sigma * load_in(a.in, xs + lx, ys + ly, zs + 4 * q + 2);
            v.w = sigma * load_in(a.in, xs + lx, ys + ly, zs + 4 * q + 3);
            *reinterpret_cast<float4 *>(SL + YP * ly + XP * lx + 4 * q) = v;
        }
        __syncthreads();

        for (int dy = -rs; dy <= rs; ++dy)
            for (int dx = -rs; dx <= rs; ++dx) {
                // ---- phase 1: two rows, direct x sums ----
                {
                    const float *sh = SL + YP * (p_y + dy + rs) + XP * (rs + dx) + p_z;
                    float e0[C::EX], e1[C::EX];
#pragma unroll
                    for (int i = 0; i < C::EX; ++i) {
                        const float2 v = *reinterpret_cast<const float2 *>(sh + XP * i);
                        e0[i] = cen0[i] - v.x;
                        e1[i] = cen1[i] - v.y;
                    }
                    float *dst = SX + C::SYP * p_y + p_z;
#pragma unroll
                    for (int j = 0; j < C::TX; ++j) {
                        float s0 = e0[j] * e0[j], s1 = e1[j] * e1[j];
#pragma unroll
                        for (int i = 1; i <= 2 * RP; ++i) {
                            s0 = fmaf(e0[j + i], e0[j + i], s0);
                            s1 = fmaf(e1[j + i], e1[j + i], s1);
                        }
                        *reinterpret_cast<float2 *>(dst + j * C::SXP) = make_float2(s0, s1);
                    }
                }
                __syncthreads();

                // ---- phase 2: column of s_x (8 x b128), DPP y box ----
                float sxy[C::EZ];
                const float4 *col = reinterpret_cast<const float4 *>(SX + cx * C::SXP + cye * C::SYP);
#pragma unroll
                for (int m = 0; m < C::EZ / 4; ++m) {
                    const float4 v = col[m];
                    sxy[4 * m + 0] = dpp_box<RP>(v.x);
                    sxy[4 * m + 1] = dpp_box<RP>(v.y);
                    sxy[4 * m + 2] = dpp_box<RP>(v.z);
                    sxy[4 * m + 3] = dpp_box<RP>(v.w);
                }

                // ---- phase 3: direct z sums, weights, accumulation ----
                const float4 *cz4 = reinterpret_cast<const float4 *>(
                    SL + YP * (cye + dy + rs) + XP * (cx + dx + RP + rs));
                float cz[C::EZ];
#pragma unroll
                for (int m = 0; m < C::EZ / 4; ++m) {
                    const float4 v = cz4[m];
                    cz[4 * m + 0] = v.x;
                    cz[4 * m + 1] = v.y;
                    cz[4 * m + 2] = v.z;
                    cz[4 * m + 3] = v.w;
                }
#pragma unroll
                for (int t = 0; t < C::TZ; ++t) {
                    float S = sxy[t];
#pragma unroll
                    for (int k = 1; k <= 2 * RP; ++k) S = S + sxy[t + k];
                    const float w = exp2_fast(-S);
                    Zd[t] = Zd[t] + w;
                    Yd[t] = fmaf(w, cz[t + RP], Yd[t]);
                }
                __syncthreads();  // s_x is rewritten by the next offset's phase 1
            }
#pragma unroll
        for (int t = 0; t < C::TZ; ++t) {
            Zacc[t] = Zacc[t] + Zd[t];
            Yacc[t] = Yacc[t] + Yd[t];
        }
    }

    const int gx = x0 + cx, gy = y0 + cye - RP;
    if (cye >= RP && cye < C::EY - RP && gx < a.in.nx && gy < a.in.ny) {
        const float inv_sigma = 1.0f / sigma;
#pragma unroll
        for (int t = 0; t < C::TZ; ++t) {
            const int gz = z0 + t;
            if (gz < a.in.nz) {
                const long long o = gx + (long long)a.in.nx * (gy + (long long)a.in.ny * gz);
                a.out[o] = (Yacc[t] / Zacc[t]) * inv_sigma;
                if (a.Zout) a.Zout[o] = Zacc[t];
            }
        }
    }
}

// Tile variants.
// kVRows / kVRowsShort / kVRowsBig are the padded-copy variants of
// nlm3d_rows.h (rows: r_p 1-4; short: 16 x 12 x 12 tiles, r_p = 2 only; big:
// 16 x (32 - 2 r_p)^2 tiles on 1024 threads, r_p 4-8, per-volume launches only).
enum Fwd3Variant {
    kV16x16x16 = 0,
    kV16x16x8 = 1,
    kVSlab = 2,
    kVSlab2 = 3,
    kVRows = 4,
    kVRowsShort = 5,
    kVRowsBig = 6,
    kNumFwd3Variants = 7
};

__host__ __device__ constexpr bool fwd3_is_rows(int v) { return v == kVRows || v == kVRowsShort; }

__host__ __device__ constexpr bool fwd3_is_slab(int v) { return v == kVSlab || v == kVSlab2; }

// Tile extents of a variant for patch radius rp (kVRows: RowsCfg).
__host__ __device__ constexpr int fwd3_tx(int v, int rp) { return (void)v, (void)rp, 16; }
__host__ __device__ constexpr int fwd3_ty(int v, int rp)
{
    return v == kVRowsBig ? 32 - 2 * rp : (fwd3_is_rows(v) ? 16 - 2 * rp : (fwd3_is_slab(v) ? 16 - 2 * rp : 16));
}
__host__ __device__ constexpr int fwd3_tz(int v, int rp)
{
    return v == kVRowsBig ? 32 - 2 * rp : v == kVRowsShort ? 12
                            : (v == kVRows ? (rp == 1 ? 14 : (rp == 2 ? 28 : (rp == 3 ? 10 : 8)))
                                           : (fwd3_is_slab(v) ? 32 - 2 * rp : (v == kV16x16x8 ? 8 : 16)));
}

}  // namespace gnlm
