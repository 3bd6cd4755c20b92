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
    float k;
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
    for (int dz = -a.rs; dz <= a.rs; ++dz)
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
                    Zacc[z] = Zacc[z] + w;
                    Yacc[z] = fmaf(w, cand[z * LXY], Yacc[z]);
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

// Tile variants: (TX, TY, TZ).  TX * TY = 256 always.
enum Fwd3Variant { kV16x16x16 = 0, kV16x16x8 = 1, kNumFwd3Variants = 2 };

__host__ __device__ constexpr int fwd3_tz(int v) { return v == kV16x16x8 ? 8 : 16; }

}  // namespace gnlm
