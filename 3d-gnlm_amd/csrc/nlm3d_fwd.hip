// nlm3d_fwd.hip -- tiled 3D NLM forward for gfx950.
//
// Computes, for every voxel x of an nx*ny*nz volume (reference math:
// 3DNLM/main.cu:52-68 patch SSD over the full offset cube, :107-111 weights,
// :125 normalisation):
//   S_d(x) = sum_{p in [-r_p,r_p]^3} (u~(x+p) - u~(x+d+p))^2
//   w_d(x) = exp(-(S_d(x)/P)/h^2) = exp2(S_d(x) * k)
//   y(x)   = sum_d w_d(x) u~(x+d) / sum_d w_d(x),   d in [-r_s,r_s]^3
//
// Design (one workgroup of 256 threads = 4 wave64 per 16x16x16 output tile):
//   * the tile's input block with its r_p+r_s halo is staged once into LDS,
//     applying reflect-101 on the fly (no padded copy in HBM);
//   * for each offset d the box sum of squared differences is separable:
//       phase 1  rows of s_x (x-direction sums) over the tile extended by
//                r_p in y and z.  A thread owns whole rows; the row's centre
//                samples never change across offsets and stay in registers,
//                only the shifted row is read from LDS.  Results go to an
//                LDS buffer laid out x-major with an odd row pitch so that
//                both the row-wise writes and the column-wise reads of
//                phase 2 are bank-conflict free;
//       phase 2  each thread owns one (x, y) column: s_xy for the tile's
//                z-extent plus halo, summed from LDS into registers;
//       phase 3  z sums in registers, exp2 weight, Z/Y accumulation in
//                registers (32 accumulators per thread).
//     The s_x buffer is double-buffered so each offset costs one barrier.
//   * Every partial sum is a direct sum in a fixed order (x fastest, then y,
//     then z, each left to right; squared differences fused with fma), the
//     same order the CPU oracle uses, so the patch distances agree bit for
//     bit with it and with the generic path, and results do not depend on
//     where a tile boundary falls.
#include "gnlm_internal.h"

namespace gnlm {
namespace {

constexpr int NT = 256;  // threads per workgroup

template <int RP>
struct G3 {
    static constexpr int TX = 16, TY = 16, TZ = 16;
    static constexpr int EX = TX + 2 * RP;  // extended row length (samples)
    static constexpr int EY = TY + 2 * RP;  // rows per plane in phase 1
    static constexpr int EZ = TZ + 2 * RP;  // planes in phase 1 / phase 2
    static constexpr int NR = EY * EZ;      // phase-1 rows per offset
    static constexpr int NRP = NR | 1;      // odd pitch of the s_x buffer
    static constexpr int ROWS = (NR + NT - 1) / NT;
    static constexpr int SXN = TX * NRP;    // floats per s_x buffer
};

struct Fwd3Args {
    InDesc in;
    float *out;
    float *Zout;
    int rs;
    float k;
    int ntx, nty, ntz;
};

__host__ __device__ constexpr int lds_x(int TX, int R) { return (TX + 2 * R) | 1; }

template <int RP>
size_t fwd3_lds_bytes(int rs)
{
    using G = G3<RP>;
    const int R = RP + rs;
    const size_t u = (size_t)lds_x(G::TX, R) * (G::TY + 2 * R) * (G::TZ + 2 * R);
    return (u + 2 * (size_t)G::SXN) * sizeof(float);
}

template <int RP>
__global__ __launch_bounds__(NT) void nlm3d_fwd_tiled(Fwd3Args a)
{
    using G = G3<RP>;
    extern __shared__ float lds[];
    const int R = RP + a.rs;
    const int LXr = G::TX + 2 * R;
    const int LX = lds_x(G::TX, R);
    const int LY = G::TY + 2 * R;
    const int LZ = G::TZ + 2 * R;
    const int LXY = LX * LY;
    float *U = lds;
    float *SXB = lds + LXY * LZ;

    const int nblocks = a.ntx * a.nty * a.ntz;
    const int t = xcd_tile(blockIdx.x, nblocks);
    const int x0 = (t % a.ntx) * G::TX;
    const int y0 = ((t / a.ntx) % a.nty) * G::TY;
    const int z0 = (t / (a.ntx * a.nty)) * G::TZ;
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
    float cen[G::ROWS][G::EX];
    int rowbase[G::ROWS];
#pragma unroll
    for (int q = 0; q < G::ROWS; ++q) {
        const int r = tid + q * NT;
        const int ry = r % G::EY, rz = r / G::EY;
        rowbase[q] = (R - RP + rz) * LXY + (R - RP + ry) * LX + (R - RP);
        if (r < G::NR) {
#pragma unroll
            for (int i = 0; i < G::EX; ++i) cen[q][i] = U[rowbase[q] + i];
        }
    }

    // ---- phase-2/3 ownership: column (cx, cy), all TZ planes ----
    const int cx = tid % G::TX;
    const int cy = tid / G::TX;
    float Zacc[G::TZ], Yacc[G::TZ];
#pragma unroll
    for (int z = 0; z < G::TZ; ++z) {
        Zacc[z] = 0.0f;
        Yacc[z] = 0.0f;
    }
    const int candbase = R * LXY + (R + cy) * LX + (R + cx);

    int it = 0;
    for (int dz = -a.rs; dz <= a.rs; ++dz)
        for (int dy = -a.rs; dy <= a.rs; ++dy)
            for (int dx = -a.rs; dx <= a.rs; ++dx, ++it) {
                float *SX = SXB + (it & 1) * G::SXN;
                const int doff = dz * LXY + dy * LX + dx;

                // phase 1: s_x rows
#pragma unroll
                for (int q = 0; q < G::ROWS; ++q) {
                    const int r = tid + q * NT;
                    if (r < G::NR) {
                        const float *sh = U + rowbase[q] + doff;
                        float e[G::EX];
#pragma unroll
                        for (int i = 0; i < G::EX; ++i) e[i] = cen[q][i] - sh[i];
#pragma unroll
                        for (int j = 0; j < G::TX; ++j) {
                            float acc = 0.0f;
#pragma unroll
                            for (int i = 0; i <= 2 * RP; ++i) acc = fmaf(e[j + i], e[j + i], acc);
                            SX[j * G::NRP + r] = acc;
                        }
                    }
                }
                __syncthreads();

                // phase 2: s_xy for planes pz (column-wise y sums)
                float sxy[G::EZ];
#pragma unroll
                for (int pz = 0; pz < G::EZ; ++pz) {
                    const float *col = SX + cx * G::NRP + pz * G::EY + cy;
                    float s = 0.0f;
#pragma unroll
                    for (int j = 0; j <= 2 * RP; ++j) s = s + col[j];
                    sxy[pz] = s;
                }

                // phase 3: z sums, weights, accumulation
                const float *cand = U + candbase + doff;
#pragma unroll
                for (int z = 0; z < G::TZ; ++z) {
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
        for (int z = 0; z < G::TZ; ++z) {
            const int gz = z0 + z;
            if (gz < a.in.nz) {
                const long long o = gx + (long long)a.in.nx * (gy + (long long)a.in.ny * gz);
                a.out[o] = Yacc[z] / Zacc[z];
                if (a.Zout) a.Zout[o] = Zacc[z];
            }
        }
    }
}

template <int RP>
hipError_t launch_rp(const InDesc &in, float *out, float *Zout, int rs, float k, hipStream_t s)
{
    using G = G3<RP>;
    Fwd3Args a;
    a.in = in;
    a.out = out;
    a.Zout = Zout;
    a.rs = rs;
    a.k = k;
    a.ntx = (in.nx + G::TX - 1) / G::TX;
    a.nty = (in.ny + G::TY - 1) / G::TY;
    a.ntz = (in.nz + G::TZ - 1) / G::TZ;
    const size_t lds = fwd3_lds_bytes<RP>(rs);
    hipError_t e = hipFuncSetAttribute((const void *)nlm3d_fwd_tiled<RP>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int nblocks = a.ntx * a.nty * a.ntz;
    hipLaunchKernelGGL(nlm3d_fwd_tiled<RP>, dim3(nblocks), dim3(NT), lds, s, a);
    return hipGetLastError();
}

constexpr size_t kLdsLimit = 160 * 1024;

}  // namespace

bool fwd3d_tiled_supported(int r_p, int r_s)
{
    if (r_s < 0) return false;
    switch (r_p) {
    case 0: return fwd3_lds_bytes<0>(r_s) <= kLdsLimit;
    case 1: return fwd3_lds_bytes<1>(r_s) <= kLdsLimit;
    case 2: return fwd3_lds_bytes<2>(r_s) <= kLdsLimit;
    case 3: return fwd3_lds_bytes<3>(r_s) <= kLdsLimit;
    default: return false;
    }
}

hipError_t fwd3d_tiled_launch(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k,
                              hipStream_t s)
{
    switch (r_p) {
    case 0: return launch_rp<0>(in, out, Zout, r_s, k, s);
    case 1: return launch_rp<1>(in, out, Zout, r_s, k, s);
    case 2: return launch_rp<2>(in, out, Zout, r_s, k, s);
    case 3: return launch_rp<3>(in, out, Zout, r_s, k, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace gnlm
