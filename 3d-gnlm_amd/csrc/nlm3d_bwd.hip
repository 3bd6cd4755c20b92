// nlm3d_bwd.hip -- hand-written adjoint of the 3D NLM forward (gfx950).
//
// Forward (3DNLM/main.cu:52-68,107-111,125 restated):
//   S_d(x) = sum_p (u~(x+p) - u~(x+d+p))^2,  w_d = exp(-c S_d),  c = 1/(P h^2)
//   Z(x) = sum_d w_d(x),   y(x) = sum_d w_d(x) u~(x+d) / Z(x)
// Given g = dL/dy, with G = g/Z and GY = g*y/Z (both zero outside the
// volume), the gradient with respect to the reflect-101 extension u~ is, in
// gather form (every term evaluated around the receiving position z):
//   grad(z) = sum_d [ G(z+d) w_d(z) + 2 boxE_d(z) (u~(z) - u~(z+d)) ]
//   E_d(x)  = -c w_d(x) [ G(x) u~(x+d) - GY(x) + G(x+d) u~(x) - GY(x+d) ]
//   boxE_d  = E_d summed over the patch,
// obtained from the scatter form of the chain rule by substituting d -> -d in
// the terms that land on x+d and using S_{-d}(x+d) = S_d(x).  Then
//   dL/dh = (2c/h) sum_{x in volume} sum_d w_d(x) S_d(x) (G(x) u~(x+d) - GY(x))
// and dL/du(i) = sum of grad(z) over the positions z of the extended domain
// [-R, n+R) (R = r_p + r_s) that reflect-101 maps onto voxel i.
//
// Nothing of size voxels x offsets is stored: patch distances are recomputed
// per offset, and only y and Z from the forward are read (as G and GY,
// formed once by a prologue kernel).
//
// Kernel: one workgroup (256 threads) per gradient tile T of
// (16 - 2 r_p) x (16 - 2 r_p) x 8 positions of the extended domain.  The tile
// grown by r_p, Q = 16 x 16 x (8 + 2 r_p), is where S, w and E live, so one
// thread owns one (x, y) column of Q.  The tile's u~ block (halo 2 r_p + r_s)
// is staged in LDS once; G, GY and u~ of a thread's own Q column sit in
// registers; G and GY at shifted positions are read through the L1/L2 cache.
// Per offset:
//   S1  x sums of squared differences on rows of Q grown by r_p in y and z
//       (one row per thread, centre samples in registers)        -> LDS
//   S2  per Q column: y sums, z sums -> S; w = exp2(S k); E; dL/dh terms
//       on the tile's own volume voxels                           -> LDS (E, w)
//   E1  x sums of E on rows of Q, outputs on T's x range          -> LDS
//   E2  per T column: y and z sums of E, gradient accumulation in registers
// with three barriers.  All sums are direct, in the CPU oracle's order, so S
// matches the forward's direct-sum variants exactly; every thread owns fixed
// positions and the dL/dh partials are reduced in a fixed order, so the
// results are deterministic.
#include "gnlm_internal.h"

namespace gnlm {
namespace {

constexpr int NT = 256;

template <int RP_>
struct B2 {
    static constexpr int RP = RP_;
    static constexpr int QX = 16, QY = 16;            // grown tile, x and y
    static constexpr int TX = QX - 2 * RP, TY = QY - 2 * RP;
    static constexpr int TZ = 8;
    static constexpr int QZ = TZ + 2 * RP;
    static constexpr int DX = QX + 2 * RP;            // samples per S1 row
    static constexpr int DY = QY + 2 * RP, DZ = QZ + 2 * RP;
    static constexpr int NR1 = DY * DZ;               // S1 rows
    static constexpr int ROWS1 = (NR1 + NT - 1) / NT;
    static constexpr int NR1P = NR1 | 1;              // odd pitch of the s_x buffer
    static constexpr int NRE = QY * QZ;               // E1 rows
    static constexpr int NREP = NRE | 1;
    static constexpr int QN = QX * QY * QZ;           // floats in E / W buffers
};

struct BwdArgs {
    const float *u;     // unpadded input
    const float *GG;    // (G, GY) interleaved per voxel
    float *gext;        // gradient on the extended domain, (n+2R)^3
    float *dh_part;     // one partial per block
    int nx, ny, nz;
    int rs;
    float k;            // exp2 scale: w = exp2(S * k)
    float negc;         // -c
    int nbx, nby, nbz;  // blocks per axis over the extended domain
};

__device__ __forceinline__ bool in_vol(int x, int y, int z, int nx, int ny, int nz)
{
    return (unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny && (unsigned)z < (unsigned)nz;
}

__device__ __forceinline__ float2 load_gg(const BwdArgs &a, int x, int y, int z)
{
    if (!in_vol(x, y, z, a.nx, a.ny, a.nz)) return make_float2(0.0f, 0.0f);
    const long long o = x + (long long)a.nx * (y + (long long)a.ny * z);
    return *reinterpret_cast<const float2 *>(a.GG + 2 * o);
}

template <int RP>
size_t bwd_lds_floats(int rs)
{
    using C = B2<RP>;
    const int HU = 2 * RP + rs;  // u~ halo around T
    const size_t lu = (size_t)((C::TX + 2 * HU) | 1) * (C::TY + 2 * HU) * (C::TZ + 2 * HU);
    size_t n = lu + (size_t)C::QX * C::NR1P + 2 * (size_t)C::QN + (size_t)C::TX * C::NREP;
    if (n < NT) n = NT;
    return n;
}

template <int RP>
__global__ __launch_bounds__(NT) void nlm3d_bwd_tiled(BwdArgs a)
{
    using C = B2<RP>;
    extern __shared__ float lds[];
    const int rs = a.rs;
    const int HU = 2 * RP + rs;
    const int LXr = C::TX + 2 * HU;
    const int LX = LXr | 1, LY = C::TY + 2 * HU, LZ = C::TZ + 2 * HU;
    const int LXY = LX * LY;
    float *U = lds;                               // u~ block around T, halo HU
    float *SX = U + LXY * LZ;                     // S1 output: [x_q][row] pitch NR1P
    float *EB = SX + C::QX * C::NR1P;             // E on Q: [z][y][x]
    float *WB = EB + C::QN;                       // w on Q: [z][y][x]
    float *EX = WB + C::QN;                       // E1 output: [x_t][row] pitch NREP

    const int R = RP + rs;
    const int b = blockIdx.x;
    // T origin in volume coordinates; the extended domain starts at -R
    const int X0 = (b % a.nbx) * C::TX - R;
    const int Y0 = ((b / a.nbx) % a.nby) * C::TY - R;
    const int Z0 = (b / (a.nbx * a.nby)) * C::TZ - R;
    const int tid = threadIdx.x;

    // ---- stage u~ around T ----
    for (int i = tid; i < LXY * LZ; i += NT) {
        const int lx = i % LX, r = i / LX, ly = r % LY, lz = r / LY;
        float v = 0.0f;
        if (lx < LXr) {
            const int x = reflect101(X0 - HU + lx, a.nx), y = reflect101(Y0 - HU + ly, a.ny),
                      z = reflect101(Z0 - HU + lz, a.nz);
            v = a.u[x + (long long)a.nx * (y + (long long)a.ny * z)];
        }
        U[i] = v;
    }
    __syncthreads();
    // LDS index of T-local (x, y, z) (T-local 0 = X0)
    auto uidx = [&](int x, int y, int z) { return (x + HU) + LX * ((y + HU) + LY * (z + HU)); };

    // ---- S1 rows: D-row (y_q in [-RP, QY+RP), z_q in [-RP, QZ+RP)) of Q;
    //      Q-local coordinate q = T-local + RP ----
    float cen[C::ROWS1][C::DX];
    int rowbase[C::ROWS1];
#pragma unroll
    for (int j = 0; j < C::ROWS1; ++j) {
        const int r = tid + j * NT;
        const int ry = r % C::DY, rz = r / C::DY;
        // samples x_q in [-RP, QX+RP) -> T-local x = x_q - RP in [-2RP, ...)
        rowbase[j] = uidx(-2 * RP, ry - 2 * RP, rz - 2 * RP);
        if (r < C::NR1) {
#pragma unroll
            for (int i = 0; i < C::DX; ++i) cen[j][i] = U[rowbase[j] + i];
        }
    }

    // ---- S2 ownership: Q column (qx, qy); own G, GY, u~ for z_q in [0, QZ) ----
    const int qx = tid % C::QX, qy = tid / C::QX;
    const int cxT = qx - RP, cyT = qy - RP;  // T-local coordinates of the column
    float Gc[C::QZ], GYc[C::QZ], uc[C::QZ];
#pragma unroll
    for (int z = 0; z < C::QZ; ++z) {
        const int zT = z - RP;
        const float2 g = load_gg(a, X0 + cxT, Y0 + cyT, Z0 + zT);
        Gc[z] = g.x;
        GYc[z] = g.y;
        uc[z] = U[uidx(cxT, cyT, zT)];
    }
    // is the column's position x of T (and of the volume) for dL/dh?
    const bool col_in_T = (cxT >= 0 && cxT < C::TX && cyT >= 0 && cyT < C::TY);

    // ---- E2 ownership: T column (tx, ty) if tid < TX*TY ----
    const bool e2_active = tid < C::TX * C::TY;
    const int tx = tid % C::TX, ty = tid / C::TX;
    float gacc[C::TZ];
#pragma unroll
    for (int z = 0; z < C::TZ; ++z) gacc[z] = 0.0f;
    float dh = 0.0f;

    for (int dz = -rs; dz <= rs; ++dz)
        for (int dy = -rs; dy <= rs; ++dy)
            for (int dx = -rs; dx <= rs; ++dx) {
                const int du = dx + LX * (dy + LY * dz);
                // ---- S1: x sums of squared differences ----
#pragma unroll
                for (int j = 0; j < C::ROWS1; ++j) {
                    const int r = tid + j * NT;
                    if (r < C::NR1) {
                        const float *sh = U + rowbase[j] + du;
                        float e[C::DX];
#pragma unroll
                        for (int i = 0; i < C::DX; ++i) e[i] = cen[j][i] - sh[i];
#pragma unroll
                        for (int xq = 0; xq < C::QX; ++xq) {
                            float acc = 0.0f;
#pragma unroll
                            for (int i = 0; i <= 2 * RP; ++i) acc = fmaf(e[xq + i], e[xq + i], acc);
                            SX[xq * C::NR1P + r] = acc;
                        }
                    }
                }
                __syncthreads();

                // ---- S2: y and z sums on the Q column, w, E, dL/dh ----
                {
                    float sxy[C::DZ];
#pragma unroll
                    for (int pz = 0; pz < C::DZ; ++pz) {
                        const float *col = SX + qx * C::NR1P + pz * C::DY + qy;
                        float s = 0.0f;
#pragma unroll
                        for (int j = 0; j <= 2 * RP; ++j) s = s + col[j];
                        sxy[pz] = s;
                    }
                    const float *ush = U + uidx(cxT + dx, cyT + dy, -RP + dz);
#pragma unroll
                    for (int z = 0; z < C::QZ; ++z) {
                        float S = 0.0f;
#pragma unroll
                        for (int kk = 0; kk <= 2 * RP; ++kk) S = S + sxy[z + kk];
                        const float w = exp2_fast(S * a.k);
                        const float ud = ush[z * LXY];
                        const float2 gd = load_gg(a, X0 + cxT + dx, Y0 + cyT + dy, Z0 + z - RP + dz);
                        const float t1 = Gc[z] * ud - GYc[z];
                        const float t2 = gd.x * uc[z] - gd.y;
                        const int qi = qx + C::QX * (qy + C::QY * z);
                        WB[qi] = w;
                        EB[qi] = a.negc * w * (t1 + t2);
                        const int zT = z - RP;
                        if (col_in_T && zT >= 0 && zT < C::TZ &&
                            in_vol(X0 + cxT, Y0 + cyT, Z0 + zT, a.nx, a.ny, a.nz))
                            dh = fmaf(w * S, t1, dh);
                    }
                }
                __syncthreads();

                // ---- E1: x sums of E over rows of Q, outputs on T's x range ----
                for (int r = tid; r < C::NRE; r += NT) {
                    const float *row = EB + C::QX * r;  // row r = (y_q, z_q), x_q fastest
                    float ev[C::QX];
#pragma unroll
                    for (int i = 0; i < C::QX; ++i) ev[i] = row[i];
#pragma unroll
                    for (int xt = 0; xt < C::TX; ++xt) {
                        float s = 0.0f;
#pragma unroll
                        for (int i = 0; i <= 2 * RP; ++i) s = s + ev[xt + i];
                        EX[xt * C::NREP + r] = s;
                    }
                }
                __syncthreads();

                // ---- E2: y and z sums of E on the T column, accumulate ----
                if (e2_active) {
                    float exy[C::QZ];
#pragma unroll
                    for (int zq = 0; zq < C::QZ; ++zq) {
                        const float *col = EX + tx * C::NREP + zq * C::QY + ty;
                        float s = 0.0f;
#pragma unroll
                        for (int j = 0; j <= 2 * RP; ++j) s = s + col[j];
                        exy[zq] = s;
                    }
                    const float *u0 = U + uidx(tx, ty, 0);
#pragma unroll
                    for (int z = 0; z < C::TZ; ++z) {
                        float bE = 0.0f;
#pragma unroll
                        for (int kk = 0; kk <= 2 * RP; ++kk) bE = bE + exy[z + kk];
                        const float w = WB[(tx + RP) + C::QX * ((ty + RP) + C::QY * (z + RP))];
                        const float2 gd = load_gg(a, X0 + tx + dx, Y0 + ty + dy, Z0 + z + dz);
                        const float uz = u0[z * LXY], uzd = u0[z * LXY + du];
                        gacc[z] = fmaf(gd.x, w, gacc[z]);
                        gacc[z] = fmaf(2.0f * bE, uz - uzd, gacc[z]);
                    }
                }
                // The next offset's S1 writes SX only (read in S2 before the
                // second barrier); EB/WB are rewritten in S2 after the next
                // S1 barrier, EX in E1 after the next S2 barrier.
            }

    // ---- write the gradient of the tile's extended-domain positions ----
    if (e2_active) {
        const long long ex = a.nx + 2LL * R, ey = a.ny + 2LL * R, ez = a.nz + 2LL * R;
        const long long x = X0 + tx + R, y = Y0 + ty + R;
#pragma unroll
        for (int z = 0; z < C::TZ; ++z) {
            const long long zz = Z0 + z + R;
            if (x < ex && y < ey && zz < ez) a.gext[x + ex * (y + ey * zz)] = gacc[z];
        }
    }
    // ---- deterministic block reduction of dL/dh ----
    __syncthreads();
    float *red = lds;
    red[tid] = dh;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = red[tid] + red[tid + s];
        __syncthreads();
    }
    if (tid == 0) a.dh_part[b] = red[0];
}

// G = g / Z and GY = g * y / Z, interleaved.
__global__ void make_gg(const float *g, const float *y, const float *Z, float *GG, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float G = g[i] / Z[i];
        GG[2 * i] = G;
        GG[2 * i + 1] = G * y[i];
    }
}

// dL/du(i) = sum of gext over the preimages of i under reflect-101 within the
// extended domain [-R, n+R) on each axis.
__device__ int preimages(int i, int n, int R, int *out)
{
    int cnt = 0;
    if (n == 1) {
        for (int j = -R; j < 1 + R; ++j) out[cnt++] = j;
        return cnt;
    }
    const int per = 2 * (n - 1);
    // j = i + k*per  or  j = -i + k*per
    const int kmin = (-R - n) / per - 1, kmax = (n + R + n) / per + 1;
    for (int k = kmin; k <= kmax; ++k) {
        int j1 = i + k * per;
        if (j1 >= -R && j1 < n + R) out[cnt++] = j1;
        int j2 = -i + k * per;
        if (j2 != j1 && j2 >= -R && j2 < n + R && reflect101(j2, n) == i) {
            bool dup = false;
            for (int q = 0; q < cnt; ++q) dup |= (out[q] == j2);
            if (!dup) out[cnt++] = j2;
        }
    }
    return cnt;
}

constexpr int MAXPRE = 64;

__global__ void fold_grad(const float *gext, float *gu, int nx, int ny, int nz, int R)
{
    const long long n = (long long)nx * ny * nz;
    const long long ex = nx + 2LL * R, ey = ny + 2LL * R;
    for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n;
         o += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(o % nx), y = (int)((o / nx) % ny), z = (int)(o / ((long long)nx * ny));
        int px[MAXPRE], py[MAXPRE], pz[MAXPRE];
        const int cx = preimages(x, nx, R, px), cy = preimages(y, ny, R, py), cz = preimages(z, nz, R, pz);
        float s = 0.0f;
        for (int k = 0; k < cz; ++k)
            for (int j = 0; j < cy; ++j)
                for (int i = 0; i < cx; ++i)
                    s = s + gext[(px[i] + R) + ex * ((py[j] + R) + ey * (long long)(pz[k] + R))];
        gu[o] = s;
    }
}

// Preimages of i when R < n - 1 (one reflection at each end): i itself, -i for
// 1 <= i <= R, then 2(n-1) - i for n-1-R <= i <= n-2.  This is the order in
// which preimages() lists them, so both folds add the same terms in the same
// order.
__device__ inline int preimages_near(int i, int n, int R, int *out)
{
    const bool neg = i >= 1 && i <= R, mir = i <= n - 2 && i >= n - 1 - R;
    out[0] = i;
    out[1] = neg ? -i : 2 * (n - 1) - i;
    out[2] = 2 * (n - 1) - i;
    return 1 + (int)neg + (int)mir;
}

// fold_grad for R < n - 1 on every axis and an extended plane of fewer than
// 2^31 positions.  Blocks stride over the z planes (blockIdx.y) and over the
// x-y plane in 32-bit indices; most voxels have one preimage (one load, one
// store), and the preimage lists (at most 3 per axis) stay in registers.
__global__ __launch_bounds__(256) void fold_grad_near(const float *__restrict__ gext, float *__restrict__ gu, int nx,
                                                      int ny, int nz, int R)
{
    const unsigned plane = (unsigned)nx * (unsigned)ny;
    const unsigned ex = (unsigned)(nx + 2 * R), eplane = ex * (unsigned)(ny + 2 * R);
    for (int z = blockIdx.y; z < nz; z += gridDim.y) {
        int pz[3];
        const int cz = preimages_near(z, nz, R, pz);
        float *gz = gu + (long long)z * plane;
        for (unsigned p = blockIdx.x * 256u + threadIdx.x; p < plane; p += gridDim.x * 256u) {
            const int y = (int)(p / (unsigned)nx), x = (int)(p - (unsigned)y * (unsigned)nx);
            int px[3], py[3];
            const int cx = preimages_near(x, nx, R, px), cy = preimages_near(y, ny, R, py);
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k < cz) {
                    const float *gk = gext + (long long)(pz[k] + R) * eplane;
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        if (j < cy) {
                            const float *gj = gk + (unsigned)(py[j] + R) * ex;
#pragma unroll
                            for (int i = 0; i < 3; ++i)
                                if (i < cx) s = s + gj[px[i] + R];
                        }
                    }
                }
            }
            gz[p] = s;
        }
    }
}

// Fixed-order reduction of the per-block dL/dh partials, times 2c/h.
__global__ void reduce_dh(const float *part, int n, float scale, float *out)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s = s + part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

size_t bwd_lds_bytes_rp(int rp, int rs)
{
    switch (rp) {
    case 0: return bwd_lds_floats<0>(rs) * sizeof(float);
    case 1: return bwd_lds_floats<1>(rs) * sizeof(float);
    case 2: return bwd_lds_floats<2>(rs) * sizeof(float);
    case 3: return bwd_lds_floats<3>(rs) * sizeof(float);
    default: return (size_t)-1;
    }
}

template <int RP>
hipError_t launch_bwd(BwdArgs a, size_t lds, hipStream_t s)
{
    using C = B2<RP>;
    a.nbx = (int)((a.nx + 2LL * (RP + a.rs) + C::TX - 1) / C::TX);
    a.nby = (int)((a.ny + 2LL * (RP + a.rs) + C::TY - 1) / C::TY);
    a.nbz = (int)((a.nz + 2LL * (RP + a.rs) + C::TZ - 1) / C::TZ);
    hipError_t e = hipFuncSetAttribute((const void *)nlm3d_bwd_tiled<RP>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nlm3d_bwd_tiled<RP>, dim3(a.nbx * a.nby * a.nbz), dim3(NT), lds, s, a);
    return hipGetLastError();
}

template <int RP>
int bwd_blocks(int nx, int ny, int nz, int rs)
{
    using C = B2<RP>;
    const long long R = RP + rs;
    return (int)(((nx + 2 * R + C::TX - 1) / C::TX) * ((ny + 2 * R + C::TY - 1) / C::TY) *
                 ((nz + 2 * R + C::TZ - 1) / C::TZ));
}

}  // namespace

bool bwd3d_block_supported(int r_p, int r_s)
{
    return r_p >= 0 && r_p <= 3 && r_s >= 0 && bwd_lds_bytes_rp(r_p, r_s) <= kLdsLimitBytes;
}

bool bwd3d_supported(int r_p, int r_s) { return bwd3d_rows_supported(r_p, r_s) || bwd3d_block_supported(r_p, r_s); }

int bwd3d_route(int nx, int ny, int nz, int r_p, int r_s, int *kernel)
{
    // the rows kernels address their padded copies with 32-bit offsets; the
    // LDS-block tile indexes in 64 bits and takes the volumes they cannot
    if (bwd3d_rows_supported(r_p, r_s)) {
        const bool sym = bwd3d_rows_symmetric_form();
        if (bwd3d_rows_fits(nx, ny, nz, r_p, r_s, sym)) {
            *kernel = sym ? GNLM_KERNEL_BWD_ROWS_SYMMETRIC : GNLM_KERNEL_BWD_ROWS_GATHER;
            return GNLM_OK;
        }
    }
    if (bwd3d_block_supported(r_p, r_s)) {
        *kernel = GNLM_KERNEL_BWD_BLOCK;
        return GNLM_OK;
    }
    // any radii: per-offset passes through HBM (the dL/dh partials are
    // reduced with an int count of voxels)
    if ((long long)nx * ny * nz < (1LL << 31)) {
        *kernel = GNLM_KERNEL_BWD_GENERIC;
        return GNLM_OK;
    }
    return GNLM_EUNSUPPORTED;
}

hipError_t bwd3d_fold_launch(const float *gext, float *gu, int nx, int ny, int nz, int R, hipStream_t s)
{
    const long long eplane = (nx + 2LL * R) * (ny + 2LL * R);
    if (R < nx - 1 && R < ny - 1 && R < nz - 1 && eplane < (1LL << 31)) {
        const long long plane = (long long)nx * ny;
        const unsigned gx = (unsigned)std::min<long long>((plane + 255) / 256, 1024);
        const unsigned gy = (unsigned)std::min<long long>(nz, 65535);
        hipLaunchKernelGGL(fold_grad_near, dim3(gx, gy), dim3(256), 0, s, gext, gu, nx, ny, nz, R);
        return hipGetLastError();
    }
    const long long nv = (long long)nx * ny * nz;
    long long gblk = (nv + 255) / 256;
    if (gblk > 65536) gblk = 65536;
    hipLaunchKernelGGL(fold_grad, dim3((unsigned)gblk), dim3(256), 0, s, gext, gu, nx, ny, nz, R);
    return hipGetLastError();
}

hipError_t bwd3d_reduce_dh_launch(const float *part, int n, float scale, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_dh, dim3(1), dim3(256), 0, s, part, n, scale, out);
    return hipGetLastError();
}

hipError_t bwd3d_launch(const float *u, const float *y, const float *Z, const float *g, int nx, int ny, int nz,
                        int rp, int rs, float h, float *gu, float *gh, hipStream_t s)
{
    int kernel = 0;
    if (bwd3d_route(nx, ny, nz, rp, rs, &kernel) != GNLM_OK) return hipErrorNotSupported;
    if (kernel == GNLM_KERNEL_BWD_GENERIC) return bwd3d_generic_launch(u, y, Z, g, nx, ny, nz, rp, rs, h, gu, gh, s);
    if (kernel != GNLM_KERNEL_BWD_BLOCK) return bwd3d_rows_launch(u, y, Z, g, nx, ny, nz, rp, rs, h, gu, gh, s);
    const int R = rp + rs;
    const long long nv = (long long)nx * ny * nz;
    const long long ext = (nx + 2LL * R) * (ny + 2LL * R) * (nz + 2LL * R);
    const int nblocks = rp == 0   ? bwd_blocks<0>(nx, ny, nz, rs)
                        : rp == 1 ? bwd_blocks<1>(nx, ny, nz, rs)
                        : rp == 2 ? bwd_blocks<2>(nx, ny, nz, rs)
                                  : bwd_blocks<3>(nx, ny, nz, rs);
    BwdArgs a;
    a.u = u;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.rs = rs;
    const int P = (2 * rp + 1) * (2 * rp + 1) * (2 * rp + 1);
    const double c = 1.0 / ((double)P * (double)h * (double)h);
    a.k = (float)weight_scale(P, h);
    a.negc = (float)(-c);
    float *ws = nullptr;
    hipError_t e = hipMallocAsync((void **)&ws, sizeof(float) * (size_t)(ext + nblocks + 2 * nv + 2), s);
    if (e != hipSuccess) return e;
    a.gext = ws;
    a.dh_part = ws + ext;
    float *GG = a.dh_part + nblocks;
    // keep the interleaved (G, GY) 8-byte aligned
    if (reinterpret_cast<uintptr_t>(GG) & 7) GG += 1;
    a.GG = GG;
    long long gblk = (nv + 255) / 256;
    if (gblk > 65536) gblk = 65536;
    hipLaunchKernelGGL(make_gg, dim3((unsigned)gblk), dim3(256), 0, s, g, y, Z, GG, nv);
    const size_t lds = bwd_lds_bytes_rp(rp, rs);
    if (rp == 0)
        e = launch_bwd<0>(a, lds, s);
    else if (rp == 1)
        e = launch_bwd<1>(a, lds, s);
    else if (rp == 2)
        e = launch_bwd<2>(a, lds, s);
    else
        e = launch_bwd<3>(a, lds, s);
    if (e == hipSuccess) {
        e = bwd3d_fold_launch(a.gext, gu, nx, ny, nz, R, s);
        if (e == hipSuccess && gh) {
            hipLaunchKernelGGL(reduce_dh, dim3(1), dim3(256), 0, s, (const float *)a.dh_part, nblocks,
                               (float)(2.0 * c / (double)h), gh);
            e = hipGetLastError();
        }
    }
    hipError_t e2 = hipFreeAsync(ws, s);
    return e != hipSuccess ? e : e2;
}

}  // namespace gnlm
