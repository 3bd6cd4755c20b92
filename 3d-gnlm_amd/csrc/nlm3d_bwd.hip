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
// that reflect-101 maps onto voxel i.
//
// Nothing of size voxels x offsets is stored: patch distances are recomputed
// per offset from the input, and only y and Z from the forward are read.
//
// Kernel structure: one workgroup (256 threads) per 16x8x8 block of gradient
// positions of the extended domain [-R, n+R)^3 (R = r_p + r_s).  The block's
// u~ (halo 2 r_p + r_s) and (G, GY) (halo r_p + r_s) are staged in LDS once;
// for each offset six short phases separated by barriers form S on the block
// grown by r_p, the weights and E, the patch box of E, and accumulate the
// gradient in registers (4 positions per thread).  All sums are direct, in
// the canonical order (x, then y, then z), so S matches the forward exactly.
// Each thread owns fixed positions for the whole offset loop and the per-
// block dL/dh partials are reduced in a fixed order: results are
// deterministic.
#include "gnlm_internal.h"

namespace gnlm {
namespace {

constexpr int NT = 256;
constexpr int BX = 16, BY = 8, BZ = 8;   // gradient positions per block
constexpr int PER = BX * BY * BZ / NT;    // positions per thread (4)

struct BwdArgs {
    const float *u;     // unpadded input
    const float *y;     // forward output
    const float *Z;     // forward normaliser
    const float *g;     // dL/dy
    float *gext;        // gradient on the extended domain, (n+2R)^3
    float *dh_part;     // one partial per block
    int nx, ny, nz;
    int rp, rs;
    float k;            // exp2 scale: w = exp2(S * k)
    float negc;         // -c
    int nbx, nby, nbz;  // blocks per axis over the extended domain
};

struct BwdLds {
    // region sizes (elements per axis)
    int ux, uy, uz;     // u~ block: B + 2(2rp+rs)
    int gx, gy, gz;     // (G, GY) block: B + 2(rp+rs)
};

__device__ __forceinline__ bool in_vol(int x, int y, int z, int nx, int ny, int nz)
{
    return (unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny && (unsigned)z < (unsigned)nz;
}

__global__ __launch_bounds__(NT) void nlm3d_bwd_tiled(BwdArgs a)
{
    extern __shared__ float lds[];
    const int rp = a.rp, rs = a.rs;
    const int HU = 2 * rp + rs, HG = rp + rs;
    const int ux = BX + 2 * HU, uy = BY + 2 * HU, uz = BZ + 2 * HU;
    const int gx = BX + 2 * HG, gy = BY + 2 * HG, gz = BZ + 2 * HG;
    // work regions (grown block): Q = B + 2rp per axis
    const int qx = BX + 2 * rp, qy = BY + 2 * rp, qz = BZ + 2 * rp;
    const int dy2 = BY + 4 * rp, dz2 = BZ + 4 * rp;  // rows needed by the x-stage

    float *U = lds;                               // ux*uy*uz
    float *GG = U + ux * uy * uz;                 // 2*gx*gy*gz  (G, GY interleaved)
    float *SX = GG + 2 * gx * gy * gz;            // qx * dy2 * dz2
    float *SXY = SX + qx * dy2 * dz2;             // qx * qy * dz2
    float *W = SXY + qx * qy * dz2;               // qx*qy*qz
    float *EE = W + qx * qy * qz;                 // qx*qy*qz
    float *EX = EE + qx * qy * qz;                // BX * qy * qz
    float *EXY = EX + BX * qy * qz;               // BX * BY * qz

    const int R = rp + rs;
    const int b = blockIdx.x;
    const int X0 = (b % a.nbx) * BX - R;
    const int Y0 = ((b / a.nbx) % a.nby) * BY - R;
    const int Z0 = (b / (a.nbx * a.nby)) * BZ - R;
    const int tid = threadIdx.x;

    // ---- stage u~ ----
    for (int i = tid; i < ux * uy * uz; i += NT) {
        const int lx = i % ux, r = i / ux, ly = r % uy, lz = r / uy;
        const int x = reflect101(X0 - HU + lx, a.nx), y = reflect101(Y0 - HU + ly, a.ny),
                  z = reflect101(Z0 - HU + lz, a.nz);
        U[i] = a.u[x + (long long)a.nx * (y + (long long)a.ny * z)];
    }
    // ---- stage G = g/Z, GY = g*y/Z (zero outside the volume) ----
    for (int i = tid; i < gx * gy * gz; i += NT) {
        const int lx = i % gx, r = i / gx, ly = r % gy, lz = r / gy;
        const int x = X0 - HG + lx, y = Y0 - HG + ly, z = Z0 - HG + lz;
        float G = 0.0f, GY = 0.0f;
        if (in_vol(x, y, z, a.nx, a.ny, a.nz)) {
            const long long o = x + (long long)a.nx * (y + (long long)a.ny * z);
            G = a.g[o] / a.Z[o];
            GY = G * a.y[o];
        }
        GG[2 * i] = G;
        GG[2 * i + 1] = GY;
    }
    __syncthreads();

    // fixed ownership of gradient positions: p = tid + j*NT
    float gacc[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) gacc[j] = 0.0f;
    float dh = 0.0f;

    // index helpers: U index of block-local (x,y,z) (block-local 0 = X0)
    auto uidx = [&](int x, int y, int z) { return (x + HU) + ux * ((y + HU) + uy * (z + HU)); };
    auto gidx = [&](int x, int y, int z) { return 2 * ((x + HG) + gx * ((y + HG) + gy * (z + HG))); };

    for (int dz = -rs; dz <= rs; ++dz)
        for (int dy = -rs; dy <= rs; ++dy)
            for (int dx = -rs; dx <= rs; ++dx) {
                const int du = dx + ux * (dy + uy * dz);
                const int dg = 2 * (dx + gx * (dy + gy * dz));
                // A1: x-stage of squared differences over x in Q, y,z in B+4rp
                for (int i = tid; i < qx * dy2 * dz2; i += NT) {
                    const int lx = i % qx, r = i / qx, ly = r % dy2, lz = r / dy2;
                    const int c = uidx(lx - rp - rp, ly - 2 * rp, lz - 2 * rp);
                    float acc = 0.0f;
                    for (int p = 0; p <= 2 * rp; ++p) {
                        const float e = U[c + p] - U[c + p + du];
                        acc = fmaf(e, e, acc);
                    }
                    SX[i] = acc;
                }
                __syncthreads();
                // A2: y-stage over x,y in Q, z in B+4rp
                for (int i = tid; i < qx * qy * dz2; i += NT) {
                    const int lx = i % qx, r = i / qx, ly = r % qy, lz = r / qy;
                    const float *col = SX + lx + qx * (ly + dy2 * lz);
                    float s = 0.0f;
                    for (int p = 0; p <= 2 * rp; ++p) s = s + col[qx * p];
                    SXY[i] = s;
                }
                __syncthreads();
                // A3: S, w, E on Q; dL/dh terms on the block's own volume voxels
                for (int i = tid; i < qx * qy * qz; i += NT) {
                    const int lx = i % qx, r = i / qx, ly = r % qy, lz = r / qy;
                    const float *col = SXY + lx + qx * (ly + qy * lz);
                    float S = 0.0f;
                    for (int p = 0; p <= 2 * rp; ++p) S = S + col[qx * qy * p];
                    const float w = exp2_fast(S * a.k);
                    const int x = lx - rp, y = ly - rp, z = lz - rp;  // block-local
                    const int ci = uidx(x, y, z);
                    const int gi = gidx(x, y, z);
                    const float uc = U[ci], ud = U[ci + du];
                    const float t1 = GG[gi] * ud - GG[gi + 1];
                    const float t2 = GG[gi + dg] * uc - GG[gi + dg + 1];
                    W[i] = w;
                    EE[i] = a.negc * w * (t1 + t2);
                }
                __syncthreads();
                // A4: x-stage of E over x in B, y,z in Q
                for (int i = tid; i < BX * qy * qz; i += NT) {
                    const int lx = i % BX, r = i / BX, ly = r % qy, lz = r / qy;
                    const float *row = EE + lx + qx * (ly + qy * lz);
                    float s = 0.0f;
                    for (int p = 0; p <= 2 * rp; ++p) s = s + row[p];
                    EX[i] = s;
                }
                __syncthreads();
                // A5: y-stage of E over x,y in B, z in Q
                for (int i = tid; i < BX * BY * qz; i += NT) {
                    const int lx = i % BX, r = i / BX, ly = r % BY, lz = r / BY;
                    const float *col = EX + lx + BX * (ly + qy * lz);
                    float s = 0.0f;
                    for (int p = 0; p <= 2 * rp; ++p) s = s + col[BX * p];
                    EXY[i] = s;
                }
                __syncthreads();
                // A6: z-stage of E and gradient accumulation at owned positions
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    const int i = tid + j * NT;
                    const int lx = i % BX, r = i / BX, ly = r % BY, lz = r / BY;
                    const float *col = EXY + lx + BX * (ly + BY * lz);
                    float bE = 0.0f;
                    for (int p = 0; p <= 2 * rp; ++p) bE = bE + col[BX * BY * p];
                    const int wi = (lx + rp) + qx * ((ly + rp) + qy * (lz + rp));
                    const int ci = uidx(lx, ly, lz);
                    const int gi = gidx(lx, ly, lz);
                    const float w = W[wi];
                    gacc[j] = fmaf(GG[gi + dg], w, gacc[j]);
                    gacc[j] = fmaf(2.0f * bE, U[ci] - U[ci + du], gacc[j]);
                    // dL/dh: x = this position, if it is a volume voxel
                    const int x = X0 + lx, y = Y0 + ly, z = Z0 + lz;
                    if (in_vol(x, y, z, a.nx, a.ny, a.nz)) {
                        // recompute S at x exactly as A3 did for the grown-block
                        // position (lx+rp, ly+rp, lz+rp): SXY planes lz+rp .. lz+3rp
                        const float *c2 = SXY + (lx + rp) + qx * ((ly + rp) + qy * (lz + rp));
                        float S = 0.0f;
                        for (int p = 0; p <= 2 * rp; ++p) S = S + c2[qx * qy * p];
                        const float t1 = GG[gi] * U[ci + du] - GG[gi + 1];
                        dh = fmaf(w * S, t1, dh);
                    }
                }
                // A1 of the next offset writes SX only; the buffers read here
                // (W, EXY, SXY) are rewritten after at least one more barrier.
            }

    // ---- write the gradient of the block's extended-domain positions ----
    const long long ex = a.nx + 2LL * R, ey = a.ny + 2LL * R, ez = a.nz + 2LL * R;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = tid + j * NT;
        const int lx = i % BX, r = i / BX, ly = r % BY, lz = r / BY;
        const long long x = X0 + lx + R, y = Y0 + ly + R, z = Z0 + lz + R;
        if (x < ex && y < ey && z < ez) a.gext[x + ex * (y + ey * z)] = gacc[j];
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

size_t bwd_lds_bytes(int rp, int rs)
{
    const int HU = 2 * rp + rs, HG = rp + rs;
    const size_t ux = BX + 2 * HU, uy = BY + 2 * HU, uz = BZ + 2 * HU;
    const size_t gx = BX + 2 * HG, gy = BY + 2 * HG, gz = BZ + 2 * HG;
    const size_t qx = BX + 2 * rp, qy = BY + 2 * rp, qz = BZ + 2 * rp;
    const size_t dy2 = BY + 4 * rp, dz2 = BZ + 4 * rp;
    size_t n = ux * uy * uz + 2 * gx * gy * gz + qx * dy2 * dz2 + qx * qy * dz2 + 2 * qx * qy * qz +
               BX * qy * qz + BX * BY * qz;
    if (n < NT) n = NT;
    return n * sizeof(float);
}

}  // namespace

bool bwd3d_supported(int r_p, int r_s)
{
    return r_p >= 0 && r_s >= 0 && bwd_lds_bytes(r_p, r_s) <= 160 * 1024;
}

hipError_t bwd3d_launch(const float *u, const float *y, const float *Z, const float *g, int nx, int ny, int nz,
                        int rp, int rs, float h, float *gu, float *gh, hipStream_t s)
{
    const int R = rp + rs;
    const long long ext = (nx + 2LL * R) * (ny + 2LL * R) * (nz + 2LL * R);
    BwdArgs a;
    a.u = u;
    a.y = y;
    a.Z = Z;
    a.g = g;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.rp = rp;
    a.rs = rs;
    const int P = (2 * rp + 1) * (2 * rp + 1) * (2 * rp + 1);
    const double c = 1.0 / ((double)P * (double)h * (double)h);
    a.k = (float)weight_scale(P, h);
    a.negc = (float)(-c);
    a.nbx = (int)((nx + 2LL * R + BX - 1) / BX);
    a.nby = (int)((ny + 2LL * R + BY - 1) / BY);
    a.nbz = (int)((nz + 2LL * R + BZ - 1) / BZ);
    const int nblocks = a.nbx * a.nby * a.nbz;
    float *ws = nullptr;
    hipError_t e = hipMallocAsync((void **)&ws, sizeof(float) * (size_t)(ext + nblocks), s);
    if (e != hipSuccess) return e;
    a.gext = ws;
    a.dh_part = ws + ext;
    const size_t lds = bwd_lds_bytes(rp, rs);
    e = hipFuncSetAttribute((const void *)nlm3d_bwd_tiled, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(nlm3d_bwd_tiled, dim3(nblocks), dim3(NT), lds, s, a);
        long long nv = (long long)nx * ny * nz;
        long long gblk = (nv + 255) / 256;
        if (gblk > 65536) gblk = 65536;
        hipLaunchKernelGGL(fold_grad, dim3((unsigned)gblk), dim3(256), 0, s, (const float *)a.gext, gu, nx, ny, nz, R);
        if (gh) hipLaunchKernelGGL(reduce_dh, dim3(1), dim3(256), 0, s, (const float *)a.dh_part, nblocks,
                                   (float)(2.0 * c / (double)h), gh);
        e = hipGetLastError();
    }
    hipError_t e2 = hipFreeAsync(ws, s);
    return e != hipSuccess ? e : e2;
}

}  // namespace gnlm
