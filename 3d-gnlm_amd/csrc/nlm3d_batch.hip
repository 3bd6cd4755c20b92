// nlm3d_batch.hip -- batched 3D NLM: many volumes of uneven size, each with
// its own h and radii, in ONE persistent launch captured in a HIP graph.
//
// No counterpart in the reference (3DNLM/main.cu:7607-7680 processes one
// compile-time volume per call).  Design:
//   * at plan creation every volume is cut into tiles of the variant the
//     autotuner picked for its (r_p, r_s).  Each volume's tile list (x
//     fastest) is split into 8 contiguous ranges, one per XCD, and XCD j's
//     work queue holds range j of every volume, volumes ordered longest-first
//     by tile cost (tile voxels x offsets).  Tiles that share halo samples
//     thus run on one XCD and share its L2 (a single global queue spreads
//     neighbouring tiles over all 8 L2s and their working sets overflow
//     them); every XCD gets an eighth of every volume, so the queues carry
//     equal work;
//   * the persistent kernel runs as many workgroups as fit on the device at
//     once; workgroup b serves the queue of XCD b % 8 (the dispatcher's
//     round-robin placement), taking the next tile index with one atomicAdd
//     broadcast through LDS, and steals from the other queues once its own
//     is empty.  It processes a tile with the same device function the
//     per-volume launch uses, so the outputs are bit-identical to per-volume
//     launches.  Every workgroup exits once all queues are exhausted;
//   * the graph holds [padding pass -> counter reset kernel -> persistent
//     kernel]; replaying it re-runs the whole batch with one launch (no
//     hipMemset in the graph: the counters are cleared by a kernel).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>
#include "gnlm_internal.h"
#include "nlm3d_rows.h"
#include "nlm3d_tile.h"

namespace gnlm {
namespace {

struct BatchVol {
    InDesc in;
    float *out;
    float *Zout;
    int rp, rs;
    float k;
    float sigma;
    int variant;
    int ntx, nty;
    // rows-variant volumes: padded copy (filled by batch_pad inside the graph);
    // Zacc = Zout (NULL when the caller wants no Z: the tiles write y alone)
    PadVol pv;
    float *Zacc;
    int px, py, pz, lp, halo;
};

struct BatchTile {
    int vol;
    int tx, ty, tz;
};

constexpr int kNumXcd = 8;

__global__ void reset_counters(int *c)
{
    if (threadIdx.x < kNumXcd) c[threadIdx.x] = 0;
}

// Each (variant, r_p) tile body is a separate, non-inlined function: inlined
// together into the one persistent kernel, their register demands add up and
// push values into AGPRs inside the hot loops; called, each body keeps the
// allocation it has in its per-volume kernel.
template <int RP>
__device__ __noinline__ void tile_slab2(const Fwd3Tile &a, float *lds) { fwd3_slab2_tile<RP>(a, lds, a.sigma); }
template <int RP>
__device__ __noinline__ void tile_slab(const Fwd3Tile &a, float *lds) { fwd3_slab_tile<RP>(a, lds, a.sigma); }
template <class C>
__device__ __noinline__ void tile_direct(const Fwd3Tile &a, float *lds) { fwd3_tile<C>(a, lds); }
// LIGHT instances are called from the three-waves-per-SIMD kernel only (a
// separate function, so that the register limit of that kernel is theirs).
template <int RP, int RS, bool SH, bool LIGHT>
__device__ __noinline__ void tile_rows(const RowsTile &a, float *lds)
{
    // volumes without a Z buffer pass Zacc = NULL: the tiles must hold their
    // totals in registers
    static_assert(RowsCfg<RP, RS, SH>::REG_TOTALS, "batched rows tiles keep their totals in registers");
    rows_tile<RowsCfg<RP, RS, SH>>(a, lds);
}

// kVRows / kVRowsShort tile: r_s is a template parameter of the tile body (its
// dx loop is unrolled), so dispatch on it; the cases match nlm3d_rows.hip.
// LIGHT: only the bodies that fit three waves per SIMD (r_p = 1, 3 and 4
// rows, r_p = 2 short rows).
template <int RP, bool LIGHT>
__device__ __forceinline__ void run_rows_rp(const BatchVol &v, const BatchTile &t, float *lds)
{
    RowsTile a;
    a.v = v.pv;
    a.out = v.out;
    a.Zacc = v.Zacc;
    a.nx = v.in.nx;
    a.ny = v.in.ny;
    a.nz = v.in.nz;
    a.k = v.k;
    a.x0 = t.tx * 16;
    a.y0 = t.ty * fwd3_ty(v.variant, RP);
    a.z0 = t.tz * fwd3_tz(v.variant, RP);
    if constexpr (RP == 1) {
        switch (v.rs) {
        case 1: tile_rows<1, 1, false, LIGHT>(a, lds); break;
        case 2: tile_rows<1, 2, false, LIGHT>(a, lds); break;
        case 3: tile_rows<1, 3, false, LIGHT>(a, lds); break;
        case 4: tile_rows<1, 4, false, LIGHT>(a, lds); break;
        case 5: tile_rows<1, 5, false, LIGHT>(a, lds); break;
        case 6: tile_rows<1, 6, false, LIGHT>(a, lds); break;
        case 7: tile_rows<1, 7, false, LIGHT>(a, lds); break;
        default: tile_rows<1, 8, false, LIGHT>(a, lds); break;
        }
    } else if constexpr (RP == 2) {
        if (LIGHT || v.variant == kVRowsShort) {
            switch (v.rs) {
            case 1: tile_rows<2, 1, true, LIGHT>(a, lds); break;
            case 2: tile_rows<2, 2, true, LIGHT>(a, lds); break;
            case 3: tile_rows<2, 3, true, LIGHT>(a, lds); break;
            case 4: tile_rows<2, 4, true, LIGHT>(a, lds); break;
            case 5: tile_rows<2, 5, true, LIGHT>(a, lds); break;
            case 6: tile_rows<2, 6, true, LIGHT>(a, lds); break;
            default: tile_rows<2, 7, true, LIGHT>(a, lds); break;
            }
            return;
        }
        if constexpr (!LIGHT) {
            switch (v.rs) {
            case 1: tile_rows<2, 1, false, false>(a, lds); break;
            case 2: tile_rows<2, 2, false, false>(a, lds); break;
            case 3: tile_rows<2, 3, false, false>(a, lds); break;
            case 4: tile_rows<2, 4, false, false>(a, lds); break;
            case 5: tile_rows<2, 5, false, false>(a, lds); break;
            case 6: tile_rows<2, 6, false, false>(a, lds); break;
            default: tile_rows<2, 7, false, false>(a, lds); break;
            }
        }
    } else if constexpr (RP == 4) {
        switch (v.rs) {
        case 1: tile_rows<4, 1, false, LIGHT>(a, lds); break;
        case 2: tile_rows<4, 2, false, LIGHT>(a, lds); break;
        case 3: tile_rows<4, 3, false, LIGHT>(a, lds); break;
        case 4: tile_rows<4, 4, false, LIGHT>(a, lds); break;
        case 5: tile_rows<4, 5, false, LIGHT>(a, lds); break;
        case 6: tile_rows<4, 6, false, LIGHT>(a, lds); break;
        default: tile_rows<4, 7, false, LIGHT>(a, lds); break;
        }
    } else if constexpr (RP == 3) {
        switch (v.rs) {
        case 1: tile_rows<3, 1, false, LIGHT>(a, lds); break;
        case 2: tile_rows<3, 2, false, LIGHT>(a, lds); break;
        case 3: tile_rows<3, 3, false, LIGHT>(a, lds); break;
        case 4: tile_rows<3, 4, false, LIGHT>(a, lds); break;
        case 5: tile_rows<3, 5, false, LIGHT>(a, lds); break;
        case 6: tile_rows<3, 6, false, LIGHT>(a, lds); break;
        default: tile_rows<3, 7, false, LIGHT>(a, lds); break;
        }
    }
}

// Padded copies of the rows-variant volumes: blockIdx.y picks the volume.
// Grid z walks the planes and grid x the samples of a plane, so the index
// arithmetic is 32-bit (each copy is below 2^30 floats).
__global__ void batch_pad(const BatchVol *__restrict__ vols, const int *__restrict__ ids)
{
    const BatchVol &v = vols[ids[blockIdx.y]];
    float *dst = const_cast<float *>(v.pv.p);
    const int r = v.halo, nxy = v.px * v.py;
    for (int zp = blockIdx.z; zp < v.pz; zp += gridDim.z) {
        const int z = zp - r;
        float *plane = dst + (long long)zp * nxy;
        const bool zin = z >= -r && z < v.in.nz + r;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nxy; i += gridDim.x * blockDim.x) {
            const int x = i % v.px - v.lp;
            const int y = i / v.px - r;
            float val = 0.0f;
            if (zin && x >= -r && x < v.in.nx + r && y >= -r && y < v.in.ny + r) val = load_in(v.in, x, y, z);
            plane[i] = val;
        }
    }
}

template <int RP>
__device__ __forceinline__ void run_tile_rp(const BatchVol &v, const BatchTile &t, float *lds)
{
    Fwd3Tile a;
    a.in = v.in;
    a.out = v.out;
    a.Zout = v.Zout;
    a.rs = v.rs;
    a.k = v.k;
    a.sigma = v.sigma;
    if (fwd3_is_rows(v.variant)) {
        run_rows_rp<RP, false>(v, t, lds);
    } else if (v.variant == kVSlab2) {
        using C = SlabCfg<RP>;
        a.x0 = t.tx * C::TX;
        a.y0 = t.ty * C::TY;
        a.z0 = t.tz * C::TZ;
        tile_slab2<RP>(a, lds);
    } else if (v.variant == kVSlab) {
        using C = SlabCfg<RP>;
        a.x0 = t.tx * C::TX;
        a.y0 = t.ty * C::TY;
        a.z0 = t.tz * C::TZ;
        tile_slab<RP>(a, lds);
    } else if (v.variant == kV16x16x8) {
        a.x0 = t.tx * 16;
        a.y0 = t.ty * 16;
        a.z0 = t.tz * 8;
        tile_direct<Fwd3Cfg<RP, 16, 16, 8>>(a, lds);
    } else {
        a.x0 = t.tx * 16;
        a.y0 = t.ty * 16;
        a.z0 = t.tz * 16;
        tile_direct<Fwd3Cfg<RP, 16, 16, 16>>(a, lds);
    }
}

// Volume v can run in the LIGHT kernel (three waves per SIMD).
__host__ __device__ inline bool batch_light(int rp, int variant)
{
    return ((rp == 1 || rp >= 3) && variant == kVRows) || (rp == 2 && variant == kVRowsShort);
}

// LIGHT: every volume of the plan satisfies batch_light; the kernel then holds
// only those tile bodies and runs three waves per SIMD.
template <bool LIGHT>
__global__ __launch_bounds__(kFwdThreads) __attribute__((amdgpu_waves_per_eu(LIGHT ? 3 : 2))) void nlm3d_fwd_persistent(
    const BatchVol *__restrict__ vols, const BatchTile *__restrict__ tiles, const int *__restrict__ qstart,
    int *counters)
{
    extern __shared__ __align__(16) float lds[];  // 128-bit LDS accesses in the tile bodies
    __shared__ int next;
    const int xcd = blockIdx.x % kNumXcd;
    int tried = 0;  // queues found empty so far (thread 0 only)
    for (;;) {
        if (threadIdx.x == 0) {
            int t = -1;
            while (tried < kNumXcd) {
                const int q = (xcd + tried) % kNumXcd;
                const int i = qstart[q] + atomicAdd(&counters[q], 1);
                if (i < qstart[q + 1]) {
                    t = i;
                    break;
                }
                ++tried;
            }
            next = t;
        }
        __syncthreads();
        const int t = next;
        __syncthreads();  // all threads have read `next` and finished the previous tile's LDS use
        if (t < 0) break;
        const BatchTile tl = tiles[t];
        const BatchVol v = vols[tl.vol];
        if constexpr (LIGHT) {
            if (v.rp == 1) run_rows_rp<1, true>(v, tl, lds);
            else if (v.rp == 2) run_rows_rp<2, true>(v, tl, lds);
            else if (v.rp == 3) run_rows_rp<3, true>(v, tl, lds);
            else run_rows_rp<4, true>(v, tl, lds);
        } else {
            switch (v.rp) {
            case 0: run_tile_rp<0>(v, tl, lds); break;
            case 1: run_tile_rp<1>(v, tl, lds); break;
            case 2: run_tile_rp<2>(v, tl, lds); break;
            case 3: run_tile_rp<3>(v, tl, lds); break;
            default: run_rows_rp<4, false>(v, tl, lds); break;  // r_p = 4: rows tiles only
            }
        }
    }
}

}  // namespace
}  // namespace gnlm

using namespace gnlm;

struct gnlm_batch_plan {
    int nvol = 0, ntiles = 0, nblocks = 0;
    size_t lds = 0;
    BatchVol *d_vols = nullptr;
    BatchTile *d_tiles = nullptr;
    int *d_counter = nullptr;  // kNumXcd queue counters, then kNumXcd + 1 queue starts
    int npad = 0;              // kVRows / kVRowsShort volumes
    int *d_padids = nullptr;   // their indices
    int padmax_xy = 0;         // largest plane of a padded copy (floats)
    int padmax_z = 0;          // most planes of a padded copy
    float *d_ws = nullptr;     // padded copies
    bool light = false;        // every volume fits the three-waves-per-SIMD kernel
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void plan_free(gnlm_batch_plan *p)
{
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_vols) (void)hipFree(p->d_vols);
    if (p->d_tiles) (void)hipFree(p->d_tiles);
    if (p->d_counter) (void)hipFree(p->d_counter);
    if (p->d_padids) (void)hipFree(p->d_padids);
    if (p->d_ws) (void)hipFree(p->d_ws);
    delete p;
}

int st_of(hipError_t e)
{
    if (e == hipSuccess) return GNLM_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return GNLM_ENOMEM;
    return GNLM_EHIP;
}

// The launches of one batch run: [padded copies] -> counter reset -> persistent kernel.
hipError_t enqueue_batch(const gnlm_batch_plan *p, hipStream_t s)
{
    if (p->npad > 0) {
        const int gx = std::min((p->padmax_xy + 255) / 256, 256), gz = std::min(p->padmax_z, 32);
        hipLaunchKernelGGL(batch_pad, dim3((unsigned)gx, (unsigned)p->npad, (unsigned)gz), dim3(256), 0, s,
                           (const BatchVol *)p->d_vols, (const int *)p->d_padids);
    }
    hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, s, p->d_counter);
    if (p->light)
        hipLaunchKernelGGL(nlm3d_fwd_persistent<true>, dim3(p->nblocks), dim3(kFwdThreads), p->lds, s,
                           (const BatchVol *)p->d_vols, (const BatchTile *)p->d_tiles,
                           (const int *)(p->d_counter + kNumXcd), p->d_counter);
    else
        hipLaunchKernelGGL(nlm3d_fwd_persistent<false>, dim3(p->nblocks), dim3(kFwdThreads), p->lds, s,
                           (const BatchVol *)p->d_vols, (const BatchTile *)p->d_tiles,
                           (const int *)(p->d_counter + kNumXcd), p->d_counter);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int gnlm_batch_plan_create(const gnlm_batch_volume *vols, int n, gnlm_batch_plan **out)
{
    if (!out) return GNLM_EINVAL;
    *out = nullptr;
    if (!vols || n < 1) return GNLM_EINVAL;
    std::vector<BatchVol> hv((size_t)n);
    std::vector<std::vector<BatchTile>> vt;  // tiles of each volume, x fastest
    std::vector<double> cost;                 // per-tile cost of each volume
    std::vector<int> padids;
    std::vector<long long> wsoff((size_t)n, -1);
    long long wsn = 0;
    int padmax_xy = 0, padmax_z = 0;
    size_t lds = 0;
    for (int i = 0; i < n; ++i) {
        const gnlm_batch_volume &v = vols[i];
        if (!v.d_in || !v.d_out || v.nx < 1 || v.ny < 1 || v.nz < 1 || v.r_p < 0 || v.r_s < 0 ||
            !(v.h > 0.0f) || !std::isfinite(v.h))
            return GNLM_EINVAL;
        if (v.r_p > 4 || !fwd3d_tiled_supported(v.r_p, v.r_s)) return GNLM_EUNSUPPORTED;
        BatchVol &b = hv[(size_t)i];
        b.in = InDesc{v.d_in, v.nx, v.ny, v.nz, 0, 0, 0};
        b.out = v.d_out;
        b.Zout = v.d_Z;
        b.rp = v.r_p;
        b.rs = v.r_s;
        const int P = (2 * v.r_p + 1) * (2 * v.r_p + 1) * (2 * v.r_p + 1);
        b.k = (float)weight_scale(P, v.h);
        b.sigma = std::sqrt(-b.k);
        b.variant = fwd3d_pick_variant(v.r_p, v.r_s);
        if (v.r_p == 4 && b.variant != kVRows) return GNLM_EUNSUPPORTED;  // big tiles are per-volume only
        const int tx = fwd3_tx(b.variant, v.r_p), ty = fwd3_ty(b.variant, v.r_p), tz = fwd3_tz(b.variant, v.r_p);
        b.ntx = (v.nx + tx - 1) / tx;
        b.nty = (v.ny + ty - 1) / ty;
        const int ntz = (v.nz + tz - 1) / tz;
        lds = std::max(lds, fwd3d_lds_bytes(v.r_p, v.r_s, b.variant));
        b.Zacc = nullptr;
        b.pv = PadVol{nullptr, 0, 0, 0u};
        b.px = b.py = b.pz = b.lp = b.halo = 0;
        if (fwd3_is_rows(b.variant)) {
            // geometry of the padded copy, as in nlm3d_rows.hip
            b.halo = v.r_p + v.r_s;
            b.lp = 4 * ((b.halo + 3) / 4);
            b.px = b.ntx * tx + 2 * b.lp;
            b.py = b.nty * ty + 2 * b.halo;
            b.pz = ntz * tz + 2 * b.halo;
            const long long np = (long long)b.px * b.py * b.pz;
            const long long nv = (long long)v.nx * v.ny * v.nz;
            if (np >= (1LL << 30) || nv >= (1LL << 31)) return GNLM_EUNSUPPORTED;
            b.pv.px = b.px;
            b.pv.pxy = b.px * b.py;
            b.pv.bytes = (unsigned)(sizeof(float) * np);
            wsoff[(size_t)i] = wsn;
            wsn += (np + 63) / 64 * 64;
            padmax_xy = std::max(padmax_xy, b.px * b.py);
            padmax_z = std::max(padmax_z, b.pz);
            padids.push_back(i);
        }
        cost.push_back((double)tx * ty * tz * std::pow(2.0 * v.r_s + 1.0, 3));
        vt.emplace_back();
        for (int z = 0; z < ntz; ++z)
            for (int y = 0; y < b.nty; ++y)
                for (int x = 0; x < b.ntx; ++x) vt.back().push_back(BatchTile{i, x, y, z});
    }
    // volumes longest-first by tile cost (stable); queue j = range j of each volume
    std::vector<int> vorder((size_t)n);
    for (int i = 0; i < n; ++i) vorder[(size_t)i] = i;
    std::stable_sort(vorder.begin(), vorder.end(), [&](int a, int b) { return cost[(size_t)a] > cost[(size_t)b]; });
    std::vector<BatchTile> sorted;
    std::vector<int> qstart(kNumXcd + 1, 0);
    for (int q = 0; q < kNumXcd; ++q) {
        qstart[(size_t)q] = (int)sorted.size();
        for (int i : vorder) {
            const std::vector<BatchTile> &t = vt[(size_t)i];
            const size_t lo = t.size() * (size_t)q / kNumXcd, hi = t.size() * (size_t)(q + 1) / kNumXcd;
            sorted.insert(sorted.end(), t.begin() + (long)lo, t.begin() + (long)hi);
        }
    }
    qstart[kNumXcd] = (int)sorted.size();

    gnlm_batch_plan *p = new (std::nothrow) gnlm_batch_plan();
    if (!p) return GNLM_ENOMEM;
    p->nvol = n;
    p->ntiles = (int)sorted.size();
    p->lds = lds;
    p->light = true;
    for (const BatchVol &b : hv) p->light = p->light && batch_light(b.rp, b.variant);
    const void *kfn = p->light ? (const void *)nlm3d_fwd_persistent<true> : (const void *)nlm3d_fwd_persistent<false>;
    hipError_t e = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int dev = 0, ncu = 0, occ = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kfn, kFwdThreads, lds);
    if (e == hipSuccess && occ < 1) e = hipErrorInvalidConfiguration;
    if (e == hipSuccess) p->nblocks = std::min(ncu * occ, p->ntiles);
    const int occ_max = occ;
    p->npad = (int)padids.size();
    p->padmax_xy = padmax_xy;
    p->padmax_z = padmax_z;
    if (e == hipSuccess && wsn > 0) e = hipMalloc((void **)&p->d_ws, sizeof(float) * (size_t)wsn);
    if (e == hipSuccess && p->npad > 0) e = hipMalloc((void **)&p->d_padids, sizeof(int) * padids.size());
    if (e == hipSuccess && p->npad > 0)
        e = hipMemcpy(p->d_padids, padids.data(), sizeof(int) * padids.size(), hipMemcpyHostToDevice);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        BatchVol &b = hv[(size_t)i];
        if (!fwd3_is_rows(b.variant)) continue;
        b.pv.p = p->d_ws + wsoff[(size_t)i];
        b.Zacc = b.Zout;  // NULL: y alone (the tiles keep their totals in registers)
    }
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_vols, sizeof(BatchVol) * hv.size());
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_tiles, sizeof(BatchTile) * sorted.size());
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_counter, sizeof(int) * (2 * kNumXcd + 1));
    if (e == hipSuccess)
        e = hipMemcpy(p->d_counter + kNumXcd, qstart.data(), sizeof(int) * (kNumXcd + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_vols, hv.data(), sizeof(BatchVol) * hv.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_tiles, sorted.data(), sizeof(BatchTile) * sorted.size(), hipMemcpyHostToDevice);
    hipStream_t cs = nullptr;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    // Launch-shape tuning (GNLM_TUNE_LAZY): the persistent kernel's
    // workgroups per CU, the most that fit or one fewer (fewer workgroups
    // contend less for each CU's LDS and L1 and leave queue stealing more
    // room).  Any count processes every tile with the same tile body, so the
    // outputs are the same bits; each candidate runs the whole batch eagerly
    // (one warm-up, two timed) on the plan's own buffers.
    if (e == hipSuccess && fwd3d_tuning_lazy() && occ_max > 1) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        int best = 0;
        float best_ms = 0.0f;
        for (int per_cu = occ_max; per_cu >= occ_max - 1 && e == hipSuccess; --per_cu) {
            p->nblocks = std::min(ncu * per_cu, p->ntiles);
            e = enqueue_batch(p, cs);
            if (e == hipSuccess) e = hipEventRecord(e0, cs);
            for (int r = 0; r < 2 && e == hipSuccess; ++r) e = enqueue_batch(p, cs);
            if (e == hipSuccess) e = hipEventRecord(e1, cs);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess && (best == 0 || ms < best_ms)) {
                best = per_cu;
                best_ms = ms;
            }
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (e == hipSuccess) p->nblocks = std::min(ncu * best, p->ntiles);
    }
    // capture [reset counter -> persistent kernel] on a private stream
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        e = enqueue_batch(p, cs);
        hipError_t e2 = hipStreamEndCapture(cs, &p->graph);
        if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (cs) (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
        plan_free(p);
        return st_of(e);
    }
    *out = p;
    return GNLM_OK;
}

int gnlm_batch_run(gnlm_batch_plan *plan, void *stream)
{
    if (!plan || !plan->exec) return GNLM_ESTATE;
    return st_of(hipGraphLaunch(plan->exec, (hipStream_t)stream));
}

int gnlm_batch_run_eager(gnlm_batch_plan *plan, void *stream)
{
    if (!plan || !plan->d_counter) return GNLM_ESTATE;
    return st_of(enqueue_batch(plan, (hipStream_t)stream));
}

int gnlm_batch_plan_info(const gnlm_batch_plan *plan, int *ntiles, int *nblocks)
{
    if (!plan) return GNLM_EINVAL;
    if (ntiles) *ntiles = plan->ntiles;
    if (nblocks) *nblocks = plan->nblocks;
    return GNLM_OK;
}

int gnlm_batch_plan_destroy(gnlm_batch_plan *plan)
{
    if (!plan) return GNLM_EINVAL;
    plan_free(plan);
    return GNLM_OK;
}

}  // extern "C"
