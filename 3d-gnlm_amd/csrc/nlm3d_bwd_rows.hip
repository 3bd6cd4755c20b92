// nlm3d_bwd_rows.hip -- adjoint of the 3D NLM for patch radii 0, 1 and 2,
// organised like the rows forward (nlm3d_rows.h).
//
// Same gather form as nlm3d_bwd.hip (3DNLM/main.cu:52-68,107-111,125 is the
// forward being differentiated):
//   grad(z) = sum_d [ G(z+d) w_d(z) + 2 boxE_d(z) (u~(z) - u~(z+d)) ]
//   E_d(x)  = -c w_d(x) [ G(x) u~(x+d) - GY(x) + G(x+d) u~(x) - GY(x+d) ]
//   dL/dh   = (2c/h) sum_x sum_d w_d(x) S_d(x) (G(x) u~(x+d) - GY(x))
// with G = g/Z and GY = g*y/Z zero outside the volume, on the extended
// domain [-R, n+R) (R = r_p + r_s), folded back onto the volume through the
// reflect-101 preimages afterwards.
//
// Organisation (one workgroup of 256 threads per gradient tile T of
// (16 - 2 r_p) x (16 - 2 r_p) x (QZ - 2 r_p) positions; Q = T grown by r_p =
// 16 x 16 x QZ is where S, w and E live, QZ = 8 in the symmetric form for
// r_p <= 1 (T = 14 x 14 x 6 at r_p = 1) and 12 otherwise; D = Q grown by
// r_p holds the squared differences):
//   * a pre-pass writes a padded copy of u~ and a padded, interleaved
//     (G, GY, u~, 0) array ((G, GY) zero outside the volume); the tile reads
//     both with buffer loads whose per-lane offsets are fixed for the tile
//     (one 128-bit load per column position and offset);
//   * per offset d (dz, dy, dx):
//       phase 1  x box sums of squared differences on the rows of D
//                (18 x 10 at r_p = 1) -> LDS; the centre samples of a row
//                stay in registers, the shifted ones are loaded once per
//                (dy, dz) and dx is unrolled;
//       phase 2  y box sums -> LDS (the forward's phase 2);
//       phase 3  one thread per Q column (x = lane in a 16-lane row, y = row
//                of lanes): z box sums -> S, w, E, the G(z+d) w term and the
//                dL/dh terms; the x box of E with DPP row shifts, its z box in
//                registers -> LDS;
//       phase 4  the same thread: y box of that from LDS and the
//                2 boxE (u~(z) - u~(z+d)) term.
//     Three barriers per offset.  G, GY and u~ of a thread's own column stay
//     in registers; the values at x + d come with buffer loads issued before
//     phase 1, so their latency hides behind phases 1 and 2.
// Nothing of size voxels x offsets is stored.  Every position of the
// extended domain belongs to exactly one tile; dL/dh partials are summed per
// lane, then per workgroup in a fixed order, so results are deterministic.
#include <atomic>
#include "gnlm_internal.h"
#include "nlm3d_rows.h"
#include "nlm3d_tile.h"

namespace gnlm {
namespace {

constexpr int NT = 256;

template <int RP_, int RS_, bool SYM_>
struct BwdRowsCfg {
    static constexpr int RP = RP_, RS = RS_, W = 2 * RP + 1, R = RP + RS;
    static_assert(RP >= 0 && RP <= 2, "patch radius 0, 1 or 2");
    static constexpr bool SYM = SYM_;
    static constexpr int H = 2 * R + RP;                  // halo of the padded copies (all axes)
    // gradient tile T, grown tile Q, phase-1 rows D
    // The symmetric form keeps a scatter accumulator in LDS.  For r_p <= 1 its
    // tiles are 8 positions deep in z, so that LDS (about 47 KB at r_s 4) and
    // registers fit three workgroups per CU; otherwise 12 deep, two per CU.
    static constexpr int QY = 16, QZ = (SYM && RP <= 1) ? 8 : 12;
    static constexpr int WAVES = SYM ? (QZ == 8 ? 3 : 2) : 3;  // waves per SIMD the registers are budgeted for
    static constexpr int GX = 16 - 2 * RP, GY = QY - 2 * RP, GZ = QZ - 2 * RP;
    static constexpr int DY = QY + 2 * RP, DZ = QZ + 2 * RP;
    // names the shared phase-1/2 code uses
    static constexpr int TX = 16;                         // x outputs per phase-1 row (Q's width)
    static constexpr int EX = TX + 2 * RP;                // centre samples per row
    static constexpr int NROWS = DY * DZ;
    // one phase-1 row per thread; rows beyond NT (r_p = 2: 320 rows) are cut
    // into x quarters (4 outputs), one quarter per thread
    static constexpr int NR1 = NROWS < NT ? NROWS : NT;
    static constexpr int NRX = NROWS - NR1;
    static_assert(NRX == 0 || NRX * 4 == NT, "the extra rows fill one quarter-row round");
    static constexpr int VX = 2, RY = 8, NXV = TX / VX;
    static constexpr int P2N = NXV * DZ * (QY / RY);
    static_assert(P2N <= NT, "one phase-2 item per thread");
    // LDS: s_x [z_d][y_d][x] (row pitch 20, plane pitch 400 == 16 mod 64),
    //      s_xy [z_d][y_q][x] (plane pitch 272); the boxed E [z_t][y_q][x]
    //      shares s_xy's layout and memory (below)
    static constexpr int SXR = 20, SXP = (SXR * DY - 16 + 63) / 64 * 64 + 16, SXN = SXP * DZ;
    static constexpr int SYP = 16 * QY + 16, SYN = SYP * DZ;
    static constexpr int EZP = SYP;
    // SYM: scatter accumulator over T grown by RS in x and y and by RS towards
    // +z (the visited half of the offsets has dz >= 0), [z][y][x]; the plane
    // pitch is a multiple of 64 floats so that a thread's z pairs share one
    // read2/write2 instruction
    static constexpr int SCX = GX + 2 * RS, SCY = GY + 2 * RS, SCZ = GZ + RS;
    static constexpr int SCP = (SCX * SCY + 63) / 64 * 64;
    static constexpr int SCN = SYM ? SCP * SCZ : 0;
    static constexpr int LDS_FLOATS = SXN + SYN + SCN;
};

struct BwdRowsArgs {
    PadVol pu;       // padded u~: ext position (x, y, z) at (z+H)*pxy + (y+H)*px + (x+H)
    PadVol pg;       // padded (G, GY, u~, 0), same geometry, float4 elements
    float *gext;     // gradient on the extended domain, (n + 2R)^3 (gather form)
    float *gsc;      // symmetric form: each tile's scatter region (SCN floats per tile)
    float *dh_part;  // one partial per workgroup
    int nx, ny, nz;
    float k;         // w = exp2(S * k)
    float negc2;     // -2c: E carries the factor 2 of the B term
    int ntx, nty, ntz;
    int nparts;      // SYM: workgroups per tile, each over a contiguous run of offset rows (1: gather form)
};

__device__ __forceinline__ float2 pad_load2(const PadRsrc &r, unsigned voff, int soff)
{
    const auto t = __builtin_amdgcn_raw_buffer_load_b64(r.r, (int)voff, soff, 0);
    return make_float2(__uint_as_float(t[0]), __uint_as_float(t[1]));
}

__device__ __forceinline__ float4 pad_load4(const PadRsrc &r, unsigned voff, int soff)
{
    const auto t = __builtin_amdgcn_raw_buffer_load_b128(r.r, (int)voff, soff, 0);
    return make_float4(__uint_as_float(t[0]), __uint_as_float(t[1]), __uint_as_float(t[2]), __uint_as_float(t[3]));
}

template <int RP_, int RS, bool SYM>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(BwdRowsCfg<RP_, RS, SYM>::WAVES))) void nlm3d_bwd_rows(
    BwdRowsArgs a)
{
    using C = BwdRowsCfg<RP_, RS, SYM>;
    constexpr int RP = C::RP, QZ = C::QZ, GZ = C::GZ;
    extern __shared__ __align__(16) float lds[];
    float *SX = lds;
    float *SXY = SX + C::SXN;
    // The boxed E of plane z and column (x_q, y_q) sits in s_xy's slot of the
    // same column and plane.  Phase 3 reads only its own column of s_xy, all
    // of it before writing that column's boxed E; phase 4 reads neighbouring
    // columns after the barrier that follows phase 3; the next phase 2 writes
    // s_xy only after the next offset's first barrier, which every thread
    // reaches after its phase 4.
    float *EZB = SXY;
    float *SC = SXY + C::SYN;  // SYM only
    const int tid = threadIdx.x;
    const PadRsrc ru = pad_rsrc(a.pu);
    const PadRsrc rg = pad_rsrc(a.pg);
    const int px = __builtin_amdgcn_readfirstlane(a.pu.px), pxy = __builtin_amdgcn_readfirstlane(a.pu.pxy);
    const float kw = a.k, negc2 = a.negc2;

    // unit = (tile, part): the parts of a tile are adjacent, so they share an XCD's L2
    const int np = a.nparts;
    const int unit = xcd_tile(blockIdx.x, a.ntx * a.nty * a.ntz * np);
    const int t = unit / np, part = unit - t * np;
    // gradient tile origin on the extended domain
    const int Xt = -C::R + (t % a.ntx) * C::GX;
    const int Yt = -C::R + ((t / a.ntx) % a.nty) * C::GY;
    const int Zt = -C::R + (t / (a.ntx * a.nty)) * C::GZ;
    auto pidx = [&](int x, int y, int z) { return (z + C::H) * pxy + (y + C::H) * px + (x + C::H); };

    // ---- phase-1 row (y_d, z_d) of D; element m of a loaded row is x = Xt - RP - R + m ----
    // Waves with no row skip phase 1 (a scalar branch); in a partly filled
    // wave the lanes past the last row repeat it, and only their LDS stores
    // are masked, so phase 1 has no divergent control flow.
    const bool p1own = tid < C::NROWS;
    const bool p1wav = C::NROWS >= NT || __builtin_amdgcn_readfirstlane(tid >> 6) * 64 < C::NROWS;
    const int r1 = p1own ? tid : C::NROWS - 1;
    const int yd = r1 % C::DY, zd = r1 / C::DY;
    // per-lane offsets sit at d = (-RS, -RS, -RS): the scalar offsets added
    // below (unsigned in the buffer address) are then all >= 0
    const unsigned rbase = 4u * (unsigned)pidx(Xt - RP - C::R, Yt - 2 * RP + yd - RS, Zt - 2 * RP + zd - RS);
    const int sxw = zd * C::SXP + yd * C::SXR;
    float cen[C::EX];
    if (p1wav) {
#pragma unroll
        for (int i = 0; i < C::EX; ++i) cen[i] = pad_load(ru, rbase, 4 * (RS * pxy + RS * px + RS + i));
    }
    // quarter of an extra row (rows NR1 ... NROWS - 1): 4 x outputs from 4 + 2 RP centres
    constexpr int EXQ = 4 + 2 * RP;
    const int xr = C::NR1 + tid / 4, xj = tid % 4;
    const int ydx = xr % C::DY, zdx = xr / C::DY;
    const unsigned rbasex =
        C::NRX > 0 ? 4u * (unsigned)pidx(Xt - RP - C::R + 4 * xj, Yt - 2 * RP + ydx - RS, Zt - 2 * RP + zdx - RS) : 0u;
    const int sxwx = zdx * C::SXP + ydx * C::SXR + 4 * xj;
    float cenx[EXQ];
    if constexpr (C::NRX > 0) {
#pragma unroll
        for (int i = 0; i < EXQ; ++i) cenx[i] = pad_load(ru, rbasex, 4 * (RS * pxy + RS * px + RS + i));
    }

    // ---- phase-2 item ----
    const bool p2ok = tid < C::P2N;
    const int it2 = p2ok ? tid : 0;
    const int xv = it2 % C::NXV, zp2 = (it2 / C::NXV) % C::DZ, yr2 = it2 / (C::NXV * C::DZ);
    const int p2src = zp2 * C::SXP + yr2 * C::RY * C::SXR + xv * C::VX;
    const int p2dst = zp2 * C::SYP + yr2 * C::RY * 16 + xv * C::VX;

    // ---- phase-3/4 column: x_q = lane within a 16-lane row, y_q = row ----
    const int xq = tid % 16, yq = tid / 16;
    const bool inT = xq >= RP && xq < 16 - RP && yq >= RP && yq < C::QY - RP;
    const int cx = Xt - RP + xq, cy = Yt - RP + yq, cz0 = Zt - RP;
    const unsigned colG = 16u * (unsigned)pidx(cx - RS, cy - RS, cz0 - RS);
    const int s0 = RS * pxy + RS * px + RS;  // d = 0
    float uc[QZ], Gc[QZ], GYc[QZ];
#pragma unroll
    for (int z = 0; z < QZ; ++z) {
        const float4 g = pad_load4(rg, colG, 16 * (s0 + z * pxy));
        Gc[z] = g.x;
        GYc[z] = g.y;
        uc[z] = g.z;
    }
    const int sxyr = yq * 16 + xq;  // s_xy column base
    const int ezw = yq * 16 + xq;   // boxed-E column base
    int ezr[C::W];  // y neighbours yq - RP ... yq + RP (clamped in the halo rows, unused there)
#pragma unroll
    for (int j = 0; j < C::W; ++j) {
        const int yy = yq - RP + j;
        ezr[j] = (yy < 0 ? 0 : (yy >= C::QY ? C::QY - 1 : yy)) * 16 + xq;
    }
    // y box of the boxed E at plane z (pairwise tree over the 2 RP + 1 rows)
    auto ybox = [&](int z) {
        const float *e = EZB + z * C::EZP;
        if constexpr (RP == 0) return e[ezr[0]];
        else if constexpr (RP == 1) return (e[ezr[0]] + e[ezr[1]]) + e[ezr[2]];
        else return ((e[ezr[0]] + e[ezr[1]]) + e[ezr[2]]) + (e[ezr[3]] + e[ezr[4]]);
    };
    // SYM visits only the offsets after d = 0 in (dz, dy, dx) order; d = 0
    // itself contributes G(z) (w_0 = 1, u~(z) - u~(z) = 0)
    float gacc[GZ];
#pragma unroll
    for (int z = 0; z < GZ; ++z) gacc[z] = (SYM && part == 0) ? Gc[z + RP] : 0.0f;
    // sum of S E over the tile's positions: sum_d w S t1 = (1/2) sum_d S E / (-2c)
    // summed over the whole extended domain (t2_d(x) = t1_-d(x+d))
    float dh = 0.0f;
    // T-local scatter base: position (x_t, y_t, z_t = 0) of the accumulator
    const int scb = (yq - RP + RS) * C::SCX + (xq - RP + RS);
    if constexpr (SYM) {
        for (int i = tid; i < C::SCN; i += NT) SC[i] = 0.0f;  // ordered by the loop's barriers
    }

    constexpr int NW = 2 * RS + 1, NOFF = NW * NW * NW;
    constexpr int O0 = SYM ? (NOFF + 1) / 2 : 0;  // first offset visited
    // part p visits the offset rows (dy, dz) [OR0 + p NRW / np, OR0 + (p + 1) NRW / np);
    // part 0 also takes the partial row before OR0.  Every part then starts its
    // shifted-row windows at a row start (or at O0), as one part over all rows does.
    constexpr int OR0 = (O0 + NW - 1) / NW, NRW = NW * NW - OR0;
    const int ob = part == 0 ? O0 : (OR0 + part * NRW / np) * NW;
    const int oe = (OR0 + (part + 1) * NRW / np) * NW;
    // d of the first offset: the scalar buffer offset and the scatter offset follow
    const int dxb = ob % NW - RS, dyb = (ob / NW) % NW - RS, dzb = ob / (NW * NW) - RS;
    // scalar offset of d relative to d = (-RS, -RS, -RS) (always >= 0), stepped
    // through the offsets in (dz, dy, dx) order
    auto sd_next = [&](int sd, int o) {  // offset index o -> o + 1
        const int ix = o % NW;
        if (ix < NW - 1) return sd + 1;
        const int iy = (o / NW) % NW;
        if (iy < NW - 1) return sd + px - (NW - 1);
        return sd + pxy - (NW - 1) * px - (NW - 1);
    };
    auto load_col = [&](int sd, float *ud, float *Gd, float *GYd) {
#pragma unroll
        for (int z = 0; z < QZ; ++z) {
            const float4 g = pad_load4(rg, colG, 16 * (sd + z * pxy));
            Gd[z] = g.x;
            GYd[z] = g.y;
            ud[z] = g.z;
        }
    };
    // The shifted samples of this thread's phase-1 row for all dx of one
    // (dy, dz): loaded once at the row start (a lane reads its own row, so
    // these loads touch a cache line per lane).  dx is unrolled, so offset j
    // of a row reads win[i + j] at a static register index.
    constexpr int NWIN = C::EX + 2 * RS, NWINX = EXQ + 2 * RS;
    float win[NWIN];
    float winx[NWINX];
    // the column values at x + d, one offset ahead
    float ud[QZ], Gd[QZ], GYd[QZ];
    int sd = s0 + dzb * pxy + dyb * px + dxb;      // 0 at d = (-RS, -RS, -RS)
    int scd = dzb * C::SCP + dyb * C::SCX + dxb;   // SYM: scatter offset of d (dz >= 0)
    load_col(sd, ud, Gd, GYd);

    for (int row = ob / NW; row * NW < oe; ++row) {
        // this part's offsets of the row: j in [jb, je) (wave-uniform)
        const int jb = ob > row * NW ? ob - row * NW : 0;
        const int je = oe < row * NW + NW ? oe - row * NW : NW;
        if (p1wav) {
            const int sr = sd - jb;  // scalar offset of the row start (dx = -RS)
#pragma unroll
            for (int m = 0; m < NWIN; m += 2) {  // row starts are 8-byte aligned (even padded x)
                const float2 t = pad_load2(ru, rbase, 4 * (sr + m));
                win[m] = t.x;
                win[m + 1] = t.y;
            }
            if constexpr (C::NRX > 0) {
#pragma unroll
                for (int m = 0; m < NWINX; m += 2) {
                    const float2 t = pad_load2(ru, rbasex, 4 * (sr + m));
                    winx[m] = t.x;
                    winx[m + 1] = t.y;
                }
            }
        }
        static_for<0, NW - 1>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < jb || j >= je) return;
            const int o = row * NW + j;
            const int sn = (o + 1 < oe) ? sd_next(sd, o) : sd;  // next offset (reloads the last one at the end)
            // phase 1: x sums of squared differences on the rows of D
            if (p1wav) {
                float qv[C::EX];
#pragma unroll
                for (int i = 0; i < C::EX; ++i) {
                    const float e = cen[i] - win[i + j];
                    qv[i] = e * e;
                }
                float sx[C::TX];
                box_tree<C::W, C::TX>(qv, sx);
                float4 *dst = reinterpret_cast<float4 *>(SX + sxw);
                if (p1own) {
#pragma unroll
                    for (int q = 0; q < C::TX / 4; ++q)
                        dst[q] = make_float4(sx[4 * q], sx[4 * q + 1], sx[4 * q + 2], sx[4 * q + 3]);
                }
                if constexpr (C::NRX > 0) {
                    float qx[EXQ];
#pragma unroll
                    for (int i = 0; i < EXQ; ++i) {
                        const float e = cenx[i] - winx[i + j];
                        qx[i] = e * e;
                    }
                    float s4[4];
                    box_tree<C::W, 4>(qx, s4);
                    *reinterpret_cast<float4 *>(SX + sxwx) = make_float4(s4[0], s4[1], s4[2], s4[3]);
                }
            }
            lds_barrier();
            // phase 2: y sums
            if (p2ok) rows_phase2<C>(SX + p2src, SXY + p2dst);
            lds_barrier();
            // phase 3: z sums, weights, E, the G(z+d) w term, dL/dh terms
            float v[QZ + 2 * RP];
#pragma unroll
            for (int i = 0; i < QZ + 2 * RP; ++i) v[i] = SXY[i * C::SYP + sxyr];
            float S[QZ];
            box_tree<C::W, QZ>(v, S);
            float EX[QZ], du[GZ], Gw[GZ];
#pragma unroll
            for (int z = 0; z < QZ; ++z) {
                const float w = exp2_fast(S[z] * kw);
                const float t1 = fmaf(Gc[z], ud[z], -GYc[z]);
                const float t2 = fmaf(Gd[z], uc[z], -GYd[z]);
                // The empty asm makes E a value of its own: otherwise its
                // product is fused into the x box's first add (an fma whose
                // addend is the shifted lane), and that add can no longer take
                // the DPP operand itself -- one extra move per z.  Every use of
                // E reads the pinned value, so the pin itself needs no copy.
                float E = (negc2 * w) * (t1 + t2);
                asm("" : "+v"(E));
                if (z >= RP && z < QZ - RP) {
                    gacc[z - RP] = fmaf(Gd[z], w, gacc[z - RP]);
                    dh = fmaf(S[z], E, dh);
                    du[z - RP] = uc[z] - ud[z];
                    if constexpr (SYM) Gw[z - RP] = Gc[z] * w;
                }
                // x box across the 16-lane row
                EX[z] = dpp_box<RP>(E);
            }
            // The next offset's column values fly across phase 4 and the next
            // phases 1-2.  Every value computed from this offset's ones is
            // pinned before the loads (volatile asm keeps the loads below it),
            // so the loads can land in the same registers: no copies.
#pragma unroll
            for (int z = 0; z < GZ; ++z) {
                asm volatile("" : "+v"(gacc[z]), "+v"(du[z]));
                if constexpr (SYM) asm volatile("" : "+v"(Gw[z]));
            }
#pragma unroll
            for (int z = 0; z < QZ; ++z) asm volatile("" : "+v"(EX[z]));
            asm volatile("" : "+v"(dh));
            load_col(sn, ud, Gd, GYd);
            float EZ[GZ];
            box_tree<C::W, GZ>(EX, EZ);
#pragma unroll
            for (int z = 0; z < GZ; ++z) EZB[z * C::EZP + ezw] = EZ[z];
            lds_barrier();
            // phase 4: y box of the boxed E, the 2 boxE (u~(z) - u~(z+d))
            // term; SYM: the -d term of position x + d, G(x) w_d(x) + boxE_d(x)
            // (u~(x+d) - u~(x)), added into the accumulator (one add per
            // position and offset: the order of the adds is fixed)
            if constexpr (SYM) {
                float A[GZ];
#pragma unroll
                for (int z = 0; z < GZ; ++z) {  // all reads first: the adds below may alias them
                    const float bE = ybox(z);
                    const float pz = bE * du[z];
                    gacc[z] = gacc[z] + pz;
                    A[z] = Gw[z] - pz;
                }
                // Plain read-add-write: within one offset the targets x + d of
                // distinct positions are distinct, and the barriers between
                // offsets order the updates (LDS float atomics cost ~12x as much).
                if (inT) {
                    float *sc = SC + scb + scd;
                    float cur[GZ];
#pragma unroll
                    for (int z = 0; z < GZ; ++z) cur[z] = sc[z * C::SCP];
#pragma unroll
                    for (int z = 0; z < GZ; ++z) sc[z * C::SCP] = cur[z] + A[z];
                }
                // scatter offset of the next d: dz SCP + dy SCX + dx
                scd += j < NW - 1 ? 1
                                  : ((o / NW) % NW < NW - 1 ? C::SCX - (NW - 1)
                                                            : C::SCP - (NW - 1) * C::SCX - (NW - 1));
            } else {
#pragma unroll
                for (int z = 0; z < GZ; ++z) {
                    const float bE = ybox(z);
                    gacc[z] = fmaf(bE, du[z], gacc[z]);
                }
            }
            sd = sn;
        });
    }

    if constexpr (SYM) {
        // the tile's own positions get their gathered terms, then the whole
        // scatter region goes to the tile's slot; bwd_sym_combine sums the
        // overlapping regions in a fixed order
        __syncthreads();
        if (inT) {
#pragma unroll
            for (int z = 0; z < GZ; ++z) SC[scb + z * C::SCP] += gacc[z];
        }
        __syncthreads();
        float *dst = a.gsc + (size_t)unit * C::SCN;
        for (int i = tid; i < C::SCN; i += NT) dst[i] = SC[i];
    } else if (inT) {  // ---- gradient of the tile's extended-domain positions ----
        const long long ex = a.nx + 2LL * C::R, ey = a.ny + 2LL * C::R, ez = a.nz + 2LL * C::R;
        const long long x = cx + C::R, y = cy + C::R;
        if (x < ex && y < ey) {
#pragma unroll
            for (int z = 0; z < GZ; ++z) {
                const long long zz = cz0 + RP + z + C::R;
                if (zz < ez) a.gext[x + ex * (y + ey * zz)] = gacc[z];
            }
        }
    }
    // ---- dL/dh: positions of T only (G, GY vanish outside the volume) ----
    __syncthreads();
    float *red = lds;
    red[tid] = inT ? dh : 0.0f;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] = red[tid] + red[tid + s];
        __syncthreads();
    }
    if (tid == 0) a.dh_part[blockIdx.x] = red[0];
}

// Padded u~ (reflect-101 on [-R, n+R), 0 beyond) and the interleaved
// (G, GY, u~, 0) with (G, GY) = (g/Z, g*y/Z) inside the volume and (0, 0)
// outside; both with origin offset H.
__global__ __launch_bounds__(256) void bwd_pad(const float *u, const float *g, const float *y, const float *Z, int nx,
                                               int ny, int nz, float *pu, float4 *pg, int px, int py, int pz, int H, int R)
{
    // The host checks 16 px py pz < 2^32, so every index here fits 32 bits.  Blocks stride over
    // the padded z planes (blockIdx.y) and over the x-y plane.
    const unsigned plane = (unsigned)px * (unsigned)py, nxy = (unsigned)nx * (unsigned)ny;
    for (int zp = blockIdx.y; zp < pz; zp += gridDim.y) {
        const int z = zp - H;
        const bool zin = z >= -R && z < nz + R, zc = z >= 0 && z < nz;
        const unsigned zo = zin ? nxy * (unsigned)reflect101(z, nz) : 0u;
        float *puz = pu + (unsigned)zp * plane;
        float4 *pgz = pg + (unsigned)zp * plane;
        for (unsigned p = blockIdx.x * 256u + threadIdx.x; p < plane; p += gridDim.x * 256u) {
            const int yp = (int)(p / (unsigned)px);
            const int x = (int)(p - (unsigned)yp * (unsigned)px) - H, yy = yp - H;
            float v = 0.0f;
            float2 gg = make_float2(0.0f, 0.0f);
            if (zin && x >= -R && x < nx + R && yy >= -R && yy < ny + R) {
                const unsigned o = zo + (unsigned)reflect101(x, nx) + (unsigned)nx * (unsigned)reflect101(yy, ny);
                v = u[o];
                if (zc && x >= 0 && x < nx && yy >= 0 && yy < ny) {
                    const float G = g[o] / Z[o];
                    gg = make_float2(G, G * y[o]);
                }
            }
            puz[p] = v;
            pgz[p] = make_float4(gg.x, gg.y, v, 0.0f);
        }
    }
}

// Symmetric form: the gradient on the extended domain is the sum of the
// scatter regions of the (up to 2 x 2 x 2) tiles that cover a position, added
// in a fixed tile order, and within a tile its np part slots in part order.
// Tile (tx, ty, tz) covers ext positions (xr, yr, zr) (origin at -R) with
// lx = xr - tx GX + RS in [0, SCX), ly = yr - ty GY + RS in [0, SCY),
// lz = zr - tz GZ in [0, SCZ), at lz SCP + ly SCX + lx of each of its slots.
template <int RP, int RS>
__global__ __launch_bounds__(256) void bwd_sym_combine(const float *__restrict__ gsc, float *__restrict__ gext, int ex,
                                                       int ey, int ez, int ntx, int nty, int ntz, int np)
{
    using C = BwdRowsCfg<RP, RS, true>;
    // one extended z plane per blockIdx.y step; ex ey < 2^32 (the host's padded-copy check bounds it)
    const unsigned plane = (unsigned)ex * (unsigned)ey;
    for (int zr = blockIdx.y; zr < ez; zr += gridDim.y) {
        const int tz1 = min(zr / C::GZ, ntz - 1);
        float *gz = gext + (long long)zr * plane;
        for (unsigned p = blockIdx.x * 256u + threadIdx.x; p < plane; p += gridDim.x * 256u) {
            const int yr = (int)(p / (unsigned)ex), xr = (int)(p - (unsigned)yr * (unsigned)ex);
            const int tx1 = min((xr + RS) / C::GX, ntx - 1), ty1 = min((yr + RS) / C::GY, nty - 1);
            float acc = 0.0f;
            for (int tz = max(tz1 - 2, 0); tz <= tz1; ++tz) {
                const int lz = zr - tz * C::GZ;
                if (lz >= C::SCZ) continue;
                for (int ty = max(ty1 - 2, 0); ty <= ty1; ++ty) {
                    const int ly = yr - ty * C::GY + RS;
                    if (ly >= C::SCY) continue;
                    for (int tx = max(tx1 - 2, 0); tx <= tx1; ++tx) {
                        const int lx = xr - tx * C::GX + RS;
                        if (lx >= C::SCX) continue;
                        const long long tile = tx + (long long)ntx * (ty + (long long)nty * tz);
                        const float *slot = gsc + tile * np * C::SCN + lz * C::SCP + ly * C::SCX + lx;
                        for (int p = 0; p < np; ++p) acc += slot[(long long)p * C::SCN];
                    }
                }
            }
            gz[p] = acc;
        }
    }
}

template <int RP, int RS, bool SYM>
constexpr size_t bwd_rows_lds()
{
    return BwdRowsCfg<RP, RS, SYM>::LDS_FLOATS * sizeof(float);
}

// Workgroups per tile for the symmetric form: a launch of t tiles on a device
// that holds `slots` of them at once takes about ceil(np t / slots) / np
// rounds of one tile's work when each tile's offset rows are split over np
// workgroups.  More parts are taken only when they cut that by at least 5%
// (each part repeats the tile's loads and writes its own scatter slot).
int bwd_sym_parts(long long t, long long slots, int nrows)
{
    int best = 1;
    double cbest = (double)((t + slots - 1) / slots);
    for (int np = 2; np <= 4 && np <= nrows; ++np) {
        const double c = (double)((np * t + slots - 1) / slots) / np;
        if (c < 0.95 * cbest) {
            best = np;
            cbest = c;
        }
    }
    return best;
}

// The (G, GY, u~, 0) copy spans [-H, nt * G + RP) on each axis at 16 bytes per
// position, addressed with 32-bit byte offsets: it must stay below 4 GiB.
template <int RP, int RS, bool SYM>
bool bwd_rows_fits_t(int nx, int ny, int nz)
{
    using C = BwdRowsCfg<RP, RS, SYM>;
    const long long ntx = (nx + 2LL * C::R + C::GX - 1) / C::GX, nty = (ny + 2LL * C::R + C::GY - 1) / C::GY,
                    ntz = (nz + 2LL * C::R + C::GZ - 1) / C::GZ;
    const long long px = ntx * C::GX + C::RP + C::H, py = nty * C::GY + C::RP + C::H,
                    pz = ntz * C::GZ + C::RP + C::H;
    return 16 * px * py * pz < (1LL << 32) && ntx * nty * ntz * 4 < (1LL << 31);
}

// 0: parts per tile chosen by bwd_sym_parts; 1-4: forced (gnlm_set_backward_parts)
std::atomic<int> g_bwd_parts{0};

template <int RP, int RS, bool SYM>
hipError_t bwd_rows_t(const float *u, const float *y, const float *Z, const float *g, int nx, int ny, int nz,
                      float h, float *gu, float *gh, hipStream_t s)
{
    using C = BwdRowsCfg<RP, RS, SYM>;
    if (!bwd_rows_fits_t<RP, RS, SYM>(nx, ny, nz)) return hipErrorNotSupported;
    BwdRowsArgs a;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.ntx = (nx + 2 * C::R + C::GX - 1) / C::GX;
    a.nty = (ny + 2 * C::R + C::GY - 1) / C::GY;
    a.ntz = (nz + 2 * C::R + C::GZ - 1) / C::GZ;
    // positions read: [-H, nt * G + RP) on each axis
    const int px = a.ntx * C::GX + C::RP + C::H, py = a.nty * C::GY + C::RP + C::H,
              pz = a.ntz * C::GZ + C::RP + C::H;
    const long long npad = (long long)px * py * pz;
    const long long ext = (nx + 2LL * C::R) * (ny + 2LL * C::R) * (nz + 2LL * C::R);
    const int nblocks = a.ntx * a.nty * a.ntz;
    const int P = C::W * C::W * C::W;
    const double c = 1.0 / ((double)P * (double)h * (double)h);
    a.k = (float)weight_scale(P, h);
    a.negc2 = (float)(-2.0 * c);
    const size_t lds = bwd_rows_lds<RP, RS, SYM>();
    hipError_t e = hipFuncSetAttribute((const void *)nlm3d_bwd_rows<RP, RS, SYM>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    a.nparts = 1;
    if (SYM) {
        constexpr int NW = 2 * RS + 1, O0 = (NW * NW * NW + 1) / 2, NRW = NW * NW - (O0 + NW - 1) / NW;
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)nlm3d_bwd_rows<RP, RS, SYM>, NT,
                                                         lds) == hipSuccess &&
            per_cu > 0)
            a.nparts = bwd_sym_parts(nblocks, (long long)per_cu * device_cus(), NRW);
        const int forced = g_bwd_parts.load(std::memory_order_relaxed);
        if (forced > 0) a.nparts = std::min(forced, NRW);
    }
    const int nunits = nblocks * a.nparts;
    float *ws = nullptr;
    // layout: (G, GY, u~, 0) copy | u~ copy | gradient on the extended domain | dL/dh partials
    //         | scatter regions (SYM, one per unit)
    const size_t nfl = (size_t)(4 * npad + npad + ext + nunits) + (size_t)nunits * C::SCN;
    e = hipMallocAsync((void **)&ws, sizeof(float) * nfl, s);
    if (e != hipSuccess) return e;
    float4 *pg = reinterpret_cast<float4 *>(ws);
    float *pu = ws + 4 * npad;
    a.gext = pu + npad;
    a.dh_part = a.gext + ext;
    a.gsc = SYM ? a.dh_part + nunits : nullptr;
    a.pu = PadVol{pu, px, px * py, (unsigned)(4 * npad)};
    a.pg = PadVol{reinterpret_cast<const float *>(pg), px, px * py, (unsigned)(16 * npad)};
    const unsigned pgx = (unsigned)std::min<long long>(((long long)px * py + 255) / 256, 1024);
    hipLaunchKernelGGL(bwd_pad, dim3(pgx, (unsigned)std::min(pz, 65535)), dim3(256), 0, s, u, g, y, Z, nx, ny, nz, pu,
                       pg, px, py, pz, C::H, C::R);
    e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL((nlm3d_bwd_rows<RP, RS, SYM>), dim3(nunits), dim3(NT), lds, s, a);
        e = hipGetLastError();
    }
    if (SYM && e == hipSuccess) {
        const int ex = nx + 2 * C::R, ey = ny + 2 * C::R, ez = nz + 2 * C::R;
        const unsigned cgx = (unsigned)std::min<long long>(((long long)ex * ey + 255) / 256, 1024);
        hipLaunchKernelGGL((bwd_sym_combine<RP, RS>), dim3(cgx, (unsigned)std::min(ez, 65535)), dim3(256), 0, s, a.gsc,
                           a.gext, ex, ey, ez, a.ntx, a.nty, a.ntz, a.nparts);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = bwd3d_fold_launch(a.gext, gu, nx, ny, nz, C::R, s);
    // the partials sum S E: dL/dh = (2c/h) sum_all-d w S t1 = -(1/(2h)) sum_all-d S E,
    // and the symmetric form visits half of the offsets (the other half has the same sum)
    if (e == hipSuccess && gh)
        e = bwd3d_reduce_dh_launch(a.dh_part, nunits, (float)((SYM ? -1.0 : -0.5) / (double)h), gh, s);
    const hipError_t e2 = hipFreeAsync(ws, s);
    return e != hipSuccess ? e : e2;
}

#define GNLM_BWD_ROWS_CASES(X)                                                                \
    X(0, 1) X(0, 2) X(0, 3) X(0, 4) X(0, 5) X(0, 6) X(0, 7) X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(1, 5) \
    X(1, 6) X(1, 7) X(2, 1) X(2, 2) X(2, 3) X(2, 4) X(2, 5) X(2, 6) X(2, 7)

std::atomic<int> g_bwd_form{kBwdAuto};

}  // namespace

int bwd3d_set_form(int form)
{
    if (form < kBwdAuto || form > kBwdSymmetric) return -1;
    g_bwd_form.store(form, std::memory_order_relaxed);
    return 0;
}

bool bwd3d_rows_supported(int r_p, int r_s)
{
#define GNLM_BR_SUP(RP, RS)                                                             \
    if (r_p == RP && r_s == RS)                                                         \
        return bwd_rows_lds<RP, RS, false>() <= kLdsLimitBytes && bwd_rows_lds<RP, RS, true>() <= kLdsLimitBytes;
    GNLM_BWD_ROWS_CASES(GNLM_BR_SUP)
#undef GNLM_BR_SUP
    return false;
}

int bwd3d_set_parts(int nparts)
{
    if (nparts < 0 || nparts > 4) return -1;
    g_bwd_parts.store(nparts, std::memory_order_relaxed);
    return 0;
}

bool bwd3d_rows_symmetric_form() { return g_bwd_form.load(std::memory_order_relaxed) != kBwdGather; }

bool bwd3d_rows_fits(int nx, int ny, int nz, int r_p, int r_s, bool sym)
{
#define GNLM_BR_FITS(RP, RS) \
    if (r_p == RP && r_s == RS) return sym ? bwd_rows_fits_t<RP, RS, true>(nx, ny, nz) : bwd_rows_fits_t<RP, RS, false>(nx, ny, nz);
    GNLM_BWD_ROWS_CASES(GNLM_BR_FITS)
#undef GNLM_BR_FITS
    return false;
}

hipError_t bwd3d_rows_launch(const float *u, const float *y, const float *Z, const float *g, int nx, int ny, int nz,
                             int r_p, int r_s, float h, float *gu, float *gh, hipStream_t s)
{
    const int form = g_bwd_form.load(std::memory_order_relaxed);
#define GNLM_BR_RUN(RP, RS)                                                                        \
    if (r_p == RP && r_s == RS) {                                                                  \
        if (form != kBwdGather)                                                                    \
            return bwd_rows_t<RP, RS, true>(u, y, Z, g, nx, ny, nz, h, gu, gh, s);                 \
        return bwd_rows_t<RP, RS, false>(u, y, Z, g, nx, ny, nz, h, gu, gh, s);                    \
    }
    GNLM_BWD_ROWS_CASES(GNLM_BR_RUN)
#undef GNLM_BR_RUN
    return hipErrorInvalidValue;
}

}  // namespace gnlm
