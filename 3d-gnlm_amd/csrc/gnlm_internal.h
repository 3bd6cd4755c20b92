// gnlm_internal.h -- launchers shared between the kernel files and the C-ABI
// layer (gnlm_api.cpp).  Host-side only; no torch types anywhere.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gnlm_device.h"
#include "../../include/gnlm.h"

namespace gnlm {

// Exponent scale of the weight: w = exp(-(S/P)/h^2) = exp2(S * k).
double weight_scale(int P, float h);

// Device scratch handed to a launch by a caller that must not allocate
// (gnlm3d_forward_noalloc): kernels that need a workspace take it from here
// instead of hipMallocAsync, or fail with hipErrorOutOfMemory when it is too
// small.  NULL everywhere else: those launches allocate on their stream.
struct Workspace {
    void *p;
    size_t bytes;
};

constexpr size_t kLdsLimitBytes = 160 * 1024;  // LDS per workgroup on gfx950

// ---- tiled 3D forward (nlm3d_fwd.hip) --------------------------------------
// Tile variants are enumerated in nlm3d_tile.h (Fwd3Variant).  A variant
// handles (r_p, r_s) when r_p is in 0..3 and the LDS image of its tile with
// the (r_p + r_s) halo fits in 160 KiB.
size_t fwd3d_lds_bytes(int r_p, int r_s, int variant);
bool fwd3d_variant_supported(int r_p, int r_s, int variant);
bool fwd3d_tiled_supported(int r_p, int r_s);
hipError_t fwd3d_tiled_launch_variant(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k,
                                      int variant, hipStream_t s, const Workspace *ws = nullptr, int shape = -1);
// Launch shapes of a variant for (r_p, r_s): the bit-identical alternatives the
// autotuner chooses between (1 where there is no choice).
int fwd3d_num_shapes(int r_p, int r_s, int variant);
// Whether shape runs anything other than shape 0 on this volume (the dz-split
// shape falls back to shape 0 where the split does not apply): the tuner
// times only the shapes that differ.
bool fwd3d_shape_distinct(const InDesc &in, int r_p, int r_s, int variant, int shape);
// Whether a rows-family variant's padded copy of an nx*ny*nz volume fits its
// 32-bit offsets.
bool fwd3d_rows_fits(int nx, int ny, int nz, int r_p, int r_s, int variant);
// The kernel a forward of this geometry runs (host logic only): the tile
// variant from the autotuner's table, or, where a rows-family variant's
// 32-bit offsets cannot address the volume, the first 64-bit tile variant
// that fits, else GNLM_KERNEL_GENERIC.  GNLM_EUNSUPPORTED when path is
// GNLM_PATH_TILED and no tile variant can run it.
int fwd3d_route(int nx, int ny, int nz, int r_p, int r_s, int path, int *kernel);

// ---- "rows" variant (nlm3d_rows.hip) ---------------------------------------
// r_p in {1, 2, 3, 4} with r_s up to 8 / 7 / 7 / 7; pads the input into a stream-ordered
// workspace first (caller-padded input needs pad >= r_p + r_s).
// Large patches (r_p 4-8): 16 x (32 - 2 r_p)^2 tiles on 1024 threads.
bool fwd3d_rows_big_supported(int r_p, int r_s);
size_t fwd3d_rows_big_lds_bytes(int r_p, int r_s);
hipError_t fwd3d_rows_big_launch(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k, hipStream_t s,
                                 const Workspace *ws = nullptr);
// short_tiles selects the 16 x 12 x 12 tiles (r_p = 2 only).
bool fwd3d_rows_supported(int r_p, int r_s, bool short_tiles);
size_t fwd3d_rows_lds_bytes(int r_p, int r_s, bool short_tiles);
hipError_t fwd3d_rows_launch(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k, bool short_tiles,
                             hipStream_t s, const Workspace *ws = nullptr, int shape = -1);
// Launch shapes of the rows tiles for (r_p, short): bit-identical builds and
// tile depths (nlm3d_rows.hip lists them); shape -1 = the default rule.
int fwd3d_rows_num_shapes(int r_p, bool short_tiles);
// The dz split would run for this volume (no caller workspace).
bool fwd3d_rows_split_applies(const InDesc &in, int r_p, int r_s, bool short_tiles);

// ---- autotuner (gnlm_tune.cpp) ---------------------------------------------
// Variant for (r_p, r_s): the tuned one if gnlm_autotune ran, otherwise the
// first supported variant.  Thread-safe.
int fwd3d_pick_variant(int r_p, int r_s);
// Launch shape for this volume and variant: the one recorded for its (device,
// radii, variant, size class), else -- when may_tune, tuning is on and the
// stream is not capturing -- the fastest of the variant's bit-identical shapes
// timed on these buffers (recorded), else -1 (the default rule).
int fwd3d_pick_shape(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k, int variant,
                     hipStream_t s, bool may_tune);
bool fwd3d_tuning_lazy();  // gnlm_set_tuning(GNLM_TUNE_LAZY) in force
void fwd3d_set_dz_split(bool on);  // gnlm_set_dz_split

// ---- tiled 2D forward (nlm2d_fwd.hip) --------------------------------------
bool fwd2d_tiled_supported(int r_p, int r_s);
hipError_t fwd2d_tiled_launch(const InDesc &in, float *out, float *Zout, int r_p, int r_s, float k,
                              hipStream_t s);
void fwd2d_set_dy_split(bool on);  // gnlm_set_dz_split (the 2D dy split)
// the 8-bit build: samples from in8, laid out as in describes (in.data unused)
hipError_t fwd2d_tiled_u8_launch(const InDesc &in, const unsigned char *in8, float *out, float *Zout, int r_p,
                                 int r_s, float k, hipStream_t s);

// ---- generic per-offset forward (nlm_generic.hip) --------------------------
// Any radii; 2D when rpz = rsz = 0 and nz = 1.  Workspace is allocated inside.
hipError_t fwd_generic_launch(const InDesc &in, float *out, float *Zout, int rp, int rs, int rpz, int rsz,
                              float k, hipStream_t s);

// ---- adjoint (nlm3d_bwd.hip, nlm3d_bwd_rows.hip) ----------------------------
bool bwd3d_supported(int r_p, int r_s);
bool bwd3d_block_supported(int r_p, int r_s);  // the LDS-block adjoint (64-bit indexing)
// The adjoint kernel for this geometry (host logic only): the rows kernels in
// the selected form where their 32-bit offsets address the volume, else the
// LDS-block tile; GNLM_EUNSUPPORTED when neither can run it.
int bwd3d_route(int nx, int ny, int nz, int r_p, int r_s, int *kernel);
// any-radius adjoint: per-offset passes through HBM scratch (nlm3d_bwd_generic.hip)
hipError_t bwd3d_generic_launch(const float *u, const float *y, const float *Z, const float *g, int nx, int ny,
                                int nz, int rp, int rs, float h, float *gu, float *gh, hipStream_t s);
hipError_t bwd3d_launch(const float *u, const float *y, const float *Z, const float *g, int nx, int ny, int nz,
                        int rp, int rs, float h, float *gu, float *gh, hipStream_t s);
// rows-organised adjoint for r_p 0-2, r_s in 1..7 (bwd3d_launch dispatches to it)
bool bwd3d_rows_supported(int r_p, int r_s);
hipError_t bwd3d_rows_launch(const float *u, const float *y, const float *Z, const float *g, int nx, int ny, int nz,
                             int r_p, int r_s, float h, float *gu, float *gh, hipStream_t s);
// shared tail kernels: fold the extended-domain gradient onto the volume
// through the reflect-101 preimages; fixed-order sum of dL/dh partials x scale
// Form of the rows adjoint (r_p 0-2): gather over all offsets, or the symmetric
// form (half of the offsets, each also scattered to x + d); auto is the
// symmetric form, the faster one at every instantiated r_s on MI355X.
enum { kBwdAuto = 0, kBwdGather = 1, kBwdSymmetric = 2 };
int bwd3d_set_form(int form);  // -1 for an unknown form
bool bwd3d_rows_symmetric_form();
int bwd3d_set_parts(int nparts);  // 0 = automatic, 1-4 forced; -1 if out of range
bool bwd3d_rows_fits(int nx, int ny, int nz, int r_p, int r_s, bool sym);
hipError_t bwd3d_fold_launch(const float *gext, float *gu, int nx, int ny, int nz, int R, hipStream_t s);

// Compute units of the current device (queried once per device; 256 if the query fails).
int device_cus();
hipError_t bwd3d_reduce_dh_launch(const float *part, int n, float scale, float *out, hipStream_t s);

// ---- 8-bit forward with integer patch distances (nlm3d_u8.hip) -------------
// r_p 6-8 with r_s 1-7 and r_p 5 with r_s 3-7; the input is the caller's
// 8-bit volume (in_pad = 0: unpadded, reflect-101 on device; in_pad >= r_p +
// r_s: padded with strides sy, sz in elements).  fwd3d_u8_fits: the radii are instantiated and the
// volume's copy and records fit 32-bit byte offsets.
bool fwd3d_u8_supported(int r_p, int r_s);
bool fwd3d_u8_fits(int nx, int ny, int nz, int r_p, int r_s);
hipError_t fwd3d_u8_launch(const unsigned char *d_in, int in_pad, long long sy, long long sz, float *out, float *Zout,
                           int nx, int ny, int nz, int r_p, int r_s, float k, hipStream_t s);

// ---- misc (gnlm_misc.hip) --------------------------------------------------
hipError_t u8_to_f32_launch(const uint8_t *in, float *out, long long n, hipStream_t s);
// any alignment (one element per thread)
hipError_t u8_to_f32_any_launch(const uint8_t *in, float *out, long long n, hipStream_t s);

}  // namespace gnlm
