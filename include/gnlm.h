/*
 * gnlm.h -- C-ABI of the MI355X (gfx950) non-local-means library libgnlm.so.
 *
 * This is the drop-in boundary for the 3D-GNLM NLM path.  The reference's
 * host launchers are C++ static members that own all device memory for one
 * synchronous call:
 *   3D: NLMeansProcessor::NL_Means(unsigned char*, float*)   3DNLM/main.h:6,
 *       body 3DNLM/main.cu:7607-7680 (geometry from 3DNLM/config.h:4-9)
 *   2D: NLMeansProcessor::NL_Means(unsigned char*, float*, int W, int H,
 *       int sr_size, int nb_size, int IMAGE_SIZE_X, int IMAGE_SIZE_Y)
 *       "2D NLM/11x11/main.h:6", body "2D NLM/11x11/main.cu:472-539".
 * The C++ drop-in with those exact signatures is include/gnlm_main.h; it is
 * built on the host-pointer entry points at the end of this file.
 *
 * Conventions of the device-pointer entry points:
 *   - all pointers are device pointers owned by the caller; nothing is freed;
 *   - volumes are x-fastest, [nz][ny][nx] FP32, the reference's output layout
 *     (3DNLM/main.cu:125);
 *   - the input is unpadded; samples outside the volume use reflect-101, the
 *     padding the reference's driver builds on the host (3DNLM/nlm.cpp:36-41);
 *   - r_p / r_s are patch / search RADII: NS = 2 r_p + 1, SS = 2 r_s + 1;
 *   - h is the filtering parameter: w = exp(-(S / NS^d) / h^2) with S the
 *     patch sum of squared differences (3DNLM/main.cu:107-111);
 *   - every call is stream-ordered on `stream` (a hipStream_t, NULL = the
 *     default stream) and returns a gnlm_status; nothing throws across the ABI;
 *   - calls on distinct streams may run concurrently from different threads.
 * The reference swallows errors (3DNLM/cuda_utility.h:9-16 prints and
 * returns a cudaError_t that its callers ignore); this ABI returns them.
 */
#ifndef GNLM_H
#define GNLM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gnlm_status {
    GNLM_OK = 0,
    GNLM_EINVAL = 1,        /* bad dimension, radius, h or NULL pointer      */
    GNLM_ENOMEM = 2,        /* device or host allocation failed              */
    GNLM_EUNSUPPORTED = 3,  /* parameter combination not implemented         */
    GNLM_EHIP = 4,          /* HIP runtime error (launch or copy)            */
    GNLM_ESTATE = 5         /* object used in the wrong state                */
} gnlm_status;

/* Text for a status code (static storage).
 * Replaces the cudaGetErrorString printing of 3DNLM/cuda_utility.h:9-16. */
const char *gnlm_status_string(int status);

/* Library version, major*10000 + minor*100 + patch. */
int gnlm_version(void);

/* Forward-path selection for the *_ex entry points. */
enum { GNLM_PATH_AUTO = 0, GNLM_PATH_TILED = 1, GNLM_PATH_GENERIC = 2 };

/* 3D forward.  d_Z (nullable) receives the normaliser Z(x) = sum_d w_d(x),
 * which the backward pass reuses.
 * Replaces the kernel launches of 3DNLM/main.cu:7622-7667. */
int gnlm3d_forward(const float *d_in, float *d_out, float *d_Z, int nx, int ny, int nz,
                   int r_p, int r_s, float h, void *stream);

/* 3D forward with an explicit input layout and path.
 * in_pad == 0: d_in is the unpadded volume (reflect-101 on device);
 * in_pad  > 0: d_in is padded by in_pad >= r_p + r_s voxels per side, with
 *              row stride in_sy and plane stride in_sz (elements; 0 selects
 *              the dense values nx+2*in_pad and in_sy*(ny+2*in_pad)) --
 *              the layout the reference's launcher receives
 *              (3DNLM/main.cu:7613-7620).
 * path: GNLM_PATH_AUTO / _TILED / _GENERIC (tiled returns EUNSUPPORTED for
 * radii it does not handle). */
int gnlm3d_forward_ex(const float *d_in, int in_pad, long long in_sy, long long in_sz, float *d_out,
                      float *d_Z, int nx, int ny, int nz, int r_p, int r_s, float h, int path,
                      void *stream);

/* 3D forward that allocates nothing -- no counterpart in the reference; for
 * callers that cannot call hipMalloc (code under stream capture, tight
 * serving loops, harnesses that forbid allocation).  Same arguments, kernels
 * and bits as gnlm3d_forward.  The workspace the kernels need (the padded
 * copy of the input; for the large-patch tiles of r_p 5-8 also the running Z
 * total when d_Z is NULL) comes from a fixed device array of
 * GNLM_NOALLOC_SCRATCH_BYTES that each device the library runs on holds
 * until the process ends: about 240^3 voxels at patch 5^3 / search 11^3.
 * GNLM_ENOMEM (before anything is launched) for a volume that needs more; GNLM_EUNSUPPORTED where the route
 * is the generic per-offset path, which allocates.  All calls on one device
 * share its array, so calls that overlap on one device (other streams, other
 * host threads) race: serialise them. */
#define GNLM_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int gnlm3d_forward_noalloc(const float *d_in, float *d_out, float *d_Z, int nx, int ny, int nz, int r_p,
                           int r_s, float h, void *stream);

/* Which kernel a call would run, and whether it can run at all: pure host
 * logic, no device work and no allocation (no counterpart in the reference,
 * whose launcher runs one fixed kernel set per compile-time geometry,
 * 3DNLM/main.cu:7622-7667).  gnlm3d_forward_ex and gnlm3d_backward take
 * exactly this route and return the same status before touching the device.
 *   forward: *kernel = the tile variant (0-6, see gnlm_autotune) or
 *            GNLM_KERNEL_GENERIC.  The table's variant is used unless it is a
 *            rows-family variant (4-6) whose padded copy would pass 2^30
 *            floats (32-bit offsets); then the first of variants 3, 2, 1, 0
 *            that fits the radii (they read the input in place with 64-bit
 *            offsets), else the generic per-offset path.  GNLM_EUNSUPPORTED
 *            when path = GNLM_PATH_TILED and no tile variant can run it, or
 *            beyond 2^40 voxels.
 *   backward: *kernel = GNLM_KERNEL_BWD_ROWS_* (the rows adjoint in the
 *            selected form) while its 16-byte-per-position copy stays below
 *            4 GiB, else GNLM_KERNEL_BWD_BLOCK (the LDS-block adjoint, 64-bit
 *            indexing) where the radii fit its LDS image, else
 *            GNLM_KERNEL_BWD_GENERIC (any radii: six passes per offset
 *            through HBM scratch -- slow, but it differentiates every
 *            geometry, the reference's NS 17 / SS 13 default included);
 *            GNLM_EUNSUPPORTED at 2^31 voxels or more.
 * GNLM_EINVAL for bad dims, radii, path or a NULL kernel pointer. */
enum {
    GNLM_KERNEL_GENERIC = 100,
    GNLM_KERNEL_BWD_ROWS_GATHER = 201,
    GNLM_KERNEL_BWD_ROWS_SYMMETRIC = 202,
    GNLM_KERNEL_BWD_BLOCK = 203,
    GNLM_KERNEL_BWD_GENERIC = 204
};
int gnlm3d_forward_route(int nx, int ny, int nz, int r_p, int r_s, int path, int *kernel);
int gnlm3d_backward_route(int nx, int ny, int nz, int r_p, int r_s, int *kernel);

/* 2D forward on an nx * ny image (x fastest).
 * Replaces the kernel launches of "2D NLM/11x11/main.cu:486-525". */
int gnlm2d_forward(const float *d_in, float *d_out, int nx, int ny, int r_p, int r_s, float h,
                   void *stream);

/* 2D forward with explicit input layout and path (see gnlm3d_forward_ex;
 * in_sy is the row stride of a padded input). */
int gnlm2d_forward_ex(const float *d_in, int in_pad, long long in_sy, float *d_out, float *d_Z,
                      int nx, int ny, int r_p, int r_s, float h, int path, void *stream);

/* 2D forward of an 8-bit image, the reference's input type: its launcher
 * takes the padded unsigned char image ("2D NLM/11x11/main.cu:472-539"), h
 * on the 0-255 scale.  Layout as gnlm2d_forward_ex (in_pad = 0: unpadded,
 * reflect-101 on the device; in_pad >= r_p + r_s: padded, row stride in_sy
 * bytes).  The tiled kernel (r_p <= 5) reads the bytes itself; every partial
 * patch sum of squared 8-bit differences is an exact integer in FP32, so its
 * box sums run as sliding windows, and the result is bit-identical to
 * gnlm2d_forward_ex on the same samples as floats.  Other radii run the FP32
 * path on a float copy made in a stream-ordered workspace. */
int gnlm2d_forward_u8(const unsigned char *d_in, int in_pad, long long in_sy, float *d_out, float *d_Z,
                      int nx, int ny, int r_p, int r_s, float h, void *stream);

/* 3D adjoint (no counterpart in the reference, which has no backward pass;
 * it differentiates the forward above).  Given the forward's input d_in,
 * output d_out and normaliser d_Z (from gnlm3d_forward) and d_gout = dL/dy,
 * writes d_gin = dL/d(input) (reflect-101 copies folded back onto their
 * source voxels) and, if d_gh is not NULL, one float dL/dh.  Patch distances
 * are recomputed; nothing of size voxels x offsets is stored.  Results are
 * deterministic.  Scratch (padded copies, the gradient on the extended
 * domain [-R, n+R)^3 with R = r_p + r_s) comes from a stream-ordered
 * workspace released on `stream`.  Supported: r_p 0, 1, 2 with r_s <= 7 (the
 * rows-organised kernels, in the form gnlm_set_backward_form selects; larger
 * r_s as the LDS image of the fallback tile allows) and r_p = 3 with r_s <= 3
 * (the fallback tile) run on chip; every other radius pair, and volumes too
 * large for the rows kernels' 32-bit offsets where the fallback tile cannot
 * take them, run the any-radius per-offset passes (gnlm3d_backward_route). */
int gnlm3d_backward(const float *d_in, const float *d_out, const float *d_Z, const float *d_gout, int nx,
                    int ny, int nz, int r_p, int r_s, float h, float *d_gin, float *d_gh, void *stream);

/* Form of the rows-organised adjoint (r_p 0, 1 and 2; no counterpart in the reference), process-
 * wide: GNLM_BWD_GATHER visits every offset d and gathers its terms;
 * GNLM_BWD_SYMMETRIC visits half of them and, since S_d(x) = S_-d(x + d),
 * scatters each one's mirror term to x + d through an LDS accumulator (half
 * the arithmetic, a larger LDS image); GNLM_BWD_AUTO (default) is the
 * symmetric form, the faster one at every r_s (1.1-1.9x).  Both give
 * the gradient to FP32 rounding; each is deterministic.  GNLM_EINVAL for an
 * unknown form. */
enum { GNLM_BWD_AUTO = 0, GNLM_BWD_GATHER = 1, GNLM_BWD_SYMMETRIC = 2 };
int gnlm_set_backward_form(int form);

/* Workgroups per tile of the symmetric adjoint (no counterpart in the
 * reference), process-wide: 0 (default) lets the launch choose 1-4 from the
 * tile count and the device's occupancy, so that the last round of tiles is
 * not mostly idle; 1-4 forces that count (capped by the tile's offset rows).
 * Each part sums its share of the offsets and the parts are added in a fixed
 * order, so the gradient is deterministic for a given device and volume, but
 * its last bits depend on the part count -- and hence, under the default, on
 * the device's compute-unit count and the volume size.  GNLM_EINVAL outside
 * 0-4. */
int gnlm_set_backward_parts(int nparts);

/* dz-split forward of small volumes (no counterpart in the reference: its
 * launcher runs one thread block per tile of the volume, 3DNLM/main.cu:15-21),
 * process-wide.  With 1 (default), a rows-tile forward (r_p 1-3) whose tiles
 * fill less than half of the GPU's workgroup slots spreads each tile over
 * several workgroups, one per range of dz, which write per-dz partial sums
 * into the stream-ordered workspace (8 (2 r_s + 1) bytes per voxel, at most
 * 256 MiB); a combine pass adds them in dz order, the additions one workgroup
 * makes, so the output bits are those of the unsplit launch.  0 turns it off
 * (the unsplit launch).  The same switch covers the 2D tiled forward, whose
 * small images split each tile's dy range the same way (8 (2 r_s + 1) bytes
 * per pixel).  GNLM_EINVAL for other values. */
int gnlm_set_dz_split(int enable);

/* ---- batched persistent forward --------------------------------------
 * No counterpart in the reference (its launcher processes one compile-time
 * volume per call, 3DNLM/main.cu:7607-7680).  Many volumes of uneven size,
 * each with its own h and radii, run in ONE persistent launch replayed from a
 * HIP graph; workgroups pull tiles from a device work queue (longest first).
 * Outputs are bit-identical to per-volume gnlm3d_forward calls.
 * The plan keeps the descriptors (device pointers included): buffers must
 * stay allocated while the plan is used.  r_p <= 4 (r_p = 4 on the rows
 * tiles only) and the tiled path's radius limits apply (GNLM_EUNSUPPORTED
 * otherwise). */
typedef struct gnlm_batch_volume {
    const float *d_in;  /* unpadded input, x fastest        */
    float *d_out;       /* output, same layout              */
    float *d_Z;         /* nullable: normaliser per voxel   */
    int nx, ny, nz;
    int r_p, r_s;
    float h;
} gnlm_batch_volume;

typedef struct gnlm_batch_plan gnlm_batch_plan;

/* Builds the tile list, uploads it, and captures the graph (on the current
 * device).  Tile shapes come from the autotuner table (gnlm_autotune).  With
 * GNLM_TUNE_LAZY (gnlm_set_tuning) creation also times the persistent
 * kernel at the most workgroups per CU that fit and at one fewer -- running
 * the whole batch three times per count, on the plan's own buffers -- and
 * captures the faster; the count never changes the outputs. */
int gnlm_batch_plan_create(const gnlm_batch_volume *vols, int n, gnlm_batch_plan **plan);
/* Replays the captured graph on `stream` (one graph launch). */
int gnlm_batch_run(gnlm_batch_plan *plan, void *stream);
/* Same work as gnlm_batch_run, launched without the graph. */
int gnlm_batch_run_eager(gnlm_batch_plan *plan, void *stream);
int gnlm_batch_plan_info(const gnlm_batch_plan *plan, int *ntiles, int *nblocks);
int gnlm_batch_plan_destroy(gnlm_batch_plan *plan);

/* ---- autotuner --------------------------------------------------------
 * No counterpart in the reference (tile sizes are #defines there,
 * 3DNLM/main.cu:15-21).  Times every tile variant of the tiled 3D forward
 * that fits for (r_p, r_s) on a probe_n^3 synthetic volume (0 = 64) on
 * `stream`, records the fastest in a process-wide, mutex-guarded table used
 * by gnlm3d_forward and the batch plans, and returns it in *variant (nullable);
 * ms[v] (nullable, gnlm_num_variants() entries) receives each variant's time
 * (-1 if it does not fit).  A volume is always processed with the variant in
 * the table, so per-volume and batched launches give identical bits. */
int gnlm_autotune(int r_p, int r_s, int probe_n, void *stream, int *variant, float *ms);
int gnlm_tuned_variant(int r_p, int r_s);            /* -1 if unsupported */
int gnlm_set_variant(int r_p, int r_s, int variant);  /* manual override    */
int gnlm_num_variants(void);
void gnlm_reset_tuning(void);                        /* forget all choices  */
/* Variants: 0 = 16x16x16 direct sums, 1 = 16x16x8 direct sums (both bit-
 * identical to the generic path and the CPU oracle's FP32 order),
 * 2 = "slab" (16 x (16-2r_p) x (32-2r_p), DPP y sums, direct x/z sums,
 * pre-scaled input; agrees with the others to FP32 rounding), 3 = "slab2"
 * (variant 2 with 64/128-bit LDS accesses; bit-identical to 2), 4 = "rows"
 * (r_p 1 with r_s <= 8, r_p 2-4 with r_s <= 7: a padded aligned copy of the
 * input in a stream-ordered workspace, per-row register reuse across dx,
 * pairwise-tree box sums; agrees with the others to FP32 rounding),
 * 5 = "rows, short tiles" (variant 4 on 16 x 12 x 12 tiles, r_p 2 only:
 * fewer registers per thread, more workgroups per CU), 6 = "rows, big tiles"
 * (r_p 4-8 with r_s <= 7, the large patches of the reference's default
 * NS = 17: 16 x (32 - 2 r_p)^2 tiles on 1024 threads; per-volume launches
 * only; batch plans run r_p = 4 on variant 4).
 * Untuned default: 5, else 4, else 6 where they are instantiated, else 3
 * where it fits, else 2, else the first direct variant. */

/* ---- launch-shape tuning (automatic) -----------------------------------
 * No counterpart in the reference (block and grid shapes are #defines there,
 * 3DNLM/main.cu:15-21).  Within the variant the table selects, a forward has
 * up to three launch shapes that compute the SAME bits: register budgets
 * (waves per SIMD) and tile depths in z (the share of halo rows per tile):
 *   r_p = 1: 16x14x14 tiles at 3 or 4 per CU, or 16x14x28 tiles at 2 per CU;
 *   r_p = 2 (short tiles): 16x12x12 at 4 per CU, or 16x12x24 at 2 per CU;
 *   r_p = 3: 16x10x10 at 4 per CU (128 VGPRs) or 3 per CU (168 VGPRs);
 *   r_p = 4: 16x8x8 tiles or 16x24x24 tiles.
 * With GNLM_TUNE_LAZY (the default) the first gnlm3d_forward of a (device,
 * r_p, r_s, variant, size class) -- size class = half octaves of the voxel
 * count -- times each shape on the caller's own volume and buffers (it
 * waits for them once) and keeps the fastest for that class; later calls and
 * gnlm3d_forward_noalloc use it.  Calls under stream capture never tune.
 * GNLM_TUNE_OFF uses the default rule for untuned classes (r_p = 1: 4 per CU
 * when that saves a round of tiles, else 3; otherwise shape 0).  Outputs never
 * depend on which shape runs.  gnlm_reset_tuning forgets these choices too. */
enum { GNLM_TUNE_LAZY = 0, GNLM_TUNE_OFF = 1 };
int gnlm_set_tuning(int mode);
/* *shape = the shape recorded for this volume's class (-1: none yet, the
 * default rule applies), *nshapes = the candidates (1: no choice). */
int gnlm_forward_shape(int nx, int ny, int nz, int r_p, int r_s, int *shape, int *nshapes);
/* Record a shape for this volume's class (-1 forgets it). */
int gnlm_set_forward_shape(int nx, int ny, int nz, int r_p, int r_s, int shape);

/* uint8 -> float32 element conversion (n elements), for 8-bit inputs such as
 * the reference's (3DNLM/main.cu:7620 copies unsigned char volumes). */
int gnlm_u8_to_f32(const unsigned char *d_in, float *d_out, long long n, void *stream);

/* 3D forward of an 8-bit volume (the reference's input type: its launcher
 * copies an unsigned char volume to the device, 3DNLM/main.cu:7620, and its
 * kernels difference the samples as ints, 3DNLM/main.cu:59).  Replaces the
 * kernel launches of 3DNLM/main.cu:7622-7667 for 8-bit data.  Layout and
 * arguments as gnlm3d_forward_ex (in_pad, in_sy, in_sz in bytes = elements;
 * h on the samples' 0-255 scale, the reference's PARAMETER); d_out and d_Z
 * are FP32.  For r_p 6-8 (NS 13-17) with r_s 1-7, and r_p 5 (NS 11) with
 * r_s 3-7, the patch distances are formed exactly in 32-bit integers (an
 * 8-bit patch SSD is an integer below 2^31 for NS <= 17) and only the
 * weights are floating point; other radii,
 * and volumes past that kernel's 32-bit offsets (about 800^3), run the FP32
 * kernels on a float copy made in a stream-ordered workspace.  On 8-bit
 * samples both give the same definition to FP32 rounding. */
int gnlm3d_forward_u8(const unsigned char *d_in, int in_pad, long long in_sy, long long in_sz, float *d_out,
                      float *d_Z, int nx, int ny, int nz, int r_p, int r_s, float h, void *stream);

/* Which kernel gnlm3d_forward_u8 runs for this geometry (host logic only, no
 * device work; no counterpart in the reference): GNLM_KERNEL_U8_INTEGER for
 * the integer-distance kernel, else what gnlm3d_forward_route gives for the
 * FP32 kernels on the float copy.  Same statuses as gnlm3d_forward_route. */
enum { GNLM_KERNEL_U8_INTEGER = 110 };
int gnlm3d_forward_u8_route(int nx, int ny, int nz, int r_p, int r_s, int *kernel);

/* ---- host-pointer, synchronous entry points (the drop-in's backing) ----
 * They allocate device memory for one call, copy in, run, copy out and free,
 * as the reference's launchers do (3DNLM/main.cu:7609-7676).  (Parameter
 * names avoid the reference's config.h macros SS, NS, depth, IMAGE_SIZE_*.)
 *
 * 3D: h_in is the padded uint8 volume of (nx+ss+ns-2)*(ny+ss+ns-2)*(nz+ss+ns-2)
 * bytes, x fastest (3DNLM/nlm.cpp:48-61); h_out receives nx*ny*nz floats.
 * ss and ns are the search and patch window SIZES (odd; the reference's SS
 * and NS).  h is on the 0-255 scale (the reference hard-wires
 * PARAMETER = 10, 3DNLM/main.cu:13). */
int gnlm_nl_means_3d_u8(const unsigned char *h_in, float *h_out, int nx, int ny, int nz, int ss, int ns,
                        float h);

/* 2D: h_in is the padded uint8 image whose row stride is W (the reference
 * kernels index it with stride W, "2D NLM/11x11/main.cu:36"), padded by
 * (sr + nb - 2)/2 per side; h_out receives rows of IMAGE_SIZE_Y floats
 * (the reference's output stride, "2D NLM/11x11/main.cu:64") of which the
 * first IMAGE_SIZE_X are written, for IMAGE_SIZE_Y rows.  H is the padded
 * row count. */
int gnlm_nl_means_2d_u8(const unsigned char *h_in, float *h_out, int W, int H, int sr_size,
                        int nb_size, int image_size_x, int image_size_y, float h);

#ifdef __cplusplus
}
#endif

#endif /* GNLM_H */
