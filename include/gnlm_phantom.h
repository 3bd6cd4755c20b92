/*
 * gnlm_phantom.h -- seeded synthetic inputs, exported by libgnlm.so.
 *
 * No counterpart in the reference, whose driver reads 8-bit BMP files
 * (3DNLM/nlm.cpp:31, "2D NLM/11x11/nlm.cpp:10").  The benchmark and test
 * workloads are synthetic phantoms instead; these generators make them
 * reproducible from a seed on any host (own uniform and Box-Muller
 * transforms over std::mt19937_64, so the values do not depend on the C++
 * library's distribution implementations).
 */
#ifndef GNLM_PHANTOM_H
#define GNLM_PHANTOM_H

#ifdef __cplusplus
extern "C" {
#endif

/* 3D ellipsoid phantom, x fastest.  `background` everywhere, then
 * n_ellipsoids axis-aligned ellipsoids (centre uniform over the volume,
 * semi-axes uniform in [axis_lo, axis_hi] voxels per axis, intensity uniform
 * in [0.1, 0.9]; later ones overwrite earlier ones), then i.i.d. Gaussian
 * noise of standard deviation sigma (no clipping).  h_clean (nullable)
 * receives the noise-free phantom.  Returns 0, or 1 on bad arguments. */
int gnlm_phantom3d(float *h_out, float *h_clean, int nx, int ny, int nz, int n_ellipsoids, float axis_lo,
                   float axis_hi, float background, float sigma, unsigned long long seed);

/* 2D piecewise-constant phantom: background 0, n_shapes axis-aligned
 * rectangles or discs (even odds; half-size uniform in
 * [size_lo, size_hi] * max(nx, ny); intensity uniform in [0, 1]), plus
 * sine_amp * sin(2 pi (x cos t + y sin t) / period) with period uniform in
 * [period_lo, period_hi] pixels and t uniform in [0, 2 pi), plus Gaussian
 * noise sigma.  Returns 0, or 1 on bad arguments. */
int gnlm_phantom2d(float *h_out, float *h_clean, int nx, int ny, int n_shapes, float size_lo, float size_hi,
                   float sine_amp, float period_lo, float period_hi, float sigma, unsigned long long seed);

/* FNV-1a 64-bit hash of n floats' bytes, for logging input identity. */
unsigned long long gnlm_hash_f32(const float *h, long long n);

#ifdef __cplusplus
}
#endif

#endif /* GNLM_PHANTOM_H */
