/*
 * nlm_oracle.c -- CPU restatement of the 3D-GNLM non-local-means path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * 3d-gnlm_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product path calls it.
 *
 * What it restates (reference paths relative to the reference repository):
 *   - padding:  3DNLM/nlm.cpp:36-41 and "2D NLM/11x11/nlm.cpp:15-19" pad the
 *     input with OpenCV BORDER_REFLECT_101 by (SS+NS-2)/2 = r_s + r_p per
 *     side.  oracle_reflect101() is that index rule (i<0 -> -i,
 *     i>=n -> 2n-2-i, folded again until inside, as OpenCV's
 *     borderInterpolate does for borders wider than the image).
 *   - layout:   x fastest, then y, then z (3DNLM/nlm.cpp:55-61,
 *     3DNLM/main.cu:125 writes out[ix + iy*X + iz*X*Y]).
 *   - math:     3DNLM/main.cu:58-68 (squared differences summed over the
 *     NS x NS x NS patch), :107-111 (w = exp(-(S/NS^3)/PARAMETER^2),
 *     gmid += w, result += w * candidate centre), :125 (out = result/gmid);
 *     the offset loops :52-54 run the full (2r_s+1)^3 cube, self included.
 *     "2D NLM/11x11/main.cu:36-63" is the same in 2D with nb x nb patches.
 *
 * What it deliberately does not reproduce: the reference kernels' mapping
 * artefacts (rows skipped when Y % 64 != 0, the odd-slice accumulator seeded
 * from the even slice at 3DNLM/main.cu:141, uint8-only input).  The oracle is
 * the mathematical definition those kernels implement.
 *
 * FP32 arithmetic order (the "canonical order", shared with the GPU
 * kernels so that the SSD stage agrees bit for bit):
 *   sx  = fma(e_{-r},e_{-r}, 0) then fma(e_i, e_i, sx) for i = -r+1 .. r
 *         (x fastest, left to right), e = u(x+p) - u(x+d+p)
 *   sxy = ((sx_{y-r} + sx_{y-r+1}) + ...) + sx_{y+r}
 *   S   = ((sxy_{z-r} + ...) + sxy_{z+r})
 *   w   = expf(-(S / P) / (h*h))            (P = (2r_p+1)^3, as main.cu:108)
 *   two-level accumulation: for each dz (3D) / dy (2D) a partial
 *     Zd = Zd + w,  Yd = fma(w, u(x+d), Yd)  over the inner offsets in
 *     (dy, dx) / (dx) order, then Z = Z + Zd, Y = Y + Yd
 *   out = Y / Z
 * (Summing 1331 weights into one running FP32 total lets its rounding grow
 * with the total; blocking by dz cuts that error about tenfold and makes the
 * FP32 result agree with FP64 to ~1e-6.)
 * The file is compiled with -ffp-contract=off so that only the explicit
 * fmaf() calls fuse.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* reflect-101 index rule (OpenCV BORDER_REFLECT_101), see header note  */
/* ------------------------------------------------------------------ */
long oracle_reflect101(long i, long n)
{
    if (n <= 1) return 0;
    long period = 2 * (n - 1);
    long m = i % period;
    if (m < 0) m += period;
    return (m < n) ? m : period - m;
}

static inline long R101(long i, long n) { return oracle_reflect101(i, n); }

/* ------------------------------------------------------------------ */
/* brute force 3D, FP64: the definition, no reuse at all                */
/* ------------------------------------------------------------------ */
void oracle_nlm3d_brute_f64(const double *u, int nx, int ny, int nz,
                            int rp, int rs, double h, double *y, double *Zout)
{
    const double P = (double)(2 * rp + 1) * (2 * rp + 1) * (2 * rp + 1);
    const double inv = 1.0 / (h * h);
#define U64(X, Y, Z) u[R101(X, nx) + (long)nx * (R101(Y, ny) + (long)ny * R101(Z, nz))]
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int yy = 0; yy < ny; ++yy)
            for (int x = 0; x < nx; ++x) {
                double Z = 0.0, Y = 0.0;
                for (int dz = -rs; dz <= rs; ++dz)
                    for (int dy = -rs; dy <= rs; ++dy)
                        for (int dx = -rs; dx <= rs; ++dx) {
                            double S = 0.0;
                            for (int pz = -rp; pz <= rp; ++pz)
                                for (int py = -rp; py <= rp; ++py)
                                    for (int px = -rp; px <= rp; ++px) {
                                        double e = U64(x + px, yy + py, z + pz) -
                                                   U64(x + dx + px, yy + dy + py, z + dz + pz);
                                        S += e * e;
                                    }
                            double w = exp(-(S / P) * inv);
                            Z += w;
                            Y += w * U64(x + dx, yy + dy, z + dz);
                        }
                long o = x + (long)nx * (yy + (long)ny * z);
                y[o] = Y / Z;
                if (Zout) Zout[o] = Z;
            }
#undef U64
}

/* ------------------------------------------------------------------ */
/* brute force 3D, FP32, canonical order                                */
/* ------------------------------------------------------------------ */
void oracle_nlm3d_brute_f32(const float *u, int nx, int ny, int nz,
                            int rp, int rs, float h, float *y, float *Zout)
{
    const float P = (float)((2 * rp + 1) * (2 * rp + 1) * (2 * rp + 1));
    const float hh = h * h;
#define U32(X, Y, Z) u[R101(X, nx) + (long)nx * (R101(Y, ny) + (long)ny * R101(Z, nz))]
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int yy = 0; yy < ny; ++yy)
            for (int x = 0; x < nx; ++x) {
                float Z = 0.0f, Y = 0.0f;
                for (int dz = -rs; dz <= rs; ++dz) {
                    float Zd = 0.0f, Yd = 0.0f;
                    for (int dy = -rs; dy <= rs; ++dy)
                        for (int dx = -rs; dx <= rs; ++dx) {
                            float S = 0.0f;
                            for (int pz = -rp; pz <= rp; ++pz) {
                                float sxy = 0.0f;
                                for (int py = -rp; py <= rp; ++py) {
                                    float sx = 0.0f;
                                    for (int px = -rp; px <= rp; ++px) {
                                        float e = U32(x + px, yy + py, z + pz) -
                                                  U32(x + dx + px, yy + dy + py, z + dz + pz);
                                        sx = fmaf(e, e, sx);
                                    }
                                    sxy = sxy + sx;
                                }
                                S = S + sxy;
                            }
                            float w = expf(-(S / P) / hh);
                            Zd = Zd + w;
                            Yd = fmaf(w, U32(x + dx, yy + dy, z + dz), Yd);
                        }
                    Z = Z + Zd;
                    Y = Y + Yd;
                }
                long o = x + (long)nx * (yy + (long)ny * z);
                y[o] = Y / Z;
                if (Zout) Zout[o] = Z;
            }
#undef U32
}

/* ------------------------------------------------------------------ */
/* brute force 2D (FP64 and FP32 canonical order)                       */
/* "2D NLM/11x11/main.cu:36-63": w = exp(-(S/nb^2)/PARAMETER^2)         */
/* ------------------------------------------------------------------ */
void oracle_nlm2d_brute_f64(const double *u, int nx, int ny, int rp, int rs,
                            double h, double *y, double *Zout)
{
    const double P = (double)(2 * rp + 1) * (2 * rp + 1);
    const double inv = 1.0 / (h * h);
#define U2(X, Y) u[R101(X, nx) + (long)nx * R101(Y, ny)]
#pragma omp parallel for schedule(static)
    for (int yy = 0; yy < ny; ++yy)
        for (int x = 0; x < nx; ++x) {
            double Z = 0.0, Y = 0.0;
            for (int dy = -rs; dy <= rs; ++dy)
                for (int dx = -rs; dx <= rs; ++dx) {
                    double S = 0.0;
                    for (int py = -rp; py <= rp; ++py)
                        for (int px = -rp; px <= rp; ++px) {
                            double e = U2(x + px, yy + py) - U2(x + dx + px, yy + dy + py);
                            S += e * e;
                        }
                    double w = exp(-(S / P) * inv);
                    Z += w;
                    Y += w * U2(x + dx, yy + dy);
                }
            long o = x + (long)nx * yy;
            y[o] = Y / Z;
            if (Zout) Zout[o] = Z;
        }
#undef U2
}

void oracle_nlm2d_brute_f32(const float *u, int nx, int ny, int rp, int rs,
                            float h, float *y, float *Zout)
{
    const float P = (float)((2 * rp + 1) * (2 * rp + 1));
    const float hh = h * h;
#define U2(X, Y) u[R101(X, nx) + (long)nx * R101(Y, ny)]
#pragma omp parallel for schedule(static)
    for (int yy = 0; yy < ny; ++yy)
        for (int x = 0; x < nx; ++x) {
            float Z = 0.0f, Y = 0.0f;
            for (int dy = -rs; dy <= rs; ++dy) {
                float Zd = 0.0f, Yd = 0.0f;
                for (int dx = -rs; dx <= rs; ++dx) {
                    float S = 0.0f;
                    for (int py = -rp; py <= rp; ++py) {
                        float sx = 0.0f;
                        for (int px = -rp; px <= rp; ++px) {
                            float e = U2(x + px, yy + py) - U2(x + dx + px, yy + dy + py);
                            sx = fmaf(e, e, sx);
                        }
                        S = S + sx;
                    }
                    float w = expf(-(S / P) / hh);
                    Zd = Zd + w;
                    Yd = fmaf(w, U2(x + dx, yy + dy), Yd);
                }
                Z = Z + Zd;
                Y = Y + Yd;
            }
            long o = x + (long)nx * yy;
            y[o] = Y / Z;
            if (Zout) Zout[o] = Z;
        }
#undef U2
}

/* ------------------------------------------------------------------ */
/* separable 3D path: the timed CPU baseline ("nlm.cpp CPU path")       */
/*                                                                      */
/* Same canonical FP32 order as the brute force above (so the two agree */
/* bit for bit), computed per offset with separable direct box sums on  */
/* z-slabs; OpenMP over slabs.  Outputs planes [z0, z1) of the volume.  */
/* Returns 0 on success, -1 on allocation failure.                      */
/* ------------------------------------------------------------------ */
#define SLAB 4
int oracle_nlm3d_separable_f32(const float *u, int nx, int ny, int nz,
                               int rp, int rs, float h, float *y,
                               int z0, int z1, int nthreads)
{
    const int R = rp + rs;
    const long ex = nx + 2L * R, ey = ny + 2L * R, ez = nz + 2L * R;
    /* padded copy: ũ on [-R, n+R) per axis */
    float *pu = (float *)malloc(sizeof(float) * ex * ey * ez);
    if (!pu) return -1;
    for (long k = 0; k < ez; ++k)
        for (long j = 0; j < ey; ++j)
            for (long i = 0; i < ex; ++i)
                pu[i + ex * (j + ey * k)] =
                    u[R101(i - R, nx) + (long)nx * (R101(j - R, ny) + (long)ny * R101(k - R, nz))];
    const float P = (float)((2 * rp + 1) * (2 * rp + 1) * (2 * rp + 1));
    const float hh = h * h;
    const int nslab = (z1 - z0 + SLAB - 1) / SLAB;
    int fail = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        /* per-thread work buffers over the slab's extended region */
        const long wx = nx, wy = ny + 2L * rp, wz = SLAB + 2L * rp;
        float *sx = (float *)malloc(sizeof(float) * wx * wy * wz);
        float *sxy = (float *)malloc(sizeof(float) * wx * ny * wz);
        float *Zs = (float *)malloc(sizeof(float) * (long)nx * ny * SLAB);
        float *Ys = (float *)malloc(sizeof(float) * (long)nx * ny * SLAB);
        float *Zd = (float *)malloc(sizeof(float) * (long)nx * ny * SLAB);
        float *Yd = (float *)malloc(sizeof(float) * (long)nx * ny * SLAB);
        if (!sx || !sxy || !Zs || !Ys || !Zd || !Yd) {
#pragma omp atomic write
            fail = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int s = 0; s < nslab; ++s) {
                const int za = z0 + s * SLAB;
                const int zb = (za + SLAB < z1) ? za + SLAB : z1;
                const int tz = zb - za;
                memset(Zs, 0, sizeof(float) * (long)nx * ny * SLAB);
                memset(Ys, 0, sizeof(float) * (long)nx * ny * SLAB);
                for (int dz = -rs; dz <= rs; ++dz) {
                    memset(Zd, 0, sizeof(float) * (long)nx * ny * SLAB);
                    memset(Yd, 0, sizeof(float) * (long)nx * ny * SLAB);
                    for (int dy = -rs; dy <= rs; ++dy)
                        for (int dx = -rs; dx <= rs; ++dx) {
                            /* x-stage on rows y in [-rp, ny+rp), z in [za-rp, zb+rp) */
                            for (int kz = 0; kz < tz + 2 * rp; ++kz) {
                                const long zz = za - rp + kz + R;
                                for (int jy = 0; jy < ny + 2 * rp; ++jy) {
                                    const long yy = jy - rp + R;
                                    const float *c = pu + ex * (yy + ey * zz) + R;
                                    const float *q = pu + ex * (yy + dy + ey * (zz + dz)) + R + dx;
                                    float *o = sx + wx * (jy + wy * kz);
                                    for (int x = 0; x < nx; ++x) {
                                        float acc = 0.0f;
                                        for (int px = -rp; px <= rp; ++px) {
                                            float e = c[x + px] - q[x + px];
                                            acc = fmaf(e, e, acc);
                                        }
                                        o[x] = acc;
                                    }
                                }
                            }
                            /* y-stage */
                            for (int kz = 0; kz < tz + 2 * rp; ++kz)
                                for (int jy = 0; jy < ny; ++jy) {
                                    float *o = sxy + wx * (jy + (long)ny * kz);
                                    const float *a = sx + wx * (jy + wy * kz);
                                    for (int x = 0; x < nx; ++x) o[x] = 0.0f;
                                    for (int py = 0; py <= 2 * rp; ++py) {
                                        const float *b = a + wx * py;
                                        for (int x = 0; x < nx; ++x) o[x] = o[x] + b[x];
                                    }
                                }
                            /* z-stage, weights, accumulation into the dz partials */
                            for (int kz = 0; kz < tz; ++kz)
                                for (int jy = 0; jy < ny; ++jy) {
                                    const float *cand = pu + ex * (jy + R + dy + ey * (za + kz + R + dz)) + R + dx;
                                    float *Zr = Zd + (long)nx * (jy + (long)ny * kz);
                                    float *Yr = Yd + (long)nx * (jy + (long)ny * kz);
                                    for (int x = 0; x < nx; ++x) {
                                        float S = 0.0f;
                                        for (int pz = 0; pz <= 2 * rp; ++pz)
                                            S = S + sxy[x + wx * (jy + (long)ny * (kz + pz))];
                                        float w = expf(-(S / P) / hh);
                                        Zr[x] = Zr[x] + w;
                                        Yr[x] = fmaf(w, cand[x], Yr[x]);
                                    }
                                }
                        }
                    for (long i = 0; i < (long)nx * ny * tz; ++i) {
                        Zs[i] = Zs[i] + Zd[i];
                        Ys[i] = Ys[i] + Yd[i];
                    }
                }
                for (int kz = 0; kz < tz; ++kz)
                    for (int jy = 0; jy < ny; ++jy)
                        for (int x = 0; x < nx; ++x) {
                            long i = x + (long)nx * (jy + (long)ny * kz);
                            y[x + (long)nx * (jy + (long)ny * (za + kz - z0))] = Ys[i] / Zs[i];
                        }
            }
        }
        free(sx); free(sxy); free(Zs); free(Ys); free(Zd); free(Yd);
    }
    free(pu);
    return fail ? -1 : 0;
}

/* FP64 separable version (same structure, plain double sums): used for
 * full-size FP64 re-runs where the brute force would take too long.
 * Zout and Aout (nullable, planes [z0, z1) like y) receive the normaliser
 * Z(x) = sum_d w_d(x) and A(x) = sum_d w_d(x) |u~(x+d)|; A/Z is the scale of
 * the rounding error of any FP32 evaluation of y = sum_d w_d u~(x+d) / Z (the
 * sum of the magnitudes of its terms), which the parity tests use to bound
 * the error at outputs near zero.  Bout (nullable) receives
 * B(x) = sum_d w_d(x) a_d(x) (|u~(x+d)| + |y(x)|), a_d = S_d / (P h^2) the
 * exponent of w_d = exp(-a_d): an FP32 evaluation of a_d carries a relative
 * error of a few 2^-24, so w_d one of a few a_d 2^-24, and y moves by up to
 * sum_d w_d a_d |u~(x+d) - y| / Z <= B / Z of those units (large for 8-bit
 * data, where a_d reaches 80 and more). */
int oracle_nlm3d_separable_f64_za(const float *u, int nx, int ny, int nz,
                                  int rp, int rs, double h, double *y, double *Zout, double *Aout,
                                  double *Bout, int z0, int z1, int nthreads);

int oracle_nlm3d_separable_f64(const float *u, int nx, int ny, int nz,
                               int rp, int rs, double h, double *y,
                               int z0, int z1, int nthreads)
{
    return oracle_nlm3d_separable_f64_za(u, nx, ny, nz, rp, rs, h, y, NULL, NULL, NULL, z0, z1, nthreads);
}

int oracle_nlm3d_separable_f64_za(const float *u, int nx, int ny, int nz,
                                  int rp, int rs, double h, double *y, double *Zout, double *Aout,
                                  double *Bout, int z0, int z1, int nthreads)
{
    const int R = rp + rs;
    const long ex = nx + 2L * R, ey = ny + 2L * R, ez = nz + 2L * R;
    double *pu = (double *)malloc(sizeof(double) * ex * ey * ez);
    if (!pu) return -1;
    for (long k = 0; k < ez; ++k)
        for (long j = 0; j < ey; ++j)
            for (long i = 0; i < ex; ++i)
                pu[i + ex * (j + ey * k)] =
                    u[R101(i - R, nx) + (long)nx * (R101(j - R, ny) + (long)ny * R101(k - R, nz))];
    const double P = (double)(2 * rp + 1) * (2 * rp + 1) * (2 * rp + 1);
    const double inv = 1.0 / (h * h);
    const int nslab = (z1 - z0 + SLAB - 1) / SLAB;
    int fail = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        const long wx = nx, wy = ny + 2L * rp, wz = SLAB + 2L * rp;
        double *sx = (double *)malloc(sizeof(double) * wx * wy * wz);
        double *sxy = (double *)malloc(sizeof(double) * wx * ny * wz);
        double *Zs = (double *)malloc(sizeof(double) * (long)nx * ny * SLAB);
        double *Ys = (double *)malloc(sizeof(double) * (long)nx * ny * SLAB);
        double *As = (double *)malloc(sizeof(double) * (long)nx * ny * SLAB);
        double *WAs = (double *)malloc(sizeof(double) * (long)nx * ny * SLAB);
        double *WUs = (double *)malloc(sizeof(double) * (long)nx * ny * SLAB);
        if (!sx || !sxy || !Zs || !Ys || !As || !WAs || !WUs) {
#pragma omp atomic write
            fail = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int s = 0; s < nslab; ++s) {
                const int za = z0 + s * SLAB;
                const int zb = (za + SLAB < z1) ? za + SLAB : z1;
                const int tz = zb - za;
                memset(Zs, 0, sizeof(double) * (long)nx * ny * SLAB);
                memset(Ys, 0, sizeof(double) * (long)nx * ny * SLAB);
                memset(As, 0, sizeof(double) * (long)nx * ny * SLAB);
                memset(WAs, 0, sizeof(double) * (long)nx * ny * SLAB);
                memset(WUs, 0, sizeof(double) * (long)nx * ny * SLAB);
                for (int dz = -rs; dz <= rs; ++dz)
                    for (int dy = -rs; dy <= rs; ++dy)
                        for (int dx = -rs; dx <= rs; ++dx) {
                            for (int kz = 0; kz < tz + 2 * rp; ++kz) {
                                const long zz = za - rp + kz + R;
                                for (int jy = 0; jy < ny + 2 * rp; ++jy) {
                                    const long yy = jy - rp + R;
                                    const double *c = pu + ex * (yy + ey * zz) + R;
                                    const double *q = pu + ex * (yy + dy + ey * (zz + dz)) + R + dx;
                                    double *o = sx + wx * (jy + wy * kz);
                                    for (int x = 0; x < nx; ++x) {
                                        double acc = 0.0;
                                        for (int px = -rp; px <= rp; ++px) {
                                            double e = c[x + px] - q[x + px];
                                            acc += e * e;
                                        }
                                        o[x] = acc;
                                    }
                                }
                            }
                            for (int kz = 0; kz < tz + 2 * rp; ++kz)
                                for (int jy = 0; jy < ny; ++jy) {
                                    double *o = sxy + wx * (jy + (long)ny * kz);
                                    const double *a = sx + wx * (jy + wy * kz);
                                    for (int x = 0; x < nx; ++x) o[x] = 0.0;
                                    for (int py = 0; py <= 2 * rp; ++py) {
                                        const double *b = a + wx * py;
                                        for (int x = 0; x < nx; ++x) o[x] += b[x];
                                    }
                                }
                            for (int kz = 0; kz < tz; ++kz)
                                for (int jy = 0; jy < ny; ++jy) {
                                    const double *cand = pu + ex * (jy + R + dy + ey * (za + kz + R + dz)) + R + dx;
                                    double *Zr = Zs + (long)nx * (jy + (long)ny * kz);
                                    double *Yr = Ys + (long)nx * (jy + (long)ny * kz);
                                    double *Ar = As + (long)nx * (jy + (long)ny * kz);
                                    double *WAr = WAs + (long)nx * (jy + (long)ny * kz);
                                    double *WUr = WUs + (long)nx * (jy + (long)ny * kz);
                                    for (int x = 0; x < nx; ++x) {
                                        double S = 0.0;
                                        for (int pz = 0; pz <= 2 * rp; ++pz)
                                            S += sxy[x + wx * (jy + (long)ny * (kz + pz))];
                                        const double arg = (S / P) * inv;
                                        double w = exp(-arg);
                                        Zr[x] += w;
                                        Yr[x] += w * cand[x];
                                        Ar[x] += w * fabs(cand[x]);
                                        WAr[x] += w * arg;
                                        WUr[x] += w * arg * fabs(cand[x]);
                                    }
                                }
                        }
                for (int kz = 0; kz < tz; ++kz)
                    for (int jy = 0; jy < ny; ++jy)
                        for (int x = 0; x < nx; ++x) {
                            long i = x + (long)nx * (jy + (long)ny * kz);
                            const long o = x + (long)nx * (jy + (long)ny * (za + kz - z0));
                            y[o] = Ys[i] / Zs[i];
                            if (Zout) Zout[o] = Zs[i];
                            if (Aout) Aout[o] = As[i];
                            if (Bout) Bout[o] = WUs[i] + fabs(y[o]) * WAs[i];
                        }
            }
        }
        free(sx); free(sxy); free(Zs); free(Ys); free(As); free(WAs); free(WUs);
    }
    free(pu);
    return fail ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* 2D separable FP64 (plain double sums), OpenMP over rows; Zout, Aout, */
/* Bout (nullable) as in oracle_nlm3d_separable_f64_za.                 */
/* ------------------------------------------------------------------ */
int oracle_nlm2d_separable_f64(const float *u, int nx, int ny, int rp, int rs,
                               double h, double *y, double *Zout, double *Aout, double *Bout, int nthreads)
{
    const int R = rp + rs;
    const long ex = nx + 2L * R, ey = ny + 2L * R;
    double *pu = (double *)malloc(sizeof(double) * ex * ey);
    if (!pu) return -1;
    for (long j = 0; j < ey; ++j)
        for (long i = 0; i < ex; ++i)
            pu[i + ex * j] = u[R101(i - R, nx) + (long)nx * R101(j - R, ny)];
    const double P = (double)((2 * rp + 1) * (2 * rp + 1));
    const double inv = 1.0 / (h * h);
    const int BAND = 16;
    const int nband = (ny + BAND - 1) / BAND;
    int fail = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *sx = (double *)malloc(sizeof(double) * nx * (BAND + 2 * rp));
        double *Zs = (double *)malloc(sizeof(double) * nx * BAND);
        double *Ys = (double *)malloc(sizeof(double) * nx * BAND);
        double *As = (double *)malloc(sizeof(double) * nx * BAND);
        double *WAs = (double *)malloc(sizeof(double) * nx * BAND);
        double *WUs = (double *)malloc(sizeof(double) * nx * BAND);
        if (!sx || !Zs || !Ys || !As || !WAs || !WUs) {
#pragma omp atomic write
            fail = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int b = 0; b < nband; ++b) {
                const int ya = b * BAND;
                const int yb = (ya + BAND < ny) ? ya + BAND : ny;
                const int ty = yb - ya;
                memset(Zs, 0, sizeof(double) * nx * BAND);
                memset(Ys, 0, sizeof(double) * nx * BAND);
                memset(As, 0, sizeof(double) * nx * BAND);
                memset(WAs, 0, sizeof(double) * nx * BAND);
                memset(WUs, 0, sizeof(double) * nx * BAND);
                for (int dy = -rs; dy <= rs; ++dy)
                    for (int dx = -rs; dx <= rs; ++dx) {
                        for (int jy = 0; jy < ty + 2 * rp; ++jy) {
                            const long yy = ya - rp + jy + R;
                            const double *c = pu + ex * yy + R;
                            const double *q = pu + ex * (yy + dy) + R + dx;
                            double *o = sx + (long)nx * jy;
                            for (int x = 0; x < nx; ++x) {
                                double acc = 0.0;
                                for (int px = -rp; px <= rp; ++px) {
                                    double e = c[x + px] - q[x + px];
                                    acc += e * e;
                                }
                                o[x] = acc;
                            }
                        }
                        for (int jy = 0; jy < ty; ++jy) {
                            const double *cand = pu + ex * (ya + jy + R + dy) + R + dx;
                            for (int x = 0; x < nx; ++x) {
                                double S = 0.0;
                                for (int py = 0; py <= 2 * rp; ++py) S += sx[x + (long)nx * (jy + py)];
                                const double arg = (S / P) * inv;
                                const double w = exp(-arg);
                                const long i = x + (long)nx * jy;
                                Zs[i] += w;
                                Ys[i] += w * cand[x];
                                As[i] += w * fabs(cand[x]);
                                WAs[i] += w * arg;
                                WUs[i] += w * arg * fabs(cand[x]);
                            }
                        }
                    }
                for (int jy = 0; jy < ty; ++jy)
                    for (int x = 0; x < nx; ++x) {
                        const long i = x + (long)nx * jy, o = x + (long)nx * (ya + jy);
                        y[o] = Ys[i] / Zs[i];
                        if (Zout) Zout[o] = Zs[i];
                        if (Aout) Aout[o] = As[i];
                        if (Bout) Bout[o] = WUs[i] + fabs(y[o]) * WAs[i];
                    }
            }
        }
        free(sx); free(Zs); free(Ys); free(As); free(WAs); free(WUs);
    }
    free(pu);
    return fail ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* 2D separable FP32 path (canonical order), OpenMP over rows           */
/* ------------------------------------------------------------------ */
int oracle_nlm2d_separable_f32(const float *u, int nx, int ny, int rp, int rs,
                               float h, float *y, int nthreads)
{
    const int R = rp + rs;
    const long ex = nx + 2L * R, ey = ny + 2L * R;
    float *pu = (float *)malloc(sizeof(float) * ex * ey);
    if (!pu) return -1;
    for (long j = 0; j < ey; ++j)
        for (long i = 0; i < ex; ++i)
            pu[i + ex * j] = u[R101(i - R, nx) + (long)nx * R101(j - R, ny)];
    const float P = (float)((2 * rp + 1) * (2 * rp + 1));
    const float hh = h * h;
    const int BAND = 16;
    const int nband = (ny + BAND - 1) / BAND;
    int fail = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        float *sx = (float *)malloc(sizeof(float) * nx * (BAND + 2 * rp));
        float *Zs = (float *)malloc(sizeof(float) * nx * BAND);
        float *Ys = (float *)malloc(sizeof(float) * nx * BAND);
        float *Zd = (float *)malloc(sizeof(float) * nx * BAND);
        float *Yd = (float *)malloc(sizeof(float) * nx * BAND);
        if (!sx || !Zs || !Ys || !Zd || !Yd) {
#pragma omp atomic write
            fail = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int b = 0; b < nband; ++b) {
                const int ya = b * BAND;
                const int yb = (ya + BAND < ny) ? ya + BAND : ny;
                const int ty = yb - ya;
                memset(Zs, 0, sizeof(float) * nx * BAND);
                memset(Ys, 0, sizeof(float) * nx * BAND);
                for (int dy = -rs; dy <= rs; ++dy) {
                    memset(Zd, 0, sizeof(float) * nx * BAND);
                    memset(Yd, 0, sizeof(float) * nx * BAND);
                    for (int dx = -rs; dx <= rs; ++dx) {
                        for (int jy = 0; jy < ty + 2 * rp; ++jy) {
                            const long yy = ya - rp + jy + R;
                            const float *c = pu + ex * yy + R;
                            const float *q = pu + ex * (yy + dy) + R + dx;
                            float *o = sx + (long)nx * jy;
                            for (int x = 0; x < nx; ++x) {
                                float acc = 0.0f;
                                for (int px = -rp; px <= rp; ++px) {
                                    float e = c[x + px] - q[x + px];
                                    acc = fmaf(e, e, acc);
                                }
                                o[x] = acc;
                            }
                        }
                        for (int jy = 0; jy < ty; ++jy) {
                            const float *cand = pu + ex * (ya + jy + R + dy) + R + dx;
                            float *Zr = Zd + (long)nx * jy;
                            float *Yr = Yd + (long)nx * jy;
                            for (int x = 0; x < nx; ++x) {
                                float S = 0.0f;
                                for (int py = 0; py <= 2 * rp; ++py)
                                    S = S + sx[x + (long)nx * (jy + py)];
                                float w = expf(-(S / P) / hh);
                                Zr[x] = Zr[x] + w;
                                Yr[x] = fmaf(w, cand[x], Yr[x]);
                            }
                        }
                    }
                    for (long i = 0; i < (long)nx * ty; ++i) {
                        Zs[i] = Zs[i] + Zd[i];
                        Ys[i] = Ys[i] + Yd[i];
                    }
                }
                for (int jy = 0; jy < ty; ++jy)
                    for (int x = 0; x < nx; ++x)
                        y[x + (long)nx * (ya + jy)] = Ys[x + (long)nx * jy] / Zs[x + (long)nx * jy];
            }
        }
        free(sx); free(Zs); free(Ys); free(Zd); free(Yd);
    }
    free(pu);
    return fail ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* Adjoint, FP64, straight chain rule over the brute-force definition.  */
/*                                                                      */
/* Given g = dL/dy on the volume, returns dL/du (u unpadded: gradients  */
/* reaching reflect-101 copies are summed back onto their source voxel) */
/* and dL/dh.  Derivation (c = 1/(P h^2)):                              */
/*   dy(x)/dw_d(x) = (u~(x+d) - y(x)) / Z(x);  dw/dS = -c w;            */
/*   dw/dh = w S 2c/h;  dS_d(x)/du~(x+p) = 2 e,  /du~(x+d+p) = -2 e.    */
/* ------------------------------------------------------------------ */
void oracle_nlm3d_grad_f64(const double *u, const double *g, int nx, int ny, int nz,
                           int rp, int rs, double h, double *gu, double *gh)
{
    /* The chain rule below runs on the reflect-101 extension ue over
     * [-R, n+R)^3 (R = rp + rs), which holds every sample the forward reads;
     * the gradient reaching an extension position is added onto its source
     * voxel at the end.  Voxels are split over OpenMP threads, each with its
     * own extended-domain gradient, summed afterwards. */
    const double P = (double)(2 * rp + 1) * (2 * rp + 1) * (2 * rp + 1);
    const double c = 1.0 / (P * h * h);
    const long N = (long)nx * ny * nz;
    const int R = rp + rs;
    const long ex = nx + 2L * R, ey = ny + 2L * R, ez = nz + 2L * R, NE = ex * ey * ez;
    double *ue = (double *)malloc(sizeof(double) * NE);
    double *ge = (double *)calloc((size_t)NE, sizeof(double));
    if (!ue || !ge) {
        free(ue);
        free(ge);
        *gh = NAN;
        return;
    }
    for (long k = 0; k < ez; ++k)
        for (long j = 0; j < ey; ++j)
            for (long i = 0; i < ex; ++i)
                ue[i + ex * (j + ey * k)] = u[R101(i - R, nx) + (long)nx * (R101(j - R, ny) + (long)ny * R101(k - R, nz))];
    double dh = 0.0;
#define E3(X, Y, Z) ((X) + R + ex * ((Y) + R + ey * ((long)(Z) + R)))
#pragma omp parallel reduction(+ : dh)
    {
        double *gl = (double *)calloc((size_t)NE, sizeof(double));
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int z = 0; z < nz; ++z)
            for (int yy = 0; yy < ny; ++yy)
                for (int x = 0; x < nx && gl; ++x) {
                    const long o = x + (long)nx * (yy + (long)ny * z);
                    const double gx = g[o];
                    if (gx == 0.0) continue;
                    /* forward at x */
                    double Z = 0.0, Y = 0.0;
                    for (int dz = -rs; dz <= rs; ++dz)
                        for (int dy = -rs; dy <= rs; ++dy)
                            for (int dx = -rs; dx <= rs; ++dx) {
                                double S = 0.0;
                                for (int pz = -rp; pz <= rp; ++pz)
                                    for (int py = -rp; py <= rp; ++py)
                                        for (int px = -rp; px <= rp; ++px) {
                                            double e = ue[E3(x + px, yy + py, z + pz)] -
                                                       ue[E3(x + dx + px, yy + dy + py, z + dz + pz)];
                                            S += e * e;
                                        }
                                double w = exp(-c * S);
                                Z += w;
                                Y += w * ue[E3(x + dx, yy + dy, z + dz)];
                            }
                    const double yx = Y / Z;
                    for (int dz = -rs; dz <= rs; ++dz)
                        for (int dy = -rs; dy <= rs; ++dy)
                            for (int dx = -rs; dx <= rs; ++dx) {
                                double S = 0.0;
                                for (int pz = -rp; pz <= rp; ++pz)
                                    for (int py = -rp; py <= rp; ++py)
                                        for (int px = -rp; px <= rp; ++px) {
                                            double e = ue[E3(x + px, yy + py, z + pz)] -
                                                       ue[E3(x + dx + px, yy + dy + py, z + dz + pz)];
                                            S += e * e;
                                        }
                                const double w = exp(-c * S);
                                const double ucand = ue[E3(x + dx, yy + dy, z + dz)];
                                const double dLdw = gx * (ucand - yx) / Z;
                                const double A = -c * w * dLdw; /* dL/dS_d(x) */
                                gl[E3(x + dx, yy + dy, z + dz)] += gx * w / Z;
                                dh += dLdw * w * S * 2.0 * c / h;
                                for (int pz = -rp; pz <= rp; ++pz)
                                    for (int py = -rp; py <= rp; ++py)
                                        for (int px = -rp; px <= rp; ++px) {
                                            const long ia = E3(x + px, yy + py, z + pz);
                                            const long ib = E3(x + dx + px, yy + dy + py, z + dz + pz);
                                            const double e = ue[ia] - ue[ib];
                                            gl[ia] += 2.0 * A * e;
                                            gl[ib] -= 2.0 * A * e;
                                        }
                            }
                }
        if (gl) {
#pragma omp critical
            for (long i = 0; i < NE; ++i) ge[i] += gl[i];
            free(gl);
        } else {
            dh = NAN;
        }
    }
#undef E3
    /* fold the extension's gradient onto its source voxels */
    memset(gu, 0, sizeof(double) * N);
    for (long k = 0; k < ez; ++k)
        for (long j = 0; j < ey; ++j)
            for (long i = 0; i < ex; ++i)
                gu[R101(i - R, nx) + (long)nx * (R101(j - R, ny) + (long)ny * R101(k - R, nz))] +=
                    ge[i + ex * (j + ey * k)];
    free(ue);
    free(ge);
    *gh = dh;
}

/* MSE loss of the FP64 forward against a clean volume, for finite
 * differences: L = sum((y - clean)^2) / N. */
double oracle_nlm3d_mse_f64(const double *u, const double *clean, int nx, int ny, int nz,
                            int rp, int rs, double h)
{
    const long N = (long)nx * ny * nz;
    double *y = (double *)malloc(sizeof(double) * N);
    if (!y) return NAN;
    oracle_nlm3d_brute_f64(u, nx, ny, nz, rp, rs, h, y, NULL);
    double L = 0.0;
    for (long i = 0; i < N; ++i) {
        double r = y[i] - clean[i];
        L += r * r;
    }
    free(y);
    return L / (double)N;
}
