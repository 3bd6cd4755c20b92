/* oracle_sanitize.c -- runs every entry point of the CPU oracle
 * (oracle/nlm_oracle.c) on small ragged volumes: single-sample axes, axes
 * shorter than the halo (reflect-101 folding more than once), patch and
 * search radii 0-3, z ranges that cover part of the volume.  Built together
 * with nlm_oracle.c under AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_sanitizers.py; any out-of-bounds access, use of uninitialised
 * heap through a sanitizer trap, or undefined arithmetic aborts the run.
 * Also checks, per case, that the FP32 brute force and separable paths agree
 * bit for bit and that the FP64 paths agree to 1e-12 (exit status 1 if
 * not). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

long oracle_reflect101(long i, long n);
void oracle_nlm3d_brute_f64(const double *u, int nx, int ny, int nz, int rp, int rs, double h, double *y,
                            double *Zout);
void oracle_nlm3d_brute_f32(const float *u, int nx, int ny, int nz, int rp, int rs, float h, float *y,
                            float *Zout);
void oracle_nlm2d_brute_f64(const double *u, int nx, int ny, int rp, int rs, double h, double *y, double *Zout);
void oracle_nlm2d_brute_f32(const float *u, int nx, int ny, int rp, int rs, float h, float *y, float *Zout);
int oracle_nlm3d_separable_f32(const float *u, int nx, int ny, int nz, int rp, int rs, float h, float *y, int z0,
                               int z1, int nthreads);
int oracle_nlm3d_separable_f64_za(const float *u, int nx, int ny, int nz, int rp, int rs, double h, double *y,
                                  double *Zout, double *Aout, double *Bout, int z0, int z1, int nthreads);
int oracle_nlm2d_separable_f64(const float *u, int nx, int ny, int rp, int rs, double h, double *y, double *Zout,
                               double *Aout, double *Bout, int nthreads);
int oracle_nlm2d_separable_f32(const float *u, int nx, int ny, int rp, int rs, float h, float *y, int nthreads);
void oracle_nlm3d_grad_f64(const double *u, const double *g, int nx, int ny, int nz, int rp, int rs, double h,
                           double *gu, double *gh);
double oracle_nlm3d_mse_f64(const double *u, const double *clean, int nx, int ny, int nz, int rp, int rs,
                            double h);

static unsigned long long rng_state = 88172645463325252ull;
static double rnd(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) / 9007199254740992.0;
}

static int fails = 0;

static void case3d(int nx, int ny, int nz, int rp, int rs)
{
    const long n = (long)nx * ny * nz;
    /* exact-size heap buffers: the sanitizer flags any access past them */
    float *u32 = malloc(sizeof(float) * n), *y32 = malloc(sizeof(float) * n), *z32 = malloc(sizeof(float) * n);
    float *ys = malloc(sizeof(float) * n);
    double *u64 = malloc(sizeof(double) * n), *y64 = malloc(sizeof(double) * n), *z64 = malloc(sizeof(double) * n);
    double *ysep = malloc(sizeof(double) * n), *zsep = malloc(sizeof(double) * n), *asep = malloc(sizeof(double) * n);
    double *bsep = malloc(sizeof(double) * n);
    double *g = malloc(sizeof(double) * n), *gu = malloc(sizeof(double) * n), *cl = malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) {
        u32[i] = (float)rnd();
        u64[i] = u32[i];
        g[i] = rnd() - 0.5;
        cl[i] = rnd();
    }
    const double h = 0.3;
    oracle_nlm3d_brute_f64(u64, nx, ny, nz, rp, rs, h, y64, z64);
    oracle_nlm3d_brute_f32(u32, nx, ny, nz, rp, rs, (float)h, y32, z32);
    if (oracle_nlm3d_separable_f32(u32, nx, ny, nz, rp, rs, (float)h, ys, 0, nz, 2) != 0) ++fails;
    if (memcmp(ys, y32, sizeof(float) * n) != 0) {
        fprintf(stderr, "3D f32 separable != brute at %dx%dx%d rp %d rs %d\n", nx, ny, nz, rp, rs);
        ++fails;
    }
    /* planes [z0, z1) of the separable FP64 path, the rest of y left alone */
    const int z0 = nz / 3, z1 = nz - nz / 4;
    if (oracle_nlm3d_separable_f64_za(u32, nx, ny, nz, rp, rs, h, ysep, zsep, asep, bsep, z0, z1, 2) != 0) ++fails;
    const long off = (long)z0 * nx * ny;
    for (long i = 0; i < (long)(z1 - z0) * nx * ny; ++i) {
        if (fabs(ysep[i] - y64[off + i]) > 1e-12 * fabs(y64[off + i]) + 1e-15 ||
            fabs(zsep[i] - z64[off + i]) > 1e-12 * z64[off + i]) {
            fprintf(stderr, "3D f64 separable != brute at %dx%dx%d rp %d rs %d\n", nx, ny, nz, rp, rs);
            ++fails;
            break;
        }
    }
    double gh = 0.0;
    oracle_nlm3d_grad_f64(u64, g, nx, ny, nz, rp, rs, h, gu, &gh);
    const double loss = oracle_nlm3d_mse_f64(u64, cl, nx, ny, nz, rp, rs, h);
    if (!(loss >= 0.0) || !isfinite(gh)) ++fails;
    free(u32), free(y32), free(z32), free(ys), free(u64), free(y64), free(z64);
    free(ysep), free(zsep), free(asep), free(bsep), free(g), free(gu), free(cl);
}

static void case2d(int nx, int ny, int rp, int rs)
{
    const long n = (long)nx * ny;
    float *u32 = malloc(sizeof(float) * n), *y32 = malloc(sizeof(float) * n), *z32 = malloc(sizeof(float) * n);
    float *ys = malloc(sizeof(float) * n);
    double *u64 = malloc(sizeof(double) * n), *y64 = malloc(sizeof(double) * n), *z64 = malloc(sizeof(double) * n);
    double *ysep = malloc(sizeof(double) * n), *zsep = malloc(sizeof(double) * n), *asep = malloc(sizeof(double) * n);
    double *bsep = malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) {
        u32[i] = (float)rnd();
        u64[i] = u32[i];
    }
    const double h = 0.25;
    oracle_nlm2d_brute_f64(u64, nx, ny, rp, rs, h, y64, z64);
    oracle_nlm2d_brute_f32(u32, nx, ny, rp, rs, (float)h, y32, z32);
    if (oracle_nlm2d_separable_f32(u32, nx, ny, rp, rs, (float)h, ys, 2) != 0) ++fails;
    if (oracle_nlm2d_separable_f64(u32, nx, ny, rp, rs, h, ysep, zsep, asep, bsep, 2) != 0) ++fails;
    for (long i = 0; i < n; ++i) {
        if (fabs(ysep[i] - y64[i]) > 1e-12 * fabs(y64[i]) + 1e-15) {
            fprintf(stderr, "2D f64 separable != brute at %dx%d rp %d rs %d\n", nx, ny, rp, rs);
            ++fails;
            break;
        }
    }
    free(u32), free(y32), free(z32), free(ys), free(u64), free(y64), free(z64), free(ysep), free(zsep), free(asep),
        free(bsep);
}

int main(void)
{
    for (long n = 1; n <= 9; ++n)
        for (long i = -3 * n - 2; i < 4 * n + 2; ++i) {
            const long r = oracle_reflect101(i, n);
            if (r < 0 || r >= n) ++fails;
        }
    static const int d3[][3] = {{1, 1, 1}, {5, 1, 3}, {1, 6, 2}, {4, 3, 2}, {7, 5, 6}, {9, 4, 1}, {3, 8, 5}};
    for (unsigned c = 0; c < sizeof d3 / sizeof d3[0]; ++c)
        for (int rp = 0; rp <= 2; ++rp)
            for (int rs = 0; rs <= 3; ++rs) case3d(d3[c][0], d3[c][1], d3[c][2], rp, rs);
    static const int d2[][2] = {{1, 1}, {1, 9}, {7, 1}, {5, 6}, {13, 4}, {3, 17}};
    for (unsigned c = 0; c < sizeof d2 / sizeof d2[0]; ++c)
        for (int rp = 0; rp <= 3; ++rp)
            for (int rs = 0; rs <= 3; ++rs) case2d(d2[c][0], d2[c][1], rp, rs);
    if (fails) fprintf(stderr, "%d failures\n", fails);
    else printf("oracle sanitize: all cases ok\n");
    return fails ? 1 : 0;
}
