// phantom.cpp -- seeded synthetic phantoms (host code).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#include "../../include/gnlm_phantom.h"

namespace {

struct Rng {
    std::mt19937_64 g;
    explicit Rng(unsigned long long seed) : g(seed) {}
    // uniform in [0, 1) with 53 random bits
    double uni() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    // Box-Muller, one value per call (the pair's second value is kept)
    bool has = false;
    double spare = 0.0;
    double normal()
    {
        if (has) {
            has = false;
            return spare;
        }
        double u1 = uni();
        while (u1 <= 0.0) u1 = uni();
        const double u2 = uni();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double t = 6.283185307179586 * u2;
        spare = r * std::sin(t);
        has = true;
        return r * std::cos(t);
    }
};

}  // namespace

extern "C" {

int gnlm_phantom3d(float *h_out, float *h_clean, int nx, int ny, int nz, int n_ellipsoids, float axis_lo,
                   float axis_hi, float background, float sigma, unsigned long long seed)
{
    if (!h_out || nx < 1 || ny < 1 || nz < 1 || n_ellipsoids < 0 || !(axis_lo > 0) || axis_hi < axis_lo ||
        sigma < 0)
        return 1;
    const long long N = (long long)nx * ny * nz;
    std::vector<float> clean((size_t)N, background);
    Rng rng(seed);
    for (int e = 0; e < n_ellipsoids; ++e) {
        const double cx = rng.uni(0, nx), cy = rng.uni(0, ny), cz = rng.uni(0, nz);
        const double ax = rng.uni(axis_lo, axis_hi), ay = rng.uni(axis_lo, axis_hi), az = rng.uni(axis_lo, axis_hi);
        const float v = (float)rng.uni(0.1, 0.9);
        const int xa = std::max(0, (int)std::floor(cx - ax)), xb = std::min(nx - 1, (int)std::ceil(cx + ax));
        const int ya = std::max(0, (int)std::floor(cy - ay)), yb = std::min(ny - 1, (int)std::ceil(cy + ay));
        const int za = std::max(0, (int)std::floor(cz - az)), zb = std::min(nz - 1, (int)std::ceil(cz + az));
        for (int z = za; z <= zb; ++z) {
            const double qz = (z - cz) / az;
            for (int y = ya; y <= yb; ++y) {
                const double qy = (y - cy) / ay;
                for (int x = xa; x <= xb; ++x) {
                    const double qx = (x - cx) / ax;
                    if (qx * qx + qy * qy + qz * qz <= 1.0) clean[(size_t)(x + (long long)nx * (y + (long long)ny * z))] = v;
                }
            }
        }
    }
    for (long long i = 0; i < N; ++i) h_out[i] = (float)(clean[(size_t)i] + sigma * rng.normal());
    if (h_clean) std::memcpy(h_clean, clean.data(), sizeof(float) * (size_t)N);
    return 0;
}

int gnlm_phantom2d(float *h_out, float *h_clean, int nx, int ny, int n_shapes, float size_lo, float size_hi,
                   float sine_amp, float period_lo, float period_hi, float sigma, unsigned long long seed)
{
    if (!h_out || nx < 1 || ny < 1 || n_shapes < 0 || size_lo < 0 || size_hi < size_lo || !(period_lo > 0) ||
        period_hi < period_lo || sigma < 0)
        return 1;
    const long long N = (long long)nx * ny;
    std::vector<float> clean((size_t)N, 0.0f);
    Rng rng(seed);
    const double m = (double)std::max(nx, ny);
    for (int s = 0; s < n_shapes; ++s) {
        const bool disc = rng.uni() < 0.5;
        const double cx = rng.uni(0, nx), cy = rng.uni(0, ny);
        const double hx = rng.uni(size_lo, size_hi) * m;
        const double hy = disc ? hx : rng.uni(size_lo, size_hi) * m;
        const float v = (float)rng.uni();
        const int xa = std::max(0, (int)std::floor(cx - hx)), xb = std::min(nx - 1, (int)std::ceil(cx + hx));
        const int ya = std::max(0, (int)std::floor(cy - hy)), yb = std::min(ny - 1, (int)std::ceil(cy + hy));
        for (int y = ya; y <= yb; ++y)
            for (int x = xa; x <= xb; ++x) {
                const double qx = x - cx, qy = y - cy;
                const bool in = disc ? (qx * qx + qy * qy <= hx * hx) : (std::fabs(qx) <= hx && std::fabs(qy) <= hy);
                if (in) clean[(size_t)(x + (long long)nx * y)] = v;
            }
    }
    const double period = rng.uni(period_lo, period_hi);
    const double th = rng.uni(0.0, 6.283185307179586);
    const double c = std::cos(th), sn = std::sin(th);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            float &p = clean[(size_t)(x + (long long)nx * y)];
            p = (float)(p + sine_amp * std::sin(6.283185307179586 * (x * c + y * sn) / period));
        }
    for (long long i = 0; i < N; ++i) h_out[i] = (float)(clean[(size_t)i] + sigma * rng.normal());
    if (h_clean) std::memcpy(h_clean, clean.data(), sizeof(float) * (size_t)N);
    return 0;
}

unsigned long long gnlm_hash_f32(const float *h, long long n)
{
    unsigned long long x = 1469598103934665603ULL;
    const unsigned char *b = reinterpret_cast<const unsigned char *>(h);
    for (long long i = 0; i < 4 * n; ++i) {
        x ^= b[i];
        x *= 1099511628211ULL;
    }
    return x;
}

}  // extern "C"
