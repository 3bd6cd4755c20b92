// nlm_box.h -- box sums along one axis in registers (device code), shared by
// the tiled forward and adjoint kernels: the patch SSD S_d(x) of the
// reference (3DNLM/main.cu:52-68; "2D NLM/11x11/main.cu:36-63") is a box sum
// of squared differences, formed one axis at a time from these.
#pragma once
#include <hip/hip_runtime.h>

namespace gnlm {

// s[i] = v[i] + v[i+1] + ... + v[i+W-1] for i < N (W = 3, 5, 7), as trees of
// pairwise sums (no running add/subtract windows).  Outputs
// come in pairs (i, i+1), i even, that share their W-1 common terms m_i =
// v[i+1] + ... + v[i+W-1]: s[i] = v[i] + m_i, s[i+1] = m_i + v[i+W].  m_i is
// built from the pair sums p_j = v[j] + v[j+1] at odd j, which neighbouring
// output pairs share again (W = 7: (p_{i+1} + p_{i+3}) + p_{i+5} for i = 0
// mod 4, p_{i+1} + (p_{i+3} + p_{i+5}) for i = 2 mod 4, so the inner sums
// are common too).  Cost per output: 1.5 adds (W = 3), 2 (W = 5), 2.25
// (W = 7), against W - 1 for direct sums.
// Any odd W (9 ... 17): the same output pairs, with m_i built from blocks of
// 2^b pair sums B_b[k] = B_{b-1}[k] + B_{b-1}[k + 2^b] (k odd), one block per
// set bit of (W-1)/2; every level is shared by all output pairs, so a box
// costs about 3 adds per output for W = 13 ... 17.
template <int W, int N>
__device__ __forceinline__ void box_tree_big(const float *v, float *s)
{
    static_assert(W % 2 == 1 && W >= 3 && W <= 31, "odd patch widths up to 31");
    constexpr int M = (W - 1) / 2;  // pair sums in m_i
    constexpr int NV = N + W - 1;   // inputs
    float B[4][NV];                 // B[b][k]: sum of v[k .. k + 2^(b+1) - 1]
#pragma unroll
    for (int k = 1; k + 1 < NV; k += 2) B[0][k] = v[k] + v[k + 1];
#pragma unroll
    for (int b = 1; b < 4; ++b) {
#pragma unroll
        for (int k = 1; k + (2 << b) <= NV; k += 2) B[b][k] = B[b - 1][k] + B[b - 1][k + (1 << b)];
    }
#pragma unroll
    for (int i = 0; i < N; i += 2) {
        float m = 0.0f;
        int off = i + 1;
        bool first = true;
#pragma unroll
        for (int b = 3; b >= 0; --b) {
            if ((M >> b) & 1) {
                m = first ? B[b][off] : m + B[b][off];
                first = false;
                off += 2 << b;
            }
        }
        s[i] = v[i] + m;
        if (i + 1 < N) s[i + 1] = m + v[i + W];
    }
}

// s[i] = e[i]^2 + ... + e[i+W-1]^2 for i < N: box_tree over the squares of
// e, with the squares of the even-indexed samples folded into the sums as
// fused multiply-adds (each such square feeds one pair sum and one output, both
// of which take it as fma(e, e, .)); only the odd-indexed squares are formed,
// which saves (N + W - 1) / 2 multiplies against squaring first.
template <int W, int N>
__device__ __forceinline__ void box_tree_sq(const float *e, float *s)
{
    static_assert(W % 2 == 1 && W <= 31, "odd patch widths up to 31");
    constexpr int NV = N + W - 1;  // inputs
    if constexpr (W == 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) s[i] = e[i] * e[i];
        return;
    }
    float vo[NV];  // squares at odd indices
#pragma unroll
    for (int k = 1; k < NV; k += 2) vo[k] = e[k] * e[k];
    if constexpr (W <= 7) {
        constexpr int NP = N + W - 2;  // p_j for odd j < NP
        float p[NP];
#pragma unroll
        for (int j = 1; j < NP; j += 2) p[j] = fmaf(e[j + 1], e[j + 1], vo[j]);
#pragma unroll
        for (int i = 0; i < N; i += 2) {
            float m;
            if constexpr (W == 3) m = p[i + 1];
            else if constexpr (W == 5) m = p[i + 1] + p[i + 3];
            else m = (i % 4 == 0) ? (p[i + 1] + p[i + 3]) + p[i + 5] : p[i + 1] + (p[i + 3] + p[i + 5]);
            s[i] = fmaf(e[i], e[i], m);
            if (i + 1 < N) s[i + 1] = m + vo[i + W];
        }
    } else {
        constexpr int M = (W - 1) / 2;
        float B[4][NV];
#pragma unroll
        for (int k = 1; k + 1 < NV; k += 2) B[0][k] = fmaf(e[k + 1], e[k + 1], vo[k]);
#pragma unroll
        for (int b = 1; b < 4; ++b) {
#pragma unroll
            for (int k = 1; k + (2 << b) <= NV; k += 2) B[b][k] = B[b - 1][k] + B[b - 1][k + (1 << b)];
        }
#pragma unroll
        for (int i = 0; i < N; i += 2) {
            float m = 0.0f;
            int off = i + 1;
            bool first = true;
#pragma unroll
            for (int b = 3; b >= 0; --b) {
                if ((M >> b) & 1) {
                    m = first ? B[b][off] : m + B[b][off];
                    first = false;
                    off += 2 << b;
                }
            }
            s[i] = fmaf(e[i], e[i], m);
            if (i + 1 < N) s[i + 1] = m + vo[i + W];
        }
    }
}

template <int W, int N>
__device__ __forceinline__ void box_tree(const float *v, float *s)
{
    if constexpr (W == 1) {  // single-sample patches
#pragma unroll
        for (int i = 0; i < N; ++i) s[i] = v[i];
        return;
    }
    if constexpr (W > 7) {
        box_tree_big<W, N>(v, s);
        return;
    }
    static_assert(W % 2 == 1, "odd patch widths");
    constexpr int NP = N + W - 2;  // p_j for odd j < NP
    float p[NP];
#pragma unroll
    for (int j = 1; j < NP; j += 2) p[j] = v[j] + v[j + 1];
#pragma unroll
    for (int i = 0; i < N; i += 2) {
        float m;
        if constexpr (W == 3) m = p[i + 1];
        else if constexpr (W == 5) m = p[i + 1] + p[i + 3];
        else m = (i % 4 == 0) ? (p[i + 1] + p[i + 3]) + p[i + 5] : p[i + 1] + (p[i + 3] + p[i + 5]);
        s[i] = v[i] + m;
        if (i + 1 < N) s[i + 1] = m + v[i + W];
    }
}

// Sliding-window box sums: s[0] = v[0] + ... + v[W-1] left to right, then
// s[i] = (s[i-1] + v[i+W-1]) - v[i-1] -- 2 operations per output, against
// about 3 for the trees.  Only for integer-valued inputs whose partial sums
// all stay below 2^24 in magnitude (squared differences of 8-bit samples over
// patches of up to 11 x 11): there every sum is exact in FP32, so the result
// is the trees' (and any order's) bit for bit; on general floating-point data
// the running window would accumulate rounding and cancellation.
template <int W, int N>
__device__ __forceinline__ void box_run(const float *v, float *s)
{
    float a = v[0];
#pragma unroll
    for (int i = 1; i < W; ++i) a = a + v[i];
    s[0] = a;
#pragma unroll
    for (int i = 1; i < N; ++i) {
        a = (a + v[i + W - 1]) - v[i - 1];
        s[i] = a;
    }
}

// box_run over the squares of e (same exactness condition), the squares
// folded into fused multiply-adds: W + 2 (N - 1) operations for N outputs.
template <int W, int N>
__device__ __forceinline__ void box_run_sq(const float *e, float *s)
{
    float a = e[0] * e[0];
#pragma unroll
    for (int i = 1; i < W; ++i) a = fmaf(e[i], e[i], a);
    s[0] = a;
#pragma unroll
    for (int i = 1; i < N; ++i) {
        a = fmaf(-e[i - 1], e[i - 1], fmaf(e[i + W - 1], e[i + W - 1], a));
        s[i] = a;
    }
}

}  // namespace gnlm
