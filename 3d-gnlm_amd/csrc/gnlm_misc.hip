// gnlm_misc.hip -- small element-wise kernels.
#include "gnlm_internal.h"

namespace gnlm {
namespace {

// Four bytes per thread in, four floats out (one dword load, one 16-byte
// store); the tail is handled element by element.
__global__ void k_u8_to_f32(const uint8_t *__restrict__ in, float *__restrict__ out, long long n)
{
    const long long n4 = n / 4;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4;
         i += (long long)gridDim.x * blockDim.x) {
        const uint32_t v = reinterpret_cast<const uint32_t *>(in)[i];
        float4 f;
        f.x = (float)(v & 0xff);
        f.y = (float)((v >> 8) & 0xff);
        f.z = (float)((v >> 16) & 0xff);
        f.w = (float)(v >> 24);
        reinterpret_cast<float4 *>(out)[i] = f;
    }
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t < n - 4 * n4) out[4 * n4 + t] = (float)in[4 * n4 + t];
}

__global__ void k_u8_to_f32_any(const uint8_t *__restrict__ in, float *__restrict__ out, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

}  // namespace

hipError_t u8_to_f32_any_launch(const uint8_t *in, float *out, long long n, hipStream_t s)
{
    long long g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_u8_to_f32_any, dim3((unsigned)g), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t u8_to_f32_launch(const uint8_t *in, float *out, long long n, hipStream_t s)
{
    // The vector path needs 4-byte aligned input and 16-byte aligned output.
    if ((reinterpret_cast<uintptr_t>(in) & 3) || (reinterpret_cast<uintptr_t>(out) & 15)) {
        return hipErrorInvalidValue;
    }
    long long g = (n / 4 + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_u8_to_f32, dim3((unsigned)g), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

}  // namespace gnlm
