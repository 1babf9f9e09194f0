// pam_state.hip -- float16 copies of whole wavefield states (the checkpoint
// slots of the binomial schedule): each field of each shot is stored as
// v / m in float16 with m = its max |v|, and read back as q * m.
#include <hip/hip_fp16.h>

#include <algorithm>

#include "pam_internal.h"

namespace pam {

__global__ void k_zero_u32(unsigned int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// bits[s] = max(bits[s], max |x_s|) as float bits (non-negative floats order
// like their bit patterns); grid (blocks, nshots).  Each block reduces its
// share in registers, then across its wavefronts, and makes one atomicMax: the
// blocks of a shot are kept few (kMaxBlocksPerShot), since every atomic on
// one address is serialised (1024 blocks per shot took 304 us per call on 32
// shots of 401 x 1344 cells, most of it waiting on the atomics).
constexpr int kMaxBlocksPerShot = 64;
__global__ __launch_bounds__(256) void k_shot_maxabs(const float *x, long long sstride, long long n,
                                                     unsigned int *bits) {
    __shared__ float red[4];
    const int s = blockIdx.y;
    const float *xs = x + (long long)s * sstride;
    float m = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(xs[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(bits + s, __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))));
}

__global__ __launch_bounds__(256) void k_pack_half(const float *x, long long sstride, long long n, __half *q,
                                                   long long qsstride, const unsigned int *bits) {
    const int s = blockIdx.y;
    const float m = __uint_as_float(bits[s]);
    const float inv = scale_inv(m);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        q[(long long)s * qsstride + i] = __float2half(x[(long long)s * sstride + i] * inv);
}

__global__ __launch_bounds__(256) void k_unpack_half(const __half *q, long long qsstride, long long n,
                                                     const unsigned int *bits, float *x, long long sstride) {
    const int s = blockIdx.y;
    const float m = __uint_as_float(bits[s]);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[(long long)s * sstride + i] = __half2float(q[(long long)s * qsstride + i]) * m;
}

static unsigned grid_blocks(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

int launch_pack_half(const float *x, long long sstride, long long n, int nshots, void *q, long long qsstride,
                     unsigned int *bits, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3((nshots + 255) / 256), dim3(256), 0, st, bits, nshots);
    const unsigned mb = std::min<unsigned>(grid_blocks(n), (unsigned)kMaxBlocksPerShot);
    hipLaunchKernelGGL(k_shot_maxabs, dim3(mb, nshots), dim3(256), 0, st, x, sstride, n, bits);
    hipLaunchKernelGGL(k_pack_half, dim3(grid_blocks(n), nshots), dim3(256), 0, st, x, sstride, n,
                       reinterpret_cast<__half *>(q), qsstride, (const unsigned int *)bits);
    return (int)hipGetLastError();
}

int launch_unpack_half(const void *q, long long qsstride, long long n, int nshots, const unsigned int *bits, float *x,
                       long long sstride, hipStream_t st) {
    hipLaunchKernelGGL(k_unpack_half, dim3(grid_blocks(n), nshots), dim3(256), 0, st,
                       reinterpret_cast<const __half *>(q), qsstride, n, bits, x, sstride);
    return (int)hipGetLastError();
}

}  // namespace pam
